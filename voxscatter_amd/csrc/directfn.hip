// Singular face-pair integrals of the Galerkin N operator by the direct evaluation method (DIRECTFN:
// A. G. Polimeridis and T. V. Yioultsis, IEEE Trans. Antennas Propag. 56(9), 2008; A. G. Polimeridis and
// J. R. Mosig, Int. J. Numer. Methods Eng. 83, 2010), as the reference evaluates them
// (mexcubatures/directfn_ws_{st,ea,va}_const.cpp, combined over two triangles per square by
// cubatures/singular_{st,ea,va}_const.m with Np_1D = 20, assembly_nop_const.m:16,105-107).
//
// Each triangle is mapped to the equilateral parameter triangle (eta, xi), eta in [-1, 1], xi in [0, sqrt 3],
// with simplex coordinates x1 = (sqrt3 (1 - eta) - xi) / (2 sqrt3), x2 = (sqrt3 (1 + eta) - xi) / (2 sqrt3),
// x3 = xi / sqrt3 (Jacobian area / sqrt3).  The 4-D integral of exp(-i k R) / R over a triangle pair is
// rewritten in coordinates that cancel the singularity (self term: a relative polar system over three
// rotated sub-triangles and eight angular sub-domains; edge-adjacent: six sub-domains in (theta, psi) and two
// mirror branches; vertex-adjacent: two polar angles and a radial split) and every 1-D integral is a
// Gauss-Legendre rule of Np points on its interval.  Host code (FP64): the outer loops run on threads.
#include "vs_internal.hpp"

#include <complex>
#include <thread>

namespace vs {
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w);

namespace {

using cd = std::complex<double>;
constexpr double kSqrt3 = 1.7320508075688772935274463415059;
constexpr double kPi = 3.14159265358979323846264338327950288;

struct P3 {
  double v[3];
};

struct Tri {
  P3 a, b, c;
  // point at equilateral coordinates (eta, xi)
  void at(double eta, double xi, double out[3]) const {
    const double x1 = (kSqrt3 * (1 - eta) - xi) / (2 * kSqrt3);
    const double x2 = (kSqrt3 * (1 + eta) - xi) / (2 * kSqrt3);
    const double x3 = xi / kSqrt3;
    for (int i = 0; i < 3; ++i) out[i] = x1 * a.v[i] + x2 * b.v[i] + x3 * c.v[i];
  }
  double jac() const {
    double u[3], w[3], n[3];
    for (int i = 0; i < 3; ++i) {
      u[i] = b.v[i] - a.v[i];
      w[i] = c.v[i] - a.v[i];
    }
    n[0] = u[1] * w[2] - u[2] * w[1];
    n[1] = u[2] * w[0] - u[0] * w[2];
    n[2] = u[0] * w[1] - u[1] * w[0];
    return 0.5 * std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) / kSqrt3;
  }
};

inline cd green(double ko, const double p[3], const double q[3]) {
  const double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
  const double R = std::sqrt(dx * dx + dy * dy + dz * dz);
  return std::exp(cd(0, -ko * R)) / R;
}

struct Rule {
  std::vector<double> z, w;
  explicit Rule(int n) { gauss_legendre(n, z, w); }
  // node i and weight of the rule mapped to [a, b]
  double x(int i, double a, double b) const { return 0.5 * (b - a) * z[i] + 0.5 * (b + a); }
  int n() const { return (int)z.size(); }
};

// sum f(i) over i in [0, count) on up to `threads` threads, partial sums combined in index order
template <class F>
cd parallel_sum(int count, F&& f) {
  unsigned hw = std::thread::hardware_concurrency();
  const int nt = (int)std::max(1u, std::min<unsigned>(hw ? hw : 1, 16u));
  std::vector<cd> part(count);
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&, t] {
      for (int i = t; i < count; i += nt) part[i] = f(i);
    });
  for (auto& th : pool) th.join();
  cd s = 0;
  for (int i = 0; i < count; ++i) s += part[i];
  return s;
}

// ------------------------------------------------------------------------------------------------
// coincident triangles
// ------------------------------------------------------------------------------------------------
// rotation of the equilateral triangle onto sub-triangle kk (1: identity, 2 and 3: by +-120 degrees)
inline void rotate(int kk, double eta, double xi, double& e, double& x) {
  if (kk == 1) {
    e = eta;
    x = xi;
  } else if (kk == 2) {
    e = 0.5 - 0.5 * eta - 0.5 * kSqrt3 * xi;
    x = 0.5 * kSqrt3 + 0.5 * kSqrt3 * eta - 0.5 * xi;
  } else {
    e = -0.5 - 0.5 * eta + 0.5 * kSqrt3 * xi;
    x = 0.5 * kSqrt3 - 0.5 * kSqrt3 * eta - 0.5 * xi;
  }
}

cd st_triangle(const Tri& T, double ko, const Rule& g) {
  const int n = g.n();
  // sub-domain m = 1..8: psi interval, eta interval (as a function of psi), radial limit
  auto psi_lim = [](int m, double& a, double& b) {
    static const double lo[8] = {0, 1, 2, 2, 0, 0, 1, 2}, hi[8] = {1, 2, 3, 3, 1, 1, 2, 3};
    a = lo[m - 1] * kPi / 3;
    b = hi[m - 1] * kPi / 3;
  };
  auto eta_lim = [](int m, double psi, double& a, double& b) {
    switch (m) {
      case 1: case 2: a = 0; b = 1; break;
      case 3: case 4: {
        const double b1 = std::tan(kPi - psi) / kSqrt3, a1 = (1 - b1) / (1 + b1);
        if (m == 3) { a = a1; b = 1; } else { a = 0; b = a1; }
        break;
      }
      case 5: case 6: {
        const double b2 = std::tan(psi) / kSqrt3, a2 = -(1 - b2) / (1 + b2);
        if (m == 5) { a = a2; b = 0; } else { a = -1; b = a2; }
        break;
      }
      default: a = -1; b = 0; break;
    }
  };
  auto lam_lim = [](int m, double psi, double eta) {
    switch (m) {
      case 1: case 5: return (1 - eta) / std::cos(psi);
      case 2: case 3: return kSqrt3 * (1 - eta) / std::sin(psi);
      case 4: case 8: return -(1 + eta) / std::cos(psi);
      default: return kSqrt3 * (1 + eta) / std::sin(psi);  // 6, 7
    }
  };
  // one (kk, m, psi node) per task
  const int tasks = 3 * 8 * n;
  const cd s = parallel_sum(tasks, [&](int id) {
    const int kk = 1 + id / (8 * n), m = 1 + (id / n) % 8, ip = id % n;
    double pa, pb;
    psi_lim(m, pa, pb);
    const double psi = g.x(ip, pa, pb);
    const double sp = std::sin(psi), cp = std::cos(psi);
    double ea, eb;
    eta_lim(m, psi, ea, eb);
    cd Ie = 0;
    for (int ie = 0; ie < n; ++ie) {
      const double eta = g.x(ie, ea, eb);
      const double Lm = lam_lim(m, psi, eta);
      cd Il = 0;
      for (int il = 0; il < n; ++il) {
        const double lam = g.x(il, 0, Lm);
        const double xi_p = lam * sp;
        double pe, px, p[3];
        rotate(kk, eta, xi_p, pe, px);
        T.at(pe, px, p);
        cd Ir = 0;
        for (int ir = 0; ir < n; ++ir) {
          const double rho = g.x(ir, 0, lam);
          double qe, qx, q[3];
          rotate(kk, eta + rho * cp, xi_p - rho * sp, qe, qx);
          T.at(qe, qx, q);
          Ir += g.w[ir] * rho * green(ko, p, q);
        }
        Il += g.w[il] * (0.5 * lam) * Ir;
      }
      Ie += g.w[ie] * (0.5 * Lm) * sp * Il;
    }
    return (0.5 * (eb - ea)) * Ie * (0.5 * (pb - pa)) * g.w[ip];
  });
  const double J = T.jac();
  return J * J * s;
}

// ------------------------------------------------------------------------------------------------
// edge-adjacent triangles
// ------------------------------------------------------------------------------------------------
cd ea_pair(const Tri& P, const Tri& Q, double ko, const Rule& g) {
  const int n = g.n();
  auto theta_lim = [](int m, double& a, double& b) {
    static const double lo[6] = {0, 1.0 / 3, 1.0 / 3, 0.5, 0, 0.5}, hi[6] = {1.0 / 3, 0.5, 0.5, 1, 0.5, 1};
    a = lo[m - 1] * kPi;
    b = hi[m - 1] * kPi;
  };
  auto psi_lim = [](int m, double th, double& a, double& b) {
    const double A = std::atan(std::sin(th) - kSqrt3 * std::cos(th));
    const double B = std::atan(std::sin(th) + kSqrt3 * std::cos(th));
    switch (m) {
      case 1: a = 0; b = B; break;
      case 2: a = A; b = B; break;
      case 3: case 4: a = 0; b = A; break;
      case 5: a = B; b = kPi / 2; break;
      default: a = A; b = kPi / 2; break;
    }
  };
  // radial integral: int_0^Lam lam^2 K dlam, branch br = 1 (eta as given) or 2 (mirrored)
  auto radial = [&](double psi, double th, double eta, double Lam, int br) {
    const double sgn = br == 1 ? 1.0 : -1.0;
    cd I = 0;
    for (int il = 0; il < n; ++il) {
      const double lam = g.x(il, 0, Lam);
      double p[3], q[3];
      P.at(sgn * eta, lam * std::sin(psi), p);
      Q.at(sgn * (lam * std::cos(psi) * std::cos(th) - eta), lam * std::cos(psi) * std::sin(th), q);
      I += g.w[il] * lam * lam * green(ko, p, q);
    }
    return (0.5 * Lam) * I;
  };
  // the eta integral of sub-domain m at (psi, theta) for one branch
  auto eta_int = [&](int m, double psi, double th, int br) {
    const double c3 = std::sin(th) + kSqrt3 * std::cos(th), c1 = std::sin(th) - kSqrt3 * std::cos(th);
    auto L3 = [&](double e) { return kSqrt3 * (1 + e) / (std::cos(psi) * c3); };
    auto L2 = [&](double e) { return kSqrt3 * (1 - std::fabs(e)) / std::sin(psi); };
    auto L1 = [&](double e) { return kSqrt3 * (1 - std::fabs(e)) / (std::cos(psi) * c1); };
    cd I = 0;
    if (m <= 3) {
      // split at e0: [0, e0] radial limit L3, [e0, 1] limit L2 (m = 1, 2) or L1 (m = 3)
      double e0;
      if (m <= 2) {
        const double gam = std::tan(psi) / c3;
        e0 = (1 - gam) / (1 + gam);
      } else {
        e0 = kSqrt3 / std::tan(th);
      }
      cd I1 = 0, I2 = 0;
      for (int ie = 0; ie < n; ++ie) {
        const double e1 = g.x(ie, 0, e0), e2 = g.x(ie, e0, 1);
        I1 += g.w[ie] * radial(psi, th, e1, L3(e1), br);
        I2 += g.w[ie] * radial(psi, th, e2, m <= 2 ? L2(e2) : L1(e2), br);
      }
      I = (0.5 * e0) * I1 + (0.5 * (1 - e0)) * I2;
    } else {
      for (int ie = 0; ie < n; ++ie) {
        const double e = g.x(ie, 0, 1);
        I += g.w[ie] * radial(psi, th, e, m == 4 ? L1(e) : L2(e), br);
      }
      I = 0.5 * I;
    }
    return I;
  };
  const int tasks = 6 * n;
  const cd s = parallel_sum(tasks, [&](int id) {
    const int m = 1 + id / n, it = id % n;
    double ta, tb;
    theta_lim(m, ta, tb);
    const double th = g.x(it, ta, tb);
    double pa, pb;
    psi_lim(m, th, pa, pb);
    cd Ip = 0;
    for (int ip = 0; ip < n; ++ip) {
      const double psi = g.x(ip, pa, pb);
      Ip += g.w[ip] * std::cos(psi) * (eta_int(m, psi, th, 1) + eta_int(m, psi, th, 2));
    }
    return (0.5 * (tb - ta)) * g.w[it] * (0.5 * (pb - pa)) * Ip;
  });
  return P.jac() * Q.jac() * s;
}

// ------------------------------------------------------------------------------------------------
// vertex-adjacent triangles (common vertex: the first of P, the last of Q)
// ------------------------------------------------------------------------------------------------
cd va_pair(const Tri& P, const Tri& Q, double ko, const Rule& g) {
  const int n = g.n();
  const double tb = kPi / 3;
  const int tasks = n * n;
  const cd s = parallel_sum(tasks, [&](int id) {
    const int itp = id / n, itq = id % n;
    const double thp = g.x(itp, 0, tb), thq = g.x(itq, 0, tb);
    const double Lp = 2 * kSqrt3 / (std::sin(thp) + kSqrt3 * std::cos(thp));
    const double Lq = 2 * kSqrt3 / (std::sin(thq) + kSqrt3 * std::cos(thq));
    const double ps = std::atan(Lq / Lp);
    auto radial = [&](double psi, double Lam) {
      cd I = 0;
      for (int il = 0; il < n; ++il) {
        const double lam = g.x(il, 0, Lam);
        double p[3], q[3];
        P.at(lam * std::cos(psi) * std::cos(thp) - 1, lam * std::cos(psi) * std::sin(thp), p);
        Q.at(lam * std::sin(psi) * std::cos(thq) - 1, lam * std::sin(psi) * std::sin(thq), q);
        I += g.w[il] * lam * lam * lam * std::sin(psi) * std::cos(psi) * green(ko, p, q);
      }
      return (0.5 * Lam) * I;
    };
    cd Ip = 0;
    for (int ip = 0; ip < n; ++ip) {
      const double p1 = g.x(ip, 0, ps), p2 = g.x(ip, ps, kPi / 2);
      Ip += g.w[ip] * ((0.5 * ps) * radial(p1, Lp / std::cos(p1)) +
                       (0.5 * (kPi / 2 - ps)) * radial(p2, Lq / std::sin(p2)));
    }
    return (0.5 * tb) * g.w[itp] * (0.5 * tb) * g.w[itq] * Ip;
  });
  return P.jac() * Q.jac() * s;
}

Tri tri(const P3& a, const P3& b, const P3& c) { return Tri{a, b, c}; }

}  // namespace

// [ST, EAc, EAo, VAc, VAo] of squares of side dx (singular_st_const.m, singular_ea_const.m,
// singular_va_const.m with method 'DIRECTFN'): each square is two triangles, and each singular square pair
// the sum of its four triangle pairs.  The vertex orders of the triangle pairs are those of the reference's
// gateways (edge-adjacent: P = (r1, r2, r3), Q = (r4, r2, r1); vertex-adjacent: P = (r1, r2, r3),
// Q = (r4, r5, r1)).
void galerkin_singular_directfn(double ko, double dx, int np, c128 out[5]) {
  const Rule g(np);
  const double d = dx;
  auto ST = [&](P3 a, P3 b, P3 c) { return st_triangle(tri(a, b, c), ko, g); };
  auto EA = [&](P3 a, P3 b, P3 c, P3 e) { return ea_pair(tri(a, b, c), tri(e, b, a), ko, g); };
  auto VA = [&](P3 a, P3 b, P3 c, P3 e, P3 f) { return va_pair(tri(a, b, c), tri(e, f, a), ko, g); };
  const P3 r1{{0, 0, 0}}, r2{{d, 0, 0}}, r3{{d, d, 0}}, r4{{0, d, 0}};
  const cd st = ST(r1, r2, r3) + EA(r1, r3, r2, r4) + EA(r1, r3, r4, r2) + ST(r1, r3, r4);
  const P3 r5{{2 * d, 0, 0}}, r6{{2 * d, d, 0}}, r5o{{d, 0, d}}, r6o{{d, d, d}};
  const cd eac = EA(r2, r3, r1, r5) + VA(r3, r1, r2, r5, r6) + VA(r3, r4, r1, r2, r5) + VA(r3, r4, r1, r5, r6);
  const cd eao = EA(r2, r3, r1, r5o) + VA(r3, r1, r2, r5o, r6o) + VA(r3, r4, r1, r2, r5o) + VA(r3, r4, r1, r5o, r6o);
  const P3 v5{{2 * d, d, 0}}, v6{{2 * d, 2 * d, 0}}, v7{{d, 2 * d, 0}};
  const P3 v5o{{d, d, d}}, v6o{{d, 2 * d, d}}, v7o{{d, 2 * d, 0}};
  const cd vac = VA(r3, r1, r2, v5, v6) + VA(r3, r1, r2, v6, v7) + VA(r3, r4, r1, v5, v6) + VA(r3, r4, r1, v6, v7);
  const cd vao =
      VA(r3, r1, r2, v5o, v6o) + VA(r3, r1, r2, v6o, v7o) + VA(r3, r4, r1, v5o, v6o) + VA(r3, r4, r1, v6o, v7o);
  const cd v[5] = {st, eac, eao, vac, vao};
  for (int i = 0; i < 5; ++i) out[i] = mk<double>(v[i].real(), v[i].imag());
}

}  // namespace vs

extern "C" int vs_galerkin_singular_directfn(double ko, double dx, int np, double out[10]) {
  return vs::guarded([&] {
    VS_CHECK(out != nullptr, VS_ERR_ARG, "null pointer");
    VS_CHECK(dx > 0 && ko >= 0, VS_ERR_ARG, "dx must be positive and ko non-negative");
    VS_CHECK(np >= 1 && np <= 256, VS_ERR_ARG, "np must be between 1 and 256");
    vs::c128 v[5];
    vs::galerkin_singular_directfn(ko, dx, np, v);
    for (int i = 0; i < 5; ++i) {
      out[2 * i] = v[i].x;
      out[2 * i + 1] = v[i].y;
    }
  });
}
