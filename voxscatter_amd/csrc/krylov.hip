// Krylov solver control logic (host) over device-resident vectors.
#include "solver.hpp"

#include <cfloat>
#include <cmath>
#include <limits>

namespace vs {

template <typename T>
struct Vecs {
  DevBuf buf;
  int64_t n;
  int count;
  Vecs(int64_t n_, int cnt) : buf(sizeof(cplx<T>) * n_ * cnt), n(n_), count(cnt) {}
  cplx<T>* operator[](int i) { return buf.as<cplx<T>>() + (int64_t)i * n; }
};

static inline c128 C(double r, double i = 0) { return mk<double>(r, i); }
static inline double cabs_(c128 a) { return std::hypot(a.x, a.y); }
static inline bool cisinf(c128 a) { return std::isinf(a.x) || std::isinf(a.y); }
static inline bool cfinite(c128 a) { return std::isfinite(a.x) && std::isfinite(a.y); }

template <typename T>
static double eps_of() {
  return sizeof(T) == 8 ? DBL_EPSILON : (double)FLT_EPSILON;
}

// MATLAB bicgstab semantics: see the restatement of the built-in documented alongside the oracle.
template <typename T>
SolveResult solve_bicgstab(Engine<T>& E, const LinOp<T>& op, const cplx<T>* b, cplx<T>* x, bool has_x0, double tol,
                           int maxit) {
  using cv = cplx<T>;
  const int64_t n = E.n;
  SolveResult R;
  const double eps = eps_of<T>();
  const double n2b = E.norm(b);
  if (n2b == 0) {
    VS_HIP(hipMemsetAsync(x, 0, sizeof(cv) * n, E.ctx->stream));
    R.flag = 0;
    R.relres = 0;
    R.iter = 0;
    R.resvec = {0.0};
    return R;
  }
  if (!has_x0) VS_HIP(hipMemsetAsync(x, 0, sizeof(cv) * n, E.ctx->stream));
  Vecs<T> W(n, 10);
  cv *r = W[0], *rt = W[1], *p = W[2], *v = W[3], *ph = W[4], *s = W[5], *sh = W[6], *t = W[7], *xh = W[8],
     *xmin = W[9];
  int flag = 1;
  E.copy(x, xmin);
  double imin = 0;
  const double tolb = tol * n2b;
  // r = b - A x
  op.A(x, r);
  double normr = E.update(UPD_SUB_, b, nullptr, nullptr, nullptr, r, nullptr, C(0), C(0), true);
  double normr_act = normr;
  if (normr <= tolb) {
    R.flag = 0;
    R.relres = normr / n2b;
    R.iter = 0;
    R.resvec = {normr};
    return R;
  }
  E.copy(r, rt);
  std::vector<double> resvec(2 * (size_t)maxit + 1, 0.0);
  resvec[0] = normr;
  int stag = 0, moresteps = 0;
  const int64_t maxmsteps = std::min<int64_t>(std::min<int64_t>(n / 50, 10), n - maxit);
  const int maxstagsteps = 3;
  double normrmin = normr;
  c128 rho = C(1), omega = C(1), alpha = C(0);
  int64_t nres = -1;
  double iter = 0;
  int ii = 0;
  const cv* a1[1];
  const cv* b1[1];
  for (ii = 1; ii <= maxit; ++ii) {
    const c128 rho1 = rho;
    a1[0] = rt;
    b1[0] = r;
    E.dots(1, a1, b1, &rho);
    if ((rho.x == 0 && rho.y == 0) || cisinf(rho)) {
      flag = 4;
      nres = 2 * ii - 1;
      break;
    }
    if (ii == 1) {
      E.copy(r, p);
    } else {
      const c128 beta = (cdiv(rho, rho1)) * cdiv(alpha, omega);
      if ((beta.x == 0 && beta.y == 0) || !cfinite(beta)) {
        flag = 4;
        nres = 2 * ii - 1;
        break;
      }
      E.update(UPD_BICG_P_, r, v, nullptr, nullptr, p, nullptr, omega, beta, false);
    }
    const cv* phv = p;
    if (op.M) {
      op.M(p, ph);
      phv = ph;
    }
    op.A(phv, v);
    c128 d3[3];
    {
      const cv* aa[3] = {rt, phv, x};
      const cv* bb[3] = {v, phv, x};
      E.dots(3, aa, bb, d3);
    }
    const c128 rtv = d3[0];
    const double nph = std::sqrt(d3[1].x), nx = std::sqrt(d3[2].x);
    if (op.M && !(std::isfinite(nph))) {
      flag = 2;
      nres = 2 * ii - 1;
      break;
    }
    if ((rtv.x == 0 && rtv.y == 0) || cisinf(rtv)) {
      flag = 4;
      nres = 2 * ii - 1;
      break;
    }
    alpha = cdiv(rho, rtv);
    if (cisinf(alpha)) {
      flag = 4;
      nres = 2 * ii - 1;
      break;
    }
    if (cabs_(alpha) * nph < eps * nx)
      ++stag;
    else
      stag = 0;
    // xhalf = x + alpha ph ; s = r - alpha v
    normr = E.update(UPD_HALF_, x, phv, r, v, xh, s, alpha, C(0), true);
    normr_act = normr;
    resvec[2 * ii - 1] = normr;
    if (normr <= tolb || stag >= maxstagsteps || moresteps) {
      op.A(xh, s);
      normr_act = E.update(UPD_SUB_, b, nullptr, nullptr, nullptr, s, nullptr, C(0), C(0), true);
      resvec[2 * ii - 1] = normr_act;
      if (normr_act <= tolb) {
        E.copy(xh, x);
        flag = 0;
        iter = ii - 0.5;
        nres = 2 * ii;
        break;
      } else {
        if (stag >= maxstagsteps && moresteps == 0) stag = 0;
        ++moresteps;
        if (moresteps >= maxmsteps) {
          flag = 3;
          E.copy(xh, x);
          nres = 2 * ii;
          break;
        }
      }
    }
    if (normr_act < normrmin) {
      normrmin = normr_act;
      E.copy(xh, xmin);
      imin = ii - 0.5;
    }
    const cv* shv = s;
    if (op.M) {
      op.M(s, sh);
      shv = sh;
    }
    op.A(shv, t);
    c128 d4[4];
    {
      const cv* aa[4] = {t, t, shv, xh};
      const cv* bb[4] = {t, s, shv, xh};
      E.dots(4, aa, bb, d4);
    }
    const c128 tt = d4[0];
    const double nsh = std::sqrt(d4[2].x), nxh = std::sqrt(d4[3].x);
    if (op.M && !std::isfinite(nsh)) {
      flag = 2;
      nres = 2 * ii;
      break;
    }
    if ((tt.x == 0 && tt.y == 0) || cisinf(tt)) {
      flag = 4;
      nres = 2 * ii;
      break;
    }
    omega = cdiv(d4[1], tt);
    if (cisinf(omega)) {
      flag = 4;
      nres = 2 * ii;
      break;
    }
    if (cabs_(omega) * nsh < eps * nxh)
      ++stag;
    else
      stag = 0;
    // x = xhalf + omega sh ; r = s - omega t
    normr = E.update(UPD_HALF_, xh, shv, s, t, x, r, omega, C(0), true);
    normr_act = normr;
    resvec[2 * ii] = normr;
    if (normr <= tolb || stag >= maxstagsteps || moresteps) {
      op.A(x, r);
      normr_act = E.update(UPD_SUB_, b, nullptr, nullptr, nullptr, r, nullptr, C(0), C(0), true);
      resvec[2 * ii] = normr_act;
      if (normr_act <= tolb) {
        flag = 0;
        iter = ii;
        nres = 2 * ii + 1;
        break;
      } else {
        if (stag >= maxstagsteps && moresteps == 0) stag = 0;
        ++moresteps;
        if (moresteps >= maxmsteps) {
          flag = 3;
          nres = 2 * ii + 1;
          break;
        }
      }
    }
    if (normr_act < normrmin) {
      normrmin = normr_act;
      E.copy(x, xmin);
      imin = ii;
    }
    if (stag >= maxstagsteps) {
      flag = 3;
      nres = 2 * ii + 1;
      break;
    }
  }
  if (ii > maxit) ii = maxit;
  if (nres < 0) nres = 2 * (int64_t)maxit + 1;
  double relres;
  if (flag == 0) {
    relres = normr_act / n2b;
  } else {
    op.A(xmin, r);
    const double nrc = E.update(UPD_SUB_, b, nullptr, nullptr, nullptr, r, nullptr, C(0), C(0), true);
    if (nrc <= normr_act) {
      E.copy(xmin, x);
      iter = imin;
      relres = nrc / n2b;
    } else {
      iter = ii;
      relres = normr_act / n2b;
    }
  }
  R.flag = flag;
  R.relres = relres;
  R.iter = iter;
  resvec.resize((size_t)nres);
  R.resvec = std::move(resvec);
  VS_HIP(hipStreamSynchronize(E.ctx->stream));
  return R;
}

// Givens rotation zeroing b in (a, b): G = [conj(c) conj(s); -s c]
static void givens(c128 a, c128 b, c128& c, c128& s) {
  const double den = std::sqrt(abs2(a) + abs2(b));
  if (den == 0) {
    c = C(1);
    s = C(0);
  } else {
    c = mk<double>(a.x / den, a.y / den);
    s = mk<double>(b.x / den, b.y / den);
  }
}

// back substitution with the upper triangle of H (column-major, leading dimension ldh)
static std::vector<c128> upper_solve(const std::vector<c128>& H, int ldh, int k, const std::vector<c128>& g) {
  std::vector<c128> y(k);
  for (int i = k - 1; i >= 0; --i) {
    c128 acc = g[i];
    for (int j = i + 1; j < k; ++j) acc = acc - H[i + (size_t)ldh * j] * y[j];
    y[i] = cdiv(acc, H[i + (size_t)ldh * i]);
  }
  return y;
}

template <typename T>
SolveResult solve_gmres(Engine<T>& E, const LinOp<T>& op, const cplx<T>* b, cplx<T>* x, bool has_x0, int restart,
                        double tol, int maxit) {
  using cv = cplx<T>;
  const int64_t n = E.n;
  SolveResult R;
  if (restart > n) restart = (int)n;
  VS_CHECK(restart >= 1 && restart <= 511, VS_ERR_UNSUPPORTED, "gmres restart must be in 1..511");
  const double n2b = E.norm(b);
  if (n2b == 0) {
    VS_HIP(hipMemsetAsync(x, 0, sizeof(cv) * n, E.ctx->stream));
    R.flag = 0;
    R.resvec = {0.0};
    return R;
  }
  if (!has_x0) VS_HIP(hipMemsetAsync(x, 0, sizeof(cv) * n, E.ctx->stream));
  Vecs<T> Vb(n, restart + 1);
  Vecs<T> W(n, 3);
  cv *r = W[0], *w = W[1], *tmp = W[2];
  auto precond = [&](const cv* in, cv* out) {
    if (op.M)
      op.M(in, out);
    else
      E.copy(in, out);
  };
  // n2minv_b = ||M \ b||
  precond(b, w);
  const double n2minv_b = E.norm(w);
  const double tolb = tol * n2minv_b;
  // r = M \ (b - A x)
  op.A(x, tmp);
  E.update(UPD_SUB_, b, nullptr, nullptr, nullptr, tmp, nullptr, C(0), C(0), false);
  precond(tmp, r);
  double normr = E.norm(r);
  std::vector<double> resvec{normr};
  if (normr <= tolb) {
    R.flag = 0;
    R.relres = normr / n2minv_b;
    R.resvec = resvec;
    return R;
  }
  int flag = 1, outer = 0, inner = 0;
  const int ldh = restart + 1;
  for (outer = 1; outer <= maxit; ++outer) {
    std::vector<c128> H((size_t)ldh * restart, C(0)), cs(restart), sn(restart), g(restart + 1, C(0));
    E.update(UPD_SCALE_, r, nullptr, nullptr, nullptr, Vb[0], nullptr, C(1.0 / normr), C(0), false);
    g[0] = C(normr);
    int kd = 0;
    for (int k = 0; k < restart; ++k) {
      op.A(Vb[k], tmp);
      precond(tmp, w);
      // CGS2
      std::vector<c128> h(k + 1), h2(k + 1);
      E.mdot(Vb[0], n, k + 1, w, h.data());
      E.gemv_acc(Vb[0], n, k + 1, h.data(), -1.0, w);
      E.mdot(Vb[0], n, k + 1, w, h2.data());
      E.gemv_acc(Vb[0], n, k + 1, h2.data(), -1.0, w);
      for (int i = 0; i <= k; ++i) h[i] = h[i] + h2[i];
      const double hn = E.norm(w);
      for (int i = 0; i <= k; ++i) H[i + (size_t)ldh * k] = h[i];
      H[k + 1 + (size_t)ldh * k] = C(hn);
      if (hn != 0) E.update(UPD_SCALE_, w, nullptr, nullptr, nullptr, Vb[k + 1], nullptr, C(1.0 / hn), C(0), false);
      for (int i = 0; i < k; ++i) {
        const c128 hi = H[i + (size_t)ldh * k], hi1 = H[i + 1 + (size_t)ldh * k];
        H[i + (size_t)ldh * k] = cconj(cs[i]) * hi + cconj(sn[i]) * hi1;
        H[i + 1 + (size_t)ldh * k] = (-sn[i]) * hi + cs[i] * hi1;
      }
      const c128 a = H[k + (size_t)ldh * k], bb = H[k + 1 + (size_t)ldh * k];
      givens(a, bb, cs[k], sn[k]);
      H[k + (size_t)ldh * k] = cconj(cs[k]) * a + cconj(sn[k]) * bb;
      H[k + 1 + (size_t)ldh * k] = C(0);
      g[k + 1] = (-sn[k]) * g[k];
      g[k] = cconj(cs[k]) * g[k];
      inner = k + 1;
      kd = k + 1;
      resvec.push_back(cabs_(g[k + 1]));
      if (cabs_(g[k + 1]) <= tolb || hn == 0) break;
    }
    std::vector<c128> y = upper_solve(H, ldh, kd, g);
    E.gemv_acc(Vb[0], n, kd, y.data(), 1.0, x);
    op.A(x, tmp);
    E.update(UPD_SUB_, b, nullptr, nullptr, nullptr, tmp, nullptr, C(0), C(0), false);
    precond(tmp, r);
    normr = E.norm(r);
    resvec.back() = normr;
    if (normr <= tolb) {
      flag = 0;
      break;
    }
  }
  if (outer > maxit) outer = maxit;
  R.flag = flag;
  R.relres = normr / n2minv_b;
  R.outer = outer;
  R.inner = inner;
  R.iter = outer;
  R.resvec = resvec;
  VS_HIP(hipStreamSynchronize(E.ctx->stream));
  return R;
}

// pgmres.m:70-180 with pgmres_iter.m:22-150; side 0: M is L1 (left), side 1: M is R1 (right)
template <typename T>
SolveResult solve_pgmres(Engine<T>& E, const LinOp<T>& op, int side, const cplx<T>* b, cplx<T>* x, bool has_x0,
                         int restart, double tol, int maxit) {
  using cv = cplx<T>;
  const int64_t n = E.n;
  SolveResult R;
  VS_CHECK(restart >= 1 && restart <= 511, VS_ERR_UNSUPPORTED, "pgmres restart must be in 1..511");
  const bool left = op.M && side == 0, right = op.M && side == 1;
  Vecs<T> Vb(n, restart + 1);
  Vecs<T> W(n, 3);
  cv *r = W[0], *w = W[1], *tmp = W[2];
  if (!has_x0) VS_HIP(hipMemsetAsync(x, 0, sizeof(cv) * n, E.ctx->stream));
  // r = b - A x0 (only if x0 is nonzero), bp = b; left preconditioning of both
  if (has_x0 && E.norm(x) != 0) {
    op.A(x, r);
    E.update(UPD_SUB_, b, nullptr, nullptr, nullptr, r, nullptr, C(0), C(0), false);
  } else {
    E.copy(b, r);
  }
  double bnorm;
  if (left) {
    op.M(r, tmp);
    E.copy(tmp, r);
    op.M(b, tmp);
    bnorm = E.norm(tmp);
  } else {
    bnorm = E.norm(b);
  }
  std::vector<double> resvec{E.norm(r) / bnorm};
  int outit = 0, p = 0;
  const int ldh = restart + 1;
  while (resvec.back() > tol && outit < maxit) {
    ++outit;
    const double rn = E.norm(r);
    E.update(UPD_SCALE_, r, nullptr, nullptr, nullptr, Vb[0], nullptr, C(1.0 / rn), C(0), false);
    std::vector<c128> H((size_t)ldh * restart, C(0)), H0((size_t)ldh * restart, C(0)), cs(restart), sn(restart),
        g(restart + 1, C(0));
    g[0] = C(rn);
    int k = 0;
    for (k = 1; k <= restart; ++k) {
      const cv* in = Vb[k - 1];
      if (right) {
        op.M(in, tmp);
        op.A(tmp, w);
      } else if (left) {
        op.A(in, tmp);
        op.M(tmp, w);
      } else {
        op.A(in, w);
      }
      std::vector<c128> h(k);
      E.mdot(Vb[0], n, k, w, h.data());
      E.gemv_acc(Vb[0], n, k, h.data(), -1.0, w);
      const double hn = E.norm(w);
      for (int i = 0; i < k; ++i) H[i + (size_t)ldh * (k - 1)] = h[i];
      H[k + (size_t)ldh * (k - 1)] = C(hn);
      for (int i = 0; i <= k; ++i) H0[i + (size_t)ldh * (k - 1)] = H[i + (size_t)ldh * (k - 1)];
      E.update(UPD_SCALE_, w, nullptr, nullptr, nullptr, Vb[k], nullptr, C(1.0 / hn), C(0), false);
      // least squares by rotations (same minimiser and residual norm as H(1:k+1,1:k) \ rhs)
      const int c = k - 1;
      for (int i = 0; i < c; ++i) {
        const c128 hi = H[i + (size_t)ldh * c], hi1 = H[i + 1 + (size_t)ldh * c];
        H[i + (size_t)ldh * c] = cconj(cs[i]) * hi + cconj(sn[i]) * hi1;
        H[i + 1 + (size_t)ldh * c] = (-sn[i]) * hi + cs[i] * hi1;
      }
      const c128 a = H[c + (size_t)ldh * c], bb = H[c + 1 + (size_t)ldh * c];
      givens(a, bb, cs[c], sn[c]);
      H[c + (size_t)ldh * c] = cconj(cs[c]) * a + cconj(sn[c]) * bb;
      H[c + 1 + (size_t)ldh * c] = C(0);
      g[c + 1] = (-sn[c]) * g[c];
      g[c] = cconj(cs[c]) * g[c];
      const double res = cabs_(g[c + 1]);
      resvec.push_back(res / bnorm);
      if (res < tol * bnorm) break;
    }
    if (k > restart) k = restart;
    p = k;
    std::vector<c128> y = upper_solve(H, ldh, k, g);
    // x += R1 (V y)   (pgmres_iter.m:133-149)
    VS_HIP(hipMemsetAsync(w, 0, sizeof(cv) * n, E.ctx->stream));
    E.gemv_acc(Vb[0], n, k, y.data(), 1.0, w);
    if (right) {
      op.M(w, tmp);
      E.update(UPD_AXPBY_, tmp, nullptr, nullptr, nullptr, x, nullptr, C(1), C(1), false);
    } else {
      E.update(UPD_AXPBY_, w, nullptr, nullptr, nullptr, x, nullptr, C(1), C(1), false);
    }
    // r = V(:,1:k+1) * (rhs - Hc y)   (pgmres_iter.m:150)
    std::vector<c128> res(k + 1);
    for (int i = 0; i <= k; ++i) {
      c128 acc = i == 0 ? C(rn) : C(0);
      for (int j = 0; j < k; ++j) acc = acc - H0[i + (size_t)ldh * j] * y[j];
      res[i] = acc;
    }
    VS_HIP(hipMemsetAsync(r, 0, sizeof(cv) * n, E.ctx->stream));
    E.gemv_acc(Vb[0], n, k + 1, res.data(), 1.0, r);
  }
  op.A(x, tmp);
  const double rr = E.update(UPD_SUB_, b, nullptr, nullptr, nullptr, tmp, nullptr, C(0), C(0), true);
  R.relres = rr / E.norm(b);
  R.flag = R.relres <= tol ? 0 : 1;
  R.outer = outit;
  R.inner = p;
  R.resvec = resvec;
  VS_HIP(hipStreamSynchronize(E.ctx->stream));
  return R;
}

template SolveResult solve_bicgstab<double>(Engine<double>&, const LinOp<double>&, const c128*, c128*, bool, double, int);
template SolveResult solve_bicgstab<float>(Engine<float>&, const LinOp<float>&, const c64*, c64*, bool, double, int);
template SolveResult solve_gmres<double>(Engine<double>&, const LinOp<double>&, const c128*, c128*, bool, int, double,
                                         int);
template SolveResult solve_gmres<float>(Engine<float>&, const LinOp<float>&, const c64*, c64*, bool, int, double, int);
template SolveResult solve_pgmres<double>(Engine<double>&, const LinOp<double>&, int, const c128*, c128*, bool, int,
                                          double, int);
template SolveResult solve_pgmres<float>(Engine<float>&, const LinOp<float>&, int, const c64*, c64*, bool, int, double,
                                         int);

}  // namespace vs

namespace vs {

// Mixed precision iterative refinement: the correction equation A d = r is solved in complex FP32
// (FP32 tensor, FP32 preconditioner blocks, FP32 Krylov vectors with FP64 reductions); the residual
// r = b - A x and the update x += d are formed in FP64 with the FP64 tensor.
SolveResult solve_refine(Engine<double>& E64, Engine<float>& E32, const LinOp<double>& A64, const LinOp<float>& A32,
                         const c128* b, c128* x, bool has_x0, double tol, int max_outer, int inner_solver,
                         double inner_tol, int inner_maxit, int restart, int* inner_total) {
  const int64_t n = E64.n;
  SolveResult R;
  vs_ctx* ctx = E64.ctx;
  const double n2b = E64.norm(b);
  if (!has_x0) VS_HIP(hipMemsetAsync(x, 0, sizeof(c128) * n, ctx->stream));
  if (n2b == 0) {
    R.flag = 0;
    R.resvec = {0.0};
    return R;
  }
  DevBuf r64(sizeof(c128) * n), r32(sizeof(c64) * n), d32(sizeof(c64) * n), d64(sizeof(c128) * n);
  int tot = 0;
  double relres = 0;
  R.flag = 1;
  int outer = 0;
  for (outer = 0; outer <= max_outer; ++outer) {
    A64.A(x, r64.as<c128>());
    const double nr = E64.update(UPD_SUB_, b, nullptr, nullptr, nullptr, r64.as<c128>(), nullptr, mk<double>(0, 0),
                                 mk<double>(0, 0), true);
    relres = nr / n2b;
    R.resvec.push_back(relres);
    if (!std::isfinite(nr)) {
      R.flag = 2;
      break;
    }
    if (relres <= tol) {
      R.flag = 0;
      break;
    }
    if (outer == max_outer) break;
    // scale the residual to unit norm before rounding to FP32
    E64.update(UPD_SCALE_, r64.as<c128>(), nullptr, nullptr, nullptr, r64.as<c128>(), nullptr, mk<double>(1.0 / nr, 0),
               mk<double>(0, 0), false);
    convert_dev(ctx, r64.p, VS_FP64, r32.p, VS_FP32, n);
    SolveResult in;
    if (inner_solver == 0)
      in = solve_bicgstab<float>(E32, A32, r32.as<c64>(), d32.as<c64>(), false, inner_tol, inner_maxit);
    else
      in = solve_gmres<float>(E32, A32, r32.as<c64>(), d32.as<c64>(), false, restart, inner_tol, inner_maxit);
    tot += inner_solver == 0 ? (int)(in.resvec.size() - 1) : (int)(in.resvec.size() - 1);
    convert_dev(ctx, d32.p, VS_FP32, d64.p, VS_FP64, n);
    E64.update(UPD_AXPBY_, d64.as<c128>(), nullptr, nullptr, nullptr, x, nullptr, mk<double>(nr, 0), mk<double>(1, 0),
               false);
  }
  R.relres = relres;
  R.outer = outer;
  R.iter = outer;
  if (inner_total) *inner_total = tot;
  VS_HIP(hipStreamSynchronize(ctx->stream));
  return R;
}

}  // namespace vs
