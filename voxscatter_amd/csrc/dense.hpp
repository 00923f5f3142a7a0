// Small dense complex FP64 linear algebra on the host (column-major), for the Krylov solvers' projected
// problems: Householder QR (economy factors and least squares), Cholesky, triangular solves and the
// eigen-decomposition of a general complex matrix (Householder reduction to Hessenberg form, shifted QR
// iteration to the complex Schur form, eigenvectors by back substitution, unit 2-norm).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace vs {
namespace dense {

using zd = std::complex<double>;

struct Mat {
  int r = 0, c = 0;
  std::vector<zd> a;  // column-major
  Mat() = default;
  Mat(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_, zd(0)) {}
  zd& operator()(int i, int j) { return a[i + (size_t)r * j]; }
  const zd& operator()(int i, int j) const { return a[i + (size_t)r * j]; }
};

inline Mat matmul(const Mat& A, const Mat& B) {
  Mat C(A.r, B.c);
  for (int j = 0; j < B.c; ++j)
    for (int l = 0; l < A.c; ++l) {
      const zd b = B(l, j);
      if (b == zd(0)) continue;
      for (int i = 0; i < A.r; ++i) C(i, j) += A(i, l) * b;
    }
  return C;
}

inline Mat adjoint(const Mat& A) {
  Mat B(A.c, A.r);
  for (int j = 0; j < A.c; ++j)
    for (int i = 0; i < A.r; ++i) B(j, i) = std::conj(A(i, j));
  return B;
}

// Householder QR of a tall m x n matrix (m >= n): Q (m x n, orthonormal columns) and R (n x n)
inline void qr_econ(const Mat& A, Mat& Q, Mat& R) {
  const int m = A.r, n = A.c;
  Mat W = A;
  std::vector<std::vector<zd>> vs(n);
  for (int k = 0; k < n; ++k) {
    double nx = 0;
    for (int i = k; i < m; ++i) nx += std::norm(W(i, k));
    nx = std::sqrt(nx);
    std::vector<zd> v(m - k);
    for (int i = k; i < m; ++i) v[i - k] = W(i, k);
    if (nx == 0) {
      vs[k] = std::vector<zd>(m - k, zd(0));
      continue;
    }
    const zd x0 = v[0];
    const zd ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : zd(1);
    v[0] += ph * nx;
    double nv = 0;
    for (auto& e : v) nv += std::norm(e);
    nv = std::sqrt(nv);
    for (auto& e : v) e /= nv;
    for (int j = k; j < n; ++j) {
      zd s = 0;
      for (int i = k; i < m; ++i) s += std::conj(v[i - k]) * W(i, j);
      for (int i = k; i < m; ++i) W(i, j) -= 2.0 * v[i - k] * s;
    }
    vs[k] = v;
  }
  R = Mat(n, n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i <= j; ++i) R(i, j) = W(i, j);
  Q = Mat(m, n);
  for (int j = 0; j < n; ++j) Q(j, j) = 1;
  for (int k = n - 1; k >= 0; --k) {
    const auto& v = vs[k];
    for (int j = 0; j < n; ++j) {
      zd s = 0;
      for (int i = k; i < m; ++i) s += std::conj(v[i - k]) * Q(i, j);
      for (int i = k; i < m; ++i) Q(i, j) -= 2.0 * v[i - k] * s;
    }
  }
}

// x = R^{-1} y, R upper triangular n x n
inline std::vector<zd> upper_solve(const Mat& R, std::vector<zd> y) {
  const int n = R.c;
  for (int i = n - 1; i >= 0; --i) {
    zd s = y[i];
    for (int j = i + 1; j < n; ++j) s -= R(i, j) * y[j];
    y[i] = s / R(i, i);
  }
  y.resize(n);
  return y;
}

// least squares min ||b - A x|| for tall A of full column rank (A \ b)
inline std::vector<zd> lstsq(const Mat& A, const std::vector<zd>& b) {
  Mat Q, R;
  qr_econ(A, Q, R);
  std::vector<zd> y(A.c);
  for (int j = 0; j < A.c; ++j) {
    zd s = 0;
    for (int i = 0; i < A.r; ++i) s += std::conj(Q(i, j)) * b[i];
    y[j] = s;
  }
  return upper_solve(R, y);
}

// X = B R^{-1} (right division by an upper triangular matrix), B rows x n
inline Mat right_div_upper(const Mat& B, const Mat& R) {
  const int n = R.c;
  Mat X(B.r, n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < B.r; ++i) {
      zd s = B(i, j);
      for (int l = 0; l < j; ++l) s -= X(i, l) * R(l, j);
      X(i, j) = s / R(j, j);
    }
  return X;
}

// Cholesky A = L L^H of a Hermitian positive definite matrix
inline Mat cholesky(const Mat& A) {
  const int n = A.r;
  Mat L(n, n);
  for (int j = 0; j < n; ++j) {
    double d = std::real(A(j, j));
    for (int l = 0; l < j; ++l) d -= std::norm(L(j, l));
    if (!(d > 0)) throw std::runtime_error("matrix not positive definite");
    const double ljj = std::sqrt(d);
    L(j, j) = ljj;
    for (int i = j + 1; i < n; ++i) {
      zd s = A(i, j);
      for (int l = 0; l < j; ++l) s -= L(i, l) * std::conj(L(j, l));
      L(i, j) = s / ljj;
    }
  }
  return L;
}

// solve L X = B (lower triangular), B n x k
inline Mat lower_solve(const Mat& L, Mat B) {
  const int n = L.r;
  for (int j = 0; j < B.c; ++j)
    for (int i = 0; i < n; ++i) {
      zd s = B(i, j);
      for (int l = 0; l < i; ++l) s -= L(i, l) * B(l, j);
      B(i, j) = s / L(i, i);
    }
  return B;
}

// solve L^H X = B
inline Mat lower_adj_solve(const Mat& L, Mat B) {
  const int n = L.r;
  for (int j = 0; j < B.c; ++j)
    for (int i = n - 1; i >= 0; --i) {
      zd s = B(i, j);
      for (int l = i + 1; l < n; ++l) s -= std::conj(L(l, i)) * B(l, j);
      B(i, j) = s / std::conj(L(i, i));
    }
  return B;
}

// eigenvalues w and unit-norm eigenvectors (columns of X) of a general complex n x n matrix
inline void eig(const Mat& A0, std::vector<zd>& w, Mat& X) {
  const int n = A0.r;
  Mat H = A0, Z(n, n);
  for (int i = 0; i < n; ++i) Z(i, i) = 1;
  // Householder reduction to upper Hessenberg form, H = Z^H A Z
  for (int k = 0; k + 2 < n; ++k) {
    double nx = 0;
    for (int i = k + 1; i < n; ++i) nx += std::norm(H(i, k));
    nx = std::sqrt(nx);
    if (nx == 0) continue;
    std::vector<zd> v(n - k - 1);
    for (int i = k + 1; i < n; ++i) v[i - k - 1] = H(i, k);
    const zd ph = std::abs(v[0]) > 0 ? v[0] / std::abs(v[0]) : zd(1);
    v[0] += ph * nx;
    double nv = 0;
    for (auto& e : v) nv += std::norm(e);
    nv = std::sqrt(nv);
    for (auto& e : v) e /= nv;
    for (int j = 0; j < n; ++j) {  // H <- (I - 2 v v^H) H
      zd s = 0;
      for (int i = k + 1; i < n; ++i) s += std::conj(v[i - k - 1]) * H(i, j);
      for (int i = k + 1; i < n; ++i) H(i, j) -= 2.0 * v[i - k - 1] * s;
    }
    for (int i = 0; i < n; ++i) {  // H <- H (I - 2 v v^H), Z <- Z (I - 2 v v^H)
      zd s = 0, t = 0;
      for (int j = k + 1; j < n; ++j) {
        s += H(i, j) * v[j - k - 1];
        t += Z(i, j) * v[j - k - 1];
      }
      for (int j = k + 1; j < n; ++j) {
        H(i, j) -= 2.0 * s * std::conj(v[j - k - 1]);
        Z(i, j) -= 2.0 * t * std::conj(v[j - k - 1]);
      }
    }
    for (int i = k + 2; i < n; ++i) H(i, k) = 0;
  }
  // shifted QR iteration (Wilkinson shift, Givens rotations) on the active window [lo, hi]
  const double eps = std::numeric_limits<double>::epsilon();
  int hi = n - 1, iter = 0;
  while (hi > 0) {
    int lo = hi;
    while (lo > 0) {
      const double s = std::abs(H(lo - 1, lo - 1)) + std::abs(H(lo, lo));
      if (std::abs(H(lo, lo - 1)) <= eps * (s > 0 ? s : 1.0)) {
        H(lo, lo - 1) = 0;
        break;
      }
      --lo;
    }
    if (lo == hi) {
      --hi;
      iter = 0;
      continue;
    }
    if (++iter > 100 * n) throw std::runtime_error("eigenvalue iteration did not converge");
    // Wilkinson shift from the trailing 2 x 2 block (exceptional shift every 11 iterations)
    const zd a = H(hi - 1, hi - 1), b = H(hi - 1, hi), c = H(hi, hi - 1), d = H(hi, hi);
    zd mu;
    if (iter % 11 == 10) {
      mu = d + std::abs(c) * 0.75;
    } else {
      const zd tr = a + d, det = a * d - b * c;
      const zd disc = std::sqrt(tr * tr / 4.0 - det);
      const zd l1 = tr / 2.0 + disc, l2 = tr / 2.0 - disc;
      mu = std::abs(l1 - d) < std::abs(l2 - d) ? l1 : l2;
    }
    // QR step on the window: H - mu I = Q R, H <- R Q + mu I (as a sequence of Givens rotations)
    std::vector<zd> cs(hi - lo), sn(hi - lo);
    for (int k = lo; k <= hi; ++k) H(k, k) -= mu;
    for (int k = lo; k < hi; ++k) {
      const zd x = H(k, k), y = H(k + 1, k);
      const double r = std::sqrt(std::norm(x) + std::norm(y));
      zd cg = 1, sg = 0;
      if (r > 0) {
        cg = x / r;
        sg = y / r;
      }
      cs[k - lo] = cg;
      sn[k - lo] = sg;
      for (int j = k; j < n; ++j) {  // rows k, k+1 <- G^H [rows]
        const zd h1 = H(k, j), h2 = H(k + 1, j);
        H(k, j) = std::conj(cg) * h1 + std::conj(sg) * h2;
        H(k + 1, j) = -sg * h1 + cg * h2;
      }
    }
    for (int k = lo; k < hi; ++k) {
      const zd cg = cs[k - lo], sg = sn[k - lo];
      for (int i = 0; i <= std::min(k + 2, hi); ++i) {  // columns k, k+1 <- [cols] G
        const zd h1 = H(i, k), h2 = H(i, k + 1);
        H(i, k) = h1 * cg + h2 * sg;
        H(i, k + 1) = -h1 * std::conj(sg) + h2 * std::conj(cg);
      }
      for (int i = 0; i < n; ++i) {
        const zd z1 = Z(i, k), z2 = Z(i, k + 1);
        Z(i, k) = z1 * cg + z2 * sg;
        Z(i, k + 1) = -z1 * std::conj(sg) + z2 * std::conj(cg);
      }
    }
    for (int k = lo; k <= hi; ++k) H(k, k) += mu;
  }
  // eigenvectors of the triangular Schur factor, back to A's basis
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = H(i, i);
  double hn = 0;
  for (auto& e : H.a) hn = std::max(hn, std::abs(e));
  const double small = eps * (hn > 0 ? hn : 1.0);
  X = Mat(n, n);
  for (int k = 0; k < n; ++k) {
    std::vector<zd> y(n, zd(0));
    y[k] = 1;
    for (int j = k - 1; j >= 0; --j) {
      zd s = 0;
      for (int l = j + 1; l <= k; ++l) s += H(j, l) * y[l];
      zd den = H(j, j) - w[k];
      if (std::abs(den) < small) den = small;
      y[j] = -s / den;
    }
    double nrm = 0;
    std::vector<zd> x(n, zd(0));
    for (int i = 0; i < n; ++i) {
      zd s = 0;
      for (int l = 0; l <= k; ++l) s += Z(i, l) * y[l];
      x[i] = s;
      nrm += std::norm(s);
    }
    nrm = std::sqrt(nrm);
    for (int i = 0; i < n; ++i) X(i, k) = x[i] / nrm;
  }
}

}  // namespace dense
}  // namespace vs
