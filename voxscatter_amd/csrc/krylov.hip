// Krylov solver control logic (host) over device-resident vectors.
#include "dense.hpp"
#include "solver.hpp"

#include <cfloat>
#include <cstring>
#include <cmath>
#include <cstdlib>
#include <limits>

namespace vs {

// n-vectors carved from one of the context's reusable workspaces
template <typename T>
struct Vecs {
  cplx<T>* base;
  int64_t n;
  Vecs(vs_ctx* ctx, int slot, int64_t n_, int cnt)
      : base(static_cast<cplx<T>*>(ctx->workspace(slot, sizeof(cplx<T>) * (size_t)n_ * cnt))), n(n_) {}
  Vecs(cplx<T>* b, int64_t n_) : base(b), n(n_) {}
  cplx<T>* operator[](int i) { return base + (int64_t)i * n; }
  Vecs sub(int first) { return Vecs(base + (int64_t)first * n, n); }
};

// a piece of a workspace
struct Span {
  void* p;
  template <typename U> U* as() const { return static_cast<U*>(p); }
};

// consecutive 256-byte aligned pieces of one workspace: sizes first, then pointers
struct Carve {
  std::vector<size_t> sz;
  size_t add(size_t bytes) {
    sz.push_back(bytes);
    return sz.size() - 1;
  }
  char* base = nullptr;
  void take(vs_ctx* ctx, int slot) {
    size_t tot = 0;
    for (size_t b : sz) tot += (b + 255) & ~size_t(255);
    base = static_cast<char*>(ctx->workspace(slot, tot));
  }
  Span at(size_t i) const {
    size_t off = 0;
    for (size_t j = 0; j < i; ++j) off += (sz[j] + 255) & ~size_t(255);
    return Span{base + off};
  }
};

// events of the pipelined BiCGStab, destroyed on every exit path
struct EventPair {
  hipEvent_t e[2] = {nullptr, nullptr};
  EventPair() {
    for (auto& x : e) VS_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
  }
  ~EventPair() {
    for (auto& x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

static inline c128 C(double r, double i = 0) { return mk<double>(r, i); }
static inline double cabs_(c128 a) { return std::hypot(a.x, a.y); }
static inline bool cisinf(c128 a) { return std::isinf(a.x) || std::isinf(a.y); }
static inline bool cfinite(c128 a) { return std::isfinite(a.x) && std::isfinite(a.y); }

template <typename T>
static double eps_of() {
  return sizeof(T) == 8 ? DBL_EPSILON : (double)FLT_EPSILON;
}

// MATLAB bicgstab semantics (the restatement of the built-in documented alongside the oracle), with
// the scalars of each iteration formed on the device.  Fast path (no pending true-residual checks): both
// half steps of iteration i are launched back to back, a copy of the scalars and an event follow, and
// the first half step of iteration i+1 is launched before the host waits for that event, so the device
// never idles while the host replays MATLAB's control logic on iteration i's scalars.  The vectors that
// half step writes (p, v, xhalf, s) are kept per iteration parity, so iteration i's are intact; when the
// logic leaves the fast path (a true-residual check, a breakdown, stagnation) the speculative half step
// is cancelled by restoring the scalars read for iteration i.  Whenever the logic takes the half-step
// check branch (a recomputed true residual replaces s), the second half step is recomputed from it,
// exactly as MATLAB orders it; once a check has failed (moresteps > 0) every half step is checked and
// read back on its own.
template <typename T>
SolveResult solve_bicgstab(Engine<T>& E, const LinOp<T>& op, const cplx<T>* b, cplx<T>* x, bool has_x0, double tol,
                           int maxit, bool verify, double bnorm) {
  using cv = cplx<T>;
  const int64_t n = E.n;
  vs_ctx* ctx = E.ctx;
  SolveResult R;
  const double eps = eps_of<T>();
  const double n2b = bnorm > 0 ? bnorm : E.norm(b);
  if (n2b == 0) {
    VS_HIP(hipMemsetAsync(x, 0, sizeof(cv) * n, ctx->stream));
    R.flag = 0;
    R.relres = 0;
    R.iter = 0;
    R.resvec = {0.0};
    return R;
  }
  Vecs<T> W(ctx, vs_ctx::WS_KRYLOV, n, 14);
  cv *r = W[0], *rt = W[1], *ph = W[2], *sh = W[3], *t = W[4], *xmin = W[5];
  // p, v, xhalf, s of iteration i live in slot i & 1
  auto P = [&](int i) { return W[6 + (i & 1)]; };
  auto V = [&](int i) { return W[8 + (i & 1)]; };
  auto XH = [&](int i) { return W[10 + (i & 1)]; };
  auto S = [&](int i) { return W[12 + (i & 1)]; };
  int flag = 1;
  double imin = 0;
  const double tolb = tol * n2b;
  double normr;
  BiScal* sc = static_cast<BiScal*>(ctx->workspace(vs_ctx::WS_KRYLOV_AUX, sizeof(BiScal)));
  BiScal* hs = reinterpret_cast<BiScal*>(ctx->pinned);  // two read-back slots (iteration parity)
  BiScal h{};
  h.rho = C(1);
  h.rho1 = C(1);
  h.omega = C(1);
  h.alpha = C(0);
  h.beta = C(0);
  h.first = 1;
  h.has_pc = op.M ? 1 : 0;
  // Fused halves (operator and preconditioner permitting) on the fast path only: a first half that must
  // finish its own half step (moresteps > 0) and a second half recomputed from a replaced s use the
  // separate stages, so they never depend on the fused passes' vectors.
  const bool fused = op.F1 && op.F2 && op.M;
  BiParts bp;
  if (fused) bi_parts_take<T>(E, bp);
  if (has_x0) {
    if (verify) E.copy(x, xmin);  // the minimum-residual iterate is kept only by verified solves
    op.A(x, r);
    normr = E.update(UPD_SUB_, b, nullptr, nullptr, nullptr, r, nullptr, C(0), C(0), true);
    E.copy(r, rt);
  } else {
    // x0 = 0: r = b exactly, no matvec; x = 0, rt = r (and the minimum-residual iterate) in one pass that
    // also leaves the partials of conj(rt) . r = ||b||^2 (and of ||x||^2 = 0 for the fused halves)
    bi_start_dev<T>(E, b, x, r, rt, verify ? xmin : nullptr, fused ? &bp : nullptr);
    normr = n2b;
  }
  double normr_act = normr;
  if (normr <= tolb) {
    R.flag = 0;
    R.relres = normr / n2b;
    R.iter = 0;
    R.resvec = {normr};
    return R;
  }
  std::vector<double> resvec(2 * (size_t)maxit + 1, 0.0);
  resvec[0] = normr;
  int stag = 0, moresteps = 0;
  const int64_t maxmsteps = std::min<int64_t>(std::min<int64_t>(n / 50, 10), n - maxit);
  const int maxstagsteps = 3;
  double normrmin = normr;
  int64_t nres = -1;
  double iter = 0;
  int ii = 0;

  if (has_x0) {
    bi_set_scalars(ctx, sc, h);
    bi_stage_init<T>(E, rt, r, sc);
    if (fused) bi_stage_xnorm<T>(E, x, bp);
  } else {
    bi_set_init_scalars<T>(E, sc, h);  // the scalars set and rho reduced from the start pass's partials
  }
  EventPair ev;

  auto first_half = [&](int i, bool finish_half) {
    cv* pi = P(i);
    if (fused && !finish_half) {
      op.F1(r, P(i - 1), V(i - 1), pi, ph, V(i), rt, sc, bp);
      bi_stage_alpha_fused<T>(E, sc, bp);
      return;
    }
    bi_stage_rho<T>(E, r, V(i - 1), P(i - 1), pi, sc);
    const cv* phv = op.M ? ph : pi;
    if (op.M) op.M(pi, ph);
    op.A(phv, V(i));
    bi_stage_alpha_half<T>(E, rt, V(i), phv, x, r, XH(i), S(i), sc, finish_half);
  };
  // fold_half: the half step's ||s|| is reduced together with the omega stage
  auto second_half = [&](int i, bool fold_half) {
    if (fused && fold_half) {
      op.F2(x, ph, r, V(i), XH(i), S(i), sh, t, sc, bp);
      bi_stage_omega_full_fused<T>(E, t, S(i), sh, XH(i), rt, x, r, sc, bp);
      return;
    }
    const cv* shv = op.M ? sh : S(i);
    if (op.M) op.M(S(i), sh);
    op.A(shv, t);
    bi_stage_omega_full<T>(E, t, S(i), shv, XH(i), rt, x, r, sc, fold_half);
  };
  auto snapshot = [&](int i) {
    VS_HIP(hipMemcpyAsync(hs + (i & 1), sc, sizeof(BiScal), hipMemcpyDeviceToHost, ctx->stream));
    VS_HIP(hipEventRecord(ev.e[i & 1], ctx->stream));
  };
  auto wait = [&](int i) {
    VS_HIP(hipEventSynchronize(ev.e[i & 1]));
    h = hs[i & 1];
  };
  bool spec = false;  // the first half step of iteration ii + 1 is in flight
  auto cancel_spec = [&] {
    if (!spec) return;
    bi_set_scalars(ctx, sc, h);
    spec = false;
  };
  auto clear_flag = [&] {
    // a breakdown seen only by a speculative second half step must not block its recomputation
    bi_clear_flag(ctx, sc);
  };

  for (ii = 1; ii <= maxit; ++ii) {
    bool second_done = false;
    if (moresteps == 0) {
      if (!spec) first_half(ii, false);
      second_half(ii, true);
      snapshot(ii);
      spec = false;
      // the next first half is launched ahead unless this iteration is likely the last: a residual
      // already within 10 tolb converges in about one iteration here, and the half step launched ahead of
      // a converged one is wasted work (skipping it costs one host round trip instead)
      if (ii < maxit && !(normr <= 10 * tolb)) {
        first_half(ii + 1, false);
        spec = true;
      }
      wait(ii);
      second_done = true;
    } else {
      first_half(ii, true);
      snapshot(ii);
      wait(ii);
    }
    // ---- first half step: rho, beta, alpha (breakdowns before the half step) ----
    if (h.flag && h.brk <= 4) {
      cancel_spec();
      flag = h.flag;
      nres = 2 * ii - 1;
      break;
    }
    if (cabs_(h.alpha) * h.nph < eps * h.nx)
      ++stag;
    else
      stag = 0;
    normr = h.ns;
    normr_act = normr;
    resvec[2 * ii - 1] = normr;
    if (!verify && normr <= tolb && moresteps == 0) {
      cancel_spec();
      E.copy(XH(ii), x);
      flag = 0;
      iter = ii - 0.5;
      nres = 2 * ii;
      break;
    }
    if (normr <= tolb || stag >= maxstagsteps || moresteps) {
      cancel_spec();
      op.A(XH(ii), S(ii));
      normr_act = E.update(UPD_SUB_, b, nullptr, nullptr, nullptr, S(ii), nullptr, C(0), C(0), true);
      resvec[2 * ii - 1] = normr_act;
      if (normr_act <= tolb) {
        E.copy(XH(ii), x);
        flag = 0;
        iter = ii - 0.5;
        nres = 2 * ii;
        break;
      } else {
        if (stag >= maxstagsteps && moresteps == 0) stag = 0;
        ++moresteps;
        if (moresteps >= maxmsteps) {
          flag = 3;
          E.copy(XH(ii), x);
          nres = 2 * ii;
          break;
        }
      }
      second_done = false;  // s was replaced by the true residual
    }
    if (normr_act < normrmin) {
      normrmin = normr_act;
      if (verify) E.copy(XH(ii), xmin);
      imin = ii - 0.5;
    }
    if (!second_done) {
      if (h.flag) clear_flag();
      second_half(ii, false);
      snapshot(ii);
      wait(ii);
    }
    // ---- second half step: omega ----
    if (h.flag) {
      cancel_spec();
      flag = h.flag;
      nres = 2 * ii;
      break;
    }
    if (cabs_(h.omega) * h.nsh < eps * h.nxh)
      ++stag;
    else
      stag = 0;
    normr = h.nr;
    normr_act = normr;
    resvec[2 * ii] = normr;
    if (!verify && normr <= tolb && moresteps == 0) {
      cancel_spec();
      flag = 0;
      iter = ii;
      nres = 2 * ii + 1;
      break;
    }
    if (normr <= tolb || stag >= maxstagsteps || moresteps) {
      cancel_spec();
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
      bi_stage_init<T>(E, rt, r, sc);  // the next rho comes from the recomputed residual
    }
    if (normr_act < normrmin) {
      normrmin = normr_act;
      if (verify) E.copy(x, xmin);
      imin = ii;
    }
    if (stag >= maxstagsteps) {
      cancel_spec();
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
  } else if (!verify) {
    // unverified (inner) solves keep the last iterate: their caller checks the result itself
    iter = ii;
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
  VS_HIP(hipStreamSynchronize(ctx->stream));
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
  Vecs<T> Vb(E.ctx, vs_ctx::WS_KRYLOV, n, restart + 4);
  Vecs<T> W = Vb.sub(restart + 1);
  c128* hdev = static_cast<c128*>(E.ctx->workspace(vs_ctx::WS_KRYLOV_AUX, sizeof(c128) * 1025));
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
      // CGS2 with the projection coefficients kept on the device; one read-back per Arnoldi step
      c128* h1d = hdev;
      c128* h2d = h1d + (k + 1);
      c128* nrd = h2d + (k + 1);
      E.mdot_dev(Vb[0], n, k + 1, w, h1d);
      if (!E.gemv_mdot_dev(Vb[0], n, k + 1, h1d, -1.0, w, h2d)) {
        E.gemv_acc_dev(Vb[0], n, k + 1, h1d, -1.0, w);
        E.mdot_dev(Vb[0], n, k + 1, w, h2d);
      }
      E.gemv_acc_norm2_dev(Vb[0], n, k + 1, h2d, -1.0, w, nrd);
      E.scale_inv_norm_dev(w, Vb[k + 1], nrd);
      std::vector<c128> hh(2 * (k + 1) + 1);
      E.read(h1d, 2 * (k + 1) + 1, hh.data());
      std::vector<c128> h(k + 1);
      for (int i = 0; i <= k; ++i) h[i] = hh[i] + hh[k + 1 + i];
      const double hn = std::sqrt(hh[2 * (k + 1)].x);
      for (int i = 0; i <= k; ++i) H[i + (size_t)ldh * k] = h[i];
      H[k + 1 + (size_t)ldh * k] = C(hn);
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
  Vecs<T> Vb(E.ctx, vs_ctx::WS_KRYLOV, n, restart + 4);
  Vecs<T> W = Vb.sub(restart + 1);
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

// ------------------------------------------------------------------------------------------------
// GMRES with deflated restarting (pgmresDR.m:73-280, pgmresDR_iter.m:23-141, getHarmVecs1.m,
// getHarmVecs2.m): after each cycle the k harmonic Ritz vectors of smallest harmonic Ritz value are
// kept as U (with C = A U orthonormal); the next cycle's Arnoldi vectors are orthogonalised against C
// and the iterate is the least-squares minimiser over span[U V].  The small projected problems run on
// the host in FP64 (QR least squares, eigenvalues; eig(A, B) through a Cholesky reduction of the
// Hermitian positive definite B = H2^H H2), the bases live on the device.
// ------------------------------------------------------------------------------------------------
namespace {

using dense::Mat;
using dense::zd;

inline zd Z(c128 a) { return zd(a.x, a.y); }
inline c128 Cz(zd a) { return mk<double>(a.real(), a.imag()); }

std::vector<c128> to_c(const std::vector<zd>& v) {
  std::vector<c128> o(v.size());
  for (size_t i = 0; i < v.size(); ++i) o[i] = Cz(v[i]);
  return o;
}
std::vector<c128> to_c(const Mat& M) {
  std::vector<c128> o(M.a.size());
  for (size_t i = 0; i < M.a.size(); ++i) o[i] = Cz(M.a[i]);
  return o;
}

// indices of the eigenvalues sorted by modulus, ascending (stable)
std::vector<int> order_by_modulus(const std::vector<zd>& w) {
  std::vector<int> idx(w.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return std::abs(w[a]) < std::abs(w[b]); });
  return idx;
}

// getHarmVecs1.m: harmonic Ritz vectors of H (m+1 x m), the k of smallest |theta|
Mat harm_vecs1(int m, int k, const Mat& H) {
  Mat Hm(m, m);
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) Hm(i, j) = H(i, j);
  // last column of inv(Hm^H): solve Hm^H f = e_m
  Mat HmH = dense::adjoint(Hm);
  std::vector<zd> em(m, zd(0));
  em[m - 1] = 1;
  std::vector<zd> f = dense::lstsq(HmH, em);
  const zd h = H(m, m - 1);
  Mat Hr = Hm;
  for (int i = 0; i < m; ++i) Hr(i, m - 1) += h * h * f[i];
  std::vector<zd> w;
  Mat X;
  dense::eig(Hr, w, X);
  const std::vector<int> idx = order_by_modulus(w);
  Mat P(m, k);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < m; ++i) P(i, j) = X(i, idx[j]);
  return P;
}

// getHarmVecs2.m: eig(H2^H A, H2^H H2) with A = [C'U 0; V'U [I; 0]], the k of largest |lambda|
Mat harm_vecs2(int m, int k, const Mat& H2, const Mat& CU, const Mat& VU) {
  Mat A(m + 1, m);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < k; ++i) A(i, j) = CU(i, j);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < m + 1 - k; ++i) A(k + i, j) = VU(i, j);
  for (int i = k; i < m; ++i) A(i, i) = 1;
  const Mat H2h = dense::adjoint(H2);
  const Mat AA = dense::matmul(H2h, A), B = dense::matmul(H2h, H2);
  // B = L L^H: eig(L^{-1} AA L^{-H}) has the same eigenvalues, eigenvectors z = L^{-H} w
  const Mat L = dense::cholesky(B);
  Mat T1 = dense::lower_solve(L, AA);                                     // L^{-1} AA
  Mat T2 = dense::adjoint(dense::lower_solve(L, dense::adjoint(T1)));      // (L^{-1} (L^{-1} AA)^H)^H
  std::vector<zd> w;
  Mat Wv;
  dense::eig(T2, w, Wv);
  Mat Zv = dense::lower_adj_solve(L, Wv);
  const std::vector<int> idx = order_by_modulus(w);
  Mat P(m, k);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < m; ++i) P(i, j) = Zv(i, idx[m - k + j]);
  return P;
}

}  // namespace

template <typename T>
SolveResult solve_pgmresdr(Engine<T>& E, const LinOp<T>& op, int side, const cplx<T>* b, cplx<T>* x, bool has_x0,
                           int restart, double tol, int maxit, int ritz, Recycle* rec) {
  using cv = cplx<T>;
  const int64_t n = E.n;
  SolveResult R;
  VS_CHECK(restart >= 2 && restart <= 500, VS_ERR_UNSUPPORTED, "pgmresDR restart must be in 2..500");
  ritz = ritz <= 0 ? std::min(10, (int)std::lround(restart / 2.0)) : std::min(ritz, restart - 1);
  const bool left = op.M && side == 0, right = op.M && side == 1;
  const int m = restart;
  Vecs<T> Vb(E.ctx, vs_ctx::WS_KRYLOV, n, m + 1 + 4 * ritz + 5);
  Vecs<T> Cb = Vb.sub(m + 1), Ub = Vb.sub(m + 1 + ritz), C2 = Vb.sub(m + 1 + 2 * ritz),
          U2 = Vb.sub(m + 1 + 3 * ritz), W = Vb.sub(m + 1 + 4 * ritz);
  cv *r = W[0], *w = W[1], *tmp = W[2], *xa = W[3], *x0 = W[4];
  cv *Cp = Cb[0], *Up = Ub[0], *Cn = C2[0], *Un = U2[0];
  // r = b - A x0, bp = b, both left-preconditioned (pgmresDR.m:110-139)
  if (has_x0) {
    E.copy(x, x0);
    op.A(x0, r);
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
  VS_HIP(hipMemsetAsync(xa, 0, sizeof(cv) * n, E.ctx->stream));
  int kc = 0;  // deflation vectors in C, U

  // one Arnoldi cycle from r (pgmresDR_iter.m), orthogonalised against C(:, 1:kc); returns p and fills
  // H (p+1 x p) and Bm (kc x p); appends the projected residual norms to resvec
  auto arnoldi = [&](Mat& H, Mat& Bm) -> int {
    const double rn = E.norm(r);
    E.update(UPD_SCALE_, r, nullptr, nullptr, nullptr, Vb[0], nullptr, C(1.0 / rn), C(0), false);
    H = Mat(m + 1, m);
    Bm = Mat(kc, m);
    std::vector<zd> g(m + 1, zd(0)), cs(m), sn(m);
    g[0] = rn;
    Mat Hg(m + 1, m);
    int k = 1;
    for (; k <= m; ++k) {
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
      if (kc > 0) {
        std::vector<c128> bc(kc);
        E.mdot(Cp, n, kc, w, bc.data());
        E.gemv_acc(Cp, n, kc, bc.data(), -1.0, w);
        for (int i = 0; i < kc; ++i) Bm(i, k - 1) = Z(bc[i]);
      }
      std::vector<c128> h(k);
      E.mdot(Vb[0], n, k, w, h.data());
      E.gemv_acc(Vb[0], n, k, h.data(), -1.0, w);
      const double hn = E.norm(w);
      for (int i = 0; i < k; ++i) H(i, k - 1) = Z(h[i]);
      H(k, k - 1) = hn;
      E.update(UPD_SCALE_, w, nullptr, nullptr, nullptr, Vb[k], nullptr, C(hn > 0 ? 1.0 / hn : 0.0), C(0), false);
      // projected residual norm by Givens rotations (same as norm(rhs - H y), y = H \ rhs)
      const int c = k - 1;
      for (int i = 0; i <= k; ++i) Hg(i, c) = H(i, c);
      for (int i = 0; i < c; ++i) {
        const zd hi = Hg(i, c), hi1 = Hg(i + 1, c);
        Hg(i, c) = std::conj(cs[i]) * hi + std::conj(sn[i]) * hi1;
        Hg(i + 1, c) = -sn[i] * hi + cs[i] * hi1;
      }
      const zd a = Hg(c, c), bb = Hg(c + 1, c);
      const double rr = std::sqrt(std::norm(a) + std::norm(bb));
      cs[c] = rr > 0 ? a / rr : zd(1);
      sn[c] = rr > 0 ? bb / rr : zd(0);
      g[c + 1] = -sn[c] * g[c];
      g[c] = std::conj(cs[c]) * g[c];
      const double res = std::abs(g[c + 1]);
      resvec.push_back(res / bnorm);
      if (res < tol * bnorm) break;
    }
    const int p = std::min(k, m);
    Mat Hp(p + 1, p), Bp(kc, p);
    for (int j = 0; j < p; ++j) {
      for (int i = 0; i <= p; ++i) Hp(i, j) = H(i, j);
      for (int i = 0; i < kc; ++i) Bp(i, j) = Bm(i, j);
    }
    H = Hp;
    Bm = Bp;
    return p;
  };

  const int mode = left ? 0 : (right ? 1 : 2);
  const bool recycled = rec && rec->k > 0;
  Mat H, Bm;
  int p = 0;
  if (recycled) {
    // recycled deflation space (GCRO-DR, Parks et al. 2006, Alg. 2 lines 2-6): [C, R] = qr(A U0) by two
    // Cholesky-QR passes, U = U0 R^-1, x = U C^H r, r = r - C C^H r
    VS_CHECK(rec->n == n && rec->k == ritz && rec->mode == mode && rec->elem == (int)sizeof(cv), VS_ERR_ARG,
             "recycled space does not match this solve (unknowns, ritz, preconditioner side or precision)");
    const int k = rec->k;
    const cv* U0 = rec->U.as<cv>();
    for (int i = 0; i < k; ++i) {
      const cv* in = U0 + (int64_t)i * n;
      cv* out = Cp + (int64_t)i * n;
      if (right) {
        op.M(in, tmp);
        op.A(tmp, out);
      } else if (left) {
        op.A(in, tmp);
        op.M(tmp, out);
      } else {
        op.A(in, out);
      }
    }
    Mat Rt(k, k), I(k, k);
    for (int i = 0; i < k; ++i) Rt(i, i) = I(i, i) = 1.0;
    for (int pass = 0; pass < 2; ++pass) {
      Mat G(k, k);
      std::vector<c128> g(k);
      for (int j = 0; j < k; ++j) {
        E.mdot(Cp, n, k, Cp + (int64_t)j * n, g.data());
        for (int i = 0; i < k; ++i) G(i, j) = Z(g[i]);
      }
      const Mat Rk = dense::adjoint(dense::cholesky(G));
      E.combine(Cp, n, k, to_c(dense::right_div_upper(I, Rk)).data(), k, Cn, n, false);
      std::swap(Cp, Cn);
      Rt = dense::matmul(Rk, Rt);
    }
    E.combine(U0, n, k, to_c(dense::right_div_upper(I, Rt)).data(), k, Up, n, false);
    std::vector<c128> cr(k);
    E.mdot(Cp, n, k, r, cr.data());
    E.gemv_acc(Up, n, k, cr.data(), 1.0, xa);
    E.gemv_acc(Cp, n, k, cr.data(), -1.0, r);
    resvec.push_back(E.norm(r) / bnorm);
    kc = k;
  } else {
    // first cycle (pgmresDR.m:152-185)
    p = arnoldi(H, Bm);
    std::vector<c128> vr(p + 1);
    E.mdot(Vb[0], n, p + 1, r, vr.data());
    std::vector<zd> rhs(p + 1);
    for (int i = 0; i <= p; ++i) rhs[i] = Z(vr[i]);
    const std::vector<zd> y = dense::lstsq(H, rhs);
    E.gemv_acc(Vb[0], n, p, to_c(y).data(), 1.0, xa);
    std::vector<zd> hy(p + 1, zd(0));
    for (int j = 0; j < p; ++j)
      for (int i = 0; i <= p; ++i) hy[i] += H(i, j) * y[j];
    E.gemv_acc(Vb[0], n, p + 1, to_c(hy).data(), -1.0, r);
    if (resvec.back() > tol) {
      const int k = ritz > p - 1 ? p - 1 : ritz;
      const Mat P = harm_vecs1(p, k, H);
      Mat Q, Rq;
      dense::qr_econ(dense::matmul(H, P), Q, Rq);
      const Mat PR = dense::right_div_upper(P, Rq);  // U = V(:,1:p) P R^{-1}
      E.combine(Vb[0], n, p, to_c(PR).data(), k, Up, n, false);
      E.combine(Vb[0], n, p + 1, to_c(Q).data(), k, Cp, n, false);
      kc = k;
    }
  }
  // deflated cycles (pgmresDR.m:187-255)
  int outit = 0;
  while (resvec.back() > tol && outit < maxit) {
    ++outit;
    p = arnoldi(H, Bm);
    // U = U D, D = diag(1 / ||U(:, i)||)
    std::vector<double> d(kc);
    for (int i = 0; i < kc; ++i) {
      cv* ui = Up + (int64_t)i * n;
      d[i] = E.norm(ui);
      E.update(UPD_SCALE_, ui, nullptr, nullptr, nullptr, ui, nullptr, C(1.0 / d[i]), C(0), false);
    }
    const int mm = kc + p;
    Mat H2(mm + 1, mm);
    for (int i = 0; i < kc; ++i) H2(i, i) = 1.0 / d[i];
    for (int j = 0; j < p; ++j) {
      for (int i = 0; i < kc; ++i) H2(i, kc + j) = Bm(i, j);
      for (int i = 0; i <= p; ++i) H2(kc + i, kc + j) = H(i, j);
    }
    // y = H2 \ ([C V]' r); x += [U V] y; r -= [C V] (H2 y)
    std::vector<c128> cr(std::max(kc, 1)), vr(p + 1);
    if (kc > 0) E.mdot(Cp, n, kc, r, cr.data());
    E.mdot(Vb[0], n, p + 1, r, vr.data());
    std::vector<zd> rhs(mm + 1);
    for (int i = 0; i < kc; ++i) rhs[i] = Z(cr[i]);
    for (int i = 0; i <= p; ++i) rhs[kc + i] = Z(vr[i]);
    const std::vector<zd> y = dense::lstsq(H2, rhs);
    const std::vector<c128> yc = to_c(y);
    if (kc > 0) E.gemv_acc(Up, n, kc, yc.data(), 1.0, xa);
    E.gemv_acc(Vb[0], n, p, yc.data() + kc, 1.0, xa);
    std::vector<zd> t(mm + 1, zd(0));
    for (int j = 0; j < mm; ++j)
      for (int i = 0; i <= mm; ++i) t[i] += H2(i, j) * y[j];
    const std::vector<c128> tc = to_c(t);
    if (kc > 0) E.gemv_acc(Cp, n, kc, tc.data(), -1.0, r);
    E.gemv_acc(Vb[0], n, p + 1, tc.data() + kc, -1.0, r);
    if (resvec.back() > tol) {
      // C'U and V'U (getHarmVecs2.m:16-18)
      Mat CU(kc, kc), VU(p + 1, kc);
      for (int j = 0; j < kc; ++j) {
        std::vector<c128> a(std::max(kc, 1)), bvec(p + 1);
        E.mdot(Cp, n, kc, Up + (int64_t)j * n, a.data());
        E.mdot(Vb[0], n, p + 1, Up + (int64_t)j * n, bvec.data());
        for (int i = 0; i < kc; ++i) CU(i, j) = Z(a[i]);
        for (int i = 0; i <= p; ++i) VU(i, j) = Z(bvec[i]);
      }
      const Mat P = harm_vecs2(mm, ritz, H2, CU, VU);
      Mat Q, Rq;
      dense::qr_econ(dense::matmul(H2, P), Q, Rq);
      const Mat PR = dense::right_div_upper(P, Rq);
      // U_new = [U V(:,1:p)] P R^{-1},  C_new = [C V] Q
      const std::vector<c128> prc = to_c(PR), qc = to_c(Q);
      std::vector<c128> top((size_t)kc * ritz), bot((size_t)p * ritz), qtop((size_t)kc * ritz),
          qbot((size_t)(p + 1) * ritz);
      for (int j = 0; j < ritz; ++j) {
        for (int i = 0; i < kc; ++i) {
          top[i + (size_t)kc * j] = prc[i + (size_t)mm * j];
          qtop[i + (size_t)kc * j] = qc[i + (size_t)(mm + 1) * j];
        }
        for (int i = 0; i < p; ++i) bot[i + (size_t)p * j] = prc[kc + i + (size_t)mm * j];
        for (int i = 0; i <= p; ++i) qbot[i + (size_t)(p + 1) * j] = qc[kc + i + (size_t)(mm + 1) * j];
      }
      E.combine(Up, n, kc, top.data(), ritz, Un, n, false);
      E.combine(Vb[0], n, p, bot.data(), ritz, Un, n, true);
      E.combine(Cp, n, kc, qtop.data(), ritz, Cn, n, false);
      E.combine(Vb[0], n, p + 1, qbot.data(), ritz, Cn, n, true);
      std::swap(Up, Un);
      std::swap(Cp, Cn);
      kc = ritz;
    }
  }
  // keep the final deflation space for the next solve
  if (rec && (outit > 0 || recycled) && kc == ritz) {
    rec->U.ensure(sizeof(cv) * (size_t)n * kc);
    VS_HIP(hipMemcpyAsync(rec->U.p, Up, sizeof(cv) * (size_t)n * kc, hipMemcpyDeviceToDevice, E.ctx->stream));
    rec->n = n;
    rec->k = kc;
    rec->mode = mode;
    rec->elem = (int)sizeof(cv);
  }
  // x = x0 + R1 x (pgmresDR.m:257-271)
  if (right) {
    op.M(xa, tmp);
    E.copy(tmp, xa);
  }
  if (has_x0) E.update(UPD_AXPBY_, x0, nullptr, nullptr, nullptr, xa, nullptr, C(1), C(1), false);
  E.copy(xa, x);
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

template SolveResult solve_pgmresdr<double>(Engine<double>&, const LinOp<double>&, int, const c128*, c128*, bool, int,
                                            double, int, int, Recycle*);
template SolveResult solve_pgmresdr<float>(Engine<float>&, const LinOp<float>&, int, const c64*, c64*, bool, int,
                                           double, int, int, Recycle*);

template SolveResult solve_bicgstab<double>(Engine<double>&, const LinOp<double>&, const c128*, c128*, bool, double, int,
                                            bool, double);
template SolveResult solve_bicgstab<float>(Engine<float>&, const LinOp<float>&, const c64*, c64*, bool, double, int,
                                           bool, double);
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
  if (n2b == 0) {
    if (!has_x0) VS_HIP(hipMemsetAsync(x, 0, sizeof(c128) * n, ctx->stream));
    R.flag = 0;
    R.resvec = {0.0};
    return R;
  }
  Carve ws;
  const size_t i64a = ws.add(sizeof(c128) * n), i32a = ws.add(sizeof(c64) * n), i32b = ws.add(sizeof(c64) * n);
  ws.take(ctx, vs_ctx::WS_REFINE);
  const Span r64 = ws.at(i64a), r32 = ws.at(i32a), d32 = ws.at(i32b);
  bool x_set = has_x0;  // x holds an iterate (x0, or the first update from x0 = 0)
  int tot = 0;
  double relres = 0;
  R.flag = 1;
  int outer = 0;
  for (outer = 0; outer <= max_outer; ++outer) {
    double nr;
    const bool from_b = outer == 0 && !has_x0;  // x0 = 0: r = b exactly, no matvec
    if (from_b) {
      nr = n2b;
    } else {
      A64.A(x, r64.as<c128>());
      nr = E64.update(UPD_SUB_, b, nullptr, nullptr, nullptr, r64.as<c128>(), nullptr, mk<double>(0, 0),
                      mk<double>(0, 0), true);
    }
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
    // scale the residual to unit norm and round it to FP32 in one pass
    scale_to32_dev(ctx, from_b ? b : r64.as<c128>(), r32.as<c64>(), n, 1.0 / nr);
    SolveResult in;
    if (inner_solver == 0)
      // the right-hand side was scaled to unit norm before rounding: its norm is 1 to FP32 rounding, which
      // is all the inner tolerance needs (no reduction and host round trip for it)
      in = solve_bicgstab<float>(E32, A32, r32.as<c64>(), d32.as<c64>(), false, inner_tol, inner_maxit, false, 1.0);
    else
      in = solve_gmres<float>(E32, A32, r32.as<c64>(), d32.as<c64>(), false, restart, inner_tol, inner_maxit);
    tot += inner_solver == 0 ? (int)(in.resvec.size() - 1) : (int)(in.resvec.size() - 1);
    // x += nr d (x = nr d on the first step from x0 = 0, which skips zeroing x)
    axpy_from32_dev(ctx, d32.as<c64>(), x, n, nr, from_b);
    x_set = true;
  }
  if (!x_set) VS_HIP(hipMemsetAsync(x, 0, sizeof(c128) * n, ctx->stream));
  R.relres = relres;
  R.outer = outer;
  R.iter = outer;
  if (inner_total) *inner_total = tot;
  VS_HIP(hipStreamSynchronize(ctx->stream));
  return R;
}

}  // namespace vs

namespace vs {

// Batched BiCGStab: nrhs independent systems advance together, every matvec / preconditioner launch
// covers the running systems of the batch and the per-system scalars live on the device, so an
// iteration needs one read-back.  With VS_GRAPH=1 the iteration is captured once and replayed as a graph (on this runtime a
// graph launch costs more host time than the eager launches it replaces, so eager is the default).
// A system stops when its recursive residual meets tol at a half or full step (x =
// xhalf at a half step); breakdowns stop that system alone.  At the end the true residuals are formed;
// flag 0 requires the true relative residual <= tol, flag 1 otherwise (or when maxit is reached), 4 on
// breakdown.
template <typename T>
BatchResult solve_bicgstab_batch(Engine<T>& E, const LinOp<T>& op, int k, const cplx<T>* B, cplx<T>* X, double tol,
                                 int maxit, const c128* diag) {
  using cv = cplx<T>;
  const int64_t n = E.n;
  vs_ctx* ctx = E.ctx;
  VS_CHECK(k >= 1 && (size_t)k * sizeof(BiScal) <= sizeof(c128) * VS_PINNED_C128, VS_ERR_UNSUPPORTED,
           "too many right-hand sides for one batch");
  BatchResult R;
  R.flag.assign(k, 1);
  R.relres.assign(k, 0.0);
  R.iter.assign(k, 0.0);
  // the iterates live in xw (slot j holds system sys[j]); a finished system's solution is copied to X.
  // Once a quarter of the slots hold finished systems, the running ones are moved to the front and the
  // batch shrinks (the state carried between iterations is r, rt, p, v, xw and the scalars).
  Vecs<T> W(ctx, vs_ctx::WS_KRYLOV, n * k, 10);
  cv *r = W[0], *rt = W[1], *p = W[2], *v = W[3], *ph = W[4], *s = W[5], *sh = W[6], *t = W[7], *xh = W[8];
  cv* xw = W[9];
  const cv* phv = op.Mb ? ph : p;
  const cv* shv = op.Mb ? sh : s;
  Carve ws;
  const size_t ip = ws.add(sizeof(c128) * (size_t)E.grid * k * 4), in_ = ws.add(sizeof(c128) * k),
               isc = ws.add(sizeof(BiScal) * k), idsl = ws.add(sizeof(c128) * k), idsy = ws.add(sizeof(c128) * k);
  ws.take(ctx, vs_ctx::WS_KRYLOV_AUX);
  const Span part = ws.at(ip), nrm = ws.at(in_), dsc = ws.at(isc);
  BiScal* sc = dsc.as<BiScal>();
  // per-system diagonals: dslot follows the systems when the batch is compacted, dsys stays in system order
  c128* dslot = ws.at(idsl).as<c128>();
  c128* dsys = ws.at(idsy).as<c128>();
  VS_CHECK(!diag || op.Abd, VS_ERR_INTERNAL, "operator has no per-right-hand-side diagonal form");
  if (diag) {
    VS_HIP(hipMemcpyAsync(dslot, diag, sizeof(c128) * k, hipMemcpyHostToDevice, ctx->stream));
    VS_HIP(hipMemcpyAsync(dsys, diag, sizeof(c128) * k, hipMemcpyHostToDevice, ctx->stream));
    VS_HIP(hipStreamSynchronize(ctx->stream));
  }
  auto apply_A = [&](const cv* xin, cv* yout, int kk, const c128* d) {
    if (diag)
      op.Abd(xin, yout, kk, d);
    else
      op.Ab(xin, yout, kk);
  };
  VS_HIP(hipMemsetAsync(X, 0, sizeof(cv) * n * k, ctx->stream));
  VS_HIP(hipMemsetAsync(xw, 0, sizeof(cv) * n * k, ctx->stream));
  batch_norm2<T>(E, k, B, part.as<c128>(), nrm.as<c128>());
  std::vector<c128> nb(k);
  E.read(nrm.as<c128>(), k, nb.data());
  std::vector<double> tolb(k);
  std::vector<int> active(k), sys(k);
  BiScal* up = reinterpret_cast<BiScal*>(ctx->pinned_up);
  for (int i = 0; i < k; ++i) {
    const double nbi = std::sqrt(nb[i].x);
    tolb[i] = tol * nbi;
    active[i] = nbi > 0 ? 1 : 0;
    sys[i] = i;
    if (!active[i]) R.flag[i] = 0;
    BiScal h{};
    h.rho = C(1);
    h.rho1 = C(1);
    h.omega = C(1);
    h.first = 1;
    h.has_pc = op.Mb ? 1 : 0;
    h.active = active[i];
    up[i] = h;
  }
  VS_HIP(hipMemcpyAsync(sc, up, sizeof(BiScal) * k, hipMemcpyHostToDevice, ctx->stream));
  VS_HIP(hipMemcpyAsync(r, B, sizeof(cv) * n * k, hipMemcpyDeviceToDevice, ctx->stream));
  VS_HIP(hipMemcpyAsync(rt, B, sizeof(cv) * n * k, hipMemcpyDeviceToDevice, ctx->stream));
  BiScal* hs = reinterpret_cast<BiScal*>(ctx->pinned);
  c128* pt = part.as<c128>();
  int kc = k;  // slots in use
  auto iteration = [&] {
    bib_stage_rho<T>(E, kc, rt, r, v, p, sc, pt);
    if (op.Mb) op.Mb(p, ph, kc);
    apply_A(phv, v, kc, dslot);
    bib_stage_alpha_half<T>(E, kc, rt, v, phv, xw, r, xh, s, sc, pt);
    if (op.Mb) op.Mb(s, sh, kc);
    apply_A(shv, t, kc, dslot);
    bib_stage_omega_full<T>(E, kc, t, s, shv, xh, xw, r, sc, pt);
    VS_HIP(hipMemcpyAsync(hs, sc, sizeof(BiScal) * kc, hipMemcpyDeviceToHost, ctx->stream));
  };
  auto copy_vec = [&](const cv* src, cv* dst) {
    VS_HIP(hipMemcpyAsync(dst, src, sizeof(cv) * n, hipMemcpyDeviceToDevice, ctx->stream));
  };
  const char* ge = getenv("VS_GRAPH");
  bool use_graph = ge && ge[0] == '1';
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int nact = 0;
  for (int i = 0; i < k; ++i) nact += active[i];
  int ii = 0;
  for (ii = 1; ii <= maxit && nact > 0; ++ii) {
    if (use_graph && ii >= 2 && !gexec) {
      if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        bool ok = true;
        try {
          iteration();
        } catch (const Error&) {
          ok = false;
        }
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
        if (ok && ec == hipSuccess && g && hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0) == hipSuccess) {
          graph = g;
        } else {
          if (g) hipGraphDestroy(g);
          gexec = nullptr;
          use_graph = false;
          (void)hipGetLastError();
        }
      } else {
        use_graph = false;
        (void)hipGetLastError();
      }
    }
    if (gexec)
      VS_HIP(hipGraphLaunch(gexec, ctx->stream));
    else
      iteration();
    VS_HIP(hipStreamSynchronize(ctx->stream));
    bool changed = false;
    for (int j = 0; j < kc; ++j) {
      if (!active[j]) continue;
      const int i = sys[j];
      const BiScal& h = hs[j];
      int stop = 0;
      if (h.flag && h.brk <= 4) {
        R.flag[i] = h.flag;
        R.iter[i] = ii - 1;
        stop = 1;
      } else if (h.ns <= tolb[j]) {
        copy_vec(xh + (int64_t)j * n, X + (int64_t)i * n);
        R.flag[i] = 0;
        R.iter[i] = ii - 0.5;
        stop = 2;
      } else if (h.flag) {
        R.flag[i] = h.flag;
        R.iter[i] = ii - 1;
        stop = 1;
      } else if (h.nr <= tolb[j]) {
        R.flag[i] = 0;
        R.iter[i] = ii;
        stop = 1;
      }
      if (stop) {
        if (stop == 1) copy_vec(xw + (int64_t)j * n, X + (int64_t)i * n);
        active[j] = 0;
        --nact;
        changed = true;
        *reinterpret_cast<int*>(reinterpret_cast<char*>(ctx->pinned_up) + sizeof(int) * j) = 0;
      } else {
        R.iter[i] = ii;
      }
    }
    if (changed && !gexec && nact > 0 && 4 * nact <= 3 * kc) {
      // compact: running slots to the front (a source slot is never below its destination, and never a
      // later destination, so the copies run in increasing order)
      int jn = 0;
      for (int j = 0; j < kc; ++j) {
        if (!active[j]) continue;
        if (j != jn) {
          for (cv* vec : {r, rt, p, v, xw}) copy_vec(vec + (int64_t)j * n, vec + (int64_t)jn * n);
          VS_HIP(hipMemcpyAsync(sc + jn, sc + j, sizeof(BiScal), hipMemcpyDeviceToDevice, ctx->stream));
          if (diag)
            VS_HIP(hipMemcpyAsync(dslot + jn, dslot + j, sizeof(c128), hipMemcpyDeviceToDevice, ctx->stream));
          sys[jn] = sys[j];
          tolb[jn] = tolb[j];
          active[jn] = 1;
        }
        ++jn;
      }
      kc = jn;
      VS_HIP(hipStreamSynchronize(ctx->stream));
    } else if (changed) {
      // deactivate the finished systems on the device (the staging ints are consumed before returning)
      for (int j = 0; j < kc; ++j)
        if (!active[j])
          VS_HIP(hipMemcpyAsync(&sc[j].active, reinterpret_cast<char*>(ctx->pinned_up) + sizeof(int) * j, sizeof(int),
                                hipMemcpyHostToDevice, ctx->stream));
      VS_HIP(hipStreamSynchronize(ctx->stream));
    }
  }
  // systems still running at maxit keep their last iterate
  for (int j = 0; j < kc; ++j)
    if (active[j]) copy_vec(xw + (int64_t)j * n, X + (int64_t)sys[j] * n);
  R.iterations = ii - 1;
  if (gexec) hipGraphExecDestroy(gexec);
  if (graph) hipGraphDestroy(graph);
  // true residuals
  apply_A(X, r, k, dsys);
  batch_sub<T>(E, k, B, r);
  batch_norm2<T>(E, k, r, part.as<c128>(), nrm.as<c128>());
  std::vector<c128> nr(k);
  E.read(nrm.as<c128>(), k, nr.data());
  for (int i = 0; i < k; ++i) {
    const double nbi = std::sqrt(nb[i].x);
    R.relres[i] = nbi > 0 ? std::sqrt(nr[i].x) / nbi : 0.0;
    if (R.flag[i] == 0 && nbi > 0 && R.relres[i] > tol) R.flag[i] = 1;
  }
  VS_HIP(hipStreamSynchronize(ctx->stream));
  return R;
}

template BatchResult solve_bicgstab_batch<double>(Engine<double>&, const LinOp<double>&, int, const c128*, c128*,
                                                  double, int, const c128*);
template BatchResult solve_bicgstab_batch<float>(Engine<float>&, const LinOp<float>&, int, const c64*, c64*, double,
                                                 int, const c128*);

// batched refinement: FP64 residuals of the whole batch, FP32 batched BiCGStab for the corrections
BatchResult solve_refine_batch(Engine<double>& E64, Engine<float>& E32, const LinOp<double>& A64,
                               const LinOp<float>& A32, int k, const c128* B, c128* X, double tol, int max_outer,
                               double inner_tol, int inner_maxit, int* outer_out, int* inner_total,
                               const c128* diag) {
  const int64_t n = E64.n;
  vs_ctx* ctx = E64.ctx;
  BatchResult R;
  R.flag.assign(k, 1);
  R.relres.assign(k, 0.0);
  R.iter.assign(k, 0.0);
  Carve ws;
  const size_t i64a = ws.add(sizeof(c128) * n * k), i32a = ws.add(sizeof(c64) * n * k),
               i32b = ws.add(sizeof(c64) * n * k), i64b = ws.add(sizeof(c128) * n * k);
  ws.take(ctx, vs_ctx::WS_REFINE);
  const Span R64 = ws.at(i64a), R32 = ws.at(i32a), D32 = ws.at(i32b), D64 = ws.at(i64b);
  Carve wa;
  const size_t ip = wa.add(sizeof(c128) * (size_t)E64.grid * k), in_ = wa.add(sizeof(c128) * k),
               ic = wa.add(sizeof(double) * k), idg = wa.add(sizeof(c128) * k);
  wa.take(ctx, vs_ctx::WS_REFINE_AUX);
  const Span part = wa.at(ip), nrm = wa.at(in_), coef = wa.at(ic);
  c128* dsys = wa.at(idg).as<c128>();
  VS_CHECK(!diag || A64.Abd, VS_ERR_INTERNAL, "operator has no per-right-hand-side diagonal form");
  if (diag) {
    VS_HIP(hipMemcpyAsync(dsys, diag, sizeof(c128) * k, hipMemcpyHostToDevice, ctx->stream));
    VS_HIP(hipStreamSynchronize(ctx->stream));
  }
  VS_HIP(hipMemsetAsync(X, 0, sizeof(c128) * n * k, ctx->stream));
  batch_norm2<double>(E64, k, B, part.as<c128>(), nrm.as<c128>());
  std::vector<c128> h(k);
  E64.read(nrm.as<c128>(), k, h.data());
  std::vector<double> nb(k);
  for (int i = 0; i < k; ++i) nb[i] = std::sqrt(h[i].x);
  int tot = 0, outer = 0;
  for (outer = 0; outer <= max_outer; ++outer) {
    if (outer == 0) {
      // X = 0: R = B exactly, no matvec
      VS_HIP(hipMemcpyAsync(R64.p, B, sizeof(c128) * n * k, hipMemcpyDeviceToDevice, ctx->stream));
    } else {
      if (diag)
        A64.Abd(X, R64.as<c128>(), k, dsys);
      else
        A64.Ab(X, R64.as<c128>(), k);
      batch_sub<double>(E64, k, B, R64.as<c128>());
      batch_norm2<double>(E64, k, R64.as<c128>(), part.as<c128>(), nrm.as<c128>());
      E64.read(nrm.as<c128>(), k, h.data());
    }
    bool all = true;
    std::vector<double> sc(k), back(k);
    for (int i = 0; i < k; ++i) {
      const double nr = std::sqrt(h[i].x);
      R.relres[i] = nb[i] > 0 ? nr / nb[i] : 0.0;
      const bool done = R.relres[i] <= tol;
      R.flag[i] = done ? 0 : 1;
      all = all && done;
      sc[i] = done || nr == 0 ? 0.0 : 1.0 / nr;  // converged systems get a zero right-hand side
      back[i] = done ? 0.0 : nr;
    }
    if (all || outer == max_outer) break;
    VS_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(ctx->pinned_up, sc.data(), sizeof(double) * k);
    VS_HIP(hipMemcpyAsync(coef.p, ctx->pinned_up, sizeof(double) * k, hipMemcpyHostToDevice, ctx->stream));
    batch_scale<double>(E64, k, R64.as<c128>(), R64.as<c128>(), coef.as<double>(), false);
    convert_dev(ctx, R64.p, VS_FP64, R32.p, VS_FP32, n * k);
    BatchResult in =
        solve_bicgstab_batch<float>(E32, A32, k, R32.as<c64>(), D32.as<c64>(), inner_tol, inner_maxit, diag);
    tot += in.iterations;
    convert_dev(ctx, D32.p, VS_FP32, D64.p, VS_FP64, n * k);
    VS_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(ctx->pinned_up, back.data(), sizeof(double) * k);
    VS_HIP(hipMemcpyAsync(coef.p, ctx->pinned_up, sizeof(double) * k, hipMemcpyHostToDevice, ctx->stream));
    batch_scale<double>(E64, k, D64.as<c128>(), X, coef.as<double>(), true);
  }
  if (outer_out) *outer_out = outer;
  if (inner_total) *inner_total = tot;
  VS_HIP(hipStreamSynchronize(ctx->stream));
  return R;
}

}  // namespace vs
