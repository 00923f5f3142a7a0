// Host-side solver engine over device vectors.
#pragma once

#include <functional>

#include "handles.hpp"

namespace vs {

// Vector engine bound to a context: fused update kernels and deterministic reductions of length-n
// complex vectors of precision T.
template <typename T>
struct Engine {
  vs_ctx* ctx;
  int64_t n;
  int grid;
  Engine(vs_ctx* c, int64_t n_);
  void fetch(int K, c128* out);
  void dots(int K, const cplx<T>* const* a, const cplx<T>* const* b, c128* out);  // sum conj(a) b
  double update(int op, const cplx<T>* x0, const cplx<T>* x1, const cplx<T>* x2, const cplx<T>* x3, cplx<T>* y,
                cplx<T>* z, c128 al, c128 be, bool want_norm);
  double norm(const cplx<T>* x);
  void copy(const cplx<T>* x, cplx<T>* y);
  void mdot(const cplx<T>* V, int64_t ldv, int k, const cplx<T>* w, c128* out);
  void gemv_acc(const cplx<T>* V, int64_t ldv, int k, const c128* h_host, double sgn, cplx<T>* w);
  // variants that leave their results on the device (no host round trip)
  void mdot_dev(const cplx<T>* V, int64_t ldv, int k, const cplx<T>* w, c128* out_dev);
  void gemv_acc_dev(const cplx<T>* V, int64_t ldv, int k, const c128* h_dev, double sgn, cplx<T>* w);
  // w += sgn V h and out_dev = V^H w (of the updated w) in one pass over V; false (nothing launched) when k
  // exceeds the fused pass's limit or VS_GMDOT=0
  bool gemv_mdot_dev(const cplx<T>* V, int64_t ldv, int k, const c128* h_dev, double sgn, cplx<T>* w,
                     c128* out_dev);
  void norm2_dev(const cplx<T>* x, c128* out_dev);
  // gemv_acc_dev with ||w||^2 of the updated w in nrm2_dev (one pass)
  void gemv_acc_norm2_dev(const cplx<T>* V, int64_t ldv, int k, const c128* h_dev, double sgn, cplx<T>* w,
                          c128* nrm2_dev);
  void scale_inv_norm_dev(const cplx<T>* w, cplx<T>* y, const c128* nrm2_dev);
  void read(const c128* dev, int K, c128* host);
  // Y(:, j) (+)= sum_i X(:, i) P(i, j), P q x k column-major on the host, columns of X / Y at stride ldx / ldy
  void combine(const cplx<T>* X, int64_t ldx, int q, const c128* P_host, int k, cplx<T>* Y, int64_t ldy,
               bool accumulate);
};

// device-resident scalars of one BiCGStab iteration (read back once per iteration)
struct BiScal {
  c128 rho, rho1, alpha, omega, beta;
  c128 rho_n;                 // conj(rt) . r of the current r (the next iteration's rho)
  double ns, nr;              // ||s|| after the half step, ||r|| after the full step
  double nph, nx, nsh, nxh;   // norms of the stagnation tests
  int first;                  // 1 until the first full step (p = r)
  int flag;                   // 0 running, 2 non-finite preconditioned vector, 4 breakdown
  int brk;                    // where the breakdown was detected
  int has_pc;
  int active;                 // batched solves: 0 once this right-hand side is finished
};

// the scalars of a solve set from the host (a one-thread kernel taking them by value), and the flag cleared
void bi_set_scalars(vs_ctx* ctx, BiScal* sc, const BiScal& h);
void bi_clear_flag(vs_ctx* ctx, BiScal* sc);
// p = r + beta (p0 - omega v0) (p = r in the first iteration), out of place
template <typename T>
void bi_stage_rho(Engine<T>& E, const cplx<T>* r, const cplx<T>* v0, const cplx<T>* p0, cplx<T>* p, BiScal* sc);
// finish_half: reduce ||s|| now (otherwise the omega stage folds it in: fold_half)
template <typename T>
void bi_stage_alpha_half(Engine<T>& E, const cplx<T>* rt, const cplx<T>* v, const cplx<T>* ph, const cplx<T>* x,
                         const cplx<T>* r, cplx<T>* xh, cplx<T>* s, BiScal* sc, bool finish_half);
template <typename T>
void bi_stage_omega_full(Engine<T>& E, const cplx<T>* t, const cplx<T>* s, const cplx<T>* sh, const cplx<T>* xh,
                         const cplx<T>* rt, cplx<T>* x, cplx<T>* r, BiScal* sc, bool fold_half);
// rho_n = conj(rt) . r for an r that did not come from the full step (start, replaced residual)
template <typename T>
void bi_stage_init(Engine<T>& E, const cplx<T>* rt, const cplx<T>* r, BiScal* sc);

// Per-workgroup FP64 partial sums written by the fused BiCGStab halves (device buffers carved from the
// context's reduction area; each count is set by the launch that writes the buffer).
struct BiParts {
  c128* ph = nullptr;  // ||ph||^2 (first half, the preconditioner's last pass)
  c128* v = nullptr;   // conj(rt) . v (first half, the matvec's last pass)
  c128* sx = nullptr;  // ||s||^2 and ||xh||^2, interleaved (second half, the preconditioner's first pass)
  c128* sh = nullptr;  // ||sh||^2 (second half, the preconditioner's last pass)
  c128* t = nullptr;   // ||t||^2 and conj(t) . s, interleaved (second half, the matvec's last pass)
  c128* x = nullptr;   // ||x||^2 (the full step's update, or the start of the solve)
  int n_ph = 0, n_v = 0, n_sx = 0, n_sh = 0, n_t = 0, n_x = 0;
  int cap = 0;         // capacity of each buffer, in partials per value
};
template <typename T>
void bi_parts_take(Engine<T>& E, BiParts& P);
// ||x||^2 partials for the first fused alpha stage
template <typename T>
void bi_stage_xnorm(Engine<T>& E, const cplx<T>* x, BiParts& P);
// alpha from the partials of the fused first half (conj(rt) . v, ||ph||^2, ||x||^2)
template <typename T>
void bi_stage_alpha_fused(Engine<T>& E, BiScal* sc, const BiParts& P);
// omega (with the half step's ||s|| folded in) from the partials of the fused second half, then the full
// step x = xh + omega sh, r = s - omega t with ||r||, conj(rt) . r and ||x||^2 partials, and its scalars
template <typename T>
void bi_stage_omega_full_fused(Engine<T>& E, const cplx<T>* t, const cplx<T>* s, const cplx<T>* sh, const cplx<T>* xh,
                               const cplx<T>* rt, cplx<T>* x, cplx<T>* r, BiScal* sc, BiParts& P);

// batched forms (nrhs systems, vectors stored consecutively, one BiScal per system)
template <typename T>
void bib_stage_rho(Engine<T>& E, int nrhs, const cplx<T>* rt, const cplx<T>* r, const cplx<T>* v, cplx<T>* p,
                   BiScal* sc, c128* part);
template <typename T>
void bib_stage_alpha_half(Engine<T>& E, int nrhs, const cplx<T>* rt, const cplx<T>* v, const cplx<T>* ph,
                          const cplx<T>* x, const cplx<T>* r, cplx<T>* xh, cplx<T>* s, BiScal* sc, c128* part);
template <typename T>
void bib_stage_omega_full(Engine<T>& E, int nrhs, const cplx<T>* t, const cplx<T>* s, const cplx<T>* sh,
                          const cplx<T>* xh, cplx<T>* x, cplx<T>* r, BiScal* sc, c128* part);
template <typename T>
void batch_norm2(Engine<T>& E, int nrhs, const cplx<T>* x, c128* part, c128* out_dev);

template <typename T>
void batch_scale(Engine<T>& E, int nrhs, const cplx<T>* x, cplx<T>* y, const double* c_dev, bool accumulate);
template <typename T>
void batch_sub(Engine<T>& E, int nrhs, const cplx<T>* x, cplx<T>* y);


enum { UPD_AXPBY_ = 0, UPD_BICG_P_ = 1, UPD_HALF_ = 2, UPD_SUB_ = 3, UPD_COPY_ = 4, UPD_SCALE_ = 5 };

// device linear operator y = A x and preconditioner y = M^{-1} x (M may be empty)
template <typename T>
struct LinOp {
  std::function<void(const cplx<T>*, cplx<T>*)> A;
  std::function<void(const cplx<T>*, cplx<T>*)> M;
  // batched forms over nrhs consecutive vectors
  std::function<void(const cplx<T>*, cplx<T>*, int)> Ab;
  std::function<void(const cplx<T>*, cplx<T>*, int)> Mb;
  // batched matvec with a per-right-hand-side diagonal (device array, one complex double per vector)
  std::function<void(const cplx<T>*, cplx<T>*, int, const c128*)> Abd;
  // Fused halves of one BiCGStab iteration (one right-hand side), set when the operator is an FFT operator
  // with a slab-major matvec and the preconditioner a circulant one on the same domain; they produce the
  // same vectors as the separate stages, with the Krylov updates and inner products formed inside the
  // FFT passes.
  //   F1(r, p0, v0, p, ph, v, rt, sc, P): p = r + beta (p0 - omega v0) (p = r at the first iteration),
  //       ph = M^{-1} p, v = A ph; partials P.ph, P.v
  std::function<void(const cplx<T>*, const cplx<T>*, const cplx<T>*, cplx<T>*, cplx<T>*, cplx<T>*, const cplx<T>*,
                     const BiScal*, BiParts&)>
      F1;
  //   F2(x, ph, r, v, xh, s, sh, t, sc, P): xh = x + alpha ph, s = r - alpha v, sh = M^{-1} s, t = A sh;
  //       partials P.sx, P.sh, P.t
  std::function<void(const cplx<T>*, const cplx<T>*, const cplx<T>*, const cplx<T>*, cplx<T>*, cplx<T>*, cplx<T>*,
                     cplx<T>*, const BiScal*, BiParts&)>
      F2;
};

struct BatchResult {
  std::vector<int> flag;
  std::vector<double> relres, iter;
  int iterations = 0;
};
// diag (nullable, host, nrhs values): the diagonal Gram Mr of each system's operator, which otherwise
// share the operator's own (the DDA polarisability of each incidence, getOPERATOR_DDA.m:15-26)
template <typename T>
BatchResult solve_bicgstab_batch(Engine<T>& E, const LinOp<T>& op, int nrhs, const cplx<T>* B, cplx<T>* X, double tol,
                                 int maxit, const c128* diag = nullptr);

struct SolveResult {
  int flag = 1;
  double relres = 0;
  double iter = 0;    // bicgstab: possibly x.5
  int outer = 0, inner = 0;
  std::vector<double> resvec;
};

template <typename T>
// verify = false (inner solves of the refinement, whose result the FP64 residual checks): a recursive
// residual norm <= tol * ||b|| ends the solve without MATLAB's true-residual check, and an unconverged
// solve returns its last iterate (no minimum-residual iterate is kept)
// bnorm > 0: ||b|| known to the caller (no reduction and host round trip for it)
SolveResult solve_bicgstab(Engine<T>& E, const LinOp<T>& op, const cplx<T>* b, cplx<T>* x, bool has_x0, double tol,
                           int maxit, bool verify = true, double bnorm = -1.0);
template <typename T>
SolveResult solve_gmres(Engine<T>& E, const LinOp<T>& op, const cplx<T>* b, cplx<T>* x, bool has_x0, int restart,
                        double tol, int maxit);
template <typename T>
SolveResult solve_pgmres(Engine<T>& E, const LinOp<T>& op, int side, const cplx<T>* b, cplx<T>* x, bool has_x0,
                         int restart, double tol, int maxit);

// deflation space kept between GMRES-DR solves (device, n x k column-major, of the preconditioned
// operator of one side: 0 left, 1 right, 2 none)
struct Recycle {
  DevBuf U;
  int64_t n = 0;
  int k = 0;
  int mode = -1;
  int elem = 0;  // bytes per complex value
};

// rec (nullable): a non-empty space replaces the first plain cycle by a projection onto it; on return it
// holds the solve's final deflation space (unchanged when no deflated cycle ran)
template <typename T>
SolveResult solve_pgmresdr(Engine<T>& E, const LinOp<T>& op, int side, const cplx<T>* b, cplx<T>* x, bool has_x0,
                           int restart, double tol, int maxit, int ritz, Recycle* rec = nullptr);

SolveResult solve_refine(Engine<double>& E64, Engine<float>& E32, const LinOp<double>& A64, const LinOp<float>& A32,
                         const c128* b, c128* x, bool has_x0, double tol, int max_outer, int inner_solver,
                         double inner_tol, int inner_maxit, int restart, int* inner_total);

BatchResult solve_refine_batch(Engine<double>& E64, Engine<float>& E32, const LinOp<double>& A64,
                               const LinOp<float>& A32, int nrhs, const c128* B, c128* X, double tol, int max_outer,
                               double inner_tol, int inner_maxit, int* outer_out, int* inner_total,
                               const c128* diag = nullptr);

void convert_dev(vs_ctx* ctx, const void* in, int pin, void* out, int pout, int64_t n);
// out = (FP32) (s * in)
void scale_to32_dev(vs_ctx* ctx, const c128* in, c64* out, int64_t n, double s);
// x = a d (first) or x += a d, d in FP32
void axpy_from32_dev(vs_ctx* ctx, const c64* d, c128* x, int64_t n, double a, bool first);
// x = 0, r = rt = b, xmin = 0 (nullable), with the partials of conj(rt) . r = ||b||^2 for
// bi_set_init_scalars and, when P is given, zero partials of ||x||^2 for the fused halves
template <typename T>
void bi_start_dev(Engine<T>& E, const cplx<T>* b, cplx<T>* x, cplx<T>* r, cplx<T>* rt, cplx<T>* xmin, BiParts* P);
// the scalars set to h and rho_n reduced from bi_start_dev's partials (one kernel)
template <typename T>
void bi_set_init_scalars(Engine<T>& E, BiScal* sc, const BiScal& h);
void mvp_dev(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs);
// the same matvec with a per-right-hand-side diagonal arhs[rhs] (device, nrhs complex doubles) in place of
// the operator's Gram Mr (and Gram conj(Mr) in the transp branch)
void mvp_dev_diag(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs, const c128* arhs);
// the same matvec with the caller's scratch (scratch_bytes of device memory) instead of the operator's own
void mvp_dev_scratch(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs, void* scratch,
                     size_t scratch_bytes);
void pc_apply_dev(vs_pc* pc, const void* x, void* y, int precision, int nrhs);
void pc_blocks_dev(vs_pc* pc, int precision, int nrhs);
// The preconditioner's stages between its first and last x pass, on its work volume (2-level: y FFT, block
// product, inverse y FFT; 1-level: block product): returns the volume holding the result and, in *scale,
// the scale its inverse x transform still applies.
void* pc_mid_dev(vs_pc* pc, int precision, double* scale);
// the preconditioner's work volume for one right-hand side (the first x pass writes it)
void* pc_work_dev(vs_pc* pc, int precision);

// fused BiCGStab halves of an FFT operator and a circulant preconditioner (see LinOp::F1 / F2)
template <typename T>
bool bi_fused_available(vs_op* op, vs_pc* pc);
void bi_fused_set(bool on);
template <typename T>
void bi_fused_h1(vs_op* op, vs_pc* pc, const cplx<T>* r, const cplx<T>* p0, const cplx<T>* v0, cplx<T>* p, cplx<T>* ph,
                 cplx<T>* v, const cplx<T>* rt, const BiScal* sc, BiParts& P);
template <typename T>
void bi_fused_h2(vs_op* op, vs_pc* pc, const cplx<T>* x, const cplx<T>* ph, const cplx<T>* r, const cplx<T>* v,
                 cplx<T>* xh, cplx<T>* s, cplx<T>* sh, cplx<T>* t, const BiScal* sc, BiParts& P);

}  // namespace vs
