/*
 * voxscatter -- MI355X (gfx950) solve loop of the voxel volume-integral-equation scattering solver.
 *
 * C ABI.  Everything is plain pointers and sizes:
 *   - complex arrays are interleaved (re, im) doubles ("complex128") unless a `precision` argument
 *     says otherwise for device pointers (VS_FP32: interleaved floats);
 *   - volumes are column-major L x M x N (x fastest), the component index slowest;
 *   - index arrays are int64, 0-based (a MATLAB gateway subtracts 1 from MATLAB's indices);
 *   - the caller owns every host buffer; handles own device memory and are released by *_destroy;
 *   - each call returns VS_OK (0) or a negative VS_ERR_* code, and vs_last_error() (thread-local)
 *     describes the failure; the library never aborts;
 *   - a context is one HIP stream on one device, to be used from one host thread at a time.
 *
 * Each entry point names the reference (MATLAB) function it replaces, file:line in
 * samuelpgroth/VoxScatter.
 */
#ifndef VOXSCATTER_H
#define VOXSCATTER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VS_OK 0
#define VS_ERR_ARG (-1)
#define VS_ERR_HIP (-2)
#define VS_ERR_OOM (-3)
#define VS_ERR_NODEV (-4)
#define VS_ERR_UNSUPPORTED (-5)
#define VS_ERR_NAN (-6)
#define VS_ERR_INTERNAL (-7)

/* storage precisions (bit mask where a handle may hold both) */
#define VS_FP64 1
#define VS_FP32 2

typedef struct vs_ctx vs_ctx;
typedef struct vs_op vs_op;
typedef struct vs_pc vs_pc;

/* ---------------------------------------------------------------------------------------------- */
/* library / context                                                                               */
/* ---------------------------------------------------------------------------------------------- */
const char* vs_last_error(void);
int vs_version(void);
int vs_device_count(int* n);
int vs_ctx_create(int device, vs_ctx** out);
void vs_ctx_destroy(vs_ctx* ctx);
int vs_ctx_synchronize(vs_ctx* ctx);

/* Device buffers on the context's device (no reference counterpart: plumbing for the *_dev entry
 * points, so that callers need no second GPU runtime).  kind: 0 host->device, 1 device->host,
 * 2 device->device; copies are ordered on the context stream and complete on return. */
int vs_dev_alloc(vs_ctx* ctx, int64_t bytes, void** ptr);
int vs_dev_free(vs_ctx* ctx, void* ptr);
int vs_dev_copy(vs_ctx* ctx, void* dst, const void* src, int64_t bytes, int kind);

/* ---------------------------------------------------------------------------------------------- */
/* operator assembly                                                                               */
/* ---------------------------------------------------------------------------------------------- */

/* LDR polarisability alpha_LDR (replaces getOPERATOR_DDA.m:4-26). */
int vs_alpha_ldr(const double refInd[2], double ko, double d, const double kvec[3], const double Eo[3],
                 double alpha_out[2]);

/* Toeplitz Green dyadic, its 2L x 2M x 2N circulant embedding and the FFT of the embedding, all
 * assembled on the device in FP64 (replaces getOPERATOR_DDA.m:28-122, circulant_nop_const.m:5-40,
 * fft_operator.m:1-8).  r: L*M*N*3 voxel centres.  nearby_quad: 0 = 'off', 1 = 'on'
 * (getOPERATOR_DDA.m:47-79).  precisions: VS_FP64 and/or VS_FP32 copies of the tensor to keep. */
int vs_op_create_dda(vs_ctx* ctx, int L, int M, int N, const double* r, double ko, int nearby_quad,
                     int precisions, vs_op** out);

/* Operator from a caller-supplied circulant FFT fN (2L*2M*2N*6 complex128), e.g. a Galerkin operator
 * (the fN argument of MVP.m:1).  Toep (L*M*N*6, nullable) enables vs_pc_create_circ.  The operator keeps
 * one octant of fN and rebuilds the rest from the mirror parity of each component (xy odd in x and y, xz
 * in x and z, yz in y and z), which the circulant embedding of a dyadic tensor has; an fN without it
 * (deviation above 1e-9 of the component's largest entry) is refused with VS_ERR_ARG. */
/* Galerkin (piecewise-constant JVIE) N operator on an L x M x N grid of cubes of side dx
 * (replaces getOPERATORS.m:71-76 with op = 'N', type = 'FFTofC': assembly_nop_const.m:1-174, then
 * circulant_nop_const.m and fft_operator.m).  The kernel is exp(-i ko R)/R.  singular: the five
 * singular face-pair integrals [ST, EAc, EAo, VAc, VAo] as interleaved complex doubles, or NULL for the
 * reference's own values: the direct evaluation method at Np_1D = 20 (vs_galerkin_singular_directfn,
 * assembly_nop_const.m:16,105-107).  Grids need L, M, N >= 2. */
int vs_op_create_galerkin(vs_ctx* ctx, int L, int M, int N, double dx, double ko, const double* singular,
                          int precisions, vs_op** out);
/* The singular integrals int_S int_S' exp(-i ko R)/R dS dS' of squares of side dx: coincident (ST),
 * edge-adjacent coplanar / orthogonal (EAc, EAo), vertex-adjacent coplanar / orthogonal (VAc, VAo),
 * as 5 interleaved complex doubles (replaces singular_st_const.m, singular_ea_const.m,
 * singular_va_const.m with this library's own semi-analytic method instead of DIRECTFN). */
int vs_galerkin_singular(vs_ctx* ctx, double ko, double dx, double out[10]);
/* The same five integrals by the direct evaluation method (DIRECTFN) with np-point Gauss-Legendre rules,
 * each square split into two triangles (replaces singular_st_const.m, singular_ea_const.m and
 * singular_va_const.m with method 'DIRECTFN', mexcubatures/directfn_ws_{st,ea,va}_const.cpp).  Host
 * computation (no device needed); np = 20 gives the reference's values. */
int vs_galerkin_singular_directfn(double ko, double dx, int np, double out[10]);
/* The face-pair table of the 2 x 2 x 2 near box (replaces the matrices C and P of nearby_correct.m:1-143):
 * for cell c = mx + 2 my + 4 mz (0-based offsets) and observation / source faces i, j (cube_faces.m order
 * -x, +x, -y, +y, -z, +z), code[36 c + 6 i + j] is 1 ST, 2 EAc, 3 EAo, 4 VAc, 5 VAo or 0, and keep[...] is
 * 1 where the cubature is kept.  Host data only (no device needed). */
int vs_galerkin_near_table(int* code, int* keep);

int vs_op_create_fN(vs_ctx* ctx, int L, int M, int N, const double* fN, const double* Toep,
                    int precisions, vs_op** out);
/* Dense operator: y = A x with an n x n column-major complex128 matrix (the matrix operand of the Krylov
 * solvers, iterapp.m:41-58: bicgstab(A, b, ...) with a numeric A).  It has no FFT tensor and no unknown set:
 * every solver, vs_mvp / vs_mvp_dev (transp: A^H x) and vs_pc_create_op accept it; vs_op_set_domain and the
 * circulant preconditioners do not. */
int vs_op_create_dense(vs_ctx* ctx, int64_t n, const double* A, int precisions, vs_op** out);
void vs_op_destroy(vs_op* op);
int vs_op_get_toeplitz(vs_op* op, double* Toep); /* L*M*N*6 complex128 */
int vs_op_get_fN(vs_op* op, double* fN);         /* 2L*2M*2N*6 complex128 */

/* Unknown set and diagonal of MVP.m:1 (idx, Gram, Mr, Mc).  idx: nidx 0-based indices into the
 * 3*L*M*N unknowns.  Gram/Mr/Mc each hold 1 complex value (scalar) or L*M*N values. */
int vs_op_set_domain(vs_op* op, const int64_t* idx, int64_t nidx, const double* Gram, int64_t gram_len,
                     const double* Mr, int64_t mr_len, const double* Mc, int64_t mc_len);
int64_t vs_op_nunknowns(const vs_op* op);
/* The matvec variant used in a precision (VS_FP64 / VS_FP32): 0 five FFT passes, 1 fused padded-slab
 * kernel, 2 fused parity-split slab kernel, 3 parity-split kernel with the two ky parities in separate
 * workgroups.  Picked by timing at the operator's first domain (no reference counterpart: the reference
 * has one FFT path). */
int vs_op_matvec_variant(const vs_op* op, int precision, int* variant);

/* ---------------------------------------------------------------------------------------------- */
/* matvec (replaces MVP.m:1-145 / mv_AN_const.m, both transp_flag branches)                        */
/* ---------------------------------------------------------------------------------------------- */
/* Host buffers, complex128, length vs_op_nunknowns().  transp: 0 'notransp', 1 'transp'. */
int vs_mvp(vs_op* op, const double* x, double* y, int transp);
/* Device buffers of nrhs consecutive vectors in `precision` (VS_FP64 or VS_FP32), on the context
 * stream, asynchronous. */
int vs_mvp_dev(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs);
/* vs_mvp_dev with one diagonal per right-hand side: diag holds nrhs complex doubles in device memory, and
 * vector r uses diag[r] in place of the operator's Gram*Mr (Gram*conj(Mr) in the transp branch).  This
 * is the reference's one-operator-per-incidence DDA system (main_DDA.m:62,75: the LDR polarisability
 * depends on the incidence and polarisation, getOPERATOR_DDA.m:15-26) for many incidences sharing one
 * tensor. */
int vs_mvp_dev_diag(vs_op* op, const void* x, void* y, const void* diag, int transp, int precision, int nrhs);
/* vs_mvp_dev that allocates nothing (no counterpart in the reference: for callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops, hosts that forbid hipMalloc).  The
 * FFT scratch is a fixed 64 MiB device array, one per GPU, held until the process ends, instead of the
 * operator's own buffer.  The operator's tuned variant is used when its scratch fits, else the variant
 * with the smallest scratch; a call whose scratch exceeds 64 MiB (e.g. more right-hand sides than fit)
 * returns VS_ERR_ARG, and so does one that would need an FFT plan the operator has not built.  Every call
 * on one GPU shares that GPU's array (nothing in it needs to start at zero): calls that overlap on one
 * GPU (two contexts' streams) race. */
int vs_mvp_dev_noalloc(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs);

/* ---------------------------------------------------------------------------------------------- */
/* circulant preconditioners                                                                       */
/* ---------------------------------------------------------------------------------------------- */
/* Chan circulant along x + FFT along x (circ_1_level.m:5-90); with blocks != NULL also the L dense
 * 3MN x 3MN 1-level blocks of circ_1_level.m:92-259 ('on'). */
int vs_circ_1_level(vs_ctx* ctx, int L, int M, int N, const double* Toep, double* circ_L, double* blocks);
/* Chan along y + FFT along y, and the L*M blocks of 3N x 3N (circ_2_level.m:1-155).  blocks is
 * ordered block (i,j) at (i + L*j), each block column-major. */
int vs_circ_2_level(vs_ctx* ctx, int L, int M, int N, const double* circ_L, double* circ_M, double* blocks);
/* inv(a*I - diag(w)*B) for nblk column-major n x n blocks (main_DDA.m:105-111,
 * scatteringHumanHead.m:178-184).  w: n complex values, or NULL to use the scalar wscalar. */
int vs_invert_blocks(vs_ctx* ctx, int64_t nblk, int n, const double* blocks, const double a[2],
                     const double* w, const double wscalar[2], double* inv_out);

/* Build the preconditioner of an operator on the device: levels 1 (circ_1_level 'on' + chan_mvp_idx)
 * or 2 (circ_2_level + mvp_circ_2_level); blocks inv(a I - diag(w) C). */
int vs_pc_create_circ(vs_op* op, int levels, const double a[2], const double* w, const double wscalar[2],
                      int precisions, vs_pc** out);
/* Preconditioner from caller-supplied inverse blocks (the inv_blocks argument of
 * mvp_circ_2_level.m:2 / chan_mvp_idx.m:2); same block ordering as the builders above. */
int vs_pc_create_blocks(vs_ctx* ctx, int levels, int L, int M, int N, const double* inv_blocks,
                        const int64_t* idx, int64_t nidx, int precisions, vs_pc** out);
/* Preconditioner whose apply is the matvec of another operator, M^-1 x = mop x (a preconditioner function
 * handle, iterapp.m:41-69, e.g. the material preconditioner mv_AN_const(J, fN, 1./Mr, 1./Mr - 1, ...) of
 * JVIE_Solver_const.m:232-234: an operator on the same fN with that diagonal).  mop must outlive the
 * preconditioner; its precisions are the preconditioner's. */
int vs_pc_create_op(vs_op* mop, vs_pc** out);
void vs_pc_destroy(vs_pc* pc);
/* apply (replaces mvp_circ_2_level.m:2-22 / chan_mvp_idx.m:2-16) */
int vs_pc_apply(vs_pc* pc, const double* x, double* y);
int vs_pc_apply_dev(vs_pc* pc, const void* x, void* y, int precision, int nrhs);

/* ---------------------------------------------------------------------------------------------- */
/* Krylov solvers (device-resident; host buffers complex128 of length vs_op_nunknowns())            */
/* ---------------------------------------------------------------------------------------------- */
/* MATLAB bicgstab(A,b,tol,maxit,M1,[],x0) semantics as called at main_DDA.m:82,120.  pc may be NULL.
 * resvec must hold 2*maxit+1 doubles; *nres receives its used length (length(resvec)). */
int vs_bicgstab(vs_op* op, vs_pc* pc, const double* b, double* x, double tol, int maxit, const double* x0,
                int* flag, double* relres, double* iter, double* resvec, int* nres);
/* MATLAB gmres(A,b,restart,tol,maxit,M) (main_DDA.m:89,128): left preconditioned, restarted.
 * resvec holds restart*maxit+1 doubles; iter[2] = (outer, inner). */
int vs_gmres(vs_op* op, vs_pc* pc, const double* b, double* x, int restart, double tol, int maxit,
             const double* x0, int* flag, double* relres, int iter[2], double* resvec, int* nres);
/* pgmres(A,b,restart,tol,maxit,L1,[],R1,[],x0) of pgmres.m:70-180 (side 0: pc as L1, 1: as R1). */
int vs_pgmres(vs_op* op, vs_pc* pc, int side, const double* b, double* x, int restart, double tol,
              int maxit, const double* x0, int* flag, double* relres, int iter[2], double* resvec,
              int* nres);

/* pgmresDR(A,b,restart,tol,maxit,ritz,L1,[],R1,[],x0) of pgmresDR.m:73-280 (GMRES with deflated
 * restarting: ritz harmonic Ritz vectors carried between cycles; ritz <= 0 takes pgmresDR.m's
 * default min(10, round(restart/2))).  side 0: pc as L1, 1: as R1.  resvec holds up to
 * restart*(maxit+1)+1 values. */
int vs_pgmresdr(vs_op* op, vs_pc* pc, int side, const double* b, double* x, int restart, double tol,
                int maxit, int ritz, const double* x0, int* flag, double* relres, int iter[2],
                double* resvec, int* nres);

/* Krylov-basis checkpoint of GMRES-DR (no counterpart in the reference: pgmresDR.m keeps its deflation
 * space only between the restarts of one solve).  A vs_recycle holds, in device memory, the deflation
 * space (ritz vectors of the preconditioned operator) that a solve ended with, for the next solve with
 * the same operator, preconditioner and side -- e.g. the adjoint solve of a DDA gradient (A^T = A) after
 * the forward solve.  Memory: unknowns * ritz complex doubles. */
typedef struct vs_recycle vs_recycle;
int vs_recycle_create(vs_ctx* ctx, vs_recycle** out);
void vs_recycle_destroy(vs_recycle* rec);
int vs_recycle_size(const vs_recycle* rec); /* vectors held; 0 when empty */
int vs_recycle_clear(vs_recycle* rec);
/* vs_pgmresdr with a checkpoint (rec may be NULL: then identical to vs_pgmresdr).  A non-empty rec, made
 * with the same unknown count, ritz and side, replaces the first plain cycle by the projection onto it
 * (GCRO-DR: [C,R] = qr(A U), U = U R^-1, x = U C^H r, r = r - C C^H r; one extra resvec entry); on
 * return rec holds the final deflation space (unchanged when no deflated cycle ran).  resvec holds up to
 * restart*(maxit+1)+2 values. */
int vs_pgmresdr_rec(vs_op* op, vs_pc* pc, int side, const double* b, double* x, int restart, double tol,
                    int maxit, int ritz, const double* x0, vs_recycle* rec, int* flag, double* relres,
                    int iter[2], double* resvec, int* nres);

/* Fused BiCGStab halves (no counterpart in the reference, whose bicgstab calls MVP and the preconditioner
 * as separate functions): for one right-hand side with an FFT operator whose matvec is slab-major and a
 * circulant preconditioner built from that operator's domain, the Krylov updates and inner products run
 * inside the FFT passes (same vectors and iteration counts; inner products summed in another order).
 * vs_set_fused_krylov(0) turns the fused path off for the whole process (1: on, the default);
 * vs_fused_krylov_available reports in *on whether solves with (op, pc) in a precision take it. */
int vs_set_fused_krylov(int on);
int vs_fused_krylov_available(vs_op* op, vs_pc* pc, int precision, int* on);

/* Device-pointer forms: b, x are FP64 device vectors of vs_op_nunknowns() values. */
int vs_bicgstab_dev(vs_op* op, vs_pc* pc, const void* b, void* x, double tol, int maxit, int* flag,
                    double* relres, double* iter, int* nres);
int vs_gmres_dev(vs_op* op, vs_pc* pc, const void* b, void* x, int restart, double tol, int maxit, int* flag,
                 double* relres, int iter[2], int* nres);

/* Mixed-precision solve (no reference counterpart; the precision form of the solve loop): the
 * correction equation A d = r is solved with the complex-FP32 tensor, FP32 preconditioner blocks and
 * FP32 Krylov vectors (inner_solver 0: BiCGStab, 1: GMRES(restart)) to inner_tol; the residual
 * b - A x and the update x += d are formed in FP64 until norm(b - A x)/norm(b) <= tol.
 * op needs VS_FP64|VS_FP32 tensors, pc (nullable) VS_FP32 blocks.  resvec_outer receives the FP64
 * relative residual before each correction (max_outer+1 doubles). */
int vs_refine(vs_op* op, vs_pc* pc, const double* b, double* x, double tol, int max_outer, int inner_solver,
              double inner_tol, int inner_maxit, int restart, const double* x0, int* flag, double* relres,
              int* outer, int* inner_total, double* resvec_outer, int* nres);
int vs_refine_dev(vs_op* op, vs_pc* pc, const void* b, void* x, double tol, int max_outer, int inner_solver,
                  double inner_tol, int inner_maxit, int restart, int* flag, double* relres, int* outer,
                  int* inner_total);

/* ---------------------------------------------------------------------------------------------- */
/* fields and cross sections (no reference counterpart: README.md:35 lists them as future work)     */
/* ---------------------------------------------------------------------------------------------- */
/* Far-field amplitude F(n) = k^2 sum_j (I - n n) P_j dV exp(-i k n.r_j) for ndir unit vectors.
 * r: nvox*3 voxel centres (row per voxel); P: [Px; Py; Pz] (3*nvox complex128); F: ndir*3 complex128. */
int vs_dda_farfield(vs_ctx* ctx, int64_t nvox, const double* r, const double* P, double dV, double k, int64_t ndir,
                    const double* dirs, double* F);
/* Extinction, absorption and scattering cross sections (Draine 1988) of a DDA polarisation P with the
 * incident field Einc at the same voxels; alpha_inv: 1/alpha (1 or nvox complex values, the DDA
 * diagonal).  out = {C_ext, C_abs, C_sca = C_ext - C_abs}. */
int vs_dda_cross_sections(vs_ctx* ctx, int64_t nvox, const double* P, const double* Einc, const double* alpha_inv,
                          int64_t ainv_len, double dV, double k, double E0sq, double out[3]);

/* Plane-wave right-hand sides on the device (replaces PlaneWavePointWise.m:7-19 for a batch of incidences):
 * r: nvox*3 voxel centres (row per voxel) of the unknowns; khat: ndir unit directions (ndir*3); pol: ndir*npol
 * polarisation vectors (row (d, p) at 3 (npol d + p)); B: device array of ndir*npol vectors of 3*nvox
 * complex128, vector npol d + p = [E_x; E_y; E_z] with E = pol(d, p) exp(i k khat_d . r). */
int vs_plane_wave_batch_dev(vs_ctx* ctx, int64_t nvox, const double* r, int ndir, const double* khat, int npol,
                            const double* pol, double k, void* B);

/* Adjoint gradient assembly (no reference counterpart): g_j = -dainv_j sum_c lam_{c,j} P_{c,j}, the
 * derivative d(w^T P)/d eps_j of a holomorphic objective, given the forward solution P and the adjoint
 * solution lam of A^T lam = w (A^T = A for the DDA system: solve it with vs_bicgstab/vs_gmres on the same
 * operator); dainv_j = d(1/alpha)/d eps at voxel j.  P, lam: [x; y; z] (3*nvox); dainv, g: nvox. */
int vs_dda_grad_eps(vs_ctx* ctx, int64_t nvox, const double* P, const double* lam, const double* dainv, double* g);

/* Batched solves (no reference counterpart; the batching form of the solve loop): nrhs right-hand
 * sides B (nrhs consecutive device vectors, nrhs <= 186) solved together, every matvec and
 * preconditioner launch covering the systems still running (finished ones drop out of the batch).
 * flags/relres/iters receive nrhs values; flag 0 = true relative residual <= tol, 1 = not reached,
 * 4 = breakdown. */
int vs_bicgstab_batch_dev(vs_op* op, vs_pc* pc, int nrhs, const void* B, void* X, int precision, double tol,
                          int maxit, int* flags, double* relres, double* iters);
/* Batched FP32-inner / FP64-refinement solve; B, X are FP64 device vectors. */
int vs_refine_batch_dev(vs_op* op, vs_pc* pc, int nrhs, const void* B, void* X, double tol, int max_outer,
                        double inner_tol, int inner_maxit, int* flags, double* relres, int* outer, int* inner_total);
/* The two batched solves with one operator diagonal per right-hand side (diag: nrhs complex doubles on the
 * host, replacing Gram*Mr for that system as in vs_mvp_dev_diag): every matvec of the solve, the inner
 * FP32 ones and the FP64 residuals, uses each system's own diagonal; the preconditioner is shared. */
int vs_bicgstab_batch_diag_dev(vs_op* op, vs_pc* pc, int nrhs, const void* B, void* X, const double* diag,
                               int precision, double tol, int maxit, int* flags, double* relres, double* iters);
int vs_refine_batch_diag_dev(vs_op* op, vs_pc* pc, int nrhs, const void* B, void* X, const double* diag, double tol,
                             int max_outer, double inner_tol, int inner_maxit, int* flags, double* relres, int* outer,
                             int* inner_total);

/* Timing helpers (no reference counterpart): nrep back-to-back matvecs / preconditioner applies on
 * the context stream, bracketed by HIP events; *ms receives the mean time per call. */
int vs_bench_mvp(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs, int nrep, double* ms);
int vs_bench_pc(vs_pc* pc, const void* x, void* y, int precision, int nrhs, int nrep, double* ms);
/* The preconditioner's block-product stage alone (the kernel of mvp_circ_2_level.m:13-15 /
 * chan_mvp_idx.m:10-12 without its FFT stages), nrep times on the preconditioner's own work volume. */
int vs_bench_pc_blocks(vs_pc* pc, int precision, int nrhs, int nrep, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* VOXSCATTER_H */
