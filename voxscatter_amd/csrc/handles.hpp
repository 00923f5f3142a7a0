// Handle structures shared by the operator, preconditioner and solver translation units.
#pragma once

#include "vs_internal.hpp"

#include <vector>

namespace vs {

// Per-voxel or scalar complex coefficient of the MVP diagonal.
struct Coef {
  bool scalar = true;
  c128 s = mk<double>(0, 0);
  DevBuf v64, v32;  // L*M*N values when !scalar
};

template <typename T>
struct CoefDev {
  int scalar;
  cplx<T> s;
  const cplx<T>* v;
  __device__ __forceinline__ cplx<T> at(int64_t vox) const { return scalar ? s : v[vox]; }
};

template <typename T>
CoefDev<T> coef_dev(const Coef& c);

// launch helpers (defined in op.hip, used by the preconditioner)
template <typename T>
struct RowIO;  // forward

void launch_col_fft(vs_ctx* ctx, int prec, void* data, int64_t fast, int64_t axis_stride, int64_t nbatch,
                    int64_t batch_stride, int n_in, int n, int n_out, bool inv, double scale, int nrhs,
                    int64_t rhs_stride);
void launch_row_fft_dense(vs_ctx* ctx, int prec, void* data, int64_t nlines, int n_in, int n, int n_out,
                          int64_t ld_in, int64_t ld_out, bool inv, double scale, int nrhs, int64_t rhs_stride);
void op_autotune(vs_op* op);
void op_build_fN(vs_op* op);
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w);
// compressed vector -> dense L x (lines) volume with an x FFT of length L (forward)
void launch_row_fft_gather(vs_ctx* ctx, int prec, const void* x, const int* map3, int64_t nunk, void* out,
                           int L, int64_t nlines, int nrhs, int64_t out_rhs_stride);
// dense volume -> inverse x FFT of length L, scaled, scattered into the compressed vector
void launch_row_ifft_scatter(vs_ctx* ctx, int prec, const void* in, const int* map3, int64_t nunk, void* y,
                             int L, int64_t nlines, double scale, int nrhs, int64_t in_rhs_stride);

}  // namespace vs

struct vs_op {
  vs_ctx* ctx = nullptr;
  // dense operator (vs_op_create_dense): y = A x with an n x n column-major matrix, no FFT tensor or domain
  bool dense = false;
  vs::DevBuf a64, a32;
  int L = 0, M = 0, N = 0;
  int64_t V = 0, P = 0;
  int precisions = 0;
  bool has_toep = false;
  vs::DevBuf toep;        // L*M*N*6 complex128
  vs::DevBuf fN64;        // 6 * P complex, MATLAB layout (2L x 2M x 2N x 6), FP64 (kept for download)
  // octant of the tensor actually read by the matvec: kx <= L, ky <= M, kz <= N, 6 * Pc values; the
  // other octants follow from the parity of each component (fN(2L-kx) = s_x fN(kx), etc.)
  vs::DevBuf fc64, fc32;
  vs::DevBuf fs64, fs32;  // the octant stored slab-major, for the fused slab matvec (when its shape has one)
  vs::DevBuf fq64, fq32;  // the octant split by (iy, iz) parity, for the parity-split slab matvec
  int mv_variant[2] = {-1, -1};  // matvec variant per precision (FP64, FP32) picked by the autotuner
  bool tuned = false;
  int64_t Pc = 0;
  // domain
  int64_t nunk = 0;
  vs::DevBuf map3;  // int32[3V]: global unknown -> compressed position or -1
  std::vector<int64_t> idx_host;
  uint64_t dom_id = 0;  // identifies this domain (process-unique, renewed by every vs_op_set_domain)
  vs::Coef a, at, b;  // a = Gram*Mr, at = Gram*conj(Mr), b = Mc
  // scratch
  vs::DevBuf S;          // padded spectra 2L*2M*N*3 per rhs
  vs::DevBuf hx, hy;     // staging of host vectors
};

struct vs_pc {
  vs_ctx* ctx = nullptr;
  vs_op* mop = nullptr;  // preconditioner that is an operator's matvec (vs_pc_create_op): M^-1 x = mop x
  int levels = 2;
  int L = 0, M = 0, N = 0;
  int64_t V = 0;
  int precisions = 0;
  int nb = 0;            // block size: 3N (2-level) or 3MN (1-level)
  bool sym = false;      // 2-level blocks stored as packed upper triangles (B^T = D B D)
  bool mirror = false;   // 2-level blocks stored for kx <= L/2, ky <= M/2 only (the rest by sign symmetry)
  int Lb = 0, Mb = 0;    // 2-level stored block grid (L x M, or L/2+1 x M/2+1 when mirrored)
  vs::DevBuf inv64, inv32;  // apply layout (2-level: [ky][r][s][kx]; 1-level: [kx][r][s])
  vs::DevBuf invq32;     // mirrored 2-level FP32 blocks for the matrix-core product: full, column-major, block kx + Lb ky
  // mirrored 2-level: the last rem = Lb mod Wt stored columns of each precision (a layout tile that is mostly
  // padding, rem from mirror_rem) again as full column-major blocks, block (Lb - rem + w, ky) at
  // (ky * rem + w) * n^2; empty when not used
  vs::DevBuf invr64, invr32;
  vs::DevBuf map3;       // own copy of the unknown map
  int64_t nunk = 0;
  uint64_t dom_id = 0;   // the operator domain the map was copied from (0: a caller's unknown set)
  vs::DevBuf U, U2;      // work volumes (3V per rhs)
  vs::DevBuf hx, hy;
};
