// Chan circulant preconditioners: build (circ_1_level.m, circ_2_level.m), explicit block inverses
// (main_DDA.m:105-111, scatteringHumanHead.m:178-184) and apply (mvp_circ_2_level.m, chan_mvp_idx.m).
//
// Block orderings (identical to the reference's reshapes):
//   2-level block (kx, ky): unknown s = z + N*c, size 3N
//   1-level block kx      : unknown s = y + M*(z + N*c), size 3MN
// Apply layouts on the device:
//   2-level: inv[((ky*3N + r)*3N + s)*L + kx]   -- kx fastest, so a tile of consecutive kx streams
//   1-level: inv[(kx*3MN + r)*3MN + s]          -- row-major per block, one wavefront per row
#include "fft_device.hpp"
#include "handles.hpp"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

namespace vs {

extern __shared__ __align__(16) unsigned char vs_smem[];

// ------------------------------------------------------------------------------------------------
// Chan averaging along x and y
// ------------------------------------------------------------------------------------------------
// c(1) = A(1); c(i) = s (n+1-i)/n A(i) + (i-1)/n A(n-i+2) along x (circ_1_level.m:13-18,28-31)
__global__ void k_chan_x(const c128* __restrict__ A, c128* __restrict__ c, int L, int64_t MN) {
  const int64_t V = (int64_t)L * MN;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * V) return;
  const int p = (int)(t / V);
  const int64_t q = t - p * V;
  const int x = (int)(q % L);
  const int64_t yz = q / L;
  const double s = (p == 1 || p == 2) ? -1.0 : 1.0;
  const c128* a = A + p * V + yz * L;
  c128 v;
  if (x == 0)
    v = a[0];
  else
    v = (s * (double)(L - x) / L) * a[x] + ((double)x / L) * a[L - x];
  c[t] = v;
}

// same along y on an L x M x N x 6 volume (circ_2_level.m:13-18,26-28,62-64)
__global__ void k_chan_y(const c128* __restrict__ A, c128* __restrict__ c, int L, int M, int N) {
  const int64_t V = (int64_t)L * M * N;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * V) return;
  const int p = (int)(t / V);
  const int64_t q = t - p * V;
  const int x = (int)(q % L);
  const int y = (int)((q / L) % M);
  const int64_t z = q / ((int64_t)L * M);
  const double s = (p == 1 || p == 4) ? -1.0 : 1.0;
  const c128* a = A + p * V + z * L * M + x;
  c128 v;
  if (y == 0)
    v = a[0];
  else
    v = (s * (double)(M - y) / M) * a[(int64_t)y * L] + ((double)y / M) * a[(int64_t)(M - y) * L];
  c[t] = v;
}

__global__ void k_negate_comp(c128* __restrict__ v, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) v[t] = -v[t];
}

// ------------------------------------------------------------------------------------------------
// block assembly:  out = a*I - diag(w)*C   (with a = 0, w = -1 the raw C is produced)
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ int tcomp(int ca, int cb) {
  // [xx xy xz; xy yy yz; xz yz zz] -> 0..5
  const int lo = ca < cb ? ca : cb, hi = ca < cb ? cb : ca;
  return lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
}

// 2-level: F = y-FFT'd Chan volume (L x M x N x 6, comp 2 not negated); blocks of the points
// kx < Lb, ky < Mb, block (kx,ky) at kx + Lb*ky (Lb = L, Mb = M for the full set)
__global__ void k_assemble2(const c128* __restrict__ F, int L, int M, int N, int Lb, int Mb, c128 a,
                            const c128* __restrict__ w, c128 ws, c128* __restrict__ out) {
  const int n = 3 * N;
  const int64_t nn = (int64_t)n * n;
  const int64_t nblk = (int64_t)Lb * Mb;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nblk * nn) return;
  const int64_t blk = t / nn;
  const int e = (int)(t - blk * nn);
  const int ra = e % n, cb_ = e / n;  // row, column (column-major block)
  const int kx = (int)(blk % Lb), ky = (int)(blk / Lb);
  const int ca = ra / N, za = ra % N, cb = cb_ / N, zb = cb_ % N;
  const int p = tcomp(ca, cb);
  const int64_t V = (int64_t)L * M * N;
  const c128* f = F + p * V + kx + (int64_t)L * ky;
  const int64_t zs = (int64_t)L * M;
  const int d = za - zb;
  c128 val;
  if (p == 2 || p == 4)  // toeplitz([c0 -c(2:N)], c): T(i,j) = -c(i-j) for i > j, c(j-i) otherwise
    val = d > 0 ? -f[d * zs] : f[(-d) * zs];
  else
    val = f[(d >= 0 ? d : -d) * zs];
  const c128 wr = w ? w[ra] : ws;
  c128 o = -(wr * val);
  if (ra == cb_) o = o + a;
  out[t] = o;
}

// 1-level: cL = x-FFT'd Chan volume (L x M x N x 6); block kx, entry (i, j) of 3MN x 3MN, column-major
__global__ void k_assemble1(const c128* __restrict__ cL, int L, int M, int N, c128 a, const c128* __restrict__ w,
                            c128 ws, c128* __restrict__ out) {
  const int64_t MN = (int64_t)M * N;
  const int64_t n = 3 * MN;
  const int64_t nn = n * n;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)L * nn) return;
  const int kx = (int)(t / nn);
  const int64_t e = t - kx * nn;
  const int64_t i = e % n, j = e / n;
  const int ca = (int)(i / MN), cb = (int)(j / MN);
  int64_t il = i - ca * MN, jl = j - cb * MN;
  const int p = tcomp(ca, cb);
  double sgn = 1.0;
  // xz: antisymmetrised from the upper triangle; yz: symmetrised from the upper triangle
  if (il > jl && (p == 2 || p == 4)) {
    const int64_t tmp = il;
    il = jl;
    jl = tmp;
    if (p == 2) sgn = -1.0;
  }
  const int za = (int)(il / M), ya = (int)(il % M), zb = (int)(jl / M), yb = (int)(jl % M);
  const int dz = za > zb ? za - zb : zb - za;
  const int64_t V = (int64_t)L * MN;
  const c128* c = cL + p * V + kx + (int64_t)L * M * dz;  // c[y * L] = chan_dz coefficient y
  const int dy = ya - yb;
  c128 val;
  if (p == 1 || p == 4)  // toeplitz([c0 -c(2:M)], c)
    val = dy > 0 ? -c[(int64_t)dy * L] : c[(int64_t)(-dy) * L];
  else
    val = c[(int64_t)(dy >= 0 ? dy : -dy) * L];
  val = sgn * val;
  const c128 wr = w ? w[i] : ws;
  c128 o = -(wr * val);
  if (i == j) o = o + a;
  out[t] = o;
}

// ------------------------------------------------------------------------------------------------
// batched Gauss-Jordan inversion with partial pivoting (column-major blocks, in place)
// ------------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void wave_argmax(T& best, int& bi) {
  for (int off = 32; off >= 1; off >>= 1) {
    const T ob = __shfl_xor(best, off);
    const int oi = __shfl_xor(bi, off);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
  }
}

// whole block in LDS, one workgroup per block
__global__ __launch_bounds__(256) void k_gj_lds(c128* __restrict__ blocks, int n) {
  c128* A = reinterpret_cast<c128*>(vs_smem);
  c128* colk = A + n * n;
  int* piv = reinterpret_cast<int*>(colk + n);
  __shared__ int s_p;
  __shared__ c128 s_pinv;
  c128* g = blocks + (int64_t)blockIdx.x * n * n;
  const int tid = threadIdx.x, nth = blockDim.x;
  for (int e = tid; e < n * n; e += nth) A[e] = g[e];
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (tid < 64) {
      double best = -1.0;
      int bi = k;
      for (int i = k + tid; i < n; i += 64) {
        const double v = abs2(A[i + n * k]);
        if (v > best) { best = v; bi = i; }
      }
      wave_argmax(best, bi);
      if (tid == 0) {
        s_p = bi;
        piv[k] = bi;
      }
    }
    __syncthreads();
    const int p = s_p;
    if (p != k)
      for (int j = tid; j < n; j += nth) {
        const c128 t = A[k + n * j];
        A[k + n * j] = A[p + n * j];
        A[p + n * j] = t;
      }
    __syncthreads();
    for (int i = tid; i < n; i += nth) colk[i] = A[i + n * k];
    if (tid == 0) s_pinv = cdiv(mk<double>(1, 0), A[k + n * k]);
    __syncthreads();
    const c128 pinv = s_pinv;
    for (int j = tid; j < n; j += nth) A[k + n * j] = (j == k) ? pinv : A[k + n * j] * pinv;
    __syncthreads();
    for (int e = tid; e < n * n; e += nth) {
      const int i = e % n, j = e / n;
      if (i == k) continue;
      const c128 f = colk[i];
      if (j == k)
        A[e] = -(f * A[k + n * k]);
      else
        A[e] = A[e] - f * A[k + n * j];
    }
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; --k) {
    const int p = piv[k];
    if (p != k)
      for (int i = tid; i < n; i += nth) {
        const c128 t = A[i + n * k];
        A[i + n * k] = A[i + n * p];
        A[i + n * p] = t;
      }
    __syncthreads();
  }
  for (int e = tid; e < n * n; e += nth) g[e] = A[e];
}

// global-memory variant for blocks too large for LDS: one pivot kernel + one elimination kernel per step
__global__ __launch_bounds__(256) void k_gj_pivot(c128* __restrict__ blocks, int64_t n, int64_t k,
                                                  c128* __restrict__ colk, int* __restrict__ piv) {
  __shared__ double s_best[4];
  __shared__ int s_bi[4];
  __shared__ int s_p;
  __shared__ c128 s_pinv;
  const int64_t blk = blockIdx.x;
  c128* A = blocks + blk * n * n;
  const int tid = threadIdx.x;
  double best = -1.0;
  int bi = (int)k;
  for (int64_t i = k + tid; i < n; i += 256) {
    const double v = abs2(A[i + n * k]);
    if (v > best) { best = v; bi = (int)i; }
  }
  wave_argmax(best, bi);
  if ((tid & 63) == 0) {
    s_best[tid >> 6] = best;
    s_bi[tid >> 6] = bi;
  }
  __syncthreads();
  if (tid == 0) {
    double b0 = s_best[0];
    int i0 = s_bi[0];
    for (int q = 1; q < 4; ++q)
      if (s_best[q] > b0 || (s_best[q] == b0 && s_bi[q] < i0)) { b0 = s_best[q]; i0 = s_bi[q]; }
    s_p = i0;
    piv[blk * n + k] = i0;
  }
  __syncthreads();
  const int64_t p = s_p;
  if (p != k)
    for (int64_t j = tid; j < n; j += 256) {
      const c128 t = A[k + n * j];
      A[k + n * j] = A[p + n * j];
      A[p + n * j] = t;
    }
  __syncthreads();
  for (int64_t i = tid; i < n; i += 256) colk[blk * n + i] = A[i + n * k];
  if (tid == 0) s_pinv = cdiv(mk<double>(1, 0), A[k + n * k]);
  __syncthreads();
  const c128 pinv = s_pinv;
  for (int64_t j = tid; j < n; j += 256) A[k + n * j] = (j == k) ? pinv : A[k + n * j] * pinv;
}

__global__ void k_gj_elim(c128* __restrict__ blocks, int64_t n, int64_t k, const c128* __restrict__ colk) {
  const int64_t blk = blockIdx.y;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int64_t i = e % n, j = e / n;
  if (i == k) return;
  c128* A = blocks + blk * n * n;
  const c128 f = colk[blk * n + i];
  if (j == k)
    A[e] = -(f * A[k + n * k]);
  else
    A[e] = A[e] - f * A[k + n * j];
}

__global__ void k_gj_unswap(c128* __restrict__ blocks, int64_t n, const int* __restrict__ piv) {
  const int64_t blk = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  c128* A = blocks + blk * n * n;
  for (int64_t k = n - 1; k >= 0; --k) {
    const int64_t p = piv[blk * n + k];
    if (p != k) {
      const c128 t = A[i + n * k];
      A[i + n * k] = A[i + n * p];
      A[i + n * p] = t;
    }
  }
}

// Blocked Gauss-Jordan for blocks too large for LDS, in panels of NB = 32 columns.  Partial pivoting picks
// the same rows as the unblocked sweep.  The row swaps of a panel commute with its steps (a swap exchanges
// two rows that the steps treat alike), so the panel's steps equal one block step over its indices K on
// the row-permuted block B:
//   B_KK <- B_KK^-1,  B_KR <- B_KK^-1 B_KR,  B_RK <- -B_RK B_KK^-1,  B_RR <- B_RR - B_RK B_KK^-1 B_KR.
// k_gjb_panel performs the steps on the panel columns in LDS, which become P = [B_KK^-1; -B_RK B_KK^-1],
// and records the row permutation; k_gjb_update then sets every other column j to
//   P X_j + (row-permuted column j with rows K cleared),   X_j = rows K of the row-permuted column j,
// a complex (n x NB) x (NB x 16) product per strip of 16 columns on the FP64 matrix cores.  The columns
// are unswapped at the end as in the unblocked sweep.
constexpr int GJB_NMAX = 512;  // 8 waves x 4 row tiles of 16 in the update kernel

// one workgroup per block, one thread per row holding its kb <= NB panel entries in registers: the steps of
// the panel starting at column k0.  A step needs only the pivot (a workgroup argmax) and the two rows it
// exchanges (through LDS, double-buffered by step parity: two barriers per step)
template <int NB>
__global__ __launch_bounds__(512) void k_gjb_panel(c128* __restrict__ blocks, int n, int k0,
                                                   int* __restrict__ perm, int* __restrict__ piv) {
  __shared__ double s_best[8];
  __shared__ int s_bi[8];
  __shared__ c128 s_rk[2][NB], s_nk[2][NB];  // row k before a step, and the new row k (pivot row scaled)
  __shared__ int s_pm[GJB_NMAX];  // source row of each row after this panel's swaps
  const int64_t blk = blockIdx.x;
  c128* A = blocks + blk * (int64_t)n * n;
  const int i = threadIdx.x, nw = (int)(blockDim.x >> 6);
  const bool live = i < n;
  const int kb = min(NB, n - k0);
  c128 r[NB];
#pragma unroll
  for (int jj = 0; jj < NB; ++jj) {
    r[jj] = mk<double>(0, 0);
    if (live && jj < kb) r[jj] = A[i + (int64_t)n * (k0 + jj)];
  }
  if (live) s_pm[i] = i;
  // entry kk of the row by selects over constant indices (the step loop stays rolled, r in registers)
  // (selects written per component: a select of whole complex values becomes a select of their addresses,
  // which sends the row to scratch memory)
  auto csel = [](bool c, c128 a, c128 b) { return mk<double>(c ? a.x : b.x, c ? a.y : b.y); };
  auto pick = [&](int kk) {
    c128 v = r[0];
#pragma unroll
    for (int jj = 1; jj < NB; ++jj) v = csel(jj == kk, r[jj], v);
    return v;
  };
  for (int kk = 0; kk < kb; ++kk) {
    const int k = k0 + kk, buf = kk & 1;
    const c128 rkk = pick(kk);
    double best = (live && i >= k) ? abs2(rkk) : -1.0;
    int bi = i;
    wave_argmax(best, bi);
    if ((i & 63) == 0) {
      s_best[i >> 6] = best;
      s_bi[i >> 6] = bi;
    }
    __syncthreads();
    // every thread reduces the wave results in the same order, so all agree on the pivot row p
    double b0 = s_best[0];
    int p = s_bi[0];
    for (int q = 1; q < nw; ++q)
      if (s_best[q] > b0 || (s_best[q] == b0 && s_bi[q] < p)) { b0 = s_best[q]; p = s_bi[q]; }
    if (i == 0) {
      piv[blk * n + k] = p;
      const int t = s_pm[k];
      s_pm[k] = s_pm[p];
      s_pm[p] = t;
    }
    if (i == k) {
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) s_rk[buf][jj] = r[jj];
    }
    if (i == p) {
      const c128 pv = cdiv(mk<double>(1, 0), rkk);
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) s_nk[buf][jj] = csel(jj == kk, pv, r[jj] * pv);
    }
    __syncthreads();
    // the step on the row-swapped panel: row k <- new row k; row p <- old row k - f_p (new row k);
    // other rows <- own - f_i (new row k); column kk <- -f pinv (pinv on row k)
    const c128 pinv = s_nk[buf][kk];
    if (live) {
      const bool rowk = i == k, rowp = i == p && p != k;
      const c128 f = csel(rowp, s_rk[buf][kk], rkk);
#pragma unroll
      for (int jj = 0; jj < NB; ++jj) {
        const c128 nk = s_nk[buf][jj];
        const c128 src = csel(rowp, s_rk[buf][jj], r[jj]);
        r[jj] = csel(rowk, nk, csel(jj == kk, -(f * pinv), src - f * nk));
      }
    }
  }
#pragma unroll
  for (int jj = 0; jj < NB; ++jj)
    if (live && jj < kb) A[i + (int64_t)n * (k0 + jj)] = r[jj];
  __syncthreads();
  if (live) perm[blk * n + i] = s_pm[i];
}

typedef double vs_d4 __attribute__((ext_vector_type(4)));

// one workgroup per (strip of 16 columns, block), 8 waves; the strip is staged row-permuted in LDS (all of
// its reads precede its writes, so the update is in place), each wave takes row tiles wave, wave + 8, ...
template <int NB>
__global__ __launch_bounds__(512) void k_gjb_update(c128* __restrict__ blocks, int n, int k0,
                                                    const int* __restrict__ perm) {
  constexpr int RT = GJB_NMAX / (16 * 8), KS = NB / 4;
  const int strip = blockIdx.x;
  if (strip * 16 >= k0 && strip * 16 < k0 + NB) return;  // the panel itself
  c128* S = reinterpret_cast<c128*>(vs_smem);  // [16][n + 1] row-permuted strip, column-major
  const int ns = n + 1;
  const int64_t blk = blockIdx.y;
  c128* A = blocks + blk * (int64_t)n * n;
  const int* pm = perm + blk * n;
  const int kb = min(NB, n - k0);
  const int j0 = strip * 16;
  const int w = min(16, n - j0);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lc = lane & 15, lq = lane >> 4;
  // one row per thread (n <= 512): its source row is read once, its 16 entries are 16 coalesced loads
  if (tid < n) {
    const c128* src = A + pm[tid] + (int64_t)n * j0;
#pragma unroll
    for (int jj = 0; jj < 16; ++jj)
      if (jj < w) S[jj * ns + tid] = src[(int64_t)n * jj];
  }
  __syncthreads();
  // accumulators start from the permuted strip with rows K cleared; the FP64 16x16x4 result layout holds
  // element (row i, column j) in lane j + 16 (i % 4), register i / 4 (unlike the FP32 16x16x4 layout,
  // lane j + 16 (i / 4), register i % 4)
  vs_d4 cr[RT], ci[RT];
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    const int i0 = (wave + 8 * t) * 16;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + lq + 4 * r;
      const bool live = i < n && lc < w && (i < k0 || i >= k0 + kb);
      const c128 v = live ? S[lc * ns + i] : mk<double>(0, 0);
      cr[t][r] = v.x;
      ci[t][r] = v.y;
    }
  }
  // X operand (B layout: row kk = 4 ks + lq, column lc), read once per k step
  c128 xb[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int kk = 4 * ks + lq;
    xb[ks] = (kk < kb && lc < w) ? S[lc * ns + k0 + kk] : mk<double>(0, 0);
  }
  const c128* Pc = A + (int64_t)n * k0;  // the panel P, written by k_gjb_panel
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    const int i0 = (wave + 8 * t) * 16;
    if (i0 >= n) break;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int kk = 4 * ks + lq, ia = i0 + lc;  // A layout: row lc, column lq
      const c128 a = (ia < n && kk < kb) ? Pc[ia + (int64_t)n * kk] : mk<double>(0, 0);
      cr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, xb[ks].x, cr[t], 0, 0, 0);
      cr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, xb[ks].y, cr[t], 0, 0, 0);
      ci[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, xb[ks].y, ci[t], 0, 0, 0);
      ci[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, xb[ks].x, ci[t], 0, 0, 0);
    }
  }
  __syncthreads();  // every wave has read the strip
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    const int i0 = (wave + 8 * t) * 16;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + lq + 4 * r;
      if (i < n && lc < w) S[lc * ns + i] = mk<double>(cr[t][r], ci[t][r]);
    }
  }
  __syncthreads();
  if (tid < n) {
    c128* dst = A + tid + (int64_t)n * j0;
#pragma unroll
    for (int jj = 0; jj < 16; ++jj)
      if (jj < w) dst[(int64_t)n * jj] = S[jj * ns + tid];
  }
}

// column-major blocks (kx + L*ky) -> 2-level apply layout, in precision To
template <typename To>
__global__ void k_relayout2(const c128* __restrict__ in, cplx<To>* __restrict__ out, int L, int M, int n) {
  const int64_t nn = (int64_t)n * n;
  const int64_t tot = (int64_t)L * M * nn;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tot) return;
  // t indexes the destination: ((ky*n + r)*n + s)*L + kx
  const int kx = (int)(t % L);
  int64_t q = t / L;
  const int s = (int)(q % n);
  q /= n;
  const int r = (int)(q % n);
  const int ky = (int)(q / n);
  out[t] = ccast<To>(in[((int64_t)kx + (int64_t)L * ky) * nn + r + (int64_t)n * s]);
}

// column-major blocks -> packed upper triangles, 2-level apply layout: entry (r <= s) of block (kx, ky) at
// ((ky*NT + t)*L + kx), t = r*(2n - r + 1)/2 + (s - r), NT = n(n+1)/2.  Valid for blocks with
// B^T = D B D, D = diag(1 (x, y rows), -1 (z rows)): inv(a I - w C) for a scalar w (C^T = D C D).
template <typename To>
__global__ void k_relayout2_sym(const c128* __restrict__ in, cplx<To>* __restrict__ out, int L, int M, int n) {
  const int64_t nn = (int64_t)n * n;
  const int64_t NT = (int64_t)n * (n + 1) / 2;
  const int64_t tot = (int64_t)L * M * nn;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= tot) return;
  const int64_t blk = e / nn;
  const int q = (int)(e - blk * nn);
  const int r = q % n, sidx = q / n;  // column-major source entry (r, s)
  if (r > sidx) return;
  const int kx = (int)(blk % L), ky = (int)(blk / L);
  const int64_t t = (int64_t)r * (2 * n - r + 1) / 2 + (sidx - r);
  out[((int64_t)ky * NT + t) * L + kx] = ccast<To>(in[blk * nn + q]);
}

// column-major blocks of the Lb x Mb stored grid -> tiled mirrored apply layout: columns kx in tiles of
// Wt (128 bytes of To), entry t of block (kx, ky) at (((kx / Wt) * Mb + ky) * NT + t) * Wt + kx % Wt, with
// t = r*(2n - r + 1)/2 + (s - r) over the packed upper triangle (SYM) or t = r*n + s (full).  Every
// (tile, ky) is one contiguous run of NT rows of 128 bytes.
template <typename To, bool SYM>
__global__ void k_relayout2_tiled(const c128* __restrict__ in, cplx<To>* __restrict__ out, int Lb, int Mb, int n,
                                  int Wt) {
  const int64_t nn = (int64_t)n * n;
  const int64_t NT = SYM ? (int64_t)n * (n + 1) / 2 : nn;
  const int64_t tot = (int64_t)Lb * Mb * nn;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= tot) return;
  const int64_t blk = e / nn;
  const int q = (int)(e - blk * nn);
  const int r = q % n, sidx = q / n;  // column-major source entry (r, s)
  if (SYM && r > sidx) return;
  const int kx = (int)(blk % Lb), ky = (int)(blk / Lb);
  const int64_t t = SYM ? (int64_t)r * (2 * n - r + 1) / 2 + (sidx - r) : (int64_t)r * n + sidx;
  out[(((int64_t)(kx / Wt) * Mb + ky) * NT + t) * Wt + kx % Wt] = ccast<To>(in[blk * nn + q]);
}

// column-major blocks -> 1-level apply layout (row-major per block)
template <typename To>
__global__ void k_relayout1(const c128* __restrict__ in, cplx<To>* __restrict__ out, int64_t nblk, int64_t n) {
  const int64_t nn = n * n;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nblk * nn) return;
  const int64_t b = t / nn, e = t - b * nn;
  const int64_t r = e / n, s = e % n;
  out[t] = ccast<To>(in[b * nn + r + n * s]);
}

// ------------------------------------------------------------------------------------------------
// apply kernels
// ------------------------------------------------------------------------------------------------
// per (kx tile, ky): u <- Binv(kx,ky) u on the y-FFT'd volume, in place, for RC right-hand sides at once
// (each block entry is read once per chunk of RC right-hand sides; blockIdx.y = chunk)
template <typename T, int RC>
__global__ __launch_bounds__(256) void k_blockmv2(const cplx<T>* __restrict__ B, cplx<T>* __restrict__ U, int L, int M,
                                                  int n, int W, int64_t rhs_stride, int nrhs) {
  cplx<T>* u = reinterpret_cast<cplx<T>*>(vs_smem);  // [rc][s][w]
  const int tiles = (L + W - 1) / W;
  const int t = blockIdx.x % tiles, ky = blockIdx.x / tiles;
  const int kx0 = t * W;
  const int wv = min(W, L - kx0);
  const int r0 = blockIdx.y * RC;
  const int nr = min(RC, nrhs - r0);
  const int64_t sstride = (int64_t)L * M;
  cplx<T>* U0 = U + (int64_t)r0 * rhs_stride + kx0 + (int64_t)L * ky;
  for (int e = threadIdx.x; e < RC * n * W; e += blockDim.x) {
    const int w = e % W, q = e / W;
    const int s = q % n, rc = q / n;
    u[e] = (w < wv && rc < nr) ? U0[rc * rhs_stride + s * sstride + w] : mk<T>(0, 0);
  }
  __syncthreads();
  const cplx<T>* Bk = B + (int64_t)ky * n * n * L + kx0;
  for (int e = threadIdx.x; e < n * W; e += blockDim.x) {
    const int w = e % W, r = e / W;
    if (w >= wv) continue;
    const cplx<T>* br = Bk + (int64_t)r * n * L + w;
    cplx<T> acc[RC];
#pragma unroll
    for (int c = 0; c < RC; ++c) acc[c] = mk<T>(0, 0);
    for (int s = 0; s < n; ++s) {
      const cplx<T> b = br[(int64_t)s * L];
#pragma unroll
      for (int c = 0; c < RC; ++c) acc[c] = acc[c] + b * u[(c * n + s) * W + w];
    }
#pragma unroll
    for (int c = 0; c < RC; ++c)
      if (c < nr) U0[c * rhs_stride + r * sstride + w] = acc[c];
  }
}

// same with packed symmetric blocks: B(r,s) = B_up(r,s) for s >= r, d_r d_s B_up(s,r) otherwise, where
// d = -1 on the z rows (index >= nz0 = 2N)
template <typename T, int RC>
__global__ __launch_bounds__(256) void k_blockmv2_sym(const cplx<T>* __restrict__ B, cplx<T>* __restrict__ U, int L,
                                                      int M, int n, int W, int64_t rhs_stride, int nrhs) {
  cplx<T>* u = reinterpret_cast<cplx<T>*>(vs_smem);  // [rc][s][w]
  const int tiles = (L + W - 1) / W;
  const int t = blockIdx.x % tiles, ky = blockIdx.x / tiles;
  const int kx0 = t * W;
  const int wv = min(W, L - kx0);
  const int r0 = blockIdx.y * RC;
  const int nr = min(RC, nrhs - r0);
  const int64_t sstride = (int64_t)L * M;
  cplx<T>* U0 = U + (int64_t)r0 * rhs_stride + kx0 + (int64_t)L * ky;
  for (int e = threadIdx.x; e < RC * n * W; e += blockDim.x) {
    const int w = e % W, q = e / W;
    const int s = q % n, rc = q / n;
    u[e] = (w < wv && rc < nr) ? U0[rc * rhs_stride + s * sstride + w] : mk<T>(0, 0);
  }
  __syncthreads();
  const int64_t NT = (int64_t)n * (n + 1) / 2;
  const int nz0 = 2 * (n / 3);
  const cplx<T>* Bk = B + (int64_t)ky * NT * L + kx0;
  for (int e = threadIdx.x; e < n * W; e += blockDim.x) {
    const int w = e % W, r = e / W;
    if (w >= wv) continue;
    const bool rz = r >= nz0;
    cplx<T> acc[RC];
#pragma unroll
    for (int c = 0; c < RC; ++c) acc[c] = mk<T>(0, 0);
    // s < r: transposed entries (s, r), t(s, r) = s(2n - s + 1)/2 + (r - s)
    int64_t ts = r;
    for (int s = 0; s < r; ++s) {
      cplx<T> b = Bk[ts * L + w];
      if (rz != (s >= nz0)) b = -b;
#pragma unroll
      for (int c = 0; c < RC; ++c) acc[c] = acc[c] + b * u[(c * n + s) * W + w];
      ts += n - s - 1;
    }
    // s >= r: row r of the upper triangle, contiguous in t
    const int64_t t0 = (int64_t)r * (2 * n - r + 1) / 2;
    for (int s = r; s < n; ++s) {
      const cplx<T> b = Bk[(t0 + s - r) * L + w];
#pragma unroll
      for (int c = 0; c < RC; ++c) acc[c] = acc[c] + b * u[(c * n + s) * W + w];
    }
#pragma unroll
    for (int c = 0; c < RC; ++c)
      if (c < nr) U0[c * rhs_stride + r * sstride + w] = acc[c];
  }
}

// Mirrored 2-level apply.  The Chan eigenvalues of a component odd in x change sign under kx -> L - kx
// (c(L-i) = -c(i) for the xy and xz vectors), those of a component odd in y under ky -> M - ky (xy, yz),
// so the blocks obey B(L-kx, ky) = Dx B(kx, ky) Dx and B(kx, M-ky) = Dy B(kx, ky) Dy with Dx (Dy) = -1
// on the x (y) unknowns; the same holds for the inverses.  Only kx <= L/2, ky <= M/2 are stored (grid
// Lb x Mb), and one workgroup applies a tile of W = 2^lw stored blocks to the up to four points each
// stands for: variant q (bit 0: x mirror, bit 1: y mirror) of right-hand side rc is u[(4 rc + q)][s][w].
// The rows of a tile are split over RS sibling workgroups (out of place: U -> U2); siblings get
// consecutive logical ids, and logical ids are laid out so that consecutive ones land on the same XCD
// (hardware dispatch sends workgroup i to XCD i mod 8), so the siblings' reads of U share that XCD's L2.
// With rem > 0 the last rem stored columns are not taken from their (mostly padding) layout tile but from
// Br, full column-major blocks: the workgroups after the tiles' take one such column each (W = 1, lanes
// over rows, RSr siblings), reading consecutive rows of a block column per load.
// Index arithmetic is 32-bit within a block and the gather runs without integer division.
template <typename T, bool SYM, int RC>
__global__ __launch_bounds__(512) void k_blockmv2m(const cplx<T>* __restrict__ B, const cplx<T>* __restrict__ U,
                                                   cplx<T>* __restrict__ U2, int L, int M, int Lb, int Mb, int n,
                                                   int Wt, int lw, int RS, int64_t rhs_stride, int nrhs,
                                                   const cplx<T>* __restrict__ Br, int rem, int RSr) {
  constexpr int NV = 4 * RC;
  constexpr int BU = 8;
  cplx<T>* u = reinterpret_cast<cplx<T>*>(vs_smem);
  const int Wm = 1 << lw;
  const int tiles = rem > 0 ? Lb / Wt : (Lb + Wt - 1) / Wt, subs = Wt / Wm;
  const int64_t G = (int64_t)tiles * subs * Mb * RS;
  const int64_t GA = G + (int64_t)rem * Mb * RSr;
  const int64_t per_xcd = (GA + 7) / 8;
  const int64_t lid = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (lid >= GA) return;
  const int NT = SYM ? n * (n + 1) / 2 : n * n;
  const bool rmd = lid >= G;  // a remainder column
  int lwe, rsi, kyb, kx0, nrs;
  const cplx<T>* Bk;
  if (!rmd) {
    rsi = (int)(lid % RS);
    int64_t rest = lid / RS;
    const int sub = (int)(rest % subs);
    rest /= subs;
    const int tile = (int)(rest % tiles);
    kyb = (int)(rest / tiles);
    kx0 = tile * Wt + sub * Wm;
    lwe = lw;
    nrs = RS;
    Bk = B + ((int64_t)tile * Mb + kyb) * NT * Wt + sub * Wm;
  } else {
    const int64_t l2 = lid - G;
    rsi = (int)(l2 % RSr);
    const int64_t rest = l2 / RSr;
    const int wr = (int)(rest % rem);
    kyb = (int)(rest / rem);
    kx0 = Lb - rem + wr;
    lwe = 0;
    nrs = RSr;
    Bk = Br + ((int64_t)kyb * rem + wr) * n * n;
  }
  const int W = 1 << lwe;
  const int wv = max(0, min(W, Lb - kx0));
  const int r0 = blockIdx.y * RC;
  const int nr = min(RC, nrhs - r0);
  const int N = n / 3;
  const int64_t sstride = (int64_t)L * M;
  const bool ymir = kyb != 0 && 2 * kyb != M;
  const cplx<T>* Ui = U + (int64_t)r0 * rhs_stride;
  cplx<T>* Uo = U2 + (int64_t)r0 * rhs_stride;
  {
    // gather: lanes over the W columns, rows over s; the NV loads of a row are issued together
    // (clamped addresses, sign 0 for dead entries)
    const int w = threadIdx.x & (W - 1);
    const int kx = kx0 + w;
    const bool wl = w < wv;
    const bool xl = kx != 0 && 2 * kx != L;
    for (int s = threadIdx.x >> lwe; s < n; s += blockDim.x >> lwe) {
      const int comp = (s >= N) + (s >= 2 * N);
      cplx<T> val[NV];
      T sg[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int rc = v >> 2, xm = v & 1, ym = (v >> 1) & 1;
        const bool live = wl && rc < nr && (!xm || xl) && (!ym || ymir);
        const int gx = live ? (xm ? L - kx : kx) : 0, gy = live ? (ym ? M - kyb : kyb) : 0;
        val[v] = Ui[live ? rc * rhs_stride + s * sstride + gx + (int64_t)L * gy : 0];
        sg[v] = live ? (((xm && comp == 0) || (ym && comp == 1)) ? (T)-1 : (T)1) : (T)0;
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) u[(v * n + s) * W + w] = sg[v] * val[v];
    }
  }
  __syncthreads();
  const int nz0 = 2 * N;
  const int rows = (n + nrs - 1) / nrs;
  const int rlo = rsi * rows, rhi = min(n, rlo + rows);
  // RM: a remainder column (entry (r, s) at Bk[s n + r], no sign folding); else the tiled layout
  auto product = [&](auto rm_tag) {
    constexpr bool RM = decltype(rm_tag)::value;
    for (int e = threadIdx.x; e < (rhi - rlo) * W; e += blockDim.x) {
      const int w = e & (W - 1);
      const int r = rlo + (e >> lwe);
      if (w >= wv) continue;
      cplx<T> acc[NV];
#pragma unroll
      for (int c = 0; c < NV; ++c) acc[c] = mk<T>(0, 0);
      const int t0 = SYM ? r * (2 * n - r + 1) / 2 : r * n;
      const int rz = r >= nz0;
      // batch of BU entries s0 .. s0 + BU - 1 of row r (sign folded in, zero past n)
      auto load_batch = [&](int s0, cplx<T> (&b)[BU], T (&sg)[BU]) {
#pragma unroll
        for (int j = 0; j < BU; ++j) {
          const int s = min(s0 + j, n - 1);
          int t;
          T g = (T)1;
          if (RM) {
            t = s * n + r;
          } else if (SYM && RC == 1) {
            // s < r: transposed entry (s, r), t = s (2n - s + 1) / 2 + (r - s), negated when exactly one of r, s
            // is a z unknown; s >= r: row r of the upper triangle, t = t0 + (s - r).  Branch-free: the terms in
            // s are uniform over the workgroup, the choice is a mask (a branch here costs an exec-mask round
            // per entry; measured 20.7 -> 19.4 us on the Koch prism's FP32 blocks)
            const int lo = s < r ? -1 : 0;
            const int tlo = s * (2 * n - s + 1) / 2 - s + r, thi = t0 - r + s;
            t = thi + ((tlo - thi) & lo);
            g = (T)(1 - 2 * (lo & (rz ^ (s >= nz0 ? 1 : 0)) & 1));
          } else if (SYM) {
            // several right-hand sides: the same choice as a branch, which the 64 multiply-adds per entry hide
            // and which keeps the loads in the compiler's original order (the masked form measured 221 -> 260 us
            // for 16 right-hand sides)
            const bool lo = s < r;
            t = lo ? s * (2 * n - s + 1) / 2 + (r - s) : t0 + (s - r);
            if (lo && rz != (s >= nz0 ? 1 : 0)) g = (T)-1;
          } else {
            t = t0 + s;
          }
          // offsets within one layout tile stay below 2^31 for n < ~16000
          b[j] = RM ? Bk[t] : Bk[t * Wt + w];
          sg[j] = s0 + j < n ? g : (T)0;
        }
      };
      for (int s0 = 0; s0 < n; s0 += BU) {
        cplx<T> b[BU];
        T sg[BU];
        load_batch(s0, b, sg);
#pragma unroll
        for (int j = 0; j < BU; ++j) {
          const int s = min(s0 + j, n - 1);
          const cplx<T> bj = sg[j] * b[j];
#pragma unroll
          for (int c = 0; c < NV; ++c) acc[c] = cfma(bj, u[(c * n + s) * W + w], acc[c]);
        }
      }
      const int kx = kx0 + w;
      const int comp = r / N;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int rc = v >> 2, xm = v & 1, ym = (v >> 1) & 1;
        const bool live = rc < nr && (!xm || (kx != 0 && 2 * kx != L)) && (!ym || ymir);
        if (!live) continue;
        const int gx = xm ? L - kx : kx, gy = ym ? M - kyb : kyb;
        const bool neg = (xm && comp == 0) || (ym && comp == 1);
        Uo[rc * rhs_stride + r * sstride + gx + (int64_t)L * gy] = neg ? -acc[v] : acc[v];
      }
    }
  };
  // remainder columns exist for one right-hand side only (rem = 0 otherwise), so only that instantiation
  // carries the second role
  if constexpr (RC == 1) {
    if (rmd)
      product(std::true_type{});
    else
      product(std::false_type{});
  } else {
    product(std::false_type{});
  }
}

// Mirrored 2-level apply for large blocks (3N >= 96): the product of k_blockmv2m organised like a
// GEMM.  A workgroup of 256 lanes covers a whole layout tile of W = 128 B / sizeof(complex) stored blocks
// (one full 128-byte line per block entry) and RL x RPT rows (RL = 256 / W row lanes, RPT rows per lane).
// The vectors are staged through LDS in chunks of SC columns, and each value read from LDS serves RPT
// rows, so a lane holds RPT x 4 RC accumulators and the gathered vectors are read n / (RL RPT) times.
template <typename T, bool SYM, int RC, int RPT>
__global__ __launch_bounds__(256) void k_blockmv2k(const cplx<T>* __restrict__ B, const cplx<T>* __restrict__ U,
                                                   cplx<T>* __restrict__ U2, int L, int M, int Lb, int Mb, int n,
                                                   int RS, int64_t rhs_stride, int nrhs) {
  constexpr int NV = 4 * RC;
  constexpr int W = 128 / (int)sizeof(cplx<T>);
  constexpr int RL = 256 / W;
  constexpr int SC = 32;
  constexpr int ROWS = RL * RPT;
  cplx<T>* u = reinterpret_cast<cplx<T>*>(vs_smem);  // [NV][SC][W]
  const int tiles = (Lb + W - 1) / W;
  const int64_t G = (int64_t)tiles * Mb * RS;
  const int64_t per_xcd = (G + 7) / 8;
  const int64_t lid = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (lid >= G) return;
  const int rsi = (int)(lid % RS);
  const int64_t rest = lid / RS;
  const int tile = (int)(rest % tiles), kyb = (int)(rest / tiles);
  const int kx0 = tile * W;
  const int wv = max(0, min(W, Lb - kx0));
  const int r0 = blockIdx.y * RC;
  const int nr = min(RC, nrhs - r0);
  const int N = n / 3;
  const int64_t sstride = (int64_t)L * M;
  const bool ymir = kyb != 0 && 2 * kyb != M;
  const cplx<T>* Ui = U + (int64_t)r0 * rhs_stride;
  cplx<T>* Uo = U2 + (int64_t)r0 * rhs_stride;
  const int NT = SYM ? n * (n + 1) / 2 : n * n;
  const int nz0 = 2 * N;
  const cplx<T>* Bk = B + ((int64_t)tile * Mb + kyb) * NT * W;
  const int w = threadIdx.x & (W - 1);
  const int kx = kx0 + w;
  const bool wl = w < wv;
  const bool xl = kx != 0 && 2 * kx != L;
  // rows of this lane (clamped to n - 1 past the end; their results are dropped)
  int rr[RPT], t0[RPT];
  bool rz[RPT];
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    const int r = min(rsi * ROWS + (int)(threadIdx.x / W) + RL * i, n - 1);
    rr[i] = r;
    t0[i] = SYM ? r * (2 * n - r + 1) / 2 : r * n;
    rz[i] = r >= nz0;
  }
  cplx<T> acc[RPT][NV];
#pragma unroll
  for (int i = 0; i < RPT; ++i)
#pragma unroll
    for (int c = 0; c < NV; ++c) acc[i][c] = mk<T>(0, 0);
  for (int s0 = 0; s0 < n; s0 += SC) {
    // stage columns s0 .. s0 + SC - 1: u[(v SC + sl) W + w] is variant v of column s0 + sl (zero past n)
#pragma unroll 8
    for (int e = threadIdx.x; e < NV * SC * W; e += 256) {
      const int q = e / W;
      const int sl = q % SC, v = q / SC;
      const int s = s0 + sl;
      const int rc = v >> 2, xm = v & 1, ym = (v >> 1) & 1;
      const bool live = s < n && wl && rc < nr && (!xm || xl) && (!ym || ymir);
      const int comp = (s >= N) + (s >= 2 * N);
      const int gx = live ? (xm ? L - kx : kx) : 0, gy = live ? (ym ? M - kyb : kyb) : 0;
      const cplx<T> val = Ui[live ? rc * rhs_stride + (int64_t)s * sstride + gx + (int64_t)L * gy : 0];
      u[e] = (live ? (((xm && comp == 0) || (ym && comp == 1)) ? (T)-1 : (T)1) : (T)0) * val;
    }
    __syncthreads();
    const int sn = min(SC, n - s0);
#pragma unroll 4
    for (int sl = 0; sl < sn; ++sl) {
      const int s = s0 + sl;
      cplx<T> uv[NV];
#pragma unroll
      for (int c = 0; c < NV; ++c) uv[c] = u[(c * SC + sl) * W + w];
      const int ts = SYM ? s * (2 * n - s + 1) / 2 : 0;
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        int t;
        T g = (T)1;
        if (SYM) {
          // s < r: transposed entry (s, r) with the z sign pattern; s >= r: row r of the upper triangle
          const bool lo = s < rr[i];
          t = lo ? ts + (rr[i] - s) : t0[i] + (s - rr[i]);
          if (lo && rz[i] != (s >= nz0)) g = (T)-1;
        } else {
          t = t0[i] + s;
        }
        const cplx<T> b = g * Bk[(int64_t)t * W + w];
#pragma unroll
        for (int c = 0; c < NV; ++c) acc[i][c] = acc[i][c] + b * uv[c];
      }
    }
    __syncthreads();
  }
  if (!wl) return;
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    const int r = rsi * ROWS + (int)(threadIdx.x / W) + RL * i;
    if (r >= n) continue;
    const int comp = r / N;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int rc = v >> 2, xm = v & 1, ym = (v >> 1) & 1;
      const bool live = rc < nr && (!xm || xl) && (!ym || ymir);
      if (!live) continue;
      const int gx = xm ? L - kx : kx, gy = ym ? M - kyb : kyb;
      const bool neg = (xm && comp == 0) || (ym && comp == 1);
      Uo[rc * rhs_stride + r * sstride + gx + (int64_t)L * gy] = neg ? -acc[i][v] : acc[i][v];
    }
  }
}

// Mirrored 2-level apply on the matrix cores (FP32).  For one stored block B (kx <= L/2, ky <= M/2) the
// product is a GEMM Y = B U over the columns q = 4 rc + v: right-hand side rc, mirror variant v (bit 0:
// x mirror, bit 1: y mirror) with the sign flips of the mirror symmetry applied to U on the way in and to
// Y on the way out (as in k_blockmv2m).  Complex arithmetic as four real products on
// v_mfma_f32_16x16x4f32: Yr += Br Ur - Bi Ui, Yi += Br Ui + Bi Ur.  A wave computes 16 rows x 16 CT
// columns (CT tiles of 16x16, real and imaginary accumulators); the NW waves of a workgroup take
// consecutive 16-row tiles and share the U panel, staged through LDS in chunks of KC columns of B.  The
// blocks are full and column-major (entry (r, s) of stored block b at (b n + s) n + r), so the 16 lanes
// of a row tile read 128 contiguous bytes.  Operand layouts of the 16x16x4 instruction: A lane l holds
// (row l % 16, k l / 16), B lane l holds (k l / 16, column l % 16), D lane l holds column l % 16 of rows
// 4 (l / 16) + i, i = 0..3.
typedef float vs_f4 __attribute__((ext_vector_type(4)));
#ifndef VS_BMQ_KC
#define VS_BMQ_KC 16  // block columns per staged chunk of U
#endif

template <int CT>
__global__ __launch_bounds__(256) void k_blockmv2q(const c64* __restrict__ B, const c64* __restrict__ U,
                                                   c64* __restrict__ U2, int L, int M, int Lb, int Mb, int n,
                                                   int RG, int CG, int64_t rhs_stride, int nrhs) {
  constexpr int KC = VS_BMQ_KC;
  constexpr int QW = 16 * CT;  // columns of a workgroup
  c64* u = reinterpret_cast<c64*>(vs_smem);  // [KC][QW]
  const int64_t G = (int64_t)Lb * Mb * RG * CG;
  const int64_t per_xcd = (G + 7) / 8;
  const int64_t lid = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (lid >= G) return;
  // consecutive logical ids (the same XCD) are consecutive stored blocks kx at one (ky, row group, column
  // group): a 128-byte line of the gathered volume holds 16 consecutive kx, read by neighbouring workgroups
  const int64_t nblk = (int64_t)Lb * Mb;
  const int64_t blk = lid % nblk;
  const int64_t rest = lid / nblk;
  const int cg = (int)(rest % CG);
  const int rg = (int)(rest / CG);
  const int kx = (int)(blk % Lb), kyb = (int)(blk / Lb);
  const int NW = blockDim.x >> 6;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r0 = (rg * NW + wave) * 16;
  const int q0 = cg * QW;
  const int NQ = 4 * nrhs;
  const int N = n / 3;
  const int64_t sstride = (int64_t)L * M;
  const bool xl = kx != 0 && 2 * kx != L;
  const bool ymir = kyb != 0 && 2 * kyb != M;
  const c64* Bb = B + blk * (int64_t)n * n;
  vs_f4 accr[CT], acci[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    accr[c] = vs_f4{0.f, 0.f, 0.f, 0.f};
    acci[c] = vs_f4{0.f, 0.f, 0.f, 0.f};
  }
  const int ar = r0 + (lane & 15);  // A row of this lane
  const int ak = lane >> 4;         // A / B k index of this lane
  const int bq = lane & 15;         // B column of this lane within a tile
  // the A operands of a whole chunk (KC / 4 k-steps) are loaded together, one chunk ahead of their use
  constexpr int KS = KC / 4;
  c64 acur[KS], anxt[KS];
  auto load_a = [&](int s0, c64* dst) {
#pragma unroll
    for (int j = 0; j < KS; ++j) {
      const int s = s0 + 4 * j + ak;
      const bool al = ar < n && s < n;
      const c64 a = Bb[al ? (int64_t)s * n + ar : 0];
      dst[j] = al ? a : mk<float>(0.f, 0.f);
    }
  };
  load_a(0, acur);
  for (int s0 = 0; s0 < n; s0 += KC) {
    // stage U(s0 .. s0 + KC - 1, q0 .. q0 + QW - 1), signs applied, zero where the variant or column is
    // dead; the gathers of a batch of 8 elements per lane are issued before any of them is stored
    for (int e0 = 0; e0 < KC * QW; e0 += 8 * (int)blockDim.x) {
      c64 val[8];
      float sg[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int e = e0 + j * (int)blockDim.x + (int)threadIdx.x;
        const int ql = e % QW, sl = e / QW;
        const int s = s0 + sl, q = q0 + ql;
        const int rc = q >> 2, xm = q & 1, ym = (q >> 1) & 1;
        const bool live = e < KC * QW && s < n && q < NQ && (!xm || xl) && (!ym || ymir);
        const int comp = (s >= N) + (s >= 2 * N);
        const int gx = xm ? L - kx : kx, gy = ym ? M - kyb : kyb;
        val[j] = U[live ? rc * rhs_stride + (int64_t)s * sstride + gx + (int64_t)L * gy : 0];
        sg[j] = live ? (((xm && comp == 0) || (ym && comp == 1)) ? -1.f : 1.f) : 0.f;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int e = e0 + j * (int)blockDim.x + (int)threadIdx.x;
        if (e < KC * QW) u[e] = mk<float>(sg[j] * val[j].x, sg[j] * val[j].y);
      }
    }
    __syncthreads();
    if (s0 + KC < n) load_a(s0 + KC, anxt);
    // k-steps past the end of B multiply zero A operands and zero staged U rows
#pragma unroll
    for (int j = 0; j < KS; ++j) {
      const float a_r = acur[j].x, a_i = acur[j].y;
      c64 b[CT];
#pragma unroll
      for (int c = 0; c < CT; ++c) b[c] = u[(4 * j + ak) * QW + c * 16 + bq];
      // dependent products on one accumulator are 2 CT instructions apart
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        accr[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_r, b[c].x, accr[c], 0, 0, 0);
        acci[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_r, b[c].y, acci[c], 0, 0, 0);
      }
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        accr[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(-a_i, b[c].y, accr[c], 0, 0, 0);
        acci[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_i, b[c].x, acci[c], 0, 0, 0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < KS; ++j) acur[j] = anxt[j];
  }
  // D lane l: column q0 + 16 c + l % 16, rows r0 + 4 (l / 16) + i
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const int q = q0 + c * 16 + bq;
    const int rc = q >> 2, xm = q & 1, ym = (q >> 1) & 1;
    if (q >= NQ || (xm && !xl) || (ym && !ymir)) continue;
    const int gx = xm ? L - kx : kx, gy = ym ? M - kyb : kyb;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = r0 + 4 * (lane >> 4) + i;
      if (r >= n) continue;
      const int comp = r / N;
      const float sg = ((xm && comp == 0) || (ym && comp == 1)) ? -1.f : 1.f;
      U2[rc * rhs_stride + (int64_t)r * sstride + gx + (int64_t)L * gy] = mk<float>(sg * accr[c][i], sg * acci[c][i]);
    }
  }
}

// The same product with a workgroup per tile of WB = 8 consecutive stored blocks (one wave per block, the
// blocks numbered kx fastest, so a tile may wrap to the next ky), so that the staged U values of a (row s,
// column q) for the tile's blocks are consecutive kx in memory (the gathers read 64 contiguous bytes instead
// of 8).  Each wave computes 16 RT rows x 16 CT columns of its block; the loads of the next chunk of KC
// block columns are issued before the products of the current one (two LDS buffers) and their results are
// touched only after those products.
template <int RT, int CT, int KC, int WB, int WPB, int MINW>
__global__ __launch_bounds__(64 * WB * WPB, MINW) void k_blockmv2t(const c64* __restrict__ B, const c64* __restrict__ U,
                                                   c64* __restrict__ U2, int L, int M, int Lb, int Mb, int n,
                                                   int RG, int CG, int64_t rhs_stride, int nrhs) {
  // WB blocks per workgroup, WPB wavefronts per block (row bands of 16 RT rows each) sharing the block's
  // staged columns; MINW wavefronts per SIMD asked of the register allocation
  constexpr int KS = KC / 4, TH = 64 * WB * WPB;
  constexpr int QW = 16 * CT;
  constexpr int PER = WB * KC * QW / TH;  // staged values per lane per chunk
  c64* lds = reinterpret_cast<c64*>(vs_smem);  // [2][WB][KC][QW + 1]
  constexpr int QP = QW + 1;                    // padded column stride: rows of a block on other banks
  const int nblk = Lb * Mb;  // stored blocks, kx fastest; a tile may wrap to the next ky
  const int tiles = (nblk + WB - 1) / WB;
  const int64_t G = (int64_t)tiles * RG * CG;
  const int64_t per_xcd = (G + 7) / 8;
  const int64_t lid = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (lid >= G) return;
  // consecutive logical ids: consecutive tiles at one (row group, column group) (the other order, row and
  // column groups of one tile together, measured 4% slower on 96^3 blocks with 64 right-hand sides)
  const int64_t tk = lid % tiles, rest = lid / tiles;
  const int cg = (int)(rest % CG), rg = (int)(rest / CG);
  const int b0 = (int)tk * WB;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int bw = b0 + wave / WPB;
  const int kx = bw % Lb, kyb = bw / Lb;
  const int r0 = (rg * WPB + wave % WPB) * 16 * RT, q0 = cg * QW;
  const int NQ = 4 * nrhs;
  const int N = n / 3;
  const int64_t sstride = (int64_t)L * M;
  // staging: element e = j * TH + thread = ((sl * QW + ql) * WB + w): consecutive lanes take consecutive
  // blocks; a lane keeps one (block w, column ql) for every chunk and element (WB * QW divides TH), so that
  // its base offset, liveness and mirror flags are computed once
  static_assert(TH % (WB * QW) == 0 && KC % (TH / (WB * QW)) == 0, "staging layout");
  constexpr int SPJ = TH / (WB * QW);  // staged rows per element index j
  const int w_l = (int)threadIdx.x % WB, ql_l = ((int)threadIdx.x / WB) % QW, sl_l = (int)threadIdx.x / (WB * QW);
  int64_t base_l = 0;
  bool live_l, xm_l, ym_l;
  {
    const int q = q0 + ql_l, b = b0 + w_l;
    const int kxw = b % Lb, kyw = b / Lb;
    const int rc = q >> 2;
    xm_l = q & 1;
    ym_l = (q >> 1) & 1;
    const bool xl = kxw != 0 && 2 * kxw != L;
    const bool yl = kyw != 0 && 2 * kyw != M;
    live_l = b < nblk && q < NQ && (!xm_l || xl) && (!ym_l || yl);
    const int gx = xm_l ? L - kxw : kxw, gy = ym_l ? M - kyw : kyw;
    if (live_l) base_l = rc * rhs_stride + gx + (int64_t)L * gy;
  }
  // the loaded values stay untouched until they are stored (sign and liveness applied there), so that the
  // loads of the next chunk are still in flight during the products of the current one
  auto stage_load = [&](int s0, c64* val) {
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int s = s0 + j * SPJ + sl_l;
      val[j] = U[(live_l && s < n) ? base_l + (int64_t)s * sstride : 0];
    }
  };
  auto stage_store = [&](int buf, int s0, const c64* val) {
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int sl = j * SPJ + sl_l, s = s0 + sl;
      const bool neg = (xm_l && s < N) || (ym_l && s >= N && s < 2 * N);
      const float sg = (live_l && s < n) ? (neg ? -1.f : 1.f) : 0.f;
      lds[((buf * WB + w_l) * KC + sl) * QP + ql_l] = mk<float>(sg * val[j].x, sg * val[j].y);
    }
  };
  const bool bl = bw < nblk;
  const c64* Bb = B + (int64_t)(bl ? bw : 0) * n * n;
  const int ak = lane >> 4, am = lane & 15;
  // raw loads of the block entries; dead entries are zeroed when the chunk is taken up
  auto load_a = [&](int s0, c64 (*dst)[KS]) {
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int j = 0; j < KS; ++j) {
        const int r = r0 + 16 * t + am, s = s0 + 4 * j + ak;
        dst[t][j] = Bb[(bl && r < n && s < n) ? (int64_t)s * n + r : 0];
      }
  };
  auto take_a = [&](int s0, c64 (*dst)[KS], const c64 (*src)[KS]) {
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int j = 0; j < KS; ++j) {
        const int r = r0 + 16 * t + am, s = s0 + 4 * j + ak;
        dst[t][j] = (bl && r < n && s < n) ? src[t][j] : mk<float>(0.f, 0.f);
      }
  };
  vs_f4 accr[RT][CT], acci[RT][CT];
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      accr[t][c] = vs_f4{0.f, 0.f, 0.f, 0.f};
      acci[t][c] = vs_f4{0.f, 0.f, 0.f, 0.f};
    }
  c64 sv[PER];
  c64 acur[RT][KS], anxt[RT][KS];
  stage_load(0, sv);
  load_a(0, anxt);
  take_a(0, acur, anxt);
  stage_store(0, 0, sv);
  __syncthreads();
  int buf = 0;
  for (int s0 = 0; s0 < n; s0 += KC) {
    const bool more = s0 + KC < n;
    if (more) {
      stage_load(s0 + KC, sv);
      load_a(s0 + KC, anxt);
    }
    const c64* ub = lds + (size_t)(buf * WB + wave / WPB) * KC * QP;
#pragma unroll
    for (int j = 0; j < KS; ++j) {
      c64 b[CT];
#pragma unroll
      for (int c = 0; c < CT; ++c) b[c] = ub[(4 * j + ak) * QP + c * 16 + am];
#pragma unroll
      for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int c = 0; c < CT; ++c) {
          accr[t][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(acur[t][j].x, b[c].x, accr[t][c], 0, 0, 0);
          acci[t][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(acur[t][j].x, b[c].y, acci[t][c], 0, 0, 0);
        }
#pragma unroll
      for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int c = 0; c < CT; ++c) {
          accr[t][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(-acur[t][j].y, b[c].y, accr[t][c], 0, 0, 0);
          acci[t][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(acur[t][j].y, b[c].x, acci[t][c], 0, 0, 0);
        }
    }
    if (more) {
      stage_store(buf ^ 1, s0 + KC, sv);
      take_a(s0 + KC, acur, anxt);
    }
    __syncthreads();
    buf ^= 1;
  }
  if (!bl) return;
  const bool xl = kx != 0 && 2 * kx != L;
  const bool ymir = kyb != 0 && 2 * kyb != M;
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const int q = q0 + c * 16 + am;
    const int rc = q >> 2, xm = q & 1, ym = (q >> 1) & 1;
    if (q >= NQ || (xm && !xl) || (ym && !ymir)) continue;
    const int gx = xm ? L - kx : kx, gy = ym ? M - kyb : kyb;
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = r0 + 16 * t + 4 * ak + i;
        if (r >= n) continue;
        const int comp = r / N;
        const float sg = ((xm && comp == 0) || (ym && comp == 1)) ? -1.f : 1.f;
        U2[rc * rhs_stride + (int64_t)r * sstride + gx + (int64_t)L * gy] =
            mk<float>(sg * accr[t][c][i], sg * acci[t][c][i]);
      }
  }
}

// column-major FP64 blocks -> the same order in FP32 (the matrix-core product's layout)
__global__ void k_blocks_to_f32(const c128* __restrict__ in, c64* __restrict__ out, int64_t tot) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < tot) out[e] = ccast<float>(in[e]);
}

// per kx: U2(:, kx) = Binv(kx) U(:, kx), one wavefront per row
template <typename T>
__global__ __launch_bounds__(256) void k_blockmv1(const cplx<T>* __restrict__ B, const cplx<T>* __restrict__ U,
                                                  cplx<T>* __restrict__ U2, int L, int n, int rows_per_wg,
                                                  int64_t rhs_stride) {
  cplx<T>* u = reinterpret_cast<cplx<T>*>(vs_smem);
  const int kx = blockIdx.y;
  const cplx<T>* Ub = U + (int64_t)blockIdx.z * rhs_stride + kx;
  cplx<T>* U2b = U2 + (int64_t)blockIdx.z * rhs_stride + kx;
  for (int j = threadIdx.x; j < n; j += blockDim.x) u[j] = Ub[(int64_t)j * L];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r0 = blockIdx.x * rows_per_wg;
  for (int r = r0 + wave; r < min(n, r0 + rows_per_wg); r += 4) {
    const cplx<T>* br = B + ((int64_t)kx * n + r) * n;
    T ar = 0, ai = 0;
    for (int s = lane; s < n; s += 64) {
      const cplx<T> v = br[s] * u[s];
      ar += v.x;
      ai += v.y;
    }
    for (int off = 32; off >= 1; off >>= 1) {
      ar += __shfl_xor(ar, off);
      ai += __shfl_xor(ai, off);
    }
    if (lane == 0) U2b[(int64_t)r * L] = mk<T>(ar, ai);
  }
}

// ------------------------------------------------------------------------------------------------
// host drivers
// ------------------------------------------------------------------------------------------------
static unsigned nblocks(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// column tile of the mirrored block layout: 128 bytes of complex values
template <typename T> constexpr int mirror_tile() { return 128 / (int)sizeof(cplx<T>); }

// Stored columns of the last, mostly padded layout tile that the single-right-hand-side product takes
// from a compact copy (full column-major blocks) instead: rem = Lb mod Wt when it is at most a quarter of
// the tile (even L gives Lb = L/2 + 1, so rem = 1 whenever L/2 is a multiple of Wt), else 0.  In the tiled
// layout a workgroup on that tile keeps rem of its W lanes per row busy and, with W = 1, would read one
// useful value per 128-byte line.
template <typename T>
static int mirror_rem(int Lb) {
  if (std::getenv("VS_BMM_REM") && std::getenv("VS_BMM_REM")[0] == '0') return 0;
  const int Wt = mirror_tile<T>(), rem = Lb % Wt;
  return (Lb > Wt && rem > 0 && 4 * rem <= Wt) ? rem : 0;
}

// the last rem stored columns of the column-major FP64 blocks (block kx + Lb ky) -> full column-major
// blocks (ky * rem + w) in To
template <typename To>
__global__ void k_blocks_rem(const c128* __restrict__ in, cplx<To>* __restrict__ out, int Lb, int rem, int64_t nn,
                             int64_t tot) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= tot) return;
  const int64_t b = e / nn, q = e - b * nn;
  const int64_t w = b % rem, ky = b / rem;
  out[e] = ccast<To>(in[(ky * Lb + Lb - rem + w) * nn + q]);
}

// Toep (device, FP64) -> circ_L (device): Chan in x + FFT along x
void build_circ_L(vs_ctx* ctx, const c128* toep, c128* cL, int L, int M, int N) {
  const int64_t V = (int64_t)L * M * N;
  hipLaunchKernelGGL(k_chan_x, dim3(nblocks(6 * V)), dim3(256), 0, ctx->stream, toep, cL, L, (int64_t)M * N);
  VS_HIP(hipGetLastError());
  launch_row_fft_dense(ctx, VS_FP64, cL, (int64_t)M * N * 6, L, L, L, L, L, false, 1.0, 1, 0);
}

// circ_L -> F (device): Chan in y + FFT along y (component 2 not negated)
void build_circ_M(vs_ctx* ctx, const c128* cL, c128* F, int L, int M, int N) {
  const int64_t V = (int64_t)L * M * N;
  hipLaunchKernelGGL(k_chan_y, dim3(nblocks(6 * V)), dim3(256), 0, ctx->stream, cL, F, L, M, N);
  VS_HIP(hipGetLastError());
  launch_col_fft(ctx, VS_FP64, F, L, L, (int64_t)N * 6, (int64_t)L * M, M, M, M, false, 1.0, 1, 0);
}

void assemble2(vs_ctx* ctx, const c128* F, int L, int M, int N, int Lb, int Mb, c128 a, const c128* w, c128 ws,
               c128* out) {
  const int64_t tot = (int64_t)Lb * Mb * 9 * N * N;
  hipLaunchKernelGGL(k_assemble2, dim3(nblocks(tot)), dim3(256), 0, ctx->stream, F, L, M, N, Lb, Mb, a, w, ws, out);
  VS_HIP(hipGetLastError());
}

void assemble1(vs_ctx* ctx, const c128* cL, int L, int M, int N, c128 a, const c128* w, c128 ws, c128* out) {
  const int64_t n = 3LL * M * N;
  const int64_t tot = (int64_t)L * n * n;
  hipLaunchKernelGGL(k_assemble1, dim3(nblocks(tot)), dim3(256), 0, ctx->stream, cL, L, M, N, a, w, ws, out);
  VS_HIP(hipGetLastError());
}

// the blocked (matrix-core) inversion for blocks beyond LDS (VS_GJ_BLOCKED=0: the unblocked sweep)
static bool gj_blocked_enabled() {
  const char* e = std::getenv("VS_GJ_BLOCKED");
  return !(e && e[0] == '0');
}

void invert_blocks_dev(vs_ctx* ctx, c128* blocks, int64_t nblk, int64_t n) {
  // a whole block in the LDS of one CU (3N <= 96)
  const size_t lds = sizeof(c128) * (size_t)(n * n + n) + sizeof(int) * (size_t)n;
  if (lds <= 160 * 1024) {
    hipLaunchKernelGGL(k_gj_lds, dim3((unsigned)nblk), dim3(256), lds, ctx->stream, blocks, (int)n);
    VS_HIP(hipGetLastError());
    return;
  }
  if (n <= GJB_NMAX && nblk <= 65535 && gj_blocked_enabled()) {
    DevBuf perm(sizeof(int) * nblk * n), piv(sizeof(int) * nblk * n);
    // panels of 32 columns while the panel fits in LDS (3N <= 300), else 16
    const int NB = 32;
    const unsigned pthreads = (unsigned)((n + 63) / 64 * 64);
    const size_t lds_u = sizeof(c128) * (size_t)16 * (n + 1);
    const unsigned strips = (unsigned)((n + 15) / 16);
    for (int k0 = 0; k0 < n; k0 += NB) {
      hipLaunchKernelGGL(k_gjb_panel<32>, dim3((unsigned)nblk), dim3(pthreads), 0, ctx->stream, blocks, (int)n,
                         k0, perm.as<int>(), piv.as<int>());
      hipLaunchKernelGGL(k_gjb_update<32>, dim3(strips, (unsigned)nblk), dim3(512), lds_u, ctx->stream, blocks,
                         (int)n, k0, perm.as<const int>());
    }
    hipLaunchKernelGGL(k_gj_unswap, dim3(nblocks(n), (unsigned)nblk), dim3(256), 0, ctx->stream, blocks, n,
                       piv.as<int>());
    VS_HIP(hipGetLastError());
    VS_HIP(hipStreamSynchronize(ctx->stream));  // perm/piv are released on return
    return;
  }
  DevBuf colk(sizeof(c128) * nblk * n), piv(sizeof(int) * nblk * n);
  for (int64_t k = 0; k < n; ++k) {
    hipLaunchKernelGGL(k_gj_pivot, dim3((unsigned)nblk), dim3(256), 0, ctx->stream, blocks, n, k, colk.as<c128>(),
                       piv.as<int>());
    hipLaunchKernelGGL(k_gj_elim, dim3(nblocks(n * n), (unsigned)nblk), dim3(256), 0, ctx->stream, blocks, n, k,
                       colk.as<c128>());
  }
  hipLaunchKernelGGL(k_gj_unswap, dim3(nblocks(n), (unsigned)nblk), dim3(256), 0, ctx->stream, blocks, n,
                     piv.as<int>());
  VS_HIP(hipGetLastError());
  VS_HIP(hipStreamSynchronize(ctx->stream));  // colk/piv are released on return
}

// the matrix-core block product for mirrored FP32 blocks (VS_BMM_MFMA=0 keeps the vector-unit kernels)
static bool bmm_mfma_enabled() {
  static const bool v = [] {
    const char* e = std::getenv("VS_BMM_MFMA");
    return !(e && e[0] == '0');
  }();
  return v;
}

// column-major FP64 inverse blocks (device) -> apply layout(s) of pc
void pc_store_inverse(vs_pc* pc, const c128* inv) {
  vs_ctx* ctx = pc->ctx;
  const int64_t n = pc->nb;
  const int64_t nblk = pc->levels == 2 ? (int64_t)pc->Lb * pc->Mb : pc->L;
  const int64_t tot = nblk * n * n;
  auto go = [&](auto tag, DevBuf& dst) {
    using To = decltype(tag);
    if (pc->levels == 2 && pc->mirror) {
      const int Wt = mirror_tile<To>();
      const int64_t tiles = (pc->Lb + Wt - 1) / Wt;
      const int64_t NT = pc->sym ? n * (n + 1) / 2 : n * n;
      const size_t bytes = sizeof(cplx<To>) * (size_t)(tiles * Wt * pc->Mb * NT);
      dst.ensure(bytes);
      VS_HIP(hipMemsetAsync(dst.p, 0, bytes, ctx->stream));  // padding columns of the last tile
      if (pc->sym)
        hipLaunchKernelGGL((k_relayout2_tiled<To, true>), dim3(nblocks(tot)), dim3(256), 0, ctx->stream, inv,
                           dst.as<cplx<To>>(), pc->Lb, pc->Mb, (int)n, Wt);
      else
        hipLaunchKernelGGL((k_relayout2_tiled<To, false>), dim3(nblocks(tot)), dim3(256), 0, ctx->stream, inv,
                           dst.as<cplx<To>>(), pc->Lb, pc->Mb, (int)n, Wt);
      VS_HIP(hipGetLastError());
      const int rem = mirror_rem<To>(pc->Lb);
      DevBuf& rdst = sizeof(To) == 8 ? pc->invr64 : pc->invr32;
      if (rem > 0) {
        const int64_t rtot = (int64_t)rem * pc->Mb * n * n;
        rdst.ensure(sizeof(cplx<To>) * (size_t)rtot);
        hipLaunchKernelGGL((k_blocks_rem<To>), dim3(nblocks(rtot)), dim3(256), 0, ctx->stream, inv,
                           rdst.as<cplx<To>>(), pc->Lb, rem, n * n, rtot);
        VS_HIP(hipGetLastError());
      }
      return;
    }
    const int64_t stored = (pc->levels == 2 && pc->sym) ? nblk * n * (n + 1) / 2 : tot;
    dst.ensure(sizeof(cplx<To>) * stored);
    if (pc->levels == 2 && pc->sym)
      hipLaunchKernelGGL((k_relayout2_sym<To>), dim3(nblocks(tot)), dim3(256), 0, ctx->stream, inv,
                         dst.as<cplx<To>>(), pc->Lb, pc->Mb, (int)n);
    else if (pc->levels == 2)
      hipLaunchKernelGGL((k_relayout2<To>), dim3(nblocks(tot)), dim3(256), 0, ctx->stream, inv, dst.as<cplx<To>>(),
                         pc->Lb, pc->Mb, (int)n);
    else
      hipLaunchKernelGGL((k_relayout1<To>), dim3(nblocks(tot)), dim3(256), 0, ctx->stream, inv, dst.as<cplx<To>>(),
                         nblk, n);
    VS_HIP(hipGetLastError());
  };
  if (pc->precisions & VS_FP64) go(double(0), pc->inv64);
  if (pc->precisions & VS_FP32) go(float(0), pc->inv32);
  if ((pc->precisions & VS_FP32) && pc->levels == 2 && pc->mirror && bmm_mfma_enabled() &&
      (n >= 96 || std::getenv("VS_BMM_MFMA_MIN"))) {
    pc->invq32.ensure(sizeof(c64) * (size_t)tot);
    hipLaunchKernelGGL(k_blocks_to_f32, dim3(nblocks(tot)), dim3(256), 0, ctx->stream, inv, pc->invq32.as<c64>(), tot);
    VS_HIP(hipGetLastError());
  }
}

// which right-hand-side counts take the matrix-core product (VS_BMM_MFMA_MIN: the smallest count)
// (default: several right-hand sides on large blocks, 3N >= 96, where it beats the vector-unit kernels)
static bool bmm_mfma_use(int n, int nrhs) {
  const char* e = std::getenv("VS_BMM_MFMA_MIN");
  if (e) return nrhs >= std::max(1, std::atoi(e));
  return nrhs >= 2 && n >= 96;
}

// the tile-of-blocks variant of the matrix-core product (VS_BMM_TILE=0: one block per workgroup)
static bool bmm_tile_use(int n, int nrhs) {
  const char* e = std::getenv("VS_BMM_TILE");
  return !(e && e[0] == '0');
}

// LDS of one CU: the mirrored block product's budget for large blocks with several right-hand sides
constexpr size_t kBmmLdsMax = 160 * 1024;

// forced row split of the mirrored block product (a tuning knob; 0 = automatic)
// Row split of the mirrored block product, picked by timing: the best number of sibling workgroups per
// layout tile depends on the block size, the precision, the right-hand sides per pass and how the
// workgroups fill the CUs (measured: 3 for the 48-row FP32 blocks of a 128 x 128 x 16 grid with one
// right-hand side, 1 with four per pass, 5 instead of 6 or 3 for 192-row blocks).  Each configuration is
// timed once per process, at its first launch outside a graph capture (candidates of 16 .. n rows per
// workgroup, four launches each on the apply's own data), and the choice is kept for every later apply of
// that shape (whatever the number of passes over right-hand sides, which shrinks as batched systems
// finish).  VS_AUTOTUNE=0 keeps the default split.
struct BmmKey {
  bool fp64, sym;
  int n, Lb, Mb, W, RC, rem;
  bool operator<(const BmmKey& o) const {
    return std::tie(fp64, sym, n, Lb, Mb, W, RC, rem) < std::tie(o.fp64, o.sym, o.n, o.Lb, o.Mb, o.W, o.RC, o.rem);
  }
};

template <typename F>
static int bmm_tuned_rs(vs_ctx* ctx, const BmmKey& key, int RS0, F&& launch) {
  static std::mutex mu;
  static std::map<BmmKey, int> tuned;
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = tuned.find(key);
    if (it != tuned.end()) return it->second;
  }
  static const bool on = [] {
    const char* e = std::getenv("VS_AUTOTUNE");
    return !(e && e[0] == '0');
  }();
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (!on || hipStreamIsCapturing(ctx->stream, &st) != hipSuccess || st != hipStreamCaptureStatusNone) return RS0;
  const int n = key.n;
  std::vector<int> cand{RS0};
  for (int rows : {16, 24, 32, 40, 48, 64, 96, 128, n}) {
    if (rows > n || (int64_t)rows * key.W > 512) continue;
    const int rs = (n + rows - 1) / rows;
    const int rs_n = (n + (n + rs - 1) / rs - 1) / ((n + rs - 1) / rs);
    if (std::find(cand.begin(), cand.end(), rs_n) == cand.end()) cand.push_back(rs_n);
  }
  hipEvent_t e0, e1;
  VS_HIP(hipEventCreate(&e0));
  VS_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int best_rs = RS0;
  for (int rs : cand) {
    launch(rs);
    VS_HIP(hipEventRecord(e0, ctx->stream));
    for (int r = 0; r < 3; ++r) launch(rs);
    VS_HIP(hipEventRecord(e1, ctx->stream));
    VS_HIP(hipEventSynchronize(e1));
    float ms = 0;
    VS_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      best_rs = rs;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::lock_guard<std::mutex> g(mu);
  tuned[key] = best_rs;
  return best_rs;
}

// Tiling of the matrix-core product: 0 = 8 blocks per 512-lane workgroup, one wavefront each (48-row
// bands); 1 = 4 blocks with two wavefronts each (row bands of 96 sharing the block's staged columns, which
// are then read half as often), taken when the block size is a multiple of 96 (96^3 particle, 64
// right-hand sides: 5.27 -> 5.03 ms).  256-lane workgroups (2 x 2 or 4 x 1, registers capped for three per
// CU) measured 6.06 / 6.22 ms there (spills, shorter gather runs).  VS_BMT_CFG=0|1 forces one.
static int bmm_tile_cfg(int n) {
  const char* e = std::getenv("VS_BMT_CFG");  // read per call, like VS_BMM_TILE
  const int forced = e ? std::atoi(e) : -1;
  if (forced == 0 || forced == 1) return forced;
  return n % 96 == 0 ? 1 : 0;
}

// VS_BMM_W: at most this many stored columns per workgroup of the mirrored product (default: the tile)
static int bmm_w_max() {
  static const int v = [] {
    const char* e = std::getenv("VS_BMM_W");
    return e ? std::max(1, std::atoi(e)) : 1 << 30;
  }();
  return v;
}

static int bmm_rows_split() {
  static const int v = [] {
    const char* e = std::getenv("VS_BMM_RS");
    return e ? std::max(0, std::atoi(e)) : 0;
  }();
  return v;
}

// mode: PC_FULL the whole apply x -> y; PC_BLOCKS the block stage alone on the work volume; PC_MID the stages
// between the first and the last x pass (y FFTs and blocks).  Returns the volume that holds the result of
// the last stage run.
enum PcMode : int { PC_FULL = 0, PC_BLOCKS = 1, PC_MID = 2 };
template <typename T>
static cplx<T>* pc_apply_impl(vs_pc* pc, const cplx<T>* x, cplx<T>* y, int nrhs, int mode = PC_FULL) {
  const bool blocks_only = mode == PC_BLOCKS;
  const bool first_pass = mode == PC_FULL, y_passes = mode != PC_BLOCKS, last_pass = mode == PC_FULL;
  vs_ctx* ctx = pc->ctx;
  const int prec = sizeof(T) == 8 ? VS_FP64 : VS_FP32;
  const int L = pc->L, M = pc->M, N = pc->N;
  const int64_t V3 = 3 * pc->V;
  const cplx<T>* B = sizeof(T) == 8 ? pc->inv64.as<cplx<T>>() : pc->inv32.as<cplx<T>>();
  VS_CHECK(B != nullptr, VS_ERR_ARG, "preconditioner holds no blocks in the requested precision");
  pc->U.ensure(sizeof(cplx<T>) * V3 * nrhs);
  cplx<T>* U = pc->U.as<cplx<T>>();
  const int* map3 = pc->map3.as<int>();
  const int64_t lines = (int64_t)M * N * 3;
  if (first_pass) launch_row_fft_gather(ctx, prec, x, map3, pc->nunk, U, L, lines, nrhs, V3);
  if (pc->levels == 2) {
    if (y_passes)
      launch_col_fft(ctx, prec, U, L, L, (int64_t)N * 3, (int64_t)L * M, M, M, M, false, 1.0, nrhs, V3);
    const int n = pc->nb;
    if (pc->mirror) {
      // right-hand sides per block pass (times the four mirror points): the largest of 4, 2, 1 that keeps
      // a tile of at least 8 stored blocks within 64 KB of LDS, or, for large blocks (3N above ~40 in FP64,
      // ~80 in FP32) with several right-hand sides, a tile of 4 within the 160 KB of LDS of a CU, so that
      // one pass over the blocks still serves up to 4 of them (the blocks are read nrhs / RC times)
      const int Lb = pc->Lb, Mb = pc->Mb;
      if (sizeof(T) == 4 && pc->invq32.p && bmm_mfma_use(n, nrhs)) {
        const int NQ = 4 * nrhs;
        const int CT = NQ <= 16 ? 1 : (NQ <= 32 ? 2 : (NQ <= 64 ? 4 : 8));
        const int CG = (NQ + 16 * CT - 1) / (16 * CT);
        const int NW = std::min(4, (n + 15) / 16);
        const int RG = (n + 16 * NW - 1) / (16 * NW);
        const int64_t G = (int64_t)Lb * Mb * RG * CG;
        const size_t lds = sizeof(c64) * VS_BMQ_KC * 16 * CT;
        pc->U2.ensure(sizeof(cplx<T>) * V3 * nrhs);
        cplx<T>* U2 = pc->U2.as<cplx<T>>();
        dim3 grid((unsigned)(((G + 7) / 8) * 8));
        const c64* Bq = pc->invq32.as<c64>();
        const c64* Ui = reinterpret_cast<const c64*>(U);
        c64* Uo = reinterpret_cast<c64*>(U2);
        if (bmm_tile_use(n, nrhs)) {
          // tiles of 8 blocks: 48-row groups, 64-column groups (fewer columns: 32 / 16); chunks of 8 block
          // columns (16 measured 4% slower: 256 VGPRs)
          const int CTt = NQ >= 64 ? 4 : (NQ >= 32 ? 2 : 1);
          const int cfg = bmm_tile_cfg(n);  // 0: 8 blocks x 1 wavefront, 1: 4 blocks x 2 wavefronts
          const int WBt = cfg == 0 ? 8 : 4, WPBt = cfg == 1 ? 2 : 1;
          const int RGt = (n + 48 * WPBt - 1) / (48 * WPBt), CGt = (NQ + 16 * CTt - 1) / (16 * CTt);
          const int64_t Gt = (int64_t)((Lb * Mb + WBt - 1) / WBt) * RGt * CGt;
          const size_t ldst = sizeof(c64) * 2 * WBt * 8 * (16 * CTt + 1);
          dim3 gridt((unsigned)(((Gt + 7) / 8) * 8));
          const dim3 tht(64 * WBt * WPBt);
#define VS_BMT(WBV, WPBV, MINWV)                                                                                \
  switch (CTt) {                                                                                                \
    case 4: hipLaunchKernelGGL((k_blockmv2t<3, 4, 8, WBV, WPBV, MINWV>), gridt, tht, ldst, ctx->stream, Bq, Ui, Uo, L, M, Lb, Mb, n, RGt, CGt, V3, nrhs); break; \
    case 2: hipLaunchKernelGGL((k_blockmv2t<3, 2, 8, WBV, WPBV, MINWV>), gridt, tht, ldst, ctx->stream, Bq, Ui, Uo, L, M, Lb, Mb, n, RGt, CGt, V3, nrhs); break; \
    default: hipLaunchKernelGGL((k_blockmv2t<3, 1, 8, WBV, WPBV, MINWV>), gridt, tht, ldst, ctx->stream, Bq, Ui, Uo, L, M, Lb, Mb, n, RGt, CGt, V3, nrhs); break; \
  }
          if (cfg == 1) {
            VS_BMT(4, 2, 1)
          } else {
            VS_BMT(8, 1, 1)
          }
#undef VS_BMT
        } else switch (CT) {
          case 1: hipLaunchKernelGGL((k_blockmv2q<1>), grid, dim3(64 * NW), lds, ctx->stream, Bq, Ui, Uo, L, M, Lb, Mb, n, RG, CG, V3, nrhs); break;
          case 2: hipLaunchKernelGGL((k_blockmv2q<2>), grid, dim3(64 * NW), lds, ctx->stream, Bq, Ui, Uo, L, M, Lb, Mb, n, RG, CG, V3, nrhs); break;
          case 4: hipLaunchKernelGGL((k_blockmv2q<4>), grid, dim3(64 * NW), lds, ctx->stream, Bq, Ui, Uo, L, M, Lb, Mb, n, RG, CG, V3, nrhs); break;
          default: hipLaunchKernelGGL((k_blockmv2q<8>), grid, dim3(64 * NW), lds, ctx->stream, Bq, Ui, Uo, L, M, Lb, Mb, n, RG, CG, V3, nrhs); break;
        }
        VS_HIP(hipGetLastError());
        U = U2;
      } else if (n >= 96 && nrhs >= 2) {
        // large blocks, several right-hand sides: the tiled product, 2 or 4 right-hand sides per pass over
        // the blocks (one right-hand side: the row-split product below is faster)
        constexpr int Wk = 128 / (int)sizeof(cplx<T>);
        const int RCk = nrhs >= 4 ? 4 : 2;
        const int tiles = (Lb + Wk - 1) / Wk;
        const int chunks = (nrhs + RCk - 1) / RCk;
        const size_t lds = sizeof(cplx<T>) * (size_t)4 * RCk * 32 * Wk;
        pc->U2.ensure(sizeof(cplx<T>) * V3 * nrhs);
        cplx<T>* U2 = pc->U2.as<cplx<T>>();
#define VS_BMK(SYMV, RCV, RPTV)                                                                            \
  do {                                                                                                     \
    const int RSk = (n + (256 / Wk) * RPTV - 1) / ((256 / Wk) * RPTV);                                    \
    const int64_t Gk = (int64_t)tiles * Mb * RSk;                                                          \
    hipLaunchKernelGGL((k_blockmv2k<T, SYMV, RCV, RPTV>), dim3((unsigned)(((Gk + 7) / 8) * 8), (unsigned)chunks), \
                       dim3(256), lds, ctx->stream, B, U, U2, L, M, Lb, Mb, n, RSk, V3, nrhs);              \
  } while (0)
        // rows per lane: the accumulators (RPT x 4 RC complex values) take at most 64 VGPRs in FP32 and 128
        // in FP64
        constexpr int R4 = sizeof(T) == 4 ? 2 : 1, R2 = sizeof(T) == 4 ? 4 : 2;
        if (pc->sym) {
          if (RCk == 4) VS_BMK(true, 4, R4); else VS_BMK(true, 2, R2);
        } else {
          if (RCk == 4) VS_BMK(false, 4, R4); else VS_BMK(false, 2, R2);
        }
#undef VS_BMK
        VS_HIP(hipGetLastError());
        U = U2;
      } else {
      size_t budget = 65536;
      auto fits = [&](int rc, int w) { return (size_t)4 * rc * n * w * sizeof(cplx<T>) <= budget; };
      int RC = nrhs >= 4 ? 4 : (nrhs >= 2 ? 2 : 1);
      const bool big = RC > 1 && !fits(RC, std::min(8, Lb));
      if (big) budget = kBmmLdsMax;
      while (RC > 1 && !fits(RC, std::min(big ? 4 : 8, Lb))) RC /= 2;
      VS_CHECK(fits(RC, 1), VS_ERR_UNSUPPORTED, "2-level block too large for LDS");
      // workgroups cover W columns of a Wt-wide layout tile (W | Wt, both powers of two); rows are split over
      // RS sibling workgroups of about 256 lanes (measured best on 48-row blocks in both precisions)
      const int Wt = mirror_tile<T>();
      int lw = 0;
      while ((2 << lw) <= Wt && (2 << lw) <= (big ? 4 : Wt) && (2 << lw) <= bmm_w_max() && fits(RC, 2 << lw)) ++lw;
      const int W = 1 << lw;
      // the last rem stored columns from their compact copy (single right-hand side), when there is one
      const cplx<T>* Br = sizeof(T) == 8 ? pc->invr64.as<cplx<T>>() : pc->invr32.as<cplx<T>>();
      const int rem = (RC == 1 && Br) ? mirror_rem<T>(Lb) : 0;
      const int tiles = rem > 0 ? Lb / Wt : (Lb + Wt - 1) / Wt, subs = Wt / W;
      const int chunks = (nrhs + RC - 1) / RC;
      const size_t lds = sizeof(cplx<T>) * (size_t)4 * RC * n * W;
      pc->U2.ensure(sizeof(cplx<T>) * V3 * nrhs);
      cplx<T>* U2 = pc->U2.as<cplx<T>>();
      // one launch with the rows of a block split over RS sibling workgroups (every split computes each
      // output row in the same order, so the results do not depend on it)
      const cplx<T>* Uin = U;
      auto launch = [&](int RS) {
        const int rows = (n + RS - 1) / RS;
        RS = (n + rows - 1) / rows;
        const int threads = (int)std::min<int64_t>(512, ((int64_t)rows * W + 63) / 64 * 64);
        // a remainder column's rows over as few workgroups of the same size as hold them
        const int RSr = rem > 0 ? (n + threads - 1) / threads : 1;
        const int64_t G = (int64_t)tiles * subs * Mb * RS + (int64_t)rem * Mb * RSr;
        dim3 grid((unsigned)(((G + 7) / 8) * 8), (unsigned)chunks);
#define VS_BMM(SYMV, RCV)                                                                                 \
  do {                                                                                                    \
    if (lds > 65536)                                                                                      \
      VS_HIP(hipFuncSetAttribute((const void*)k_blockmv2m<T, SYMV, RCV>,                                  \
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));                  \
    hipLaunchKernelGGL((k_blockmv2m<T, SYMV, RCV>), grid, dim3(threads), lds, ctx->stream, B, Uin, U2, L, M, \
                       Lb, Mb, n, Wt, lw, RS, V3, nrhs, Br, rem, RSr);                                    \
  } while (0)
        if (pc->sym) {
          if (RC == 4) VS_BMM(true, 4); else if (RC == 2) VS_BMM(true, 2); else VS_BMM(true, 1);
        } else {
          if (RC == 4) VS_BMM(false, 4); else if (RC == 2) VS_BMM(false, 2); else VS_BMM(false, 1);
        }
#undef VS_BMM
        VS_HIP(hipGetLastError());
      };
      const int rows_wg = std::max(1, 256 / W);
      int RS = (n + rows_wg - 1) / rows_wg;
      if (bmm_rows_split() > 0) {
        RS = std::min(bmm_rows_split(), n);
      } else {
        const BmmKey key{sizeof(T) == 8, pc->sym, n, Lb, Mb, W, RC, rem};
        RS = bmm_tuned_rs(ctx, key, RS, launch);
      }
      launch(RS);
      U = U2;
      }
    } else {
    // right-hand sides per block-matrix pass: 1, 4 or 8
    int RC = nrhs >= 8 ? 8 : (nrhs >= 4 ? 4 : 1);
    while (RC > 1 && (size_t)RC * n * sizeof(cplx<T>) > 65536 / 4) RC = RC == 8 ? 4 : 1;
    int W = std::min(64, L);
    while (W > 1 && (size_t)RC * n * W * sizeof(cplx<T>) > 65536) W /= 2;
    while (W > 16 && (int64_t)((L + W - 1) / W) * M * ((nrhs + RC - 1) / RC) < 1024) W /= 2;
    VS_CHECK((size_t)RC * n * W * sizeof(cplx<T>) <= 65536, VS_ERR_UNSUPPORTED, "2-level block too large for LDS");
    const size_t lds = sizeof(cplx<T>) * (size_t)RC * n * W;
    const int tiles = (L + W - 1) / W;
    dim3 grid((unsigned)(tiles * M), (unsigned)((nrhs + RC - 1) / RC));
#define VS_BMV(RCV)                                                                                        \
  if (pc->sym)                                                                                             \
    hipLaunchKernelGGL((k_blockmv2_sym<T, RCV>), grid, dim3(256), lds, ctx->stream, B, U, L, M, n, W, V3, nrhs); \
  else                                                                                                     \
    hipLaunchKernelGGL((k_blockmv2<T, RCV>), grid, dim3(256), lds, ctx->stream, B, U, L, M, n, W, V3, nrhs);
    if (RC == 8) {
      VS_BMV(8)
    } else if (RC == 4) {
      VS_BMV(4)
    } else {
      VS_BMV(1)
    }
#undef VS_BMV
    VS_HIP(hipGetLastError());
    }
    if (blocks_only) return U;
    launch_col_fft(ctx, prec, U, L, L, (int64_t)N * 3, (int64_t)L * M, M, M, M, true, 1.0, nrhs, V3);
    if (last_pass) launch_row_ifft_scatter(ctx, prec, U, map3, pc->nunk, y, L, lines, 1.0 / ((double)L * M), nrhs, V3);
    return U;
  } else {
    pc->U2.ensure(sizeof(cplx<T>) * V3 * nrhs);
    cplx<T>* U2 = pc->U2.as<cplx<T>>();
    const int n = pc->nb;
    const int rows = 16;
    const size_t lds = sizeof(cplx<T>) * (size_t)n;
    hipLaunchKernelGGL((k_blockmv1<T>), dim3((unsigned)((n + rows - 1) / rows), (unsigned)L, (unsigned)nrhs), dim3(256),
                       lds, ctx->stream, B, U, U2, L, n, rows, V3);
    VS_HIP(hipGetLastError());
    if (last_pass) launch_row_ifft_scatter(ctx, prec, U2, map3, pc->nunk, y, L, lines, 1.0 / (double)L, nrhs, V3);
    return U2;
  }
}

// the block-product stage alone on the preconditioner's work volume (zeroed at the first call)
void pc_blocks_dev(vs_pc* pc, int precision, int nrhs) {
  VS_CHECK(!pc->mop, VS_ERR_UNSUPPORTED, "an operator preconditioner has no block stage");
  const size_t es = precision == VS_FP64 ? sizeof(c128) : sizeof(c64);
  const size_t need = es * 3 * (size_t)pc->V * nrhs;
  if (pc->U.bytes < need) {
    pc->U.ensure(need);
    VS_HIP(hipMemsetAsync(pc->U.p, 0, need, pc->ctx->stream));
  }
  if (precision == VS_FP64)
    pc_apply_impl<double>(pc, nullptr, nullptr, nrhs, PC_BLOCKS);
  else if (precision == VS_FP32)
    pc_apply_impl<float>(pc, nullptr, nullptr, nrhs, PC_BLOCKS);
  else
    throw Error{VS_ERR_ARG, "precision must be VS_FP64 or VS_FP32"};
}

void* pc_work_dev(vs_pc* pc, int precision) {
  const size_t es = precision == VS_FP64 ? sizeof(c128) : sizeof(c64);
  pc->U.ensure(es * 3 * (size_t)pc->V);
  return pc->U.p;
}

void* pc_mid_dev(vs_pc* pc, int precision, double* scale) {
  VS_CHECK(!pc->mop, VS_ERR_UNSUPPORTED, "an operator preconditioner has no circulant stages");
  *scale = pc->levels == 2 ? 1.0 / ((double)pc->L * pc->M) : 1.0 / (double)pc->L;
  if (precision == VS_FP64) return pc_apply_impl<double>(pc, nullptr, nullptr, 1, PC_MID);
  if (precision == VS_FP32) return pc_apply_impl<float>(pc, nullptr, nullptr, 1, PC_MID);
  throw Error{VS_ERR_ARG, "precision must be VS_FP64 or VS_FP32"};
}

void mvp_dev(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs);

void pc_apply_dev(vs_pc* pc, const void* x, void* y, int precision, int nrhs) {
  VS_CHECK(x != y, VS_ERR_ARG, "preconditioner input and output must differ");
  if (pc->mop) return mvp_dev(pc->mop, x, y, 0, precision, nrhs);
  if (precision == VS_FP64)
    pc_apply_impl<double>(pc, (const c128*)x, (c128*)y, nrhs);
  else if (precision == VS_FP32)
    pc_apply_impl<float>(pc, (const c64*)x, (c64*)y, nrhs);
  else
    throw Error{VS_ERR_ARG, "precision must be VS_FP64 or VS_FP32"};
}

}  // namespace vs
