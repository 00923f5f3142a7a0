// Operator assembly (Toeplitz Green dyadic, circulant embedding, its FFT) and the FFT matvec.
//
// Matvec data flow for one right-hand side (MVP.m:101-133), with S the padded spectrum volume of
// 2L x 2M x N x 3 complex values (only the lower N z-planes ever hold data):
//   mv_fx   : gather J from the compressed vector, zero-pad x to 2L, FFT along x   -> S(kx, y<M, z<N, c)
//   col FFT : zero-pad y to 2M, FFT along y                                         -> S(kx, ky, z<N, c)
//   mv_fzmul: zero-pad z to 2N, FFT along z, 3x3 tensor product with fN, inverse FFT along z, keep z<N
//   col IFFT: inverse FFT along y, keep y<M
//   mv_ix   : inverse FFT along x, keep x<L, y = Gram Mr J - Mc conv / P, scatter to the compressed vector
// The zero padding is never materialised in HBM: pad values are produced in LDS and cropped values are
// never written.
#include "fft_device.hpp"
#include "handles.hpp"

#include <algorithm>
#include <cstring>

namespace vs {

extern __shared__ __align__(16) unsigned char vs_smem[];

template <typename T>
CoefDev<T> coef_dev(const Coef& c) {
  CoefDev<T> d;
  d.scalar = c.scalar ? 1 : 0;
  d.s = ccast<T>(c.s);
  d.v = c.scalar ? nullptr : (sizeof(T) == 8 ? c.v64.as<cplx<T>>() : c.v32.as<cplx<T>>());
  return d;
}
template CoefDev<float> coef_dev<float>(const Coef&);
template CoefDev<double> coef_dev<double>(const Coef&);

// ------------------------------------------------------------------------------------------------
// generic tiled FFT kernels
// ------------------------------------------------------------------------------------------------

// Row FFT: lines along the contiguous axis.  IO supplies load(line, pos, rhs) / store(line, k, rhs, v).
template <typename T, class IO>
__global__ __launch_bounds__(FFT_THREADS) void k_row_fft(FftDev<T> plan, IO io, int64_t nlines, int nl,
                                                         int n_in, int n_out, int inv) {
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(vs_smem);
  const int n = plan.n;
  const int ld = nl + 1;
  const int rhs = blockIdx.y;
  const int64_t g0 = (int64_t)blockIdx.x * nl;
  for (int e = threadIdx.x; e < nl * n; e += FFT_THREADS) {
    const int l = e / n, x = e - l * n;
    const int64_t g = g0 + l;
    cplx<T> v = mk<T>(0, 0);
    if (x < n_in && g < nlines) v = io.load(g, x, rhs);
    buf[x * ld + l] = v;
  }
  __syncthreads();
  lds_fft<T>(buf, ld, nl, plan, inv != 0);
  for (int e = threadIdx.x; e < nl * n_out; e += FFT_THREADS) {
    const int l = e / n_out, k = e - l * n_out;
    const int64_t g = g0 + l;
    if (g < nlines) io.store(g, k, rhs, buf[k * ld + l]);
  }
}

// Column FFT along a strided axis of a volume, in place: element (col, pos, batch) at
// data[rhs*rhs_stride + batch*sb + pos*sa + col], col < fast.  Tiles of W consecutive columns.
template <typename T>
__global__ __launch_bounds__(FFT_THREADS) void k_col_fft(FftDev<T> plan, cplx<T>* data, int64_t fast,
                                                         int64_t sa, int64_t sb, int W, int n_in, int n_out,
                                                         int inv, T scale, int64_t rhs_stride) {
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(vs_smem);
  const int n = plan.n;
  const int64_t tiles = (fast + W - 1) / W;
  const int64_t t = blockIdx.x % tiles, b = blockIdx.x / tiles;
  const int64_t c0 = t * W;
  cplx<T>* base = data + (int64_t)blockIdx.y * rhs_stride + b * sb + c0;
  const int wv = (int)std::min<int64_t>(W, fast - c0);
  for (int e = threadIdx.x; e < W * n; e += FFT_THREADS) {
    const int pos = e / W, w = e - pos * W;
    cplx<T> v = mk<T>(0, 0);
    if (pos < n_in && w < wv) v = base[pos * sa + w];
    buf[pos * W + w] = v;
  }
  __syncthreads();
  lds_fft<T>(buf, W, W, plan, inv != 0);
  for (int e = threadIdx.x; e < W * n_out; e += FFT_THREADS) {
    const int pos = e / W, w = e - pos * W;
    if (w < wv) base[pos * sa + w] = scale * buf[pos * W + w];
  }
}

// ------------------------------------------------------------------------------------------------
// IO functors
// ------------------------------------------------------------------------------------------------
template <typename T>
struct IoDense {
  cplx<T>* in;
  cplx<T>* out;
  int64_t ld_in, ld_out, rhs_stride;
  T scale;
  __device__ __forceinline__ cplx<T> load(int64_t g, int x, int rhs) const {
    return in[rhs * rhs_stride + g * ld_in + x];
  }
  __device__ __forceinline__ void store(int64_t g, int k, int rhs, cplx<T> v) const {
    out[rhs * rhs_stride + g * ld_out + k] = scale * v;
  }
};

// compressed vector -> dense volume (x FFT of the L x (lines) volume, stored with row pitch L)
template <typename T>
struct IoGather {
  const cplx<T>* x;
  const int* map3;
  int64_t nunk;
  int L;
  cplx<T>* out;
  int64_t out_rhs;
  __device__ __forceinline__ cplx<T> load(int64_t g, int xp, int rhs) const {
    const int m = map3[g * L + xp];
    return m >= 0 ? x[rhs * nunk + m] : mk<T>(0, 0);
  }
  __device__ __forceinline__ void store(int64_t g, int k, int rhs, cplx<T> v) const {
    out[rhs * out_rhs + g * L + k] = v;
  }
};

template <typename T>
struct IoScatter {
  const cplx<T>* in;
  const int* map3;
  int64_t nunk;
  int L;
  cplx<T>* y;
  int64_t in_rhs;
  T scale;
  __device__ __forceinline__ cplx<T> load(int64_t g, int xp, int rhs) const {
    return in[rhs * in_rhs + g * L + xp];
  }
  __device__ __forceinline__ void store(int64_t g, int k, int rhs, cplx<T> v) const {
    const int m = map3[g * L + k];
    if (m >= 0) y[rhs * nunk + m] = scale * v;
  }
};

// matvec first pass: gather (+ transp pre-op Mc conj(J)), pad, x FFT -> S rows
template <typename T>
struct IoMvFx {
  const cplx<T>* x;
  const int* map3;
  int64_t nunk, V;
  int L, M;
  int transp;
  CoefDev<T> mc;
  cplx<T>* S;
  int64_t S_rhs;
  __device__ __forceinline__ cplx<T> load(int64_t g, int xp, int rhs) const {
    const int64_t gi = g * L + xp;
    const int m = map3[gi];
    if (m < 0) return mk<T>(0, 0);
    cplx<T> J = x[rhs * nunk + m];
    if (transp) J = mc.at(gi % V) * cconj(J);
    return J;
  }
  __device__ __forceinline__ void store(int64_t g, int k, int rhs, cplx<T> v) const {
    const int64_t y = g % M, zc = g / M;
    S[rhs * S_rhs + k + (int64_t)2 * L * (y + (int64_t)2 * M * zc)] = v;
  }
};

// matvec last pass: inverse x FFT, crop, diagonal term, scatter
template <typename T>
struct IoMvIx {
  const cplx<T>* x;
  cplx<T>* y;
  const int* map3;
  int64_t nunk, V;
  int L, M;
  int transp;
  CoefDev<T> a, at, b;
  const cplx<T>* S;
  int64_t S_rhs;
  T invP;
  __device__ __forceinline__ cplx<T> load(int64_t g, int kx, int rhs) const {
    const int64_t yy = g % M, zc = g / M;
    return S[rhs * S_rhs + kx + (int64_t)2 * L * (yy + (int64_t)2 * M * zc)];
  }
  __device__ __forceinline__ void store(int64_t g, int xp, int rhs, cplx<T> v) const {
    const int64_t gi = g * L + xp;
    const int m = map3[gi];
    if (m < 0) return;
    const int64_t vox = gi % V;
    const cplx<T> J = x[rhs * nunk + m];
    const cplx<T> conv = invP * v;
    cplx<T> out;
    if (!transp)
      out = a.at(vox) * J - b.at(vox) * conv;
    else
      out = at.at(vox) * J - cconj(conv);
    y[rhs * nunk + m] = out;
  }
};

// ------------------------------------------------------------------------------------------------
// fused z pass: FFT along z, 3x3 tensor product, inverse FFT along z
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(FFT_THREADS) void k_mv_fzmul(FftDev<T> plan, cplx<T>* S, const cplx<T>* fN,
                                                          int L2, int M2, int N, int64_t P, int W,
                                                          int64_t S_rhs) {
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(vs_smem);
  const int n = plan.n;  // 2N
  const int ld = 3 * W;
  const int tiles = (L2 + W - 1) / W;
  const int t = blockIdx.x % tiles, ky = blockIdx.x / tiles;
  const int kx0 = t * W;
  const int wv = min(W, L2 - kx0);
  cplx<T>* Sb = S + (int64_t)blockIdx.y * S_rhs;
  const int64_t zs = (int64_t)L2 * M2;  // z stride in S
  const int64_t cs = zs * N;            // component stride in S
  // load z < N of the three components, zero the padding
  for (int e = threadIdx.x; e < 3 * W * n; e += FFT_THREADS) {
    const int w = e % W, q = e / W;
    const int pos = q % n, c = q / n;
    cplx<T> v = mk<T>(0, 0);
    if (pos < N && w < wv) v = Sb[c * cs + pos * zs + (int64_t)ky * L2 + kx0 + w];
    buf[pos * ld + c * W + w] = v;
  }
  __syncthreads();
  lds_fft<T>(buf, ld, 3 * W, plan, false);
  for (int e = threadIdx.x; e < W * n; e += FFT_THREADS) {
    const int w = e % W, kz = e / W;
    if (w < wv) {
      const int64_t o = (int64_t)kx0 + w + (int64_t)L2 * (ky + (int64_t)M2 * kz);
      const cplx<T> f0 = fN[o], f1 = fN[P + o], f2 = fN[2 * P + o];
      const cplx<T> f3 = fN[3 * P + o], f4 = fN[4 * P + o], f5 = fN[5 * P + o];
      cplx<T>* rowp = buf + kz * ld + w;
      const cplx<T> jx = rowp[0], jy = rowp[W], jz = rowp[2 * W];
      rowp[0] = f0 * jx + f1 * jy + f2 * jz;
      rowp[W] = f1 * jx + f3 * jy + f4 * jz;
      rowp[2 * W] = f2 * jx + f4 * jy + f5 * jz;
    }
  }
  __syncthreads();
  lds_fft<T>(buf, ld, 3 * W, plan, true);
  for (int e = threadIdx.x; e < 3 * W * N; e += FFT_THREADS) {
    const int w = e % W, q = e / W;
    const int pos = q % N, c = q / N;
    if (w < wv) Sb[c * cs + pos * zs + (int64_t)ky * L2 + kx0 + w] = buf[pos * ld + c * W + w];
  }
}

// ------------------------------------------------------------------------------------------------
// operator assembly kernels (FP64)
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void green6(double dx, double dy, double dz, double ko, c128* out) {
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  const double rh[3] = {dx / r, dy / r, dz / r};
  double sn, cn;
  sincos(ko * r, &sn, &cn);
  const c128 e = mk<double>(cn / r, sn / r);         // exp(i k r) / r
  const double c1 = ko * ko;
  const c128 c2 = mk<double>(-1.0 / (r * r), ko * r / (r * r));  // (i k r - 1) / r^2
  const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int p = 0; p < 6; ++p) {
    const double dab = pa[p] == pb[p] ? 1.0 : 0.0;
    const double rr = rh[pa[p]] * rh[pb[p]];
    const c128 t = mk<double>(c1 * (dab - rr), 0.0) + (dab - 3.0 * rr) * c2;
    out[p] = e * t;
  }
}

__global__ void k_toeplitz_dda(const double* __restrict__ r, int L, int M, int N, double ko, int nearby,
                               const double* __restrict__ gx, const double* __restrict__ gw, int nq,
                               double dxv, c128* __restrict__ toep) {
  const int64_t V = (int64_t)L * M * N;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const double R0x = r[0], R0y = r[V], R0z = r[2 * V];
  const double R1x = r[v], R1y = r[V + v], R1z = r[2 * V + v];
  const double dx = R1x - R0x, dy = R1y - R0y, dz = R1z - R0z;
  const double rjk = sqrt(dx * dx + dy * dy + dz * dz);
  c128 out[6];
#pragma unroll
  for (int p = 0; p < 6; ++p) out[p] = mk<double>(0, 0);
  if (nearby && rjk < 5 * dxv && rjk > 1e-15) {
    // 10^3-point Gauss-Legendre average over the voxel (getOPERATOR_DDA.m:47-79)
    for (int i = 0; i < nq; ++i)
      for (int j = 0; j < nq; ++j)
        for (int k = 0; k < nq; ++k) {
          // meshgrid convention: X varies with the 2nd index, Y with the 1st
          const double qx = R1x + dxv / 2 * gx[j] - R0x;
          const double qy = R1y + dxv / 2 * gx[i] - R0y;
          const double qz = R1z + dxv / 2 * gx[k] - R0z;
          const double w = (0.5 * gw[j]) * (0.5 * gw[i]) * (0.5 * gw[k]);
          c128 g[6];
          green6(qx, qy, qz, ko, g);
#pragma unroll
          for (int p = 0; p < 6; ++p) out[p] = out[p] + w * g[p];
        }
  } else if (fabs(rjk) > 1e-15) {
    green6(dx, dy, dz, ko, out);
  }
#pragma unroll
  for (int p = 0; p < 6; ++p) toep[v + V * p] = out[p];
}

// circulant embedding with parity signs (circulant_nop_const.m:5-40, gperiodic_coeff_nop.m:4-31)
__global__ void k_circ_embed(const c128* __restrict__ G, int L, int M, int N, c128* __restrict__ Gp) {
  const int64_t L2 = 2 * L, M2 = 2 * M, N2 = 2 * N;
  const int64_t P = L2 * M2 * N2;
  const int64_t V = (int64_t)L * M * N;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * P) return;
  const int p = (int)(t / P);
  int64_t q = t - p * P;
  const int i = (int)(q % L2);
  q /= L2;
  const int j = (int)(q % M2);
  const int k = (int)(q / M2);
  // sign of each component under reflection of x, y, z
  const double sL[6] = {1, -1, -1, 1, 1, 1}, sM[6] = {1, -1, 1, 1, -1, 1}, sN[6] = {1, 1, -1, 1, -1, 1};
  c128 v = mk<double>(0, 0);
  if (i != L && j != M && k != N) {
    double s = 1.0;
    int si = i, sj = j, sk = k;
    if (i > L) { si = 2 * L - i; s *= sL[p]; }
    if (j > M) { sj = 2 * M - j; s *= sM[p]; }
    if (k > N) { sk = 2 * N - k; s *= sN[p]; }
    v = s * G[si + (int64_t)L * (sj + (int64_t)M * sk) + V * p];
  }
  Gp[t] = v;
}

template <typename To, typename Ti>
__global__ void k_convert(const cplx<Ti>* __restrict__ in, cplx<To>* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = ccast<To>(in[i]);
}

// ------------------------------------------------------------------------------------------------
// launch helpers
// ------------------------------------------------------------------------------------------------
// LDS tile budget in complex elements: 32 KB of payload per workgroup (plus row padding), so that every
// FFT workgroup stays within 64 KB of LDS in both precisions
static int tile_budget(size_t elem_bytes) { return elem_bytes == 16 ? 2048 : 4096; }

static int row_tile(int n, int64_t nlines, size_t eb) {
  int nl = std::max(1, tile_budget(eb) / n);
  nl = std::min(nl, 64);
  // keep at least ~512 workgroups when there are enough lines
  while (nl > 4 && (nlines + nl - 1) / nl < 512) nl /= 2;
  return nl;
}

static int col_tile(int n, int64_t fast, int64_t nbatch, size_t eb) {
  int W = std::max(1, tile_budget(eb) / n);
  W = std::min(W, 64);
  int p = 1;
  while (p * 2 <= W) p *= 2;
  W = p;
  while (W > 16 && (int64_t)((fast + W - 1) / W) * nbatch < 512) W /= 2;
  while (W > 1 && W / 2 >= fast) W /= 2;
  return W;
}

template <typename T, class IO>
static void run_rows(vs_ctx* ctx, const IO& io, int n, int64_t nlines, int n_in, int n_out, bool inv, int nrhs) {
  const FftPlanHost& ph = ctx->plan(n);
  const FftDev<T> pd = ph.dev<T>();
  VS_CHECK(n <= tile_budget(sizeof(cplx<T>)), VS_ERR_UNSUPPORTED, "FFT length exceeds the LDS tile");
  const int nl = row_tile(n, nlines, sizeof(cplx<T>));
  const size_t lds = sizeof(cplx<T>) * (size_t)(nl + 1) * n;
  dim3 grid((unsigned)((nlines + nl - 1) / nl), (unsigned)nrhs);
  hipLaunchKernelGGL((k_row_fft<T, IO>), grid, dim3(FFT_THREADS), lds, ctx->stream, pd, io, nlines, nl, n_in,
                     n_out, inv ? 1 : 0);
  VS_HIP(hipGetLastError());
}

template <typename T>
static void run_cols(vs_ctx* ctx, cplx<T>* data, int64_t fast, int64_t sa, int64_t nbatch, int64_t sb, int n_in,
                     int n, int n_out, bool inv, double scale, int nrhs, int64_t rhs_stride) {
  const FftPlanHost& ph = ctx->plan(n);
  const FftDev<T> pd = ph.dev<T>();
  VS_CHECK(n <= tile_budget(sizeof(cplx<T>)), VS_ERR_UNSUPPORTED, "FFT length exceeds the LDS tile");
  const int W = col_tile(n, fast, nbatch, sizeof(cplx<T>));
  const size_t lds = sizeof(cplx<T>) * (size_t)W * n;
  const int64_t tiles = (fast + W - 1) / W;
  dim3 grid((unsigned)(tiles * nbatch), (unsigned)nrhs);
  hipLaunchKernelGGL((k_col_fft<T>), grid, dim3(FFT_THREADS), lds, ctx->stream, pd, data, fast, sa, sb, W, n_in,
                     n_out, inv ? 1 : 0, (T)scale, rhs_stride);
  VS_HIP(hipGetLastError());
}

void launch_col_fft(vs_ctx* ctx, int prec, void* data, int64_t fast, int64_t axis_stride, int64_t nbatch,
                    int64_t batch_stride, int n_in, int n, int n_out, bool inv, double scale, int nrhs,
                    int64_t rhs_stride) {
  if (prec == VS_FP64)
    run_cols<double>(ctx, (c128*)data, fast, axis_stride, nbatch, batch_stride, n_in, n, n_out, inv, scale, nrhs,
                     rhs_stride);
  else
    run_cols<float>(ctx, (c64*)data, fast, axis_stride, nbatch, batch_stride, n_in, n, n_out, inv, scale, nrhs,
                    rhs_stride);
}

void launch_row_fft_dense(vs_ctx* ctx, int prec, void* data, int64_t nlines, int n_in, int n, int n_out,
                          int64_t ld_in, int64_t ld_out, bool inv, double scale, int nrhs, int64_t rhs_stride) {
  if (prec == VS_FP64) {
    IoDense<double> io{(c128*)data, (c128*)data, ld_in, ld_out, rhs_stride, scale};
    run_rows<double>(ctx, io, n, nlines, n_in, n_out, inv, nrhs);
  } else {
    IoDense<float> io{(c64*)data, (c64*)data, ld_in, ld_out, rhs_stride, (float)scale};
    run_rows<float>(ctx, io, n, nlines, n_in, n_out, inv, nrhs);
  }
}

void launch_row_fft_gather(vs_ctx* ctx, int prec, const void* x, const int* map3, int64_t nunk, void* out, int L,
                           int64_t nlines, int nrhs, int64_t out_rhs_stride) {
  if (prec == VS_FP64) {
    IoGather<double> io{(const c128*)x, map3, nunk, L, (c128*)out, out_rhs_stride};
    run_rows<double>(ctx, io, L, nlines, L, L, false, nrhs);
  } else {
    IoGather<float> io{(const c64*)x, map3, nunk, L, (c64*)out, out_rhs_stride};
    run_rows<float>(ctx, io, L, nlines, L, L, false, nrhs);
  }
}

void launch_row_ifft_scatter(vs_ctx* ctx, int prec, const void* in, const int* map3, int64_t nunk, void* y, int L,
                             int64_t nlines, double scale, int nrhs, int64_t in_rhs_stride) {
  if (prec == VS_FP64) {
    IoScatter<double> io{(const c128*)in, map3, nunk, L, (c128*)y, in_rhs_stride, scale};
    run_rows<double>(ctx, io, L, nlines, L, L, true, nrhs);
  } else {
    IoScatter<float> io{(const c64*)in, map3, nunk, L, (c64*)y, in_rhs_stride, (float)scale};
    run_rows<float>(ctx, io, L, nlines, L, L, true, nrhs);
  }
}

// ------------------------------------------------------------------------------------------------
// matvec driver
// ------------------------------------------------------------------------------------------------
template <typename T>
static void mvp_impl(vs_op* op, const cplx<T>* x, cplx<T>* y, int transp, int nrhs) {
  vs_ctx* ctx = op->ctx;
  const int L = op->L, M = op->M, N = op->N;
  const int L2 = 2 * L, M2 = 2 * M, N2 = 2 * N;
  const int64_t S_rhs = (int64_t)L2 * M2 * N * 3;
  op->S.ensure(sizeof(cplx<T>) * S_rhs * nrhs);
  cplx<T>* S = op->S.as<cplx<T>>();
  const cplx<T>* fN = (sizeof(T) == 8) ? op->fN64.as<cplx<T>>() : op->fN32.as<cplx<T>>();
  VS_CHECK(fN != nullptr, VS_ERR_ARG, "operator holds no tensor in the requested precision");
  const int* map3 = op->map3.as<int>();
  VS_CHECK(map3 != nullptr, VS_ERR_ARG, "operator domain not set (vs_op_set_domain)");
  const int64_t lines = (int64_t)M * N * 3;
  // 1. gather + x FFT
  IoMvFx<T> fx{x, map3, op->nunk, op->V, L, M, transp, coef_dev<T>(op->b), S, S_rhs};
  run_rows<T>(ctx, fx, L2, lines, L, L2, false, nrhs);
  // 2. y FFT: columns (kx), axis y (stride 2L), batch (z, c) stride 2L*2M, input y < M
  run_cols<T>(ctx, S, L2, L2, (int64_t)N * 3, (int64_t)L2 * M2, M, M2, M2, false, 1.0, nrhs, S_rhs);
  // 3. fused z FFT + tensor product + inverse z FFT
  {
    const FftPlanHost& ph = ctx->plan(N2);
    const FftDev<T> pd = ph.dev<T>();
    VS_CHECK(3 * N2 <= tile_budget(sizeof(cplx<T>)), VS_ERR_UNSUPPORTED, "z FFT length exceeds the LDS tile");
    int W = std::max(1, tile_budget(sizeof(cplx<T>)) / (3 * N2));
    int p2 = 1;
    while (p2 * 2 <= W) p2 *= 2;
    W = std::min(p2, 64);
    while (W > 1 && W / 2 >= L2) W /= 2;
    const size_t lds = sizeof(cplx<T>) * (size_t)3 * W * N2;
    const int tiles = (L2 + W - 1) / W;
    dim3 grid((unsigned)(tiles * M2), (unsigned)nrhs);
    hipLaunchKernelGGL((k_mv_fzmul<T>), grid, dim3(FFT_THREADS), lds, ctx->stream, pd, S, fN, L2, M2, N, op->P, W,
                       S_rhs);
    VS_HIP(hipGetLastError());
  }
  // 4. inverse y FFT, keep y < M
  run_cols<T>(ctx, S, L2, L2, (int64_t)N * 3, (int64_t)L2 * M2, M2, M2, M, true, 1.0, nrhs, S_rhs);
  // 5. inverse x FFT, crop, diagonal, scatter
  IoMvIx<T> ix{x, y, map3, op->nunk, op->V, L, M, transp, coef_dev<T>(op->a), coef_dev<T>(op->at),
               coef_dev<T>(op->b), S, S_rhs, (T)(1.0 / (double)op->P)};
  run_rows<T>(ctx, ix, L2, lines, L2, L, true, nrhs);
}

void mvp_dev(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs) {
  VS_CHECK(nrhs >= 1, VS_ERR_ARG, "nrhs must be >= 1");
  VS_CHECK(x != y, VS_ERR_ARG, "matvec input and output must differ");
  if (precision == VS_FP64)
    mvp_impl<double>(op, (const c128*)x, (c128*)y, transp, nrhs);
  else if (precision == VS_FP32)
    mvp_impl<float>(op, (const c64*)x, (c64*)y, transp, nrhs);
  else
    throw Error{VS_ERR_ARG, "precision must be VS_FP64 or VS_FP32"};
}

// Gauss-Legendre nodes/weights on [-1,1] by Newton iteration on P_n (long double)
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.resize(n);
  w.resize(n);
  const long double PI = 3.141592653589793238462643383279502884L;
  for (int i = 0; i < n; ++i) {
    long double z = cosl(PI * (i + 0.75L) / (n + 0.5L)), pp = 0;
    for (int it = 0; it < 100; ++it) {
      long double p1 = 1, p2 = 0;
      for (int j = 1; j <= n; ++j) {
        long double p3 = p2;
        p2 = p1;
        p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1);
      long double z1 = z;
      z = z1 - p1 / pp;
      if (fabsl(z - z1) < 1e-19L) break;
    }
    x[n - 1 - i] = (double)z;
    w[n - 1 - i] = (double)(2 / ((1 - z * z) * pp * pp));
  }
}

void op_build_fN(vs_op* op) {
  // circulant embedding of the Toeplitz tensor and its 3-D FFT, in FP64 (fft_operator.m:1-8)
  vs_ctx* ctx = op->ctx;
  const int L2 = 2 * op->L, M2 = 2 * op->M, N2 = 2 * op->N;
  const int64_t P = op->P;
  op->fN64.ensure(sizeof(c128) * 6 * P);
  {
    const int64_t tot = 6 * P;
    hipLaunchKernelGGL(k_circ_embed, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream,
                       op->toep.as<c128>(), op->L, op->M, op->N, op->fN64.as<c128>());
    VS_HIP(hipGetLastError());
  }
  void* d = op->fN64.p;
  launch_row_fft_dense(ctx, VS_FP64, d, (int64_t)M2 * N2 * 6, L2, L2, L2, L2, L2, false, 1.0, 1, 0);
  launch_col_fft(ctx, VS_FP64, d, L2, L2, (int64_t)N2 * 6, (int64_t)L2 * M2, M2, M2, M2, false, 1.0, 1, 0);
  launch_col_fft(ctx, VS_FP64, d, (int64_t)L2 * M2, (int64_t)L2 * M2, 6, P, N2, N2, N2, false, 1.0, 1, 0);
  if (op->precisions & VS_FP32) {
    op->fN32.ensure(sizeof(c64) * 6 * P);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(2048), dim3(256), 0, ctx->stream, op->fN64.as<c128>(),
                       op->fN32.as<c64>(), 6 * P);
    VS_HIP(hipGetLastError());
  }
}

void op_build_dda(vs_op* op, const double* r_host, double ko, int nearby) {
  vs_ctx* ctx = op->ctx;
  const int64_t V = op->V;
  DevBuf r(sizeof(double) * 3 * V);
  VS_HIP(hipMemcpyAsync(r.p, r_host, sizeof(double) * 3 * V, hipMemcpyHostToDevice, ctx->stream));
  std::vector<double> gx, gw;
  gauss_legendre(10, gx, gw);
  DevBuf dgx(sizeof(double) * 10), dgw(sizeof(double) * 10);
  VS_HIP(hipMemcpyAsync(dgx.p, gx.data(), sizeof(double) * 10, hipMemcpyHostToDevice, ctx->stream));
  VS_HIP(hipMemcpyAsync(dgw.p, gw.data(), sizeof(double) * 10, hipMemcpyHostToDevice, ctx->stream));
  const double dx = r_host[1] - r_host[0];
  op->toep.ensure(sizeof(c128) * 6 * V);
  hipLaunchKernelGGL(k_toeplitz_dda, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, ctx->stream, r.as<double>(),
                     op->L, op->M, op->N, ko, nearby, dgx.as<double>(), dgw.as<double>(), 10, dx,
                     op->toep.as<c128>());
  VS_HIP(hipGetLastError());
  op->has_toep = true;
  op_build_fN(op);
  VS_HIP(hipStreamSynchronize(ctx->stream));
  if (!(op->precisions & VS_FP64)) op->fN64.release();
}

}  // namespace vs
