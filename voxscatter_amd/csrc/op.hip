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
#include "fft2.hpp"
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

// Compressed tensor: the octant kx <= L, ky <= M, kz <= N of fN.  Mirrored points pick up the parity
// sign of each component: xy is odd in x and y, xz in x and z, yz in y and z (gperiodic_coeff_nop.m).
template <typename T>
struct TensorC {
  const cplx<T>* f;
  int L, M, N;
  int64_t Pc;
  __device__ __forceinline__ void load6(int kx, int ky, int kz, cplx<T>* out) const {
    const bool fx = kx > L, fy = ky > M, fz = kz > N;
    const int ix = fx ? 2 * L - kx : kx, iy = fy ? 2 * M - ky : ky, iz = fz ? 2 * N - kz : kz;
    const int64_t o = ix + (int64_t)(L + 1) * (iy + (int64_t)(M + 1) * iz);
    out[0] = f[o];
    out[1] = f[Pc + o];
    out[2] = f[2 * Pc + o];
    out[3] = f[3 * Pc + o];
    out[4] = f[4 * Pc + o];
    out[5] = f[5 * Pc + o];
    if (fx != fy) out[1] = -out[1];
    if (fx != fz) out[2] = -out[2];
    if (fy != fz) out[4] = -out[4];
  }
};

// ------------------------------------------------------------------------------------------------
// fused z pass: FFT along z, 3x3 tensor product, inverse FFT along z
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(FFT_THREADS) void k_mv_fzmul(FftDev<T> plan, cplx<T>* S, TensorC<T> fN, int L2,
                                                          int M2, int N, int W, int64_t S_rhs) {
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
      cplx<T> f[6];
      fN.load6(kx0 + w, ky, kz, f);
      cplx<T>* rowp = buf + kz * ld + w;
      const cplx<T> jx = rowp[0], jy = rowp[W], jz = rowp[2 * W];
      rowp[0] = f[0] * jx + f[1] * jy + f[2] * jz;
      rowp[W] = f[1] * jx + f[3] * jy + f[4] * jz;
      rowp[2 * W] = f[2] * jx + f[4] * jy + f[5] * jz;
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



template <typename To>
__global__ void k_compress_tensor(const c128* __restrict__ fN, cplx<To>* __restrict__ fc, int L, int M, int N) {
  const int64_t L1 = L + 1, M1 = M + 1, N1 = N + 1;
  const int64_t Pc = L1 * M1 * N1;
  const int64_t P = 8LL * L * M * N;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * Pc) return;
  const int p = (int)(t / Pc);
  int64_t q = t - p * Pc;
  const int64_t ix = q % L1;
  q /= L1;
  const int64_t iy = q % M1, iz = q / M1;
  fc[t] = ccast<To>(fN[p * P + ix + 2LL * L * (iy + 2LL * M * iz)]);
}

// ------------------------------------------------------------------------------------------------
// compile-time two-pass FFT kernels (lengths in VS_FFT2_SIZES)
// ------------------------------------------------------------------------------------------------
// complex elements of LDS payload per workgroup: 32 KB in both precisions
template <typename T> constexpr int lds_budget() { return sizeof(T) == 8 ? 2048 : 4096; }
constexpr int pow2_floor(int x) { return x >= 2 ? 2 * pow2_floor(x / 2) : 1; }

template <typename T, int N>
struct RowCfg {
  static constexpr int R1 = Split<N>::R1, R2 = Split<N>::R2, TPL = cmax(R1, R2);
  static constexpr int LPW0 = 256 / TPL < lds_budget<T>() / N ? 256 / TPL : lds_budget<T>() / N;
  static constexpr int LPW = LPW0 < 1 ? 1 : LPW0;  // lines per workgroup
  static constexpr int THREADS = TPL * LPW;
};

// lines along the contiguous axis; IO supplies load(line, pos, rhs) / store(line, k, rhs, v)
template <typename T, int N, bool INV, class IO>
__global__ __launch_bounds__(256) void k_row2(const cplx<T>* __restrict__ tw, IO io, int64_t nlines, int n_in,
                                              int n_out) {
  using C = RowCfg<T, N>;
  constexpr int R1 = C::R1, R2 = C::R2, TPL = C::TPL, LPW = C::LPW, LP = LPW + 1;
  __shared__ cplx<T> lds[N * LP];
  const int t = threadIdx.x;
  const int ll = t / TPL, s = t - ll * TPL;
  const int64_t g = (int64_t)blockIdx.x * LPW + ll;
  const int rhs = blockIdx.y;
  const bool live = g < nlines;
  if (s < R1) {
    cplx<T> v[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = s + R1 * n2;
      v[n2] = (live && n < n_in) ? io.load(g, n, rhs) : mk<T>(0, 0);
    }
    fft2_step1<T, N, INV>(v, s, tw);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) lds[(s * R2 + k2) * LP + ll] = v[k2];
  }
  __syncthreads();
  if (s < R2) {
    cplx<T> u[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) u[n1] = lds[(n1 * R2 + s) * LP + ll];
    fft2_step2<T, N, INV>(u);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int k = s + R2 * k1;
      if (live && k < n_out) io.store(g, k, rhs, u[k1]);
    }
  }
}

template <typename T, int N>
struct ColCfg {
  static constexpr int R1 = Split<N>::R1, R2 = Split<N>::R2, TPL = cmax(R1, R2);
  static constexpr int W0 = pow2_floor(256 / TPL < lds_budget<T>() / N ? 256 / TPL : lds_budget<T>() / N);
  static constexpr int W = W0 > 64 ? 64 : W0;  // columns per workgroup
  static constexpr int THREADS = TPL * W;
};

// in-place FFT along a strided axis: element (col, pos, batch) at data[rhs*rhs_stride + batch*sb + pos*sa + col]
template <typename T, int N, bool INV>
__global__ __launch_bounds__(256) void k_col2(const cplx<T>* __restrict__ tw, cplx<T>* data, int64_t fast, int64_t sa,
                                              int64_t sb, int n_in, int n_out, T scale, int64_t rhs_stride) {
  using C = ColCfg<T, N>;
  constexpr int R1 = C::R1, R2 = C::R2, W = C::W;
  __shared__ cplx<T> lds[N * W];
  const int t = threadIdx.x;
  const int col = t % W, s = t / W;
  const int64_t tiles = (fast + W - 1) / W;
  const int64_t tile = blockIdx.x % tiles, b = blockIdx.x / tiles;
  const int64_t c0 = tile * W;
  const bool live = c0 + col < fast;
  cplx<T>* base = data + (int64_t)blockIdx.y * rhs_stride + b * sb + c0 + col;
  if (s < R1) {
    cplx<T> v[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = s + R1 * n2;
      v[n2] = (live && n < n_in) ? base[n * sa] : mk<T>(0, 0);
    }
    fft2_step1<T, N, INV>(v, s, tw);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) lds[(s * R2 + k2) * W + col] = v[k2];
  }
  __syncthreads();
  if (s < R2) {
    cplx<T> u[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) u[n1] = lds[(n1 * R2 + s) * W + col];
    fft2_step2<T, N, INV>(u);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int k = s + R2 * k1;
      if (live && k < n_out) base[k * sa] = scale * u[k1];
    }
  }
}

template <typename T, int N2>
struct FzCfg {
  static constexpr int R1 = Split<N2>::R1, R2 = Split<N2>::R2, TPL = cmax(R1, R2);
  static constexpr int WA = 512 / (3 * TPL), WB = lds_budget<T>() / (3 * N2);
  static constexpr int W0 = pow2_floor(WA < WB ? WA : WB);
  static constexpr int W = W0 > 32 ? 32 : (W0 < 1 ? 1 : W0);  // columns (kx) per workgroup
  static constexpr int THREADS = 3 * TPL * W;
};

// z pass of the matvec: zero-pad z to 2N, FFT, 3x3 symmetric tensor product, inverse FFT, crop to N
template <typename T, int N2>
__global__ __launch_bounds__(512) void k_fz2(const cplx<T>* __restrict__ tw, cplx<T>* S, TensorC<T> fN, int L2, int M2,
                                             int N, int64_t S_rhs) {
  using C = FzCfg<T, N2>;
  constexpr int R1 = C::R1, R2 = C::R2, W = C::W;
  __shared__ cplx<T> xch[3 * N2 * W];
  __shared__ cplx<T> spec[3 * N2 * W];
  const int t = threadIdx.x;
  const int col = t % W, r = t / W;
  const int c = r % 3, s = r / 3;
  const int tiles = (L2 + W - 1) / W;
  const int tile = blockIdx.x % tiles, ky = blockIdx.x / tiles;
  const int kx0 = tile * W;
  const bool live = kx0 + col < L2;
  cplx<T>* Sb = S + (int64_t)blockIdx.y * S_rhs;
  const int64_t zs = (int64_t)L2 * M2, cs = zs * N;
  cplx<T>* gcol = Sb + c * cs + (int64_t)ky * L2 + kx0 + col;
  // forward transform of the three components
  if (s < R1) {
    cplx<T> v[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = s + R1 * n2;
      v[n2] = (live && n < N) ? gcol[n * zs] : mk<T>(0, 0);
    }
    fft2_step1<T, N2, false>(v, s, tw);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) xch[(c * N2 + s * R2 + k2) * W + col] = v[k2];
  }
  __syncthreads();
  if (s < R2) {
    cplx<T> u[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) u[n1] = xch[(c * N2 + n1 * R2 + s) * W + col];
    fft2_step2<T, N2, false>(u);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) spec[(c * N2 + s + R2 * k1) * W + col] = u[k1];
  }
  __syncthreads();
  // tensor product, one (kx, kz) per iteration
  for (int e = t; e < W * N2; e += C::THREADS) {
    const int cc = e % W, kz = e / W;
    if (kx0 + cc < L2) {
      cplx<T> f[6];
      fN.load6(kx0 + cc, ky, kz, f);
      cplx<T>* sp = spec + kz * W + cc;
      const cplx<T> jx = sp[0], jy = sp[N2 * W], jz = sp[2 * N2 * W];
      sp[0] = f[0] * jx + f[1] * jy + f[2] * jz;
      sp[N2 * W] = f[1] * jx + f[3] * jy + f[4] * jz;
      sp[2 * N2 * W] = f[2] * jx + f[4] * jy + f[5] * jz;
    }
  }
  __syncthreads();
  // inverse transform, keep z < N
  if (s < R1) {
    cplx<T> v[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) v[n2] = spec[(c * N2 + s + R1 * n2) * W + col];
    fft2_step1<T, N2, true>(v, s, tw);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) xch[(c * N2 + s * R2 + k2) * W + col] = v[k2];
  }
  __syncthreads();
  if (s < R2) {
    cplx<T> u[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) u[n1] = xch[(c * N2 + n1 * R2 + s) * W + col];
    fft2_step2<T, N2, true>(u);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int k = s + R2 * k1;
      if (live && k < N) gcol[k * zs] = u[k1];
    }
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

template <typename T, bool INV, class IO>
static void run_rows(vs_ctx* ctx, const IO& io, int n, int64_t nlines, int n_in, int n_out, int nrhs) {
  const FftPlanHost& ph = ctx->plan(n);
  const FftDev<T> pd = ph.dev<T>();
  switch (n) {
#define VS_ROW_CASE(NN)                                                                                  \
  case NN: {                                                                                             \
    using Cf = RowCfg<T, NN>;                                                                            \
    dim3 grid((unsigned)((nlines + Cf::LPW - 1) / Cf::LPW), (unsigned)nrhs);                            \
    hipLaunchKernelGGL((k_row2<T, NN, INV, IO>), grid, dim3(Cf::THREADS), 0, ctx->stream, pd.tw, io,   \
                       nlines, n_in, n_out);                                                             \
    VS_HIP(hipGetLastError());                                                                           \
    return;                                                                                              \
  }
    VS_FFT2_SIZES(VS_ROW_CASE)
#undef VS_ROW_CASE
    default: break;
  }
  VS_CHECK(n <= tile_budget(sizeof(cplx<T>)), VS_ERR_UNSUPPORTED, "FFT length exceeds the LDS tile");
  const int nl = row_tile(n, nlines, sizeof(cplx<T>));
  const size_t lds = sizeof(cplx<T>) * (size_t)(nl + 1) * n;
  dim3 grid((unsigned)((nlines + nl - 1) / nl), (unsigned)nrhs);
  hipLaunchKernelGGL((k_row_fft<T, IO>), grid, dim3(FFT_THREADS), lds, ctx->stream, pd, io, nlines, nl, n_in,
                     n_out, INV ? 1 : 0);
  VS_HIP(hipGetLastError());
}

template <typename T, bool INV>
static void run_cols(vs_ctx* ctx, cplx<T>* data, int64_t fast, int64_t sa, int64_t nbatch, int64_t sb, int n_in,
                     int n, int n_out, double scale, int nrhs, int64_t rhs_stride) {
  const FftPlanHost& ph = ctx->plan(n);
  const FftDev<T> pd = ph.dev<T>();
  switch (n) {
#define VS_COL_CASE(NN)                                                                                  \
  case NN: {                                                                                             \
    using Cf = ColCfg<T, NN>;                                                                            \
    const int64_t tiles = (fast + Cf::W - 1) / Cf::W;                                                    \
    dim3 grid((unsigned)(tiles * nbatch), (unsigned)nrhs);                                               \
    hipLaunchKernelGGL((k_col2<T, NN, INV>), grid, dim3(Cf::THREADS), 0, ctx->stream, pd.tw, data, fast, sa, \
                       sb, n_in, n_out, (T)scale, rhs_stride);                                           \
    VS_HIP(hipGetLastError());                                                                           \
    return;                                                                                              \
  }
    VS_FFT2_SIZES(VS_COL_CASE)
#undef VS_COL_CASE
    default: break;
  }
  VS_CHECK(n <= tile_budget(sizeof(cplx<T>)), VS_ERR_UNSUPPORTED, "FFT length exceeds the LDS tile");
  const int W = col_tile(n, fast, nbatch, sizeof(cplx<T>));
  const size_t lds = sizeof(cplx<T>) * (size_t)W * n;
  const int64_t tiles = (fast + W - 1) / W;
  dim3 grid((unsigned)(tiles * nbatch), (unsigned)nrhs);
  hipLaunchKernelGGL((k_col_fft<T>), grid, dim3(FFT_THREADS), lds, ctx->stream, pd, data, fast, sa, sb, W, n_in,
                     n_out, INV ? 1 : 0, (T)scale, rhs_stride);
  VS_HIP(hipGetLastError());
}

template <typename T>
static void run_cols_dir(vs_ctx* ctx, cplx<T>* data, int64_t fast, int64_t sa, int64_t nbatch, int64_t sb, int n_in,
                         int n, int n_out, bool inv, double scale, int nrhs, int64_t rhs_stride) {
  if (inv)
    run_cols<T, true>(ctx, data, fast, sa, nbatch, sb, n_in, n, n_out, scale, nrhs, rhs_stride);
  else
    run_cols<T, false>(ctx, data, fast, sa, nbatch, sb, n_in, n, n_out, scale, nrhs, rhs_stride);
}

void launch_col_fft(vs_ctx* ctx, int prec, void* data, int64_t fast, int64_t axis_stride, int64_t nbatch,
                    int64_t batch_stride, int n_in, int n, int n_out, bool inv, double scale, int nrhs,
                    int64_t rhs_stride) {
  if (prec == VS_FP64)
    run_cols_dir<double>(ctx, (c128*)data, fast, axis_stride, nbatch, batch_stride, n_in, n, n_out, inv, scale,
                         nrhs, rhs_stride);
  else
    run_cols_dir<float>(ctx, (c64*)data, fast, axis_stride, nbatch, batch_stride, n_in, n, n_out, inv, scale, nrhs,
                        rhs_stride);
}

void launch_row_fft_dense(vs_ctx* ctx, int prec, void* data, int64_t nlines, int n_in, int n, int n_out,
                          int64_t ld_in, int64_t ld_out, bool inv, double scale, int nrhs, int64_t rhs_stride) {
  if (prec == VS_FP64) {
    IoDense<double> io{(c128*)data, (c128*)data, ld_in, ld_out, rhs_stride, scale};
    if (inv) run_rows<double, true>(ctx, io, n, nlines, n_in, n_out, nrhs);
    else run_rows<double, false>(ctx, io, n, nlines, n_in, n_out, nrhs);
  } else {
    IoDense<float> io{(c64*)data, (c64*)data, ld_in, ld_out, rhs_stride, (float)scale};
    if (inv) run_rows<float, true>(ctx, io, n, nlines, n_in, n_out, nrhs);
    else run_rows<float, false>(ctx, io, n, nlines, n_in, n_out, nrhs);
  }
}

void launch_row_fft_gather(vs_ctx* ctx, int prec, const void* x, const int* map3, int64_t nunk, void* out, int L,
                           int64_t nlines, int nrhs, int64_t out_rhs_stride) {
  if (prec == VS_FP64) {
    IoGather<double> io{(const c128*)x, map3, nunk, L, (c128*)out, out_rhs_stride};
    run_rows<double, false>(ctx, io, L, nlines, L, L, nrhs);
  } else {
    IoGather<float> io{(const c64*)x, map3, nunk, L, (c64*)out, out_rhs_stride};
    run_rows<float, false>(ctx, io, L, nlines, L, L, nrhs);
  }
}

void launch_row_ifft_scatter(vs_ctx* ctx, int prec, const void* in, const int* map3, int64_t nunk, void* y, int L,
                             int64_t nlines, double scale, int nrhs, int64_t in_rhs_stride) {
  if (prec == VS_FP64) {
    IoScatter<double> io{(const c128*)in, map3, nunk, L, (c128*)y, in_rhs_stride, scale};
    run_rows<double, true>(ctx, io, L, nlines, L, L, nrhs);
  } else {
    IoScatter<float> io{(const c64*)in, map3, nunk, L, (c64*)y, in_rhs_stride, (float)scale};
    run_rows<float, true>(ctx, io, L, nlines, L, L, nrhs);
  }
}

// ------------------------------------------------------------------------------------------------
// matvec driver
// ------------------------------------------------------------------------------------------------
template <typename T>
static bool launch_fz2(vs_ctx* ctx, cplx<T>* S, const TensorC<T>& fN, int L2, int M2, int N, int nrhs,
                       int64_t S_rhs) {
  const int N2 = 2 * N;
  const FftPlanHost& ph = ctx->plan(N2);
  const FftDev<T> pd = ph.dev<T>();
  switch (N2) {
#define VS_FZ_CASE(NN)                                                                                   \
  case NN: {                                                                                             \
    using Cf = FzCfg<T, NN>;                                                                             \
    if (3 * NN * Cf::W > lds_budget<T>()) return false;                                                  \
    const int tiles = (L2 + Cf::W - 1) / Cf::W;                                                          \
    dim3 grid((unsigned)(tiles * M2), (unsigned)nrhs);                                                   \
    hipLaunchKernelGGL((k_fz2<T, NN>), grid, dim3(Cf::THREADS), 0, ctx->stream, pd.tw, S, fN, L2, M2, N, \
                       S_rhs);                                                                           \
    VS_HIP(hipGetLastError());                                                                           \
    return true;                                                                                         \
  }
    VS_FFT2_SIZES(VS_FZ_CASE)
#undef VS_FZ_CASE
    default: return false;
  }
}

template <typename T>
static void mvp_impl(vs_op* op, const cplx<T>* x, cplx<T>* y, int transp, int nrhs) {
  vs_ctx* ctx = op->ctx;
  const int L = op->L, M = op->M, N = op->N;
  const int L2 = 2 * L, M2 = 2 * M, N2 = 2 * N;
  const int64_t S_rhs = (int64_t)L2 * M2 * N * 3;
  op->S.ensure(sizeof(cplx<T>) * S_rhs * nrhs);
  cplx<T>* S = op->S.as<cplx<T>>();
  TensorC<T> fN;
  fN.f = (sizeof(T) == 8) ? op->fc64.as<cplx<T>>() : op->fc32.as<cplx<T>>();
  VS_CHECK(fN.f != nullptr, VS_ERR_ARG, "operator holds no tensor in the requested precision");
  fN.L = L;
  fN.M = M;
  fN.N = N;
  fN.Pc = op->Pc;
  const int* map3 = op->map3.as<int>();
  VS_CHECK(map3 != nullptr, VS_ERR_ARG, "operator domain not set (vs_op_set_domain)");
  const int64_t lines = (int64_t)M * N * 3;
  // 1. gather + x FFT
  IoMvFx<T> fx{x, map3, op->nunk, op->V, L, M, transp, coef_dev<T>(op->b), S, S_rhs};
  run_rows<T, false>(ctx, fx, L2, lines, L, L2, nrhs);
  // 2. y FFT: columns (kx), axis y (stride 2L), batch (z, c) stride 2L*2M, input y < M
  run_cols<T, false>(ctx, S, L2, L2, (int64_t)N * 3, (int64_t)L2 * M2, M, M2, M2, 1.0, nrhs, S_rhs);
  // 3. fused z FFT + tensor product + inverse z FFT
  if (!launch_fz2<T>(ctx, S, fN, L2, M2, N, nrhs, S_rhs)) {
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
    hipLaunchKernelGGL((k_mv_fzmul<T>), grid, dim3(FFT_THREADS), lds, ctx->stream, pd, S, fN, L2, M2, N, W, S_rhs);
    VS_HIP(hipGetLastError());
  }
  // 4. inverse y FFT, keep y < M
  run_cols<T, true>(ctx, S, L2, L2, (int64_t)N * 3, (int64_t)L2 * M2, M2, M2, M, 1.0, nrhs, S_rhs);
  // 5. inverse x FFT, crop, diagonal, scatter
  IoMvIx<T> ix{x, y, map3, op->nunk, op->V, L, M, transp, coef_dev<T>(op->a), coef_dev<T>(op->at),
               coef_dev<T>(op->b), S, S_rhs, (T)(1.0 / (double)op->P)};
  run_rows<T, true>(ctx, ix, L2, lines, L2, L, nrhs);
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

// octant copies of fN64 in the precisions the operator keeps
void op_compress_tensor(vs_op* op) {
  vs_ctx* ctx = op->ctx;
  op->Pc = (int64_t)(op->L + 1) * (op->M + 1) * (op->N + 1);
  const int64_t tot = 6 * op->Pc;
  const unsigned g = (unsigned)((tot + 255) / 256);
  if (op->precisions & VS_FP64) {
    op->fc64.ensure(sizeof(c128) * tot);
    hipLaunchKernelGGL((k_compress_tensor<double>), dim3(g), dim3(256), 0, ctx->stream, op->fN64.as<c128>(),
                       op->fc64.as<c128>(), op->L, op->M, op->N);
    VS_HIP(hipGetLastError());
  }
  if (op->precisions & VS_FP32) {
    op->fc32.ensure(sizeof(c64) * tot);
    hipLaunchKernelGGL((k_compress_tensor<float>), dim3(g), dim3(256), 0, ctx->stream, op->fN64.as<c128>(),
                       op->fc32.as<c64>(), op->L, op->M, op->N);
    VS_HIP(hipGetLastError());
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
  op_compress_tensor(op);
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
