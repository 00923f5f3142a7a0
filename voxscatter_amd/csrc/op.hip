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
//
// Fused slab path (shapes in VS_YZ_SHAPES whose padded y spectrum fits in LDS): the x pass writes the
// volume slab-major, Sx(kx)[c][z][y], and one workgroup per kx slab does the y FFT, the z FFT, the tensor
// product, the inverse z FFT and the inverse y FFT in LDS (k_yz), so HBM sees the slab once in and once
// out plus the octant of the tensor for that kx, stored slab-major as well.
#include "bi_device.hpp"
#include "fft2.hpp"
#include "fft_device.hpp"
#include "handles.hpp"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <type_traits>

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
    if (x < n_in && g < nlines) v = io_finish(io, g, x, rhs, io.load(g, x, rhs));
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
// Every IO functor also offers pre(line, k, rhs): the loads a store depends on (unknown map, solution
// value), issued for all of a thread's outputs before any store so that they are in flight together;
// store(line, k, rhs, v, pre) then only computes and writes.  Loads are branch-free (clamped addresses,
// the result selected afterwards) for the same reason.
struct NoPre {};

template <typename T>
struct IoDense {
  cplx<T>* in;
  cplx<T>* out;
  int64_t ld_in, ld_out, rhs_stride;
  T scale;
  using Pre = NoPre;
  __device__ __forceinline__ cplx<T> load(int64_t g, int x, int rhs) const {
    return in[rhs * rhs_stride + g * ld_in + x];
  }
  __device__ __forceinline__ Pre pre(int64_t, int, int) const { return {}; }
  __device__ __forceinline__ void store(int64_t g, int k, int rhs, cplx<T> v, const Pre& = {}) const {
    out[rhs * rhs_stride + g * ld_out + k] = scale * v;
  }
};

// compressed vector -> dense volume (x FFT of the L x (lines) volume, stored with row pitch L)
// v where ok, else zero, as a multiplication by 1 or 0: a select would let the compiler move the (clamped,
// always valid) load that produced v under a branch on ok, and the loads of consecutive elements would then
// wait on each other instead of being in flight together
template <typename T>
__device__ __forceinline__ cplx<T> masked(cplx<T> v, bool ok) {
  const T k = ok ? (T)1 : (T)0;
  return mk<T>(v.x * k, v.y * k);
}

template <typename T>
struct IoGather {
  const cplx<T>* x;
  const int* map3;
  int64_t nunk;
  int L;
  cplx<T>* out;
  int64_t out_rhs;
  using Pre = NoPre;
  __device__ __forceinline__ cplx<T> load(int64_t g, int xp, int rhs) const {
    const int m = map3[g * L + xp];
    return masked(x[rhs * nunk + (m >= 0 ? m : 0)], m >= 0);
  }
  __device__ __forceinline__ Pre pre(int64_t, int, int) const { return {}; }
  __device__ __forceinline__ void store(int64_t g, int k, int rhs, cplx<T> v, const Pre& = {}) const {
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
  struct Pre { int m; };
  __device__ __forceinline__ cplx<T> load(int64_t g, int xp, int rhs) const {
    return in[rhs * in_rhs + g * L + xp];
  }
  __device__ __forceinline__ Pre pre(int64_t g, int k, int) const { return {map3[g * L + k]}; }
  __device__ __forceinline__ void store(int64_t g, int k, int rhs, cplx<T> v, const Pre& p) const {
    if (p.m >= 0) y[rhs * nunk + p.m] = scale * v;
  }
  __device__ __forceinline__ void store(int64_t g, int k, int rhs, cplx<T> v) const { store(g, k, rhs, v, pre(g, k, rhs)); }
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
  using Pre = NoPre;
  __device__ __forceinline__ cplx<T> load(int64_t g, int xp, int rhs) const {
    const int64_t gi = g * L + xp;
    const int m = map3[gi];
    return masked(x[rhs * nunk + (m >= 0 ? m : 0)], m >= 0);
  }
  // the transp pre-operation Mc conj(J), applied once every input of a thread is loaded (io_finish); the
  // voxel of line g = y + M (z + N c) at x is (g mod MN) L + x
  __device__ __forceinline__ cplx<T> finish(int64_t g, int xp, int, cplx<T> J) const {
    if (!transp) return J;
    const int64_t MN = V / L;
    const int64_t gv = g - MN * ((g >= MN ? 1 : 0) + (g >= 2 * MN ? 1 : 0));
    return mc.at(gv * L + xp) * cconj(J);
  }
  __device__ __forceinline__ Pre pre(int64_t, int, int) const { return {}; }
  __device__ __forceinline__ void store(int64_t g, int k, int rhs, cplx<T> v, const Pre& = {}) const {
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
  const c128* arhs;  // nullable: per-right-hand-side diagonal (Gram Mr) replacing a / at
  struct Pre {
    int m;
    cplx<T> J;
  };
  __device__ __forceinline__ cplx<T> load(int64_t g, int kx, int rhs) const {
    const int64_t yy = g % M, zc = g / M;
    return S[rhs * S_rhs + kx + (int64_t)2 * L * (yy + (int64_t)2 * M * zc)];
  }
  __device__ __forceinline__ Pre pre(int64_t g, int xp, int rhs) const {
    const int m = map3[g * L + xp];
    return {m, x[rhs * nunk + (m >= 0 ? m : 0)]};
  }
  // the output value at an unknown (p.m >= 0) from the cropped inverse transform v
  __device__ __forceinline__ cplx<T> value(int64_t g, int xp, int rhs, cplx<T> v, const Pre& p) const {
    const int64_t vox = (g * L + xp) % V;
    const cplx<T> conv = invP * v;
    if (!transp) return (arhs ? ccast<T>(arhs[rhs]) : a.at(vox)) * p.J - b.at(vox) * conv;
    return (arhs ? ccast<T>(arhs[rhs]) : at.at(vox)) * p.J - cconj(conv);
  }
  __device__ __forceinline__ void store(int64_t g, int xp, int rhs, cplx<T> v, const Pre& p) const {
    if (p.m < 0) return;
    y[rhs * nunk + p.m] = value(g, xp, rhs, v, p);
  }
  __device__ __forceinline__ void store(int64_t g, int xp, int rhs, cplx<T> v) const {
    store(g, xp, rhs, v, pre(g, xp, rhs));
  }
};

// slab-major variants: the x pass stores Sx(kx)[g] with g = y + M (z + N c) the line index, so that a
// kx slab of 3MN values is contiguous (stores staged through LDS: consecutive threads, consecutive lines)
template <typename T>
struct IoMvFxSlab {
  IoMvFx<T> in;
  int64_t SL;  // slab length 3MN
  using Pre = NoPre;
  __device__ __forceinline__ cplx<T> load(int64_t g, int xp, int rhs) const { return in.load(g, xp, rhs); }
  __device__ __forceinline__ Pre pre(int64_t, int, int) const { return {}; }
  __device__ __forceinline__ void store(int64_t g, int k, int rhs, cplx<T> v, const Pre& = {}) const {
    in.S[rhs * in.S_rhs + k * SL + g] = v;
  }
};

// the IO's per-value step after its loads (identity but for the matvec's first pass)
template <class IO, typename T>
__device__ __forceinline__ cplx<T> io_finish(const IO&, int64_t, int, int, cplx<T> v) {
  return v;
}
template <typename T>
__device__ __forceinline__ cplx<T> io_finish(const IoMvFx<T>& io, int64_t g, int xp, int rhs, cplx<T> v) {
  return io.finish(g, xp, rhs, v);
}
template <typename T>
__device__ __forceinline__ cplx<T> io_finish(const IoMvFxSlab<T>& io, int64_t g, int xp, int rhs, cplx<T> v) {
  return io.in.finish(g, xp, rhs, v);
}

template <typename T>
struct IoMvIxSlab {
  IoMvIx<T> out;
  int64_t SL;
  using Pre = typename IoMvIx<T>::Pre;
  __device__ __forceinline__ cplx<T> load(int64_t g, int kx, int rhs) const {
    return out.S[rhs * out.S_rhs + kx * SL + g];
  }
  __device__ __forceinline__ Pre pre(int64_t g, int xp, int rhs) const { return out.pre(g, xp, rhs); }
  __device__ __forceinline__ void store(int64_t g, int xp, int rhs, cplx<T> v, const Pre& p) const {
    out.store(g, xp, rhs, v, p);
  }
  __device__ __forceinline__ void store(int64_t g, int xp, int rhs, cplx<T> v) const { out.store(g, xp, rhs, v); }
};

// As IoMvIxSlab, loading the sum of two slab volumes (the two ky-parity terms of the split slab pass)
template <typename T>
struct IoMvIxSlab2 {
  IoMvIx<T> out;
  int64_t SL;
  const cplx<T>* S2;
  using Pre = typename IoMvIx<T>::Pre;
  __device__ __forceinline__ cplx<T> load(int64_t g, int kx, int rhs) const {
    const int64_t o = rhs * out.S_rhs + kx * SL + g;
    return out.S[o] + S2[o];
  }
  __device__ __forceinline__ Pre pre(int64_t g, int xp, int rhs) const { return out.pre(g, xp, rhs); }
  __device__ __forceinline__ void store(int64_t g, int xp, int rhs, cplx<T> v, const Pre& p) const {
    out.store(g, xp, rhs, v, p);
  }
  __device__ __forceinline__ void store(int64_t g, int xp, int rhs, cplx<T> v) const { out.store(g, xp, rhs, v); }
};

// The last matvec pass of a slab-major variant (Base = IoMvIxSlab or IoMvIxSlab2) with inner products of its
// output y formed on the way, in FP64 per workgroup: KD = 1: conj(w) . y; KD = 2: ||y||^2 and conj(y) . w, w
// read at the same unknowns (one right-hand side).
template <typename T, class Base, int KD>
struct IoDot {
  Base io;
  const cplx<T>* w;
  c128* partial;
  static constexpr int KRED = KD;
  struct Pre {
    typename Base::Pre b;
    cplx<T> w;
  };
  __device__ __forceinline__ cplx<T> load(int64_t g, int kx, int rhs) const { return io.load(g, kx, rhs); }
  __device__ __forceinline__ Pre pre(int64_t g, int xp, int rhs) const {
    const typename Base::Pre b = io.pre(g, xp, rhs);
    return {b, w[rhs * io.out.nunk + (b.m >= 0 ? b.m : 0)]};
  }
  __device__ __forceinline__ void store(int64_t g, int xp, int rhs, cplx<T> v, const Pre& p, double (&re)[KD],
                                        double (&im)[KD]) const {
    if (p.b.m < 0) return;
    const cplx<T> o = io.out.value(g, xp, rhs, v, p.b);
    io.out.y[rhs * io.out.nunk + p.b.m] = o;
    const cplx<T> c = p.w;
    if (KD == 1) {
      re[0] += (double)c.x * o.x + (double)c.y * o.y;
      im[0] += (double)c.x * o.y - (double)c.y * o.x;
    } else {
      re[0] += (double)o.x * o.x + (double)o.y * o.y;
      re[KD - 1] += (double)o.x * c.x + (double)o.y * c.y;
      im[KD - 1] += (double)o.x * c.y - (double)o.y * c.x;
    }
  }
};

// number of FP64 sums an IO accumulates over its stores (0 for the plain IOs)
template <class IO, class = void>
struct RedK {
  static constexpr int K = 0;
};
template <class IO>
struct RedK<IO, std::void_t<decltype(IO::KRED)>> {
  static constexpr int K = IO::KRED;
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

// The same octant stored slab-major: value (ix, iy, iz) of component p at ((ix*6 + p)*(N+1) + iz)*(M+1) + iy,
// so that the tensor of one kx slab is contiguous and consecutive ky are consecutive addresses.
template <typename T>
struct TensorS {
  const cplx<T>* f;
  int L, M, N;
  __device__ __forceinline__ void load6(int kx, int ky, int kz, cplx<T>* out) const {
    const bool fx = kx > L, fy = ky > M, fz = kz > N;
    const int ix = fx ? 2 * L - kx : kx, iy = fy ? 2 * M - ky : ky, iz = fz ? 2 * N - kz : kz;
    const int64_t cs = (int64_t)(N + 1) * (M + 1);
    const cplx<T>* b = f + (int64_t)ix * 6 * cs + (int64_t)iz * (M + 1) + iy;
#pragma unroll
    for (int p = 0; p < 6; ++p) out[p] = b[p * cs];
    if (fx != fy) out[1] = -out[1];
    if (fx != fz) out[2] = -out[2];
    if (fy != fz) out[4] = -out[4];
  }
};

// The octant split by the parity of iy and iz (which equal the parities of ky and kz, 2M and 2N being
// even): value (ix, iy, iz) of component p at ((((ix*6 + p)*2 + iz%2)*Hz + iz/2)*2 + iy%2)*Hy + iy/2, with
// Hy = M/2 + 1, Hz = N/2 + 1.  For one kx slab and one (ky, kz) parity quarter, consecutive ky/2 are
// consecutive addresses.
template <typename T>
struct TensorQ {
  const cplx<T>* f;
  int L, M, N, Hy, Hz;
  __device__ __forceinline__ void load6(int kx, int ky, int kz, cplx<T>* out) const {
    const bool fx = kx > L, fy = ky > M, fz = kz > N;
    const int ix = fx ? 2 * L - kx : kx, iy = fy ? 2 * M - ky : ky, iz = fz ? 2 * N - kz : kz;
    const int64_t ps = (int64_t)2 * Hz * 2 * Hy;  // component stride
    const cplx<T>* b = f + (int64_t)ix * 6 * ps + ((int64_t)((iz & 1) * Hz + (iz >> 1)) * 2 + (iy & 1)) * Hy + (iy >> 1);
#pragma unroll
    for (int p = 0; p < 6; ++p) out[p] = b[p * ps];
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
      rowp[0] = cfma(f[2], jz, cfma(f[1], jy, f[0] * jx));
      rowp[W] = cfma(f[4], jz, cfma(f[3], jy, f[1] * jx));
      rowp[2 * W] = cfma(f[5], jz, cfma(f[4], jy, f[2] * jx));
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

template <typename To>
__global__ void k_compress_tensor_slab(const c128* __restrict__ fN, cplx<To>* __restrict__ fs, int L, int M, int N) {
  const int64_t L1 = L + 1, M1 = M + 1, N1 = N + 1;
  const int64_t P = 8LL * L * M * N;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * L1 * M1 * N1) return;
  // destination order ((ix*6 + p)*N1 + iz)*M1 + iy
  const int64_t iy = t % M1;
  int64_t q = t / M1;
  const int64_t iz = q % N1;
  q /= N1;
  const int p = (int)(q % 6);
  const int64_t ix = q / 6;
  fs[t] = ccast<To>(fN[p * P + ix + 2LL * L * (iy + 2LL * M * iz)]);
}

template <typename To>
__global__ void k_compress_tensor_q(const c128* __restrict__ fN, cplx<To>* __restrict__ fq, int L, int M, int N,
                                    int Hy, int Hz) {
  const int64_t P = 8LL * L * M * N;
  const int64_t tot = (int64_t)(L + 1) * 6 * 2 * Hz * 2 * Hy;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tot) return;
  const int iyh = (int)(t % Hy);
  int64_t q = t / Hy;
  const int py = (int)(q % 2);
  q /= 2;
  const int izh = (int)(q % Hz);
  q /= Hz;
  const int pz = (int)(q % 2);
  q /= 2;
  const int p = (int)(q % 6);
  const int64_t ix = q / 6;
  const int iy = 2 * iyh + py, iz = 2 * izh + pz;
  fq[t] = (iy <= M && iz <= N) ? ccast<To>(fN[p * P + ix + 2LL * L * (iy + 2LL * M * iz)]) : mk<To>(0, 0);
}

// ------------------------------------------------------------------------------------------------
// compile-time two-pass FFT kernels (lengths in VS_FFT2_SIZES)
// ------------------------------------------------------------------------------------------------
// complex elements of LDS payload per workgroup: 32 KB in both precisions
template <typename T> constexpr int lds_budget() { return sizeof(T) == 8 ? 2048 : 4096; }
constexpr int pow2_floor(int x) { return x >= 2 ? 2 * pow2_floor(x / 2) : 1; }

// Step-1 -> step-2 exchange of a line transform, laid out [line][k2][n1] with an odd pitch XP >= R1: step 1
// (thread n1) writes consecutive addresses and step 2 (thread k2) reads at the odd stride XP, so neither
// side has LDS bank conflicts.
template <int R1>
struct Xch {
  static constexpr int XP = R1 % 2 == 0 ? R1 + 1 : R1;
  static __device__ __forceinline__ int at(int line, int R2, int k2, int n1) { return (line * R2 + k2) * XP + n1; }
};

template <typename T, int N>
struct RowCfg {
  static constexpr int R1 = Split<N>::R1, R2 = Split<N>::R2, TPL = cmax(R1, R2);
  static constexpr int LPW0 = 256 / TPL < lds_budget<T>() / N ? 256 / TPL : lds_budget<T>() / N;
  static constexpr int LPW = LPW0 < 1 ? 1 : LPW0;  // lines per workgroup
  static constexpr int THREADS = TPL * LPW;
  static constexpr int LDS = cmax(N * (LPW + 1), LPW * R2 * Xch<R1>::XP);
};

// lines along the contiguous axis; IO supplies load(line, pos, rhs) / store(line, k, rhs, v).
// TIN / TOUT stage the load / store through LDS so that consecutive threads touch consecutive lines at
// one position (for line-contiguous layouts such as the slab-major matvec volume).
// PADIN: the forward transform of a line zero-padded to twice its length (n_in = N / 2); with R1 | N / 2 the
// padding is the inputs n2 >= N / (2 R1) of every thread, known at compile time, so no load is issued for it
template <typename T, int N, bool INV, class IO, bool TIN = false, bool TOUT = false, bool PADIN = false>
__global__ __launch_bounds__(256) void k_row2(const cplx<T>* __restrict__ tw, IO io, int64_t nlines, int n_in,
                                              int n_out) {
  using C = RowCfg<T, N>;
  constexpr int R1 = C::R1, R2 = C::R2, TPL = C::TPL, LPW = C::LPW, LP = LPW + 1;
  using X = Xch<R1>;
  __shared__ cplx<T> lds[C::LDS];
  const int t = threadIdx.x;
  const int ll = t / TPL, s = t - ll * TPL;
  const int64_t g0 = (int64_t)blockIdx.x * LPW;
  const int64_t g = g0 + ll;
  const int rhs = blockIdx.y;
  const bool live = g < nlines;
  constexpr int NI = N / TPL;  // staged elements per thread (THREADS = TPL * LPW, TPL divides N)
  constexpr bool SZ = PADIN && !INV && (N / 2) % R1 == 0;
  if (TIN) {
    cplx<T> st[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int e = t + i * C::THREADS;
      const int l = e % LPW, n = e / LPW;
      const bool ok = g0 + l < nlines && n < n_in;
      st[i] = io.load(ok ? g0 + l : 0, ok ? n : 0, rhs);
      if (!ok) st[i] = mk<T>(0, 0);
    }
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int e = t + i * C::THREADS;
      const int l = e % LPW, n = e / LPW;
      // positions past the input (and lines past the end) are zero: the step runs at position 0 of line 0,
      // in bounds of every per-voxel table
      const bool ok = g0 + l < nlines && n < n_in;
      lds[(e / LPW) * LP + e % LPW] = io_finish(io, ok ? g0 + l : 0, ok ? n : 0, rhs, st[i]);
    }
    __syncthreads();
  }
  cplx<T> v[R2];
  if (s < R1) {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = s + R1 * n2;
      if (TIN) {
        v[n2] = lds[n * LP + ll];
      } else if (SZ && R1 * n2 >= N / 2) {
        v[n2] = mk<T>(0, 0);
      } else {
        const bool ok = live && n < n_in;
        v[n2] = io.load(ok ? g : 0, ok ? n : 0, rhs);
        if (!ok) v[n2] = mk<T>(0, 0);
      }
    }
    if (!TIN) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) {
        if (SZ && R1 * n2 >= N / 2) continue;  // static padding: zero before and after the step
        const int n = s + R1 * n2;
        const bool ok = live && n < n_in;  // zero inputs take the step at an in-bounds position
        v[n2] = io_finish(io, ok ? g : 0, ok ? n : 0, rhs, v[n2]);
      }
    }
    fft2_step1<T, N, INV>(v, s, tw);
  }
  if (TIN) __syncthreads();  // every thread has read its inputs before the exchange overwrites them
  if (s < R1) {
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) lds[X::at(ll, R2, k2, s)] = v[k2];
  }
  __syncthreads();
  cplx<T> u[R1];
  if (s < R2) {
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) u[n1] = lds[X::at(ll, R2, s, n1)];
    fft2_step2<T, N, INV>(u);
    if (TOUT) {
      // staged below, once every thread has read the exchange
    } else {
      // the loads the stores depend on, PB outputs at a time (bounded register use); outputs grow with k0, so
      // once a batch's first one is cropped (k >= n_out: the padded inverse keeps half) no later one is
      // stored and their loads are not issued
      constexpr int PB = sizeof(T) == 8 ? 4 : 8;
#pragma unroll
      for (int k0 = 0; k0 < R1; k0 += PB) {
        if (s + R2 * k0 >= n_out) break;
        typename IO::Pre pr[PB];
#pragma unroll
        for (int j = 0; j < PB; ++j) {
          const int k = s + R2 * (k0 + j);
          const bool ok = k0 + j < R1 && live && k < n_out;
          pr[j] = io.pre(ok ? g : 0, ok ? k : 0, rhs);
        }
#pragma unroll
        for (int j = 0; j < PB; ++j) {
          const int k = s + R2 * (k0 + j);
          if (k0 + j < R1 && live && k < n_out) io.store(g, k, rhs, u[k0 + j], pr[j]);
        }
      }
    }
  }
  if (TOUT) {
    __syncthreads();
    if (s < R2) {
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) lds[(s + R2 * k1) * LP + ll] = u[k1];
    }
    __syncthreads();
    constexpr int PB = sizeof(T) == 8 ? 4 : 8;
#pragma unroll
    for (int i0 = 0; i0 < NI; i0 += PB) {
      typename IO::Pre pr[PB];
#pragma unroll
      for (int j = 0; j < PB; ++j) {
        const int e = t + (i0 + j) * C::THREADS;
        const int l = e % LPW, k = e / LPW;
        const bool ok = i0 + j < NI && g0 + l < nlines && k < n_out;
        pr[j] = io.pre(ok ? g0 + l : 0, ok ? k : 0, rhs);
      }
#pragma unroll
      for (int j = 0; j < PB; ++j) {
        const int e = t + (i0 + j) * C::THREADS;
        const int l = e % LPW, k = e / LPW;
        if (i0 + j < NI && g0 + l < nlines && k < n_out) io.store(g0 + l, k, rhs, lds[k * LP + l], pr[j]);
      }
    }
  }
}


// Zero-padded forward / cropped inverse row transform of length 2n as two length-n transforms split by the
// parity of the frequency index: X(2j) = DFT_n(x)(j), X(2j+1) = DFT_n(x w_2n^x)(j), and the inverse cropped
// to x < n, y(x) = IDFT_n(X_even)(x) + conj(w_2n^x) IDFT_n(X_odd)(x).  One thread group per (line, parity):
// twice the threads of the length-2n kernel for the same lines, and no work on the padding.  Forward: the n
// inputs through io.load, the 2n outputs staged in LDS and stored line-contiguous (slab-major volumes);
// inverse: the 2n inputs staged line-contiguous, the n outputs combined and stored through io.store.
template <typename T, int NH>
struct RowPCfg {
  static constexpr int R1 = Split<NH>::R1, R2 = Split<NH>::R2, TPL = cmax(R1, R2);
  static constexpr int LPW = sizeof(T) == 8 ? 8 : 16;  // lines per workgroup
  static constexpr int HL = 2 * LPW;                   // (line, parity) groups
  static constexpr int THREADS = TPL * HL;
  static constexpr int LP = LPW + 1, HLP = HL + 1;
  // staging [k][line], outputs [x][group] or the exchange [group][k2][n1]
  static constexpr int LDS = cmax(cmax(2 * NH * LP, NH * HLP), HL * R2 * Xch<R1>::XP);
};

template <typename T, int NH, bool INV, class IO>
__global__ __launch_bounds__(512) void k_row2p(const cplx<T>* __restrict__ tw, const cplx<T>* __restrict__ tw2, IO io,
                                               int64_t nlines) {
  using C = RowPCfg<T, NH>;
  constexpr int R1 = C::R1, R2 = C::R2, TPL = C::TPL, LPW = C::LPW, LP = C::LP, HLP = C::HLP, TH = C::THREADS;
  constexpr int N2 = 2 * NH;
  constexpr int PB = sizeof(T) == 8 ? 4 : 8;
  __shared__ cplx<T> lds[C::LDS];
  const int t = threadIdx.x;
  const int hl = t / TPL, s = t - hl * TPL;
  const int l = hl >> 1, par = hl & 1;
  const int64_t g0 = (int64_t)blockIdx.x * LPW;
  const int rhs = blockIdx.y;
  cplx<T> v[R2];
  // inverse: the loads the stores depend on (unknown map, input value, inner-product partner) are issued
  // first, so that their round trips overlap the staging loads and the transform
  constexpr int NOI = (NH * LPW + TH - 1) / TH;
  typename IO::Pre pri[INV ? NOI : 1];
  if constexpr (INV) {
#pragma unroll
    for (int j = 0; j < NOI; ++j) {
      const int e = t + j * TH;
      const int x = e % NH, ll = e / NH;
      const bool ok = e < NH * LPW && g0 + ll < nlines;
      pri[j] = io.pre(ok ? g0 + ll : 0, ok ? x : 0, rhs);
    }
  }
  if (INV) {
    constexpr int NI = (N2 * LPW + TH - 1) / TH;
    cplx<T> st[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int e = t + i * TH;
      const int ll = e % LPW, k = e / LPW;
      const bool ok = e < N2 * LPW && g0 + ll < nlines;
      st[i] = io.load(ok ? g0 + ll : 0, ok ? k : 0, rhs);
      if (!ok) st[i] = mk<T>(0, 0);
    }
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int e = t + i * TH;
      if (e < N2 * LPW) lds[(e / LPW) * LP + e % LPW] = st[i];
    }
    __syncthreads();
    if (s < R1) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) v[n2] = lds[(2 * (s + R1 * n2) + par) * LP + l];
      fft2_step1<T, NH, true>(v, s, tw);
    }
    __syncthreads();
  } else if (s < R1) {
    const bool live = g0 + l < nlines;
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = s + R1 * n2;
      v[n2] = io.load(live ? g0 + l : 0, live ? n : 0, rhs);
      if (!live) v[n2] = mk<T>(0, 0);
    }
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = s + R1 * n2;
      const cplx<T> x = io_finish(io, live ? g0 + l : 0, n, rhs, v[n2]);
      v[n2] = par ? x * tw2[n] : x;
    }
    fft2_step1<T, NH, false>(v, s, tw);
  }
  // exchange [n1 R2 + k2][group]
  if (s < R1) {
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) lds[Xch<R1>::at(hl, R2, k2, s)] = v[k2];
  }
  __syncthreads();
  cplx<T> u[R1];
  if (s < R2) {
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) u[n1] = lds[Xch<R1>::at(hl, R2, s, n1)];
    fft2_step2<T, NH, INV>(u);
  }
  __syncthreads();  // the exchange buffer is reused below
  if constexpr (!INV) {
    // X(2 j + par), j = s + R2 k1, staged at [kx][line]
    if (s < R2) {
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) lds[(2 * (s + R2 * k1) + par) * LP + l] = u[k1];
    }
    __syncthreads();
    constexpr int NO = (N2 * LPW + TH - 1) / TH;
#pragma unroll
    for (int i0 = 0; i0 < NO; i0 += PB) {
      typename IO::Pre pr[PB];
#pragma unroll
      for (int j = 0; j < PB; ++j) {
        const int e = t + (i0 + j) * TH;
        const int ll = e % LPW, k = e / LPW;
        const bool ok = i0 + j < NO && e < N2 * LPW && g0 + ll < nlines;
        pr[j] = io.pre(ok ? g0 + ll : 0, ok ? k : 0, rhs);
      }
#pragma unroll
      for (int j = 0; j < PB; ++j) {
        const int e = t + (i0 + j) * TH;
        const int ll = e % LPW, k = e / LPW;
        if (i0 + j < NO && e < N2 * LPW && g0 + ll < nlines) io.store(g0 + ll, k, rhs, lds[k * LP + ll], pr[j]);
      }
    }
  } else {
    // y_par(x), x = s + R2 k1, at [x][group]; then y = y_even + conj(w_2n^x) y_odd, x fastest over threads
    if (s < R2) {
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) lds[(s + R2 * k1) * HLP + hl] = u[k1];
    }
    __syncthreads();
    constexpr int KR = RedK<IO>::K;
    double rre[KR > 0 ? KR : 1] = {}, rim[KR > 0 ? KR : 1] = {};
#pragma unroll
    for (int j = 0; j < NOI; ++j) {
      const int e = t + j * TH;
      const int x = e % NH, ll = e / NH;
      if (e < NH * LPW && g0 + ll < nlines) {
        const cplx<T> y = lds[x * HLP + 2 * ll] + cconj(tw2[x]) * lds[x * HLP + 2 * ll + 1];
        if constexpr (KR > 0)
          io.store(g0 + ll, x, rhs, y, pri[j], rre, rim);
        else
          io.store(g0 + ll, x, rhs, y, pri[j]);
      }
    }
    // per-workgroup partials of the IO's sums (one right-hand side per launch)
    if constexpr (KR > 0) block_partials<TH, KR>(rre, rim, io.partial + (int64_t)blockIdx.x * KR);
  }
}
// Preconditioner output -> matvec input in one pass over the lines g = (y, z, c) of length NH = L: the
// inverse x transform of the preconditioner's volume U (pitch NH) scaled by `scale` and scattered into the
// compressed vector y (the preconditioned vector, as k_row2 with IoScatter), and the same values, zero off
// the domain, zero-padded to 2 NH and transformed forward into the slab-major matvec volume S (as k_row2p
// with IoMvFxSlab); one partial of ||y||^2 per workgroup.  Phase A: group hl < LPW transforms line g0 + hl;
// phase B: group (line, parity) as in k_row2p.
template <typename T, int NH>
__global__ __launch_bounds__(512) void k_row_pcmv(const cplx<T>* __restrict__ tw, const cplx<T>* __restrict__ tw2,
                                                  const cplx<T>* __restrict__ U, const int* __restrict__ map3, T scale,
                                                  cplx<T>* __restrict__ y, cplx<T>* __restrict__ S, int64_t SL,
                                                  int64_t nlines, c128* __restrict__ partial) {
  using C = RowPCfg<T, NH>;
  constexpr int R1 = C::R1, R2 = C::R2, TPL = C::TPL, LPW = C::LPW, LP = C::LP, HLP = C::HLP, TH = C::THREADS;
  constexpr int N2 = 2 * NH;
  __shared__ cplx<T> lds[C::LDS];
  const int t = threadIdx.x;
  const int hl = t / TPL, s = t - hl * TPL;
  const int64_t g0 = (int64_t)blockIdx.x * LPW;
  double acc[1] = {0.0}, acci[1] = {0.0};
  const bool la = hl < LPW;
  const bool lva = la && g0 + hl < nlines;
  const int64_t ga = lva ? g0 + hl : 0;
  // the unknown map of this thread's phase-A outputs, requested before the transform
  int mm[R1];
  if (la && s < R2) {
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) mm[k1] = lva ? map3[ga * NH + s + R2 * k1] : -1;
  }
  if (la && s < R1) {
    cplx<T> v[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = s + R1 * n2;
      v[n2] = U[ga * NH + n];
      if (!lva) v[n2] = mk<T>(0, 0);
    }
    fft2_step1<T, NH, true>(v, s, tw);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) lds[Xch<R1>::at(hl, R2, k2, s)] = v[k2];
  }
  __syncthreads();
  {
    cplx<T> u[R1];
    if (la && s < R2) {
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) u[n1] = lds[Xch<R1>::at(hl, R2, s, n1)];
      fft2_step2<T, NH, true>(u);
    }
    __syncthreads();  // the exchange area is reused by the staging below
    if (la && s < R2) {
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        const int x = s + R2 * k1;
        cplx<T> val = scale * u[k1];
        if (mm[k1] >= 0) {
          y[mm[k1]] = val;
          acc[0] += (double)val.x * val.x + (double)val.y * val.y;
        } else {
          val = mk<T>(0, 0);
        }
        lds[x * LP + hl] = val;
      }
    }
  }
  __syncthreads();
  const int l = hl >> 1, par = hl & 1;
  cplx<T> v[R2];
  if (s < R1) {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = s + R1 * n2;
      const cplx<T> xv = lds[n * LP + l];
      v[n2] = par ? xv * tw2[n] : xv;
    }
  }
  __syncthreads();  // the staging is read before the exchange overwrites it
  if (s < R1) {
    fft2_step1<T, NH, false>(v, s, tw);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) lds[Xch<R1>::at(hl, R2, k2, s)] = v[k2];
  }
  __syncthreads();
  cplx<T> u[R1];
  if (s < R2) {
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) u[n1] = lds[Xch<R1>::at(hl, R2, s, n1)];
    fft2_step2<T, NH, false>(u);
  }
  __syncthreads();
  if (s < R2) {
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) lds[(2 * (s + R2 * k1) + par) * LP + l] = u[k1];
  }
  __syncthreads();
  constexpr int NO = (N2 * LPW + TH - 1) / TH;
#pragma unroll
  for (int i = 0; i < NO; ++i) {
    const int e = t + i * TH;
    const int ll = e % LPW, k = e / LPW;
    if (e < N2 * LPW && g0 + ll < nlines) S[k * SL + g0 + ll] = lds[k * LP + ll];
  }
  block_partials<TH, 1>(acc, acci, partial + blockIdx.x);
}

// The preconditioner's first pass (gather + x transform of length N = L into the volume U, pitch L, as
// k_row2 with IoGather) on a vector formed on the way from the Krylov vectors at each unknown, with the
// scalars of the stage read from the device as the separate update kernels read them:
//   MODE 0 (start of the iteration): p = r + beta (p0 - omega v0) (p = r at the first), written to w0;
//          a0 = r, a1 = p0, a2 = v0; nothing is written after a breakdown
//   MODE 1 (half step): xh = x + alpha ph -> w0, s = r - alpha v -> w1 (transformed); a0 = x, a1 = ph, a2 = r,
//          a3 = v; partials ||s||^2 and ||xh||^2 per workgroup (nothing written once the flag is set)
template <typename T>
struct BiRowArgs {
  const int* map3;
  cplx<T>* U;
  const cplx<T>* a0;
  const cplx<T>* a1;
  const cplx<T>* a2;
  const cplx<T>* a3;
  cplx<T>* w0;
  cplx<T>* w1;
  const BiScal* sc;
  c128* partial;
};

template <typename T, int N, int MODE>
__global__ __launch_bounds__(256) void k_row_bi(const cplx<T>* __restrict__ tw, BiRowArgs<T> a, int64_t nlines) {
  using C = RowCfg<T, N>;
  constexpr int R1 = C::R1, R2 = C::R2, TPL = C::TPL, LPW = C::LPW;
  __shared__ cplx<T> lds[C::LDS];
  const int t = threadIdx.x;
  const int ll = t / TPL, s = t - ll * TPL;
  const int64_t g = (int64_t)blockIdx.x * LPW + ll;
  const bool live = g < nlines;
  const BiScal& S = *a.sc;
  bool ok = S.flag == 0;
  bool first = false;
  cplx<T> c0 = mk<T>(0, 0), c1 = mk<T>(0, 0);
  if (MODE == 0) {
    c128 rho, beta;
    if (ok && bi_rho_beta(S, rho, beta)) ok = false;
    first = S.first != 0;
    c0 = ccast<T>(beta);
    c1 = ccast<T>(S.omega);
  } else {
    c0 = ccast<T>(S.alpha);
  }
  double re[2] = {0.0, 0.0}, im[2] = {0.0, 0.0};
  cplx<T> v[R2];
  if (s < R1) {
    int mm[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) mm[n2] = live ? a.map3[g * N + s + R1 * n2] : -1;
    cplx<T> q0[R2], q1[R2], q2[R2], q3[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int m = mm[n2] >= 0 ? mm[n2] : 0;
      q0[n2] = a.a0[m];
      q1[n2] = a.a1[m];
      q2[n2] = a.a2[m];
      if (MODE == 1) q3[n2] = a.a3[m];
    }
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const bool on = ok && mm[n2] >= 0;
      cplx<T> val;
      if (MODE == 0) {
        val = first ? q0[n2] : q0[n2] + c0 * (q1[n2] - c1 * q2[n2]);
        if (on) a.w0[mm[n2]] = val;
      } else {
        const cplx<T> xh = q0[n2] + c0 * q1[n2];
        val = q2[n2] - c0 * q3[n2];
        if (on) {
          a.w0[mm[n2]] = xh;
          a.w1[mm[n2]] = val;
          re[0] += (double)val.x * val.x + (double)val.y * val.y;
          re[1] += (double)xh.x * xh.x + (double)xh.y * xh.y;
        }
      }
      v[n2] = on ? val : mk<T>(0, 0);
    }
    fft2_step1<T, N, false>(v, s, tw);
  }
  if (s < R1) {
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) lds[Xch<R1>::at(ll, R2, k2, s)] = v[k2];
  }
  __syncthreads();
  if (s < R2) {
    cplx<T> u[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) u[n1] = lds[Xch<R1>::at(ll, R2, s, n1)];
    fft2_step2<T, N, false>(u);
    if (live) {
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) a.U[g * N + s + R2 * k1] = u[k1];
    }
  }
  if (MODE == 1) block_partials<C::THREADS, 2>(re, im, a.partial + (int64_t)blockIdx.x * 2);
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

// columns (kx) per workgroup of the z pass: as many as 1024 threads and 80 KB of LDS (two workgroups per
// CU) allow, at most 32; the z lines of a column tile are read and written in runs of W consecutive kx, so
// long lines (2N = 128 ... 256) still get runs of 128 bytes (W = 16 in FP32, 8 in FP64)
template <typename T>
constexpr int fz_lds_elems() { return 80 * 1024 / (int)sizeof(cplx<T>); }
template <typename T, int N2>
struct FzCfg {
  static constexpr int R1 = Split<N2>::R1, R2 = Split<N2>::R2, TPL = cmax(R1, R2);
  static constexpr int WA = 1024 / (3 * TPL), WB = fz_lds_elems<T>() / (3 * N2);
  static constexpr int W0 = pow2_floor(WA < WB ? WA : WB);
  static constexpr int W = W0 > 32 ? 32 : (W0 < 1 ? 1 : W0);
  static constexpr int THREADS = 3 * TPL * W;
  // waves per SIMD of two resident workgroups (the LDS allows two): the register cap of the launch bound
  static constexpr int MINW = THREADS / 128 > 0 ? THREADS / 128 : 1;
};

// z pass of the matvec: zero-pad z to 2N, FFT, 3x3 symmetric tensor product, inverse FFT, crop to N.
// One LDS buffer holds the three components' lines; each stage reads all of its inputs before a barrier
// and writes after it.  The tensor stage takes kz and 2N - kz together: they share one octant entry (the
// xz and yz components change sign), so each entry is read once per (kx, ky) instead of twice.
template <typename T, int N2>
__global__ __launch_bounds__((FzCfg<T, N2>::THREADS), (FzCfg<T, N2>::MINW)) void k_fz2(const cplx<T>* __restrict__ tw, cplx<T>* S,
                                                                TensorC<T> fN, int L2, int M2, int N, int64_t S_rhs) {
  using C = FzCfg<T, N2>;
  constexpr int R1 = C::R1, R2 = C::R2, W = C::W, NH = N2 / 2;
  __shared__ cplx<T> buf[3 * N2 * W];
  cplx<T>* xch = buf;
  cplx<T>* spec = buf;
  const int t = threadIdx.x;
  const int col = t % W, r = t / W;
  const int c = r % 3, s = r / 3;
  const int tiles = (L2 + W - 1) / W;
  // logical ids: the tiles of ky, then those of its mirror 2M - ky (same tensor rows), ky in the order
  // 0, M, 1, 2M-1, 2, 2M-2, ...; laid out XCD by XCD (workgroup i runs on XCD i mod 8) so that one XCD's
  // L2 serves all tiles of a mirror pair
  const int G = tiles * M2;
  const int per_xcd = (G + 7) / 8;
  const int lid = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
  if (lid >= G) return;
  const int tile = lid % tiles, qy = lid / tiles;
  const int My = M2 / 2;
  const int ky = qy == 0 ? 0 : (qy == 1 ? My : ((qy & 1) == 0 ? 1 + (qy - 2) / 2 : M2 - 1 - (qy - 3) / 2));
  const int kx0 = tile * W;
  const bool live = kx0 + col < L2;
  cplx<T>* Sb = S + (int64_t)blockIdx.y * S_rhs;
  // element offsets within one right-hand side's volume (2L x 2M x N x 3 < 2^31) in 32 bits; N = N2 / 2,
  // so with R1 | N the padding (n >= N) is the inputs n2 >= N / R1 of every thread, known at compile time
  const unsigned zs = (unsigned)L2 * (unsigned)M2, cs = zs * (unsigned)NH;
  cplx<T>* gcol = Sb + (c * cs + (unsigned)ky * (unsigned)L2 + (unsigned)(kx0 + col));
  constexpr bool RZ = NH % R1 == 0;
  // FP32: the tensor entries of this thread's first product points (e = t, t + THREADS below: two for the
  // long lines, where every thread has at least two, one otherwise) are requested here, so that their round
  // trips overlap the forward transforms (12 registers each, within the two workgroups per CU that the LDS
  // allows)
  constexpr int NPF = sizeof(T) == 4 ? (N2 >= 192 ? 2 : 1) : 0;
  cplx<T> f0[NPF > 0 ? NPF : 1][6];
#pragma unroll
  for (int i = 0; i < NPF; ++i) {
    const int e = min(t + i * C::THREADS, W * (NH + 1) - 1);
    fN.load6(min(kx0 + e % W, L2 - 1), ky, e / W, f0[i]);
  }
  // forward transform of the three components
  if (s < R1) {
    cplx<T> v[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      const int n = s + R1 * n2;
      if (RZ && R1 * n2 >= NH)
        v[n2] = mk<T>(0, 0);
      else
        v[n2] = (live && n < NH) ? gcol[(unsigned)n * zs] : mk<T>(0, 0);
    }
    fft2_step1<T, N2, false>(v, s, tw);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) xch[(c * N2 + s * R2 + k2) * W + col] = v[k2];
  }
  __syncthreads();
  cplx<T> u[R1];
  if (s < R2) {
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) u[n1] = xch[(c * N2 + n1 * R2 + s) * W + col];
    fft2_step2<T, N2, false>(u);
  }
  __syncthreads();
  if (s < R2) {
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) spec[(c * N2 + s + R2 * k1) * W + col] = u[k1];
  }
  __syncthreads();
  // tensor product, one (kx, kz <= N) per iteration, applied at kz and at 2N - kz
  auto product = [&](int cc, int kz, const cplx<T> (&f)[6]) {
    cplx<T>* sp = spec + kz * W + cc;
    cplx<T> jx = sp[0], jy = sp[N2 * W], jz = sp[2 * N2 * W];
    sp[0] = cfma(f[2], jz, cfma(f[1], jy, f[0] * jx));
    sp[N2 * W] = cfma(f[4], jz, cfma(f[3], jy, f[1] * jx));
    sp[2 * N2 * W] = cfma(f[5], jz, cfma(f[4], jy, f[2] * jx));
    if (kz != 0 && kz != NH) {
      sp = spec + (N2 - kz) * W + cc;
      jx = sp[0];
      jy = sp[N2 * W];
      jz = sp[2 * N2 * W];
      sp[0] = cfma(-f[2], jz, cfma(f[1], jy, f[0] * jx));
      sp[N2 * W] = cfma(-f[4], jz, cfma(f[3], jy, f[1] * jx));
      sp[2 * N2 * W] = cfma(f[5], jz, cfma(-f[4], jy, (-f[2]) * jx));
    }
  };
  int e = t;
#pragma unroll
  for (int i = 0; i < NPF; ++i, e += C::THREADS)
    if (e < W * (NH + 1) && kx0 + e % W < L2) product(e % W, e / W, f0[i]);
  for (; e < W * (NH + 1); e += C::THREADS) {
    const int cc = e % W, kz = e / W;
    if (kx0 + cc < L2) {
      cplx<T> f[6];
      fN.load6(kx0 + cc, ky, kz, f);
      product(cc, kz, f);
    }
  }
  __syncthreads();
  // inverse transform, keep z < N
  {
    cplx<T> v[R2];
    if (s < R1) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) v[n2] = spec[(c * N2 + s + R1 * n2) * W + col];
      fft2_step1<T, N2, true>(v, s, tw);
    }
    __syncthreads();
    if (s < R1) {
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) xch[(c * N2 + s * R2 + k2) * W + col] = v[k2];
    }
  }
  __syncthreads();
  if (s < R2) {
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) u[n1] = xch[(c * N2 + n1 * R2 + s) * W + col];
    fft2_step2<T, N2, true>(u);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      const int k = s + R2 * k1;
      if ((NH % R2 == 0 && R2 * k1 >= NH) || !(live && k < NH)) continue;
      gcol[(unsigned)k * zs] = u[k1];
    }
  }
}

// ------------------------------------------------------------------------------------------------
// fused y-z slab pass: one workgroup per kx slab, everything between the two x passes in LDS
// ------------------------------------------------------------------------------------------------
// (2M, 2N) pairs with a compiled fused kernel; a shape runs fused when its LDS footprint also fits
#define VS_YZ_SHAPES(X) X(256, 32) X(128, 32) X(128, 16) X(64, 16) X(64, 64) X(48, 48) X(96, 96) X(192, 48)

constexpr int VS_LDS_MAX = 160 * 1024;

template <typename T, int M2, int N2>
struct YzCfg {
  static constexpr int M = M2 / 2, N = N2 / 2, NL = 3 * N;
  static constexpr int NLP = NL + 1;  // LDS pitch of one ky row of the spectrum [ky][c][z]
  static constexpr int RY1 = Split<M2>::R1, RY2 = Split<M2>::R2;
  static constexpr int RZ1 = Split<N2>::R1, RZ2 = Split<N2>::R2, TPZ = cmax(RZ1, RZ2);
  static constexpr int THREADS = 1024;
  static constexpr int IY1 = (NL * RY1 + THREADS - 1) / THREADS;  // y-FFT step-1 items per thread
  static constexpr int IY2 = (NL * RY2 + THREADS - 1) / THREADS;
  static constexpr int SPEC = M2 * NLP;
  static constexpr int INB = NL * (M + 1);  // input staging, overlapping the spectrum
  static constexpr int SPECMAX = SPEC > INB ? SPEC : INB;
  static constexpr int fit_wz(int w) {
    return (w > 1 && (int64_t)sizeof(cplx<T>) * (SPECMAX + 6 * N2 * w) > VS_LDS_MAX) ? fit_wz(w / 2) : w;
  }
  static constexpr int WZ0 = pow2_floor(THREADS / (3 * TPZ));
  static constexpr int WZ = fit_wz(WZ0 > M2 ? M2 : WZ0);  // ky columns per z pass
  static constexpr int ZB = 3 * N2 * WZ;
  static constexpr size_t BYTES = sizeof(cplx<T>) * (size_t)(SPECMAX + 2 * ZB);
};

template <typename T, int M2, int N2>
__global__ __launch_bounds__(1024) void k_yz(const cplx<T>* __restrict__ twy, const cplx<T>* __restrict__ twz,
                                             cplx<T>* S, TensorS<T> fN, int64_t S_rhs) {
  using C = YzCfg<T, M2, N2>;
  constexpr int M = C::M, N = C::N, NL = C::NL, NLP = C::NLP, RY1 = C::RY1, RY2 = C::RY2;
  constexpr int RZ1 = C::RZ1, RZ2 = C::RZ2, TPZ = C::TPZ, WZ = C::WZ, TH = C::THREADS;
  cplx<T>* sp = reinterpret_cast<cplx<T>*>(vs_smem);
  cplx<T>* xch = sp + C::SPECMAX;
  cplx<T>* spec = xch + C::ZB;
  const int t = threadIdx.x;
  const int L = fN.L, L2 = 2 * L;
  // consecutive logical ids are the mirror slabs kx, 2L - kx (same tensor octant); hardware dispatch
  // sends workgroup i to XCD i mod 8, so logical ids are laid out XCD by XCD to share that XCD's L2
  const int per_xcd = (L2 + 7) / 8;
  const int q = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
  if (q >= L2) return;
  const int kx = q == 0 ? 0 : (q == 1 ? L : ((q & 1) == 0 ? 1 + (q - 2) / 2 : L2 - 1 - (q - 3) / 2));
  cplx<T>* g = S + (int64_t)blockIdx.y * S_rhs + (int64_t)kx * NL * M;

  // 0. slab -> LDS, line-major with pitch M + 1 (all of a thread's loads in flight together)
  {
    constexpr int I0 = (NL * M + TH - 1) / TH;
    cplx<T> st[I0];
#pragma unroll
    for (int i = 0; i < I0; ++i) {
      const int e = t + i * TH;
      st[i] = g[e < NL * M ? e : 0];
    }
#pragma unroll
    for (int i = 0; i < I0; ++i) {
      const int e = t + i * TH;
      if (e < NL * M) sp[(e / M) * (M + 1) + e % M] = st[i];
    }
  }
  __syncthreads();
  // 1. y FFT of the zero-padded lines, step 1 (results held in registers across the barrier)
  cplx<T> v[C::IY1][RY2];
#pragma unroll
  for (int i = 0; i < C::IY1; ++i) {
    const int it = t + i * TH;
    if (it < NL * RY1) {
      const int line = it % NL, n1 = it / NL;
#pragma unroll
      for (int n2 = 0; n2 < RY2; ++n2) {
        const int n = n1 + RY1 * n2;
        v[i][n2] = n < M ? sp[line * (M + 1) + n] : mk<T>(0, 0);
      }
      fft2_step1<T, M2, false>(v[i], n1, twy);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < C::IY1; ++i) {
    const int it = t + i * TH;
    if (it < NL * RY1) {
      const int line = it % NL, n1 = it / NL;
#pragma unroll
      for (int k2 = 0; k2 < RY2; ++k2) sp[(n1 * RY2 + k2) * NLP + line] = v[i][k2];
    }
  }
  __syncthreads();
  // 2. step 2, in place (each item reads and writes the same locations)
#pragma unroll
  for (int i = 0; i < C::IY2; ++i) {
    const int it = t + i * TH;
    if (it < NL * RY2) {
      const int line = it % NL, k2 = it / NL;
      cplx<T> u[RY1];
#pragma unroll
      for (int n1 = 0; n1 < RY1; ++n1) u[n1] = sp[(n1 * RY2 + k2) * NLP + line];
      fft2_step2<T, M2, false>(u);
#pragma unroll
      for (int k1 = 0; k1 < RY1; ++k1) sp[(k2 + RY2 * k1) * NLP + line] = u[k1];
    }
  }
  __syncthreads();
  // 3. z stage on WZ ky columns at a time: zero-pad z to 2N, FFT, tensor product, inverse FFT, crop.
  // The tensor values of the next pass are loaded while the current one runs.
  {
    const int col = t % WZ, r = t / WZ;
    const int c = r % 3, sz = r / 3;
    const bool act = t < 3 * TPZ * WZ;
    constexpr int IZ = (WZ * N2 + TH - 1) / TH;
    cplx<T> fnext[IZ][6];
    auto load_f = [&](int ky0) {
#pragma unroll
      for (int j = 0; j < IZ; ++j) {
        const int e = t + j * TH;
        const int cc = e % WZ, kz = e / WZ;
        const int ky = min(ky0 + cc, M2 - 1), kzc = min(kz, N2 - 1);
        fN.load6(kx, ky, kzc, fnext[j]);
      }
    };
    load_f(0);
    for (int ky0 = 0; ky0 < M2; ky0 += WZ) {
      cplx<T> fcur[IZ][6];
#pragma unroll
      for (int j = 0; j < IZ; ++j)
#pragma unroll
        for (int p = 0; p < 6; ++p) fcur[j][p] = fnext[j][p];
      if (ky0 + WZ < M2) load_f(ky0 + WZ);
      const int ky = ky0 + col;
      const bool kyv = ky < M2;
      if (act && sz < RZ1) {
        cplx<T> vz[RZ2];
#pragma unroll
        for (int n2 = 0; n2 < RZ2; ++n2) {
          const int n = sz + RZ1 * n2;
          vz[n2] = (kyv && n < N) ? sp[ky * NLP + c * N + n] : mk<T>(0, 0);
        }
        fft2_step1<T, N2, false>(vz, sz, twz);
#pragma unroll
        for (int k2 = 0; k2 < RZ2; ++k2) xch[(c * N2 + sz * RZ2 + k2) * WZ + col] = vz[k2];
      }
      __syncthreads();
      if (act && sz < RZ2) {
        cplx<T> u[RZ1];
#pragma unroll
        for (int n1 = 0; n1 < RZ1; ++n1) u[n1] = xch[(c * N2 + n1 * RZ2 + sz) * WZ + col];
        fft2_step2<T, N2, false>(u);
#pragma unroll
        for (int k1 = 0; k1 < RZ1; ++k1) spec[(c * N2 + sz + RZ2 * k1) * WZ + col] = u[k1];
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < IZ; ++j) {
        const int e = t + j * TH;
        const int cc = e % WZ, kz = e / WZ;
        if (e < WZ * N2 && ky0 + cc < M2) {
          const cplx<T>* f = fcur[j];
          cplx<T>* s3 = spec + kz * WZ + cc;
          const cplx<T> jx = s3[0], jy = s3[N2 * WZ], jz = s3[2 * N2 * WZ];
          s3[0] = cfma(f[2], jz, cfma(f[1], jy, f[0] * jx));
          s3[N2 * WZ] = cfma(f[4], jz, cfma(f[3], jy, f[1] * jx));
          s3[2 * N2 * WZ] = cfma(f[5], jz, cfma(f[4], jy, f[2] * jx));
        }
      }
      __syncthreads();
      if (act && sz < RZ1) {
        cplx<T> vz[RZ2];
#pragma unroll
        for (int n2 = 0; n2 < RZ2; ++n2) vz[n2] = spec[(c * N2 + sz + RZ1 * n2) * WZ + col];
        fft2_step1<T, N2, true>(vz, sz, twz);
#pragma unroll
        for (int k2 = 0; k2 < RZ2; ++k2) xch[(c * N2 + sz * RZ2 + k2) * WZ + col] = vz[k2];
      }
      __syncthreads();
      if (act && sz < RZ2) {
        cplx<T> u[RZ1];
#pragma unroll
        for (int n1 = 0; n1 < RZ1; ++n1) u[n1] = xch[(c * N2 + n1 * RZ2 + sz) * WZ + col];
        fft2_step2<T, N2, true>(u);
#pragma unroll
        for (int k1 = 0; k1 < RZ1; ++k1) {
          const int k = sz + RZ2 * k1;
          if (kyv && k < N) sp[ky * NLP + c * N + k] = u[k1];
        }
      }
      __syncthreads();
    }
  }
  // 4. inverse y FFT, step 1 over all 2M spectral values
#pragma unroll
  for (int i = 0; i < C::IY1; ++i) {
    const int it = t + i * TH;
    if (it < NL * RY1) {
      const int line = it % NL, n1 = it / NL;
#pragma unroll
      for (int n2 = 0; n2 < RY2; ++n2) v[i][n2] = sp[(n1 + RY1 * n2) * NLP + line];
      fft2_step1<T, M2, true>(v[i], n1, twy);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < C::IY1; ++i) {
    const int it = t + i * TH;
    if (it < NL * RY1) {
      const int line = it % NL, n1 = it / NL;
#pragma unroll
      for (int k2 = 0; k2 < RY2; ++k2) sp[(n1 * RY2 + k2) * NLP + line] = v[i][k2];
    }
  }
  __syncthreads();
  // 5. step 2, keep y < M (in place)
#pragma unroll
  for (int i = 0; i < C::IY2; ++i) {
    const int it = t + i * TH;
    if (it < NL * RY2) {
      const int line = it % NL, k2 = it / NL;
      cplx<T> u[RY1];
#pragma unroll
      for (int n1 = 0; n1 < RY1; ++n1) u[n1] = sp[(n1 * RY2 + k2) * NLP + line];
      fft2_step2<T, M2, true>(u);
#pragma unroll
      for (int k1 = 0; k1 < RY1; ++k1) {
        const int y = k2 + RY2 * k1;
        if (y < M) sp[y * NLP + line] = u[k1];
      }
    }
  }
  __syncthreads();
  // 6. LDS -> slab
  for (int e = t; e < NL * M; e += TH) {
    const int line = e / M, y = e - line * M;
    g[e] = sp[y * NLP + line];
  }
}

// ------------------------------------------------------------------------------------------------
// parity-split slab pass: the zero-padded transforms as pairs of unpadded ones
// ------------------------------------------------------------------------------------------------
// With x zero beyond M, the length-2M DFT splits by the parity of the output index:
//   X(2j)   = DFT_M(x)(j),   X(2j+1) = DFT_M(x . w_2M^y)(j)                      (w_n = exp(-2 pi i/n))
// and the inverse length-2M DFT cropped to y < M is
//   y(y)    = IDFT_M(X_even)(y) + conj(w_2M^y) IDFT_M(X_odd)(y)
// (the same along z with N).  The tensor product is pointwise in (ky, kz), so one workgroup runs the two
// ky parities of its kx slab one after the other with an unpadded spectrum of M x 3N values (half the
// padded one), keeping the first half's result in registers; the z stage is one thread per (ky, c) with
// both kz parities in registers (length-N register butterflies).
// HALF: the two ky parities of a slab run in two workgroups (twice the workgroups, two resident per CU
// where one slab's workgroup alone left the CU half idle); each writes its term of the cropped inverse
// (IDFT_M(X_even), resp. conj(w_2M^y) IDFT_M(X_odd)) to a buffer of its own, and the inverse x pass adds
// the two when it loads.
#define VS_YZQ_SHAPES(X) X(128, 16) X(64, 16) X(128, 8) X(64, 8) X(32, 8) X(32, 16) X(16, 8) X(16, 4) X(64, 64)

template <typename T, int M, int N>
struct YzqCfg {
  static constexpr int NL = 3 * N, NLP = NL + 1;
  static constexpr int RY1 = Split<M>::R1, RY2 = Split<M>::R2;
  // long z columns (N > 16, the spectrum of a slab alone fills most of the LDS, one workgroup per CU): 1024
  // lanes, four wavefronts per SIMD, and half the registers per lane for the spectrum copies
  static constexpr int TH = N > 16 ? 1024 : 512;
  static constexpr int IY1 = (NL * RY1 + TH - 1) / TH, IY2 = (NL * RY2 + TH - 1) / TH;
  static constexpr int IS = (NL * M + TH - 1) / TH;  // slab values per thread
  static constexpr int IZ = (M * N + TH - 1) / TH;   // tensor-product points per thread
  static constexpr int SP = M * NLP > NL * (M + 1) ? M * NLP : NL * (M + 1);
  static constexpr int TW = M + 2 * M + 2 * N;       // twiddles of M, 2M, 2N points, after the spectrum
  static constexpr size_t BYTES = sizeof(cplx<T>) * (size_t)(SP + TW);
  static_assert(3 * M <= TH, "one thread per (ky, component) in the z stage");
};

// Latency structure: every global load of a phase is issued before the first of them is used, and the
// twiddle tables live in LDS; in FP32 the tensor of a kz parity is requested before the z transforms it
// waits for (the next parity's right after the product), so a parity costs three global round trips (FP64
// loads the tensor at the product: its registers do not allow the prefetch).
// an index the compiler must treat as unknown, so that address arithmetic derived from it is redone where
// it is used instead of being hoisted and held in registers across the whole kernel
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <typename T, int M, int N, bool HALF>
__global__ __launch_bounds__((YzqCfg<T, M, N>::TH), (HALF && sizeof(T) == 4 && N <= 16 ? 4 : 2)) void k_yzq(const cplx<T>* __restrict__ twm, const cplx<T>* __restrict__ tw2m,
                                             const cplx<T>* __restrict__ tw2n, cplx<T>* S, TensorQ<T> fN,
                                             int64_t S_rhs, cplx<T>* __restrict__ Se, cplx<T>* __restrict__ So) {
  using C = YzqCfg<T, M, N>;
  constexpr int NL = C::NL, NLP = C::NLP, RY1 = C::RY1, RY2 = C::RY2, TH = C::TH, IZ = C::IZ;
  cplx<T>* sp = reinterpret_cast<cplx<T>*>(vs_smem);
  cplx<T>* lm = sp + C::SP;  // DFT_M twiddles
  cplx<T>* l2m = lm + M;     // w_2M^y
  cplx<T>* l2n = l2m + 2 * M;  // w_2N^z
  const int t = threadIdx.x;
  const int L = fN.L, L2 = 2 * L;
  // mirror slabs kx, 2L - kx on consecutive logical ids (HALF: both parities of each), laid out XCD by XCD
  // (as in k_yz)
  const int nlog = HALF ? 2 * L2 : L2;
  const int per_xcd = (nlog + 7) / 8;
  const int lid = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
  if (lid >= nlog) return;
  const int q = HALF ? lid >> 1 : lid;
  const int kx = q == 0 ? 0 : (q == 1 ? L : ((q & 1) == 0 ? 1 + (q - 2) / 2 : L2 - 1 - (q - 3) / 2));
  cplx<T>* g = S + (int64_t)blockIdx.y * S_rhs + (int64_t)kx * NL * M;
  cplx<T> acc[HALF ? 1 : C::IS];  // result of the even-ky half, element t + i TH of the slab
  const int py0 = HALF ? (lid & 1) : 0, py1 = HALF ? py0 + 1 : 2;
  constexpr bool PREF = sizeof(T) == 4 && !HALF;
  for (int e = t; e < C::TW; e += TH) lm[e] = e < M ? twm[e] : (e < 3 * M ? tw2m[e - M] : tw2n[e - 3 * M]);
  __syncthreads();

  // tensor of this thread's product points for kz parity pz
  auto tload = [&](int py, int pz, cplx<T> (&f)[IZ][6]) {
#pragma unroll
    for (int i = 0; i < IZ; ++i) {
      const int e = min(opaque(t + i * TH), M * N - 1);
      const int jj = e % M, m = e / M;
      fN.load6(kx, 2 * jj + py, 2 * m + pz, f[i]);
    }
  };
  // load-and-apply point by point (FP64)
  auto tmul = [&](int py, int pz) {
#pragma unroll
    for (int i = 0; i < IZ; ++i) {
      const int e = opaque(t + i * TH);
      if (e < M * N) {
        const int jj = e % M, m = e / M;
        cplx<T> f[6];
        fN.load6(kx, 2 * jj + py, 2 * m + pz, f);
        cplx<T>* s3 = sp + jj * NLP + m;
        const cplx<T> jx = s3[0], jy = s3[N], jz = s3[2 * N];
        s3[0] = cfma(f[2], jz, cfma(f[1], jy, f[0] * jx));
        s3[N] = cfma(f[4], jz, cfma(f[3], jy, f[1] * jx));
        s3[2 * N] = cfma(f[5], jz, cfma(f[4], jy, f[2] * jx));
      }
    }
  };
  auto tapply = [&](const cplx<T> (&f)[IZ][6]) {
#pragma unroll
    for (int i = 0; i < IZ; ++i) {
      const int e = opaque(t + i * TH);
      if (e < M * N) {
        const int jj = e % M, m = e / M;
        cplx<T>* s3 = sp + jj * NLP + m;
        const cplx<T> jx = s3[0], jy = s3[N], jz = s3[2 * N];
        s3[0] = cfma(f[i][2], jz, cfma(f[i][1], jy, f[i][0] * jx));
        s3[N] = cfma(f[i][4], jz, cfma(f[i][3], jy, f[i][1] * jx));
        s3[2 * N] = cfma(f[i][5], jz, cfma(f[i][4], jy, f[i][2] * jx));
      }
    }
  };

  for (int py = py0; py < py1; ++py) {
    // 0. slab -> LDS [line][y] (pitch M + 1), twisted by w_2M^y for the odd half
    {
      cplx<T> st[C::IS];
#pragma unroll
      for (int i = 0; i < C::IS; ++i) st[i] = g[min(opaque(t + i * TH), NL * M - 1)];
#pragma unroll
      for (int i = 0; i < C::IS; ++i) {
        const int e = opaque(t + i * TH);
        if (e < NL * M) {
          const int line = e / M, y = e - line * M;
          sp[line * (M + 1) + y] = py ? st[i] * l2m[y] : st[i];
        }
      }
    }
    __syncthreads();
    // 1. DFT_M along y, step 1 (results in registers across the barrier)
    cplx<T> v[C::IY1][RY2];
#pragma unroll
    for (int i = 0; i < C::IY1; ++i) {
      const int it = t + i * TH;
      if (it < NL * RY1) {
        const int line = it % NL, n1 = it / NL;
#pragma unroll
        for (int n2 = 0; n2 < RY2; ++n2) v[i][n2] = sp[line * (M + 1) + n1 + RY1 * n2];
        fft2_step1<T, M, false>(v[i], n1, lm);
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < C::IY1; ++i) {
      const int it = t + i * TH;
      if (it < NL * RY1) {
        const int line = it % NL, n1 = it / NL;
#pragma unroll
        for (int k2 = 0; k2 < RY2; ++k2) sp[(n1 * RY2 + k2) * NLP + line] = v[i][k2];
      }
    }
    __syncthreads();
    // 2. step 2 in place -> spectrum [j][line], j = ky / 2
#pragma unroll
    for (int i = 0; i < C::IY2; ++i) {
      const int it = t + i * TH;
      if (it < NL * RY2) {
        const int line = it % NL, k2 = it / NL;
        cplx<T> u[RY1];
#pragma unroll
        for (int n1 = 0; n1 < RY1; ++n1) u[n1] = sp[(n1 * RY2 + k2) * NLP + line];
        fft2_step2<T, M, false>(u);
#pragma unroll
        for (int k1 = 0; k1 < RY1; ++k1) sp[(k2 + RY2 * k1) * NLP + line] = u[k1];
      }
    }
    __syncthreads();
    // 3. z stage: thread (j, c) transforms its column both ways; the tensor product runs per kz parity
    if constexpr (N <= 16) {
      const bool zt = t < 3 * M;
      const int j = t % M, c = t / M;
      cplx<T>* col = sp + j * NLP + c * N;
      cplx<T> a[N], o[N];
      {
        cplx<T> f[PREF ? IZ : 1][6];
        if constexpr (PREF) tload(py, 0, f);
        if (zt) {
#pragma unroll
          for (int z = 0; z < N; ++z) {
            a[z] = col[z];
            o[z] = a[z] * l2n[z];
          }
          Bfly<T, N>::run(a, false);
          Bfly<T, N>::run(o, false);
#pragma unroll
          for (int m = 0; m < N; ++m) col[m] = a[m];  // even kz
        }
        __syncthreads();
        if constexpr (PREF)
          tapply(f);
        else
          tmul(py, 0);
      }
      {
        cplx<T> f[PREF ? IZ : 1][6];
        if constexpr (PREF) tload(py, 1, f);
        __syncthreads();
        if (zt) {
#pragma unroll
          for (int m = 0; m < N; ++m) a[m] = col[m];
          Bfly<T, N>::run(a, true);
#pragma unroll
          for (int m = 0; m < N; ++m) col[m] = o[m];  // odd kz
        }
        __syncthreads();
        if constexpr (PREF)
          tapply(f);
        else
          tmul(py, 1);
      }
      __syncthreads();
      if (zt) {
#pragma unroll
        for (int m = 0; m < N; ++m) o[m] = col[m];
        Bfly<T, N>::run(o, true);
#pragma unroll
        for (int z = 0; z < N; ++z) col[z] = a[z] + cconj(l2n[z]) * o[z];
      }
      __syncthreads();
    }
    else {
      // long z columns (N = M): the 3M columns of N values stay in LDS and are transformed there (two-step
      // DFT_N), each kz parity in turn; every thread keeps its IZS values of the untransformed spectrum (for
      // the odd parity's twisted input) and of the first parity's result in registers.  Value e = (3 j + c) N + z sits at
      // sp[j NLP + c N + z] = sp[e + e / (3 N)].
      static_assert(N == M, "the DFT_N twiddles are the DFT_M table");
      constexpr int NC = 3 * M, RZ1 = Split<N>::R1, RZ2 = Split<N>::R2;
      constexpr int IZS = (NC * N + TH - 1) / TH;
      cplx<T> orig[IZS], res[IZS];
#pragma unroll
      for (int i = 0; i < IZS; ++i) {
        const int e = min(t + i * TH, NC * N - 1);
        orig[i] = sp[e + e / (3 * N)];
      }
      __syncthreads();  // every copy is taken before the transforms overwrite the spectrum
      // DFT_N of every column in place: the TPC lanes of a column are consecutive lanes of one wavefront,
      // so its step-1 reads precede the exchange writes and those precede the step-2 reads in program
      // order (no barrier); the columns go in rounds of TH / TPC
      constexpr int TPC = cmax(RZ1, RZ2);
      static_assert(64 % TPC == 0, "a column's lanes within one wavefront");
      auto zdft = [&](auto inv_tag) {
        constexpr bool INV = decltype(inv_tag)::value;
#pragma unroll 1
        for (int base = 0; base < NC * TPC; base += TH) {
          const int it = base + t;
          const int col = min(it / TPC, NC - 1), s = it % TPC;
          const bool live = it < NC * TPC;
          cplx<T>* cz = sp + (col / 3) * NLP + (col % 3) * N;
          cplx<T> v[RZ2];
          if (live && s < RZ1) {
#pragma unroll
            for (int n2 = 0; n2 < RZ2; ++n2) v[n2] = cz[s + RZ1 * n2];
            fft2_step1<T, N, INV>(v, s, lm);
#pragma unroll
            for (int k2 = 0; k2 < RZ2; ++k2) cz[s * RZ2 + k2] = v[k2];
          }
          cplx<T> u[RZ1];
          if (live && s < RZ2) {
#pragma unroll
            for (int n1 = 0; n1 < RZ1; ++n1) u[n1] = cz[n1 * RZ2 + s];
            fft2_step2<T, N, INV>(u);
#pragma unroll
            for (int k1 = 0; k1 < RZ1; ++k1) cz[s + RZ2 * k1] = u[k1];
          }
        }
        __syncthreads();
      };
#pragma unroll 1
      for (int pz = 0; pz < 2; ++pz) {
        if (pz == 1) {
#pragma unroll
          for (int i = 0; i < IZS; ++i) {
            const int e = t + i * TH;
            if (e < NC * N) sp[e + e / (3 * N)] = orig[i] * l2n[e % N];
          }
          __syncthreads();
        }
        zdft(std::false_type{});
        tmul(py, pz);
        __syncthreads();
        zdft(std::true_type{});
#pragma unroll
        for (int i = 0; i < IZS; ++i) {
          const int e = t + i * TH;
          if (e < NC * N) {
            const cplx<T> r = sp[e + e / (3 * N)];
            res[i] = pz == 0 ? r : res[i] + cconj(l2n[e % N]) * r;
          }
        }
        __syncthreads();
      }
#pragma unroll
      for (int i = 0; i < IZS; ++i) {
        const int e = t + i * TH;
        if (e < NC * N) sp[e + e / (3 * N)] = res[i];
      }
      __syncthreads();
    }
    // 4. inverse DFT_M along y, step 1
#pragma unroll
    for (int i = 0; i < C::IY1; ++i) {
      const int it = t + i * TH;
      if (it < NL * RY1) {
        const int line = it % NL, n1 = it / NL;
#pragma unroll
        for (int n2 = 0; n2 < RY2; ++n2) v[i][n2] = sp[(n1 + RY1 * n2) * NLP + line];
        fft2_step1<T, M, true>(v[i], n1, lm);
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < C::IY1; ++i) {
      const int it = t + i * TH;
      if (it < NL * RY1) {
        const int line = it % NL, n1 = it / NL;
#pragma unroll
        for (int k2 = 0; k2 < RY2; ++k2) sp[(n1 * RY2 + k2) * NLP + line] = v[i][k2];
      }
    }
    __syncthreads();
    // 5. step 2 in place -> [y][line]
#pragma unroll
    for (int i = 0; i < C::IY2; ++i) {
      const int it = t + i * TH;
      if (it < NL * RY2) {
        const int line = it % NL, k2 = it / NL;
        cplx<T> u[RY1];
#pragma unroll
        for (int n1 = 0; n1 < RY1; ++n1) u[n1] = sp[(n1 * RY2 + k2) * NLP + line];
        fft2_step2<T, M, true>(u);
#pragma unroll
        for (int k1 = 0; k1 < RY1; ++k1) sp[(k2 + RY2 * k1) * NLP + line] = u[k1];
      }
    }
    __syncthreads();
    // 6. even half: keep in registers; odd half: combine and write the slab.  HALF: write this parity's term
    if (HALF) {
      cplx<T>* o = (py ? So : Se) + (int64_t)blockIdx.y * S_rhs + (int64_t)kx * NL * M;
#pragma unroll
      for (int i = 0; i < C::IS; ++i) {
        const int e = opaque(t + i * TH);
        if (e < NL * M) {
          const int line = e / M, y = e - line * M;
          const cplx<T> r = sp[y * NLP + line];
          o[e] = py ? cconj(l2m[y]) * r : r;
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < C::IS; ++i) {
        const int e = opaque(t + i * TH);
        if (e < NL * M) {
          const int line = e / M, y = e - line * M;
          const cplx<T> r = sp[y * NLP + line];
          if (py == 0)
            acc[i] = r;
          else
            g[e] = acc[i] + cconj(l2m[y]) * r;
        }
      }
    }
    __syncthreads();
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

template <typename T, bool INV, class IO, bool TIN = false, bool TOUT = false, bool PADIN = false>
static void run_rows(vs_ctx* ctx, const IO& io, int n, int64_t nlines, int n_in, int n_out, int nrhs) {
  VS_CHECK(!PADIN || 2 * n_in == n, VS_ERR_INTERNAL, "padded row FFT takes half-length inputs");
  const FftPlanHost& ph = ctx->plan(n);
  const FftDev<T> pd = ph.dev<T>();
  switch (n) {
#define VS_ROW_CASE(NN)                                                                                  \
  case NN: {                                                                                             \
    using Cf = RowCfg<T, NN>;                                                                            \
    dim3 grid((unsigned)((nlines + Cf::LPW - 1) / Cf::LPW), (unsigned)nrhs);                            \
    hipLaunchKernelGGL((k_row2<T, NN, INV, IO, TIN, TOUT, PADIN>), grid, dim3(Cf::THREADS), 0,           \
                       ctx->stream, pd.tw, io, nlines, n_in, n_out);                                     \
    VS_HIP(hipGetLastError());                                                                           \
    return;                                                                                              \
  }
    VS_FFT2_SIZES(VS_ROW_CASE)
#undef VS_ROW_CASE
    default: break;
  }
  VS_CHECK(!TIN && !TOUT, VS_ERR_INTERNAL, "line-transposed row FFT needs a compiled length");
  VS_CHECK(n <= tile_budget(sizeof(cplx<T>)), VS_ERR_UNSUPPORTED, "FFT length exceeds the LDS tile");
  const int nl = row_tile(n, nlines, sizeof(cplx<T>));
  const size_t lds = sizeof(cplx<T>) * (size_t)(nl + 1) * n;
  dim3 grid((unsigned)((nlines + nl - 1) / nl), (unsigned)nrhs);
  hipLaunchKernelGGL((k_row_fft<T, IO>), grid, dim3(FFT_THREADS), lds, ctx->stream, pd, io, nlines, nl, n_in,
                     n_out, INV ? 1 : 0);
  VS_HIP(hipGetLastError());
}

// the parity-split row transform of length 2 nh (k_row2p) when nh has a compiled split; false otherwise
// (VS_ROWP=0 turns it off)
template <typename T, bool INV, class IO>
static bool run_rows_p(vs_ctx* ctx, const IO& io, int nh, int64_t nlines, int nrhs) {
  static const bool off = [] {
    const char* e = std::getenv("VS_ROWP");
    return e && e[0] == '0';
  }();
  if (off || nh < 8) return false;
  const cplx<T>* tw = ctx->plan(nh).dev<T>().tw;
  const cplx<T>* tw2 = ctx->plan(2 * nh).dev<T>().tw;
  switch (nh) {
#define VS_ROWP_CASE(NN)                                                                                 \
  case NN: {                                                                                             \
    using Cf = RowPCfg<T, NN>;                                                                           \
    if constexpr (Cf::LDS * sizeof(cplx<T>) <= 65536 && Cf::THREADS <= 512) {                            \
      dim3 grid((unsigned)((nlines + Cf::LPW - 1) / Cf::LPW), (unsigned)nrhs);                          \
      hipLaunchKernelGGL((k_row2p<T, NN, INV, IO>), grid, dim3(Cf::THREADS), 0, ctx->stream, tw, tw2, io, \
                         nlines);                                                                        \
      VS_HIP(hipGetLastError());                                                                         \
      return true;                                                                                       \
    }                                                                                                    \
    return false;                                                                                        \
  }
    VS_FFT2_SIZES(VS_ROWP_CASE)
#undef VS_ROWP_CASE
    default: return false;
  }
}

static bool fft2_length(int n) {
  switch (n) {
#define VS_LEN_CASE(NN) case NN: return true;
    VS_FFT2_SIZES(VS_LEN_CASE)
#undef VS_LEN_CASE
    default: return false;
  }
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
    if (3 * NN * Cf::W > fz_lds_elems<T>()) return false;                                                \
    const int tiles = (L2 + Cf::W - 1) / Cf::W;                                                          \
    dim3 grid((unsigned)(((tiles * M2 + 7) / 8) * 8), (unsigned)nrhs);                                   \
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

// the fused slab kernel for (2M, 2N), if compiled and within the LDS limit (and 2L has a compiled x FFT)
template <typename T>
static size_t yz_bytes(int M2, int N2) {
  switch (M2 * 4096 + N2) {
#define VS_YZ_BYTES(A, B) case A * 4096 + B: return YzCfg<T, A, B>::BYTES;
    VS_YZ_SHAPES(VS_YZ_BYTES)
#undef VS_YZ_BYTES
    default: return 0;
  }
}

template <typename T>
static size_t yzq_bytes(int M, int N) {
  switch (M * 4096 + N) {
#define VS_YZQ_BYTES(A, B) case A * 4096 + B: return YzqCfg<T, A, B>::BYTES;
    VS_YZQ_SHAPES(VS_YZQ_BYTES)
#undef VS_YZQ_BYTES
    default: return 0;
  }
}

template <typename T>
bool yzq_available(int L, int M, int N) {
  const size_t b = yzq_bytes<T>(M, N);
  return b > 0 && b <= (size_t)VS_LDS_MAX && fft2_length(2 * L);
}

template <typename T>
bool yz_fused_available(int L, int M, int N) {
  const size_t b = yz_bytes<T>(2 * M, 2 * N);
  return b > 0 && b <= (size_t)VS_LDS_MAX && fft2_length(2 * L);
}
template bool yz_fused_available<float>(int, int, int);
template bool yz_fused_available<double>(int, int, int);

// matvec variants: 0 = five passes, 1 = padded-slab fused kernel, 2 = parity-split fused kernel, 3 = the
// parity-split kernel with the two ky parities in separate workgroups.  VS_MATVEC=pass5|slab|split|split2
// forces one (when the operator holds its tensor layout); VS_FUSED=0 is pass5.  Otherwise the variant
// timed fastest for the operator at its first domain is used.
static int forced_variant() {
  const char* f = std::getenv("VS_FUSED");
  if (f && f[0] == '0') return 0;
  const char* e = std::getenv("VS_MATVEC");
  if (!e) return -1;
  if (!std::strcmp(e, "pass5")) return 0;
  if (!std::strcmp(e, "slab")) return 1;
  if (!std::strcmp(e, "split")) return 2;
  if (!std::strcmp(e, "split2")) return 3;
  return -1;
}

template <typename T>
static bool variant_available(const vs_op* op, int v) {
  if (v == 0) return true;
  if (v == 1) return (sizeof(T) == 8 ? op->fs64.p : op->fs32.p) != nullptr;
  return (sizeof(T) == 8 ? op->fq64.p : op->fq32.p) != nullptr;
}

template <typename T>
static int choose_variant(const vs_op* op) {
  const int f = forced_variant();
  if (f >= 0 && variant_available<T>(op, f)) return f;
  const int tuned = op->mv_variant[sizeof(T) == 8 ? 0 : 1];
  if (tuned >= 0 && variant_available<T>(op, tuned)) return tuned;
  for (int v = 3; v >= 1; --v)
    if (variant_available<T>(op, v)) return v;
  return 0;
}

// raise the dynamic-LDS limit of the fused kernel for (2M, 2N) above the 64 KB default (done when the
// operator's slab tensor is built, never inside a stream capture)
template <typename T>
static void yz_set_attr(int M2, int N2) {
  switch (M2 * 4096 + N2) {
#define VS_YZ_ATTR(A, B)                                                                                     \
  case A * 4096 + B:                                                                                         \
    VS_HIP(hipFuncSetAttribute((const void*)k_yz<T, A, B>, hipFuncAttributeMaxDynamicSharedMemorySize,       \
                               (int)YzCfg<T, A, B>::BYTES));                                                 \
    return;
    VS_YZ_SHAPES(VS_YZ_ATTR)
#undef VS_YZ_ATTR
    default: return;
  }
}

template <typename T>
static void yzq_set_attr(int M, int N) {
  switch (M * 4096 + N) {
#define VS_YZQ_ATTR(A, B)                                                                                    \
  case A * 4096 + B:                                                                                         \
    VS_HIP(hipFuncSetAttribute((const void*)k_yzq<T, A, B, false>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                               (int)YzqCfg<T, A, B>::BYTES));                                                \
    VS_HIP(hipFuncSetAttribute((const void*)k_yzq<T, A, B, true>, hipFuncAttributeMaxDynamicSharedMemorySize,  \
                               (int)YzqCfg<T, A, B>::BYTES));                                                \
    return;
    VS_YZQ_SHAPES(VS_YZQ_ATTR)
#undef VS_YZQ_ATTR
    default: return;
  }
}

// Se == nullptr: one workgroup per slab, result in place; otherwise one per (slab, ky parity), the two
// terms to Se and So (same layout and right-hand-side stride as S)
template <typename T>
static void launch_yzq(vs_ctx* ctx, cplx<T>* S, const TensorQ<T>& fN, int L, int M, int N, int nrhs, int64_t S_rhs,
                       cplx<T>* Se = nullptr, cplx<T>* So = nullptr) {
  const cplx<T>* twm = ctx->plan(M).dev<T>().tw;
  const cplx<T>* tw2m = ctx->plan(2 * M).dev<T>().tw;
  const cplx<T>* tw2n = ctx->plan(2 * N).dev<T>().tw;
  const bool half = Se != nullptr;
  const int per_xcd = ((half ? 4 : 2) * L + 7) / 8;
  dim3 grid((unsigned)(per_xcd * 8), (unsigned)nrhs);
  switch (M * 4096 + N) {
#define VS_YZQ_CASE(A, B)                                                                                  \
  case A * 4096 + B: {                                                                                     \
    using Cf = YzqCfg<T, A, B>;                                                                            \
    if (half)                                                                                              \
      hipLaunchKernelGGL((k_yzq<T, A, B, true>), grid, dim3(Cf::TH), Cf::BYTES, ctx->stream, twm, tw2m,    \
                         tw2n, S, fN, S_rhs, Se, So);                                                      \
    else                                                                                                   \
      hipLaunchKernelGGL((k_yzq<T, A, B, false>), grid, dim3(Cf::TH), Cf::BYTES, ctx->stream, twm, tw2m,   \
                         tw2n, S, fN, S_rhs, Se, So);                                                      \
    VS_HIP(hipGetLastError());                                                                             \
    return;                                                                                                \
  }
    VS_YZQ_SHAPES(VS_YZQ_CASE)
#undef VS_YZQ_CASE
    default: throw Error{VS_ERR_INTERNAL, "no parity-split slab kernel for this shape"};
  }
}

template <typename T>
static void launch_yz(vs_ctx* ctx, cplx<T>* S, const TensorS<T>& fN, int L, int M, int N, int nrhs, int64_t S_rhs) {
  const int M2 = 2 * M, N2 = 2 * N;
  const cplx<T>* twy = ctx->plan(M2).dev<T>().tw;
  const cplx<T>* twz = ctx->plan(N2).dev<T>().tw;
  const int per_xcd = (2 * L + 7) / 8;
  dim3 grid((unsigned)(per_xcd * 8), (unsigned)nrhs);
  switch (M2 * 4096 + N2) {
#define VS_YZ_CASE(A, B)                                                                                   \
  case A * 4096 + B: {                                                                                     \
    using Cf = YzCfg<T, A, B>;                                                                             \
    hipLaunchKernelGGL((k_yz<T, A, B>), grid, dim3(Cf::THREADS), Cf::BYTES, ctx->stream, twy, twz, S, fN,  \
                       S_rhs);                                                                             \
    VS_HIP(hipGetLastError());                                                                             \
    return;                                                                                                \
  }
    VS_YZ_SHAPES(VS_YZ_CASE)
#undef VS_YZ_CASE
    default: throw Error{VS_ERR_INTERNAL, "no fused slab kernel for this shape"};
  }
}

// scratch bytes of one matvec call of a variant: five passes 2L x 2M x N x 3 per right-hand side, slab
// variants 2L slabs of 3MN, the two-workgroup parity split three such volumes
template <typename T>
static size_t mvp_scratch_bytes(const vs_op* op, int variant, int nrhs) {
  const int64_t Sr = 2LL * op->L * op->M * op->N * 3;
  const int64_t per = variant == 0 ? 2 * Sr : (variant == 3 ? 3 * Sr : Sr);
  return sizeof(cplx<T>) * (size_t)per * (size_t)nrhs;
}

// ext != nullptr: the scratch is the caller's ext_bytes buffer (nothing is allocated), and the variant is
// the tuned one if its scratch fits, else the available variant with the smallest scratch
// right-hand sides per group of the five-pass matvec: as many as keep the group's intermediate volumes
// within VS_P5_GROUP_MB (default 192) MB; 0 = all at once
static int pass5_rhs_group(int64_t bytes_per_rhs) {
  static const int64_t budget = [] {
    const char* e = std::getenv("VS_P5_GROUP_MB");
    return (int64_t)(e ? std::max(0, std::atoi(e)) : 192) << 20;
  }();
  if (budget == 0) return 1 << 30;
  return (int)std::max<int64_t>(1, budget / std::max<int64_t>(1, bytes_per_rhs));
}

template <typename T>
static void mvp_impl(vs_op* op, const cplx<T>* x, cplx<T>* y, int transp, int nrhs, cplx<T>* ext = nullptr,
                     size_t ext_bytes = 0, const c128* arhs = nullptr) {
  vs_ctx* ctx = op->ctx;
  const int L = op->L, M = op->M, N = op->N;
  const int L2 = 2 * L, M2 = 2 * M, N2 = 2 * N;
  const int64_t S_rhs = (int64_t)L2 * M2 * N * 3;
  int variant = choose_variant<T>(op);
  if (variant == 0 && nrhs > 1) {
    // five passes over many right-hand sides: in groups whose intermediate volumes stay within the
    // 256 MB last-level cache, so that each pass reads what the previous one wrote from there
    const int ck = pass5_rhs_group(S_rhs * (int64_t)sizeof(cplx<T>));
    if (nrhs > ck) {
      for (int j0 = 0; j0 < nrhs; j0 += ck)
        mvp_impl<T>(op, x + j0 * op->nunk, y + j0 * op->nunk, transp, std::min(ck, nrhs - j0), ext, ext_bytes,
                    arhs ? arhs + j0 : nullptr);
      return;
    }
  }
  cplx<T>* S = nullptr;
  if (ext) {
    if (mvp_scratch_bytes<T>(op, variant, nrhs) > ext_bytes) {
      for (int v : {2, 1, 3, 0})
        if (variant_available<T>(op, v) && mvp_scratch_bytes<T>(op, v, nrhs) <= ext_bytes) {
          variant = v;
          break;
        }
    }
    VS_CHECK(mvp_scratch_bytes<T>(op, variant, nrhs) <= ext_bytes, VS_ERR_ARG,
             "matvec needs " + std::to_string(mvp_scratch_bytes<T>(op, variant, nrhs)) +
                 " bytes of scratch, the no-allocation scratch holds " + std::to_string(ext_bytes));
    S = ext;
  } else {
    op->S.ensure(mvp_scratch_bytes<T>(op, variant, nrhs));
    S = op->S.as<cplx<T>>();
  }
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
  const cplx<T>* fq = (sizeof(T) == 8) ? op->fq64.as<cplx<T>>() : op->fq32.as<cplx<T>>();
  const cplx<T>* fsl = (sizeof(T) == 8) ? op->fs64.as<cplx<T>>() : op->fs32.as<cplx<T>>();
  if (variant > 0) {
    // slab-major volume 2L x (3MN): x pass, fused y-z slab pass, inverse x pass
    const int64_t SL = lines, Sr = (int64_t)L2 * SL;
    IoMvFx<T> fx0{x, map3, op->nunk, op->V, L, M, transp, coef_dev<T>(op->b), S, Sr};
    if (!run_rows_p<T, false>(ctx, IoMvFxSlab<T>{fx0, SL}, L, lines, nrhs))
      run_rows<T, false, IoMvFxSlab<T>, false, true>(ctx, IoMvFxSlab<T>{fx0, SL}, L2, lines, L, L2, nrhs);
    if (variant == 3) {
      // input slabs, then the even-ky and odd-ky terms, nrhs slab volumes each
      TensorQ<T> tq{fq, L, M, N, M / 2 + 1, N / 2 + 1};
      cplx<T>* Se = S + Sr * nrhs;
      cplx<T>* So = S + 2 * Sr * nrhs;
      launch_yzq<T>(ctx, S, tq, L, M, N, nrhs, Sr, Se, So);
      IoMvIx<T> ix0{x, y, map3, op->nunk, op->V, L, M, transp, coef_dev<T>(op->a), coef_dev<T>(op->at),
                    coef_dev<T>(op->b), Se, Sr, (T)(1.0 / (double)op->P), arhs};
      if (!run_rows_p<T, true>(ctx, IoMvIxSlab2<T>{ix0, SL, So}, L, lines, nrhs))
        run_rows<T, true, IoMvIxSlab2<T>, true, false>(ctx, IoMvIxSlab2<T>{ix0, SL, So}, L2, lines, L2, L, nrhs);
      return;
    }
    if (variant == 2) {
      TensorQ<T> tq{fq, L, M, N, M / 2 + 1, N / 2 + 1};
      launch_yzq<T>(ctx, S, tq, L, M, N, nrhs, Sr);
    } else {
      TensorS<T> ts{fsl, L, M, N};
      launch_yz<T>(ctx, S, ts, L, M, N, nrhs, Sr);
    }
    IoMvIx<T> ix0{x, y, map3, op->nunk, op->V, L, M, transp, coef_dev<T>(op->a), coef_dev<T>(op->at),
                  coef_dev<T>(op->b), S, Sr, (T)(1.0 / (double)op->P), arhs};
    if (!run_rows_p<T, true>(ctx, IoMvIxSlab<T>{ix0, SL}, L, lines, nrhs))
      run_rows<T, true, IoMvIxSlab<T>, true, false>(ctx, IoMvIxSlab<T>{ix0, SL}, L2, lines, L2, L, nrhs);
    return;
  }
  // 1. gather + x FFT
  IoMvFx<T> fx{x, map3, op->nunk, op->V, L, M, transp, coef_dev<T>(op->b), S, S_rhs};
  run_rows<T, false, IoMvFx<T>, false, false, true>(ctx, fx, L2, lines, L, L2, nrhs);
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
               coef_dev<T>(op->b), S, S_rhs, (T)(1.0 / (double)op->P), arhs};
  run_rows<T, true>(ctx, ix, L2, lines, L2, L, nrhs);
}

void dense_mvp_dev(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs);

void mvp_dev(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs) {
  if (op->dense) return dense_mvp_dev(op, x, y, transp, precision, nrhs);
  VS_CHECK(nrhs >= 1, VS_ERR_ARG, "nrhs must be >= 1");
  VS_CHECK(x != y, VS_ERR_ARG, "matvec input and output must differ");
  if (precision == VS_FP64)
    mvp_impl<double>(op, (const c128*)x, (c128*)y, transp, nrhs);
  else if (precision == VS_FP32)
    mvp_impl<float>(op, (const c64*)x, (c64*)y, transp, nrhs);
  else
    throw Error{VS_ERR_ARG, "precision must be VS_FP64 or VS_FP32"};
}

void mvp_dev_diag(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs, const c128* arhs) {
  VS_CHECK(!op->dense, VS_ERR_UNSUPPORTED, "not available for a dense operator");
  VS_CHECK(nrhs >= 1, VS_ERR_ARG, "nrhs must be >= 1");
  VS_CHECK(x != y, VS_ERR_ARG, "matvec input and output must differ");
  if (precision == VS_FP64)
    mvp_impl<double>(op, (const c128*)x, (c128*)y, transp, nrhs, nullptr, 0, arhs);
  else if (precision == VS_FP32)
    mvp_impl<float>(op, (const c64*)x, (c64*)y, transp, nrhs, nullptr, 0, arhs);
  else
    throw Error{VS_ERR_ARG, "precision must be VS_FP64 or VS_FP32"};
}

void mvp_dev_scratch(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs, void* scratch,
                     size_t scratch_bytes) {
  VS_CHECK(!op->dense, VS_ERR_UNSUPPORTED, "not available for a dense operator");
  VS_CHECK(nrhs >= 1, VS_ERR_ARG, "nrhs must be >= 1");
  VS_CHECK(x != y, VS_ERR_ARG, "matvec input and output must differ");
  VS_CHECK(scratch != nullptr, VS_ERR_INTERNAL, "null scratch");
  if (precision == VS_FP64)
    mvp_impl<double>(op, (const c128*)x, (c128*)y, transp, nrhs, (c128*)scratch, scratch_bytes);
  else if (precision == VS_FP32)
    mvp_impl<float>(op, (const c64*)x, (c64*)y, transp, nrhs, (c64*)scratch, scratch_bytes);
  else
    throw Error{VS_ERR_ARG, "precision must be VS_FP64 or VS_FP32"};
}

template <typename T>
int matvec_variant(const vs_op* op) {
  return choose_variant<T>(op);
}
template int matvec_variant<float>(const vs_op*);
template int matvec_variant<double>(const vs_op*);

// ------------------------------------------------------------------------------------------------
// fused BiCGStab halves: FFT operator with a slab-major matvec and a circulant preconditioner on the same
// domain, one right-hand side.  Per half step: the preconditioner's first x pass with the Krylov update
// (k_row_bi), its middle stages (pc_mid_dev), its last x pass merged with the matvec's first (k_row_pcmv),
// the slab pass, and the matvec's last pass with the half step's inner products (k_row2p with IoDot).
// ------------------------------------------------------------------------------------------------
template <typename T, int NN>
constexpr bool bi_len_ok() {
  return NN >= 8 && RowPCfg<T, NN>::LDS * sizeof(cplx<T>) + 256 <= 65536 && RowPCfg<T, NN>::THREADS <= 512 &&
         RowCfg<T, NN>::THREADS <= 256;
}

template <typename T>
static bool bi_len_compiled(int n) {
  switch (n) {
#define VS_BI_LEN(NN) \
  case NN: return bi_len_ok<T, NN>();
    VS_FFT2_SIZES(VS_BI_LEN)
#undef VS_BI_LEN
    default: return false;
  }
}

// process-wide switch of the fused halves (VS_BIFUSED=0 in the environment starts with it off)
static std::atomic<int> g_bi_fused{-1};
static bool bi_fused_env() {
  int v = g_bi_fused.load();
  if (v < 0) {
    const char* e = std::getenv("VS_BIFUSED");
    int expect = -1;
    g_bi_fused.compare_exchange_strong(expect, (e && e[0] == '0') ? 0 : 1);
    v = g_bi_fused.load();
  }
  return v != 0;
}

void bi_fused_set(bool on) { g_bi_fused.store(on ? 1 : 0); }

template <typename T>
bool bi_fused_available(vs_op* op, vs_pc* pc) {
  if (!bi_fused_env() || !op || !pc || op->dense || pc->mop) return false;
  if (op->dom_id == 0 || pc->dom_id != op->dom_id) return false;
  if (pc->L != op->L || pc->M != op->M || pc->N != op->N || pc->nunk != op->nunk) return false;
  const int prec = sizeof(T) == 8 ? VS_FP64 : VS_FP32;
  if (!(op->precisions & prec) || !(pc->precisions & prec)) return false;
  if (!(sizeof(T) == 8 ? pc->inv64.p : pc->inv32.p)) return false;
  if (choose_variant<T>(op) == 0) return false;
  const char* rp = std::getenv("VS_ROWP");
  if (rp && rp[0] == '0') return false;
  return bi_len_compiled<T>(op->L);
}
template bool bi_fused_available<float>(vs_op*, vs_pc*);
template bool bi_fused_available<double>(vs_op*, vs_pc*);

// returns the number of workgroups (partials)
template <typename T, int MODE>
static int launch_row_bi(vs_ctx* ctx, const BiRowArgs<T>& a, int L, int64_t lines) {
  const cplx<T>* tw = ctx->plan(L).dev<T>().tw;
  switch (L) {
#define VS_BI_ROW(NN)                                                                                   \
  case NN:                                                                                              \
    if constexpr (bi_len_ok<T, NN>()) {                                                                 \
      using Cf = RowCfg<T, NN>;                                                                         \
      const int64_t g = (lines + Cf::LPW - 1) / Cf::LPW;                                                \
      hipLaunchKernelGGL((k_row_bi<T, NN, MODE>), dim3((unsigned)g), dim3(Cf::THREADS), 0, ctx->stream, tw, a, \
                         lines);                                                                        \
      VS_HIP(hipGetLastError());                                                                        \
      return (int)g;                                                                                    \
    }                                                                                                   \
    break;
    VS_FFT2_SIZES(VS_BI_ROW)
#undef VS_BI_ROW
    default: break;
  }
  throw Error{VS_ERR_INTERNAL, "fused BiCGStab pass for an unsupported line length"};
}

template <typename T>
static int launch_row_pcmv(vs_ctx* ctx, const cplx<T>* U, const int* map3, double scale, cplx<T>* y, cplx<T>* S,
                           int64_t SL, int L, int64_t lines, c128* part) {
  const cplx<T>* tw = ctx->plan(L).dev<T>().tw;
  const cplx<T>* tw2 = ctx->plan(2 * L).dev<T>().tw;
  switch (L) {
#define VS_BI_PCMV(NN)                                                                                  \
  case NN:                                                                                              \
    if constexpr (bi_len_ok<T, NN>()) {                                                                 \
      using Cf = RowPCfg<T, NN>;                                                                        \
      const int64_t g = (lines + Cf::LPW - 1) / Cf::LPW;                                                \
      hipLaunchKernelGGL((k_row_pcmv<T, NN>), dim3((unsigned)g), dim3(Cf::THREADS), 0, ctx->stream, tw, tw2, U, \
                         map3, (T)scale, y, S, SL, lines, part);                                        \
      VS_HIP(hipGetLastError());                                                                        \
      return (int)g;                                                                                    \
    }                                                                                                   \
    break;
    VS_FFT2_SIZES(VS_BI_PCMV)
#undef VS_BI_PCMV
    default: break;
  }
  throw Error{VS_ERR_INTERNAL, "fused BiCGStab pass for an unsupported line length"};
}

template <typename T>
static int row_groups(int L, int64_t lines) {
  switch (L) {
#define VS_BI_G(NN) \
  case NN: return (int)((lines + RowCfg<T, NN>::LPW - 1) / RowCfg<T, NN>::LPW);
    VS_FFT2_SIZES(VS_BI_G)
#undef VS_BI_G
    default: return 0;
  }
}

template <typename T>
static int rowp_groups(int L, int64_t lines) {
  switch (L) {
#define VS_BI_G(NN) \
  case NN: return (int)((lines + RowPCfg<T, NN>::LPW - 1) / RowPCfg<T, NN>::LPW);
    VS_FFT2_SIZES(VS_BI_G)
#undef VS_BI_G
    default: return 0;
  }
}

// Y (the preconditioner's volume before its inverse x transform, scale still to apply) -> xin = M^{-1}(...)
// with ||xin||^2 partials, and y = A xin with the inner products of IoDot<KD> against w
template <typename T, int KD>
static void bi_mv_tail(vs_op* op, const cplx<T>* Y, double scale, cplx<T>* xin, c128* part_in, int& n_in, cplx<T>* y,
                       const cplx<T>* w, c128* part_out, int& n_out, int cap) {
  vs_ctx* ctx = op->ctx;
  const int L = op->L, M = op->M, N = op->N;
  const int64_t lines = (int64_t)M * N * 3, SL = lines, Sr = (int64_t)2 * L * SL;
  const int variant = choose_variant<T>(op);
  VS_CHECK(variant > 0, VS_ERR_INTERNAL, "fused BiCGStab half without a slab-major matvec");
  const int* map3 = op->map3.as<int>();
  const int g = rowp_groups<T>(L, lines);
  VS_CHECK(g > 0 && g <= cap, VS_ERR_UNSUPPORTED, "fused BiCGStab: too many partial sums");
  op->S.ensure(mvp_scratch_bytes<T>(op, variant, 1));
  cplx<T>* S = op->S.as<cplx<T>>();
  n_in = launch_row_pcmv<T>(ctx, Y, map3, scale, xin, S, SL, L, lines, part_in);
  const cplx<T>* fq = (sizeof(T) == 8) ? op->fq64.as<cplx<T>>() : op->fq32.as<cplx<T>>();
  const cplx<T>* fsl = (sizeof(T) == 8) ? op->fs64.as<cplx<T>>() : op->fs32.as<cplx<T>>();
  IoMvIx<T> ix0{xin, y, map3, op->nunk, op->V, L, M, 0, coef_dev<T>(op->a), coef_dev<T>(op->at),
                coef_dev<T>(op->b), S, Sr, (T)(1.0 / (double)op->P), nullptr};
  bool ok;
  if (variant == 3) {
    TensorQ<T> tq{fq, L, M, N, M / 2 + 1, N / 2 + 1};
    cplx<T>* Se = S + Sr;
    cplx<T>* So = S + 2 * Sr;
    launch_yzq<T>(ctx, S, tq, L, M, N, 1, Sr, Se, So);
    ix0.S = Se;
    IoDot<T, IoMvIxSlab2<T>, KD> io{IoMvIxSlab2<T>{ix0, SL, So}, w, part_out};
    ok = run_rows_p<T, true>(ctx, io, L, lines, 1);
  } else {
    if (variant == 2) {
      TensorQ<T> tq{fq, L, M, N, M / 2 + 1, N / 2 + 1};
      launch_yzq<T>(ctx, S, tq, L, M, N, 1, Sr);
    } else {
      TensorS<T> ts{fsl, L, M, N};
      launch_yz<T>(ctx, S, ts, L, M, N, 1, Sr);
    }
    IoDot<T, IoMvIxSlab<T>, KD> io{IoMvIxSlab<T>{ix0, SL}, w, part_out};
    ok = run_rows_p<T, true>(ctx, io, L, lines, 1);
  }
  VS_CHECK(ok, VS_ERR_INTERNAL, "fused BiCGStab: no parity-split row transform for this length");
  n_out = g;
}

template <typename T>
void bi_fused_h1(vs_op* op, vs_pc* pc, const cplx<T>* r, const cplx<T>* p0, const cplx<T>* v0, cplx<T>* p, cplx<T>* ph,
                 cplx<T>* v, const cplx<T>* rt, const BiScal* sc, BiParts& P) {
  const int prec = sizeof(T) == 8 ? VS_FP64 : VS_FP32;
  const int64_t lines = (int64_t)op->M * op->N * 3;
  cplx<T>* U = static_cast<cplx<T>*>(pc_work_dev(pc, prec));
  BiRowArgs<T> a{op->map3.as<int>(), U, r, p0, v0, nullptr, p, nullptr, sc, nullptr};
  launch_row_bi<T, 0>(op->ctx, a, op->L, lines);
  double scale = 1.0;
  const cplx<T>* Y = static_cast<const cplx<T>*>(pc_mid_dev(pc, prec, &scale));
  bi_mv_tail<T, 1>(op, Y, scale, ph, P.ph, P.n_ph, v, rt, P.v, P.n_v, P.cap);
}

template <typename T>
void bi_fused_h2(vs_op* op, vs_pc* pc, const cplx<T>* x, const cplx<T>* ph, const cplx<T>* r, const cplx<T>* v,
                 cplx<T>* xh, cplx<T>* s, cplx<T>* sh, cplx<T>* t, const BiScal* sc, BiParts& P) {
  const int prec = sizeof(T) == 8 ? VS_FP64 : VS_FP32;
  const int64_t lines = (int64_t)op->M * op->N * 3;
  cplx<T>* U = static_cast<cplx<T>*>(pc_work_dev(pc, prec));
  BiRowArgs<T> a{op->map3.as<int>(), U, x, ph, r, v, xh, s, sc, P.sx};
  const int gs = row_groups<T>(op->L, lines);
  VS_CHECK(gs > 0 && gs <= P.cap, VS_ERR_UNSUPPORTED, "fused BiCGStab: too many partial sums");
  P.n_sx = launch_row_bi<T, 1>(op->ctx, a, op->L, lines);
  double scale = 1.0;
  const cplx<T>* Y = static_cast<const cplx<T>*>(pc_mid_dev(pc, prec, &scale));
  bi_mv_tail<T, 2>(op, Y, scale, sh, P.sh, P.n_sh, t, s, P.t, P.n_t, P.cap);
}

#define VS_BIF_INST(T)                                                                                          \
  template void bi_fused_h1<T>(vs_op*, vs_pc*, const cplx<T>*, const cplx<T>*, const cplx<T>*, cplx<T>*, cplx<T>*, \
                               cplx<T>*, const cplx<T>*, const BiScal*, BiParts&);                              \
  template void bi_fused_h2<T>(vs_op*, vs_pc*, const cplx<T>*, const cplx<T>*, const cplx<T>*, const cplx<T>*,      \
                               cplx<T>*, cplx<T>*, cplx<T>*, cplx<T>*, const BiScal*, BiParts&);
VS_BIF_INST(float)
VS_BIF_INST(double)
#undef VS_BIF_INST

// Autotuning: time every available matvec variant in each precision the operator holds (one warm-up and
// three timed applications on a unit vector, HIP events on the context stream) and keep the fastest.
// Done once per operator, at its first domain (the matvec needs the unknown map).
template <typename T>
static void autotune_prec(vs_op* op) {
  vs_ctx* ctx = op->ctx;
  const int slot = sizeof(T) == 8 ? 0 : 1;
  DevBuf x(sizeof(cplx<T>) * op->nunk), y(sizeof(cplx<T>) * op->nunk);
  VS_HIP(hipMemsetAsync(x.p, 0, sizeof(cplx<T>) * op->nunk, ctx->stream));
  hipEvent_t e0, e1;
  VS_HIP(hipEventCreate(&e0));
  VS_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int best_v = 0;
  for (int v = 0; v <= 3; ++v) {
    if (!variant_available<T>(op, v)) continue;
    op->mv_variant[slot] = v;
    mvp_impl<T>(op, x.as<cplx<T>>(), y.as<cplx<T>>(), 0, 1);
    VS_HIP(hipEventRecord(e0, ctx->stream));
    for (int r = 0; r < 3; ++r) mvp_impl<T>(op, x.as<cplx<T>>(), y.as<cplx<T>>(), 0, 1);
    VS_HIP(hipEventRecord(e1, ctx->stream));
    VS_HIP(hipEventSynchronize(e1));
    float ms = 0;
    VS_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      best_v = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  op->mv_variant[slot] = best_v;
}

void op_autotune(vs_op* op) {
  if (op->tuned) return;
  const char* e = std::getenv("VS_AUTOTUNE");
  if (e && e[0] == '0') return;
  if (op->precisions & VS_FP64) autotune_prec<double>(op);
  if (op->precisions & VS_FP32) autotune_prec<float>(op);
  op->tuned = true;
}

// Gauss-Legendre nodes/weights on [-1,1] by Newton iteration on P_n (long double), ascending nodes
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
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
  // slab-major copies for the fused matvec variants, in the precisions whose shape has them
  const int Hy = op->M / 2 + 1, Hz = op->N / 2 + 1;
  const int64_t totq = (int64_t)(op->L + 1) * 6 * 2 * Hz * 2 * Hy;
  const unsigned gq = (unsigned)((totq + 255) / 256);
  bool q64 = false, q32 = false;
  if ((op->precisions & VS_FP64) && yzq_available<double>(op->L, op->M, op->N)) {
    op->fq64.ensure(sizeof(c128) * totq);
    hipLaunchKernelGGL((k_compress_tensor_q<double>), dim3(gq), dim3(256), 0, ctx->stream, op->fN64.as<c128>(),
                       op->fq64.as<c128>(), op->L, op->M, op->N, Hy, Hz);
    VS_HIP(hipGetLastError());
    yzq_set_attr<double>(op->M, op->N);
    q64 = true;
  }
  if ((op->precisions & VS_FP32) && yzq_available<float>(op->L, op->M, op->N)) {
    op->fq32.ensure(sizeof(c64) * totq);
    hipLaunchKernelGGL((k_compress_tensor_q<float>), dim3(gq), dim3(256), 0, ctx->stream, op->fN64.as<c128>(),
                       op->fq32.as<c64>(), op->L, op->M, op->N, Hy, Hz);
    VS_HIP(hipGetLastError());
    yzq_set_attr<float>(op->M, op->N);
    q32 = true;
  }
  if (q64 || q32) ctx->plan(op->M);  // twiddles of the unpadded y transform, created outside any capture
  if ((op->precisions & VS_FP64) && yz_fused_available<double>(op->L, op->M, op->N)) {
    op->fs64.ensure(sizeof(c128) * tot);
    hipLaunchKernelGGL((k_compress_tensor_slab<double>), dim3(g), dim3(256), 0, ctx->stream, op->fN64.as<c128>(),
                       op->fs64.as<c128>(), op->L, op->M, op->N);
    VS_HIP(hipGetLastError());
    yz_set_attr<double>(2 * op->M, 2 * op->N);
  }
  if ((op->precisions & VS_FP32) && yz_fused_available<float>(op->L, op->M, op->N)) {
    op->fs32.ensure(sizeof(c64) * tot);
    hipLaunchKernelGGL((k_compress_tensor_slab<float>), dim3(g), dim3(256), 0, ctx->stream, op->fN64.as<c128>(),
                       op->fs32.as<c64>(), op->L, op->M, op->N);
    VS_HIP(hipGetLastError());
    yz_set_attr<float>(2 * op->M, 2 * op->N);
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
