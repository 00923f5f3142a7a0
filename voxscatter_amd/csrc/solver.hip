// Krylov vector kernels and the device-resident solvers.
//
// Reductions are deterministic: a fixed-size grid writes per-workgroup partial sums (accumulated in
// FP64 whatever the vector precision), and one workgroup combines them in a fixed order.  The host
// reads the reduced scalars once per Krylov half step; all vectors stay in HBM.
//
//  bicgstab : MATLAB bicgstab semantics (as called at main_DDA.m:82,120); see solve_bicgstab
//  gmres    : MATLAB gmres(A,b,restart,tol,maxit,M) -- left preconditioned, restarted; Arnoldi by
//             classical Gram-Schmidt applied twice, Givens rotations on the host
//  pgmres   : pgmres.m:70-180 / pgmres_iter.m:22-150 -- one-pass CGS, L1 or R1 preconditioner
#include "bi_device.hpp"
#include "handles.hpp"
#include "solver.hpp"

#include <algorithm>
#include <cfloat>
#include <cstring>

namespace vs {

constexpr int RED_THREADS = 256;
constexpr int RED_MAX_K = 8;
#ifndef VS_MDOT_JB
#define VS_MDOT_JB 4
#endif

static int red_grid(int64_t n) {
  int64_t g = (n + RED_THREADS * 4 - 1) / (RED_THREADS * 4);
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 1024));
}

// block-wide sum of K complex doubles held per thread; result valid in thread 0
template <int K>
__device__ __forceinline__ void block_sum(double (&re)[K], double (&im)[K], c128* out) {
  __shared__ double sre[K][RED_THREADS / 64], sim[K][RED_THREADS / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    for (int off = 32; off >= 1; off >>= 1) {
      re[k] += __shfl_xor(re[k], off);
      im[k] += __shfl_xor(im[k], off);
    }
    if (lane == 0) {
      sre[k][wave] = re[k];
      sim[k][wave] = im[k];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double a = 0, b = 0;
      for (int w = 0; w < RED_THREADS / 64; ++w) {
        a += sre[k][w];
        b += sim[k][w];
      }
      out[k] = mk<double>(a, b);
    }
  }
}

template <typename T>
struct DotArgs {
  const cplx<T>* a[RED_MAX_K];
  const cplx<T>* b[RED_MAX_K];
};

// partial[blk*K + k] = sum_i conj(a_k[i]) b_k[i] over this workgroup's grid-stride range
template <typename T, int K>
__global__ __launch_bounds__(RED_THREADS) void k_dots(DotArgs<T> args, int64_t n, c128* partial) {
  double re[K], im[K];
#pragma unroll
  for (int k = 0; k < K; ++k) re[k] = im[k] = 0.0;
  constexpr int U = K <= 2 ? 4 : 2;  // elements per thread whose loads are issued together
  for (int64_t i0 = (int64_t)blockIdx.x * RED_THREADS * U + threadIdx.x; i0 < n;
       i0 += (int64_t)gridDim.x * RED_THREADS * U) {
    cplx<T> a[U][K], b[U][K];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = min(i0 + u * RED_THREADS, n - 1);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        a[u][k] = args.a[k][i];
        b[u][k] = args.b[k][i];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (i0 + u * RED_THREADS >= n) break;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        re[k] += (double)a[u][k].x * b[u][k].x + (double)a[u][k].y * b[u][k].y;
        im[k] += (double)a[u][k].x * b[u][k].y - (double)a[u][k].y * b[u][k].x;
      }
    }
  }
  c128 tot[K];
  block_sum<K>(re, im, tot);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) partial[blockIdx.x * K + k] = tot[k];
}

// lane's share of a fixed-order partial sum: entries lane, lane + 64, ... of p[i * K + k], i < nblk, added in
// that order; the loads of 8 entries are issued together (out-of-range slots add zero)
__device__ __forceinline__ void lane_sum(const c128* __restrict__ p, int nblk, int K, int k, int lane, double& a,
                                         double& b) {
  constexpr int U = 8;
  for (int i0 = lane; i0 < nblk; i0 += 64 * U) {
    c128 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + 64 * u;
      v[u] = i < nblk ? p[(int64_t)i * K + k] : mk<double>(0, 0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      a += v[u].x;
      b += v[u].y;
    }
  }
}

// out[k] = sum over nblk partials, fixed order
// out[k] = sum over nblk partials, fixed order: one wavefront per k (lane l sums blocks l, l + 64, ..., then a
// shuffle tree); workgroup b covers k = 4 b ... 4 b + 3
__global__ __launch_bounds__(RED_THREADS) void k_reduce_partials(const c128* partial, int nblk, int K, c128* out) {
  const int k = blockIdx.x * (RED_THREADS / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (k >= K) return;
  double a = 0, b = 0;
  lane_sum(partial, nblk, K, k, lane, a, b);
  for (int off = 32; off >= 1; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
  }
  if (lane == 0) out[k] = mk<double>(a, b);
}

// workgroups of k_reduce_partials for K sums
static unsigned reduce_grid(int K) { return (unsigned)((K + RED_THREADS / 64 - 1) / (RED_THREADS / 64)); }

// element-wise updates, each optionally accumulating ||result||^2 partials
enum UpdOp : int {
  UPD_AXPBY = 0,    // y = al x + be y
  UPD_BICG_P = 1,   // p = r + be (p - al v)           (x0=r, x1=v, y=p)
  UPD_HALF = 2,     // y = x + al x1 ; z = x2 - al x3   (xhalf = x + alpha ph ; s = r - alpha v)
  UPD_SUB = 3,      // y = x - y                        (r = b - Ax)
  UPD_COPY = 4,     // y = x
  UPD_SCALE = 5,    // y = al x
};

template <typename T>
struct UpdArgs {
  const cplx<T>* x0;
  const cplx<T>* x1;
  const cplx<T>* x2;
  const cplx<T>* x3;
  cplx<T>* y;
  cplx<T>* z;
  cplx<T> al, be;
};

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_update(int op, UpdArgs<T> a, int64_t n, int want_norm, c128* partial) {
  double re[1] = {0.0}, im[1] = {0.0};
  // the operands of U elements per thread are loaded together (each element is read and written by one
  // thread only, so loading a group before any of its stores is safe even for in-place operands)
  constexpr int U = 4;
  const bool r0 = op != UPD_COPY && op != UPD_SCALE && op != UPD_SUB;  // reads x1
  const bool ry = op == UPD_AXPBY || op == UPD_BICG_P || op == UPD_SUB;
  for (int64_t i0 = (int64_t)blockIdx.x * RED_THREADS * U + threadIdx.x; i0 < n;
       i0 += (int64_t)gridDim.x * RED_THREADS * U) {
    cplx<T> q0[U], q1[U], q2[U], q3[U], qy[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = min(i0 + u * RED_THREADS, n - 1);
      q0[u] = a.x0[i];
      if (r0 && op != UPD_AXPBY) q1[u] = a.x1[i];
      if (op == UPD_HALF) {
        q2[u] = a.x2[i];
        q3[u] = a.x3[i];
      }
      if (ry) qy[u] = a.y[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = i0 + u * RED_THREADS;
      if (i >= n) break;
      cplx<T> out;
      switch (op) {
        case UPD_AXPBY: out = a.al * q0[u] + a.be * qy[u]; a.y[i] = out; break;
        case UPD_BICG_P: out = q0[u] + a.be * (qy[u] - a.al * q1[u]); a.y[i] = out; break;
        case UPD_HALF: {
          a.y[i] = q0[u] + a.al * q1[u];
          out = q2[u] - a.al * q3[u];
          a.z[i] = out;
          break;
        }
        case UPD_SUB: out = q0[u] - qy[u]; a.y[i] = out; break;
        case UPD_COPY: out = q0[u]; a.y[i] = out; break;
        default: out = a.al * q0[u]; a.y[i] = out; break;
      }
      re[0] += (double)out.x * out.x + (double)out.y * out.y;
    }
  }
  if (want_norm) {
    c128 tot[1];
    block_sum<1>(re, im, tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot[0];
  }
}

// GMRES helpers: partial[blk*(k) + j] = sum_i conj(V_j[i]) w[i]  for j < k
// Each lane keeps IPT of its elements of w in registers (the grid of red_grid gives a lane IPT = 4
// elements up to n = 2^20) while the basis streams past JB vectors at a time, so w is read once (per
// chunk of IPT grid strides) instead of once per JB basis vectors.  The loads are branch-free (indices
// past n clamped to n - 1, their w set to zero), so the IPT x JB loads of a group are all in flight
// together; a bounds branch per element serialises them.  Each wavefront reduces its sums across lanes
// and accumulates them in LDS slots of its own (no barrier between groups); the wavefronts' slots are
// added in order at the end.
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_mdot(const cplx<T>* V, int64_t ldv, int k, const cplx<T>* w, int64_t n,
                                                      c128* partial) {
  constexpr int JB = VS_MDOT_JB, IPT = 4, NW = RED_THREADS / 64;
  __shared__ c128 sacc[NW][512];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t stride = (int64_t)gridDim.x * RED_THREADS;
  for (int j = lane; j < k; j += 64) sacc[wave][j] = mk<double>(0, 0);
  for (int64_t i0 = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i0 - threadIdx.x < n; i0 += IPT * stride) {
    cplx<T> wv[IPT];
    int64_t ic[IPT];
#pragma unroll
    for (int t = 0; t < IPT; ++t) {
      const int64_t i = i0 + t * stride;
      const bool ok = i < n;
      ic[t] = ok ? i : n - 1;
      const cplx<T> x = w[ic[t]];
      wv[t] = mk<T>(ok ? x.x : (T)0, ok ? x.y : (T)0);
    }
    for (int j0 = 0; j0 < k; j0 += JB) {
      double re[JB], im[JB];
      const cplx<T>* v[JB];
#pragma unroll
      for (int q = 0; q < JB; ++q) {
        re[q] = im[q] = 0.0;
        v[q] = V + (int64_t)min(j0 + q, k - 1) * ldv;
      }
      cplx<T> a[IPT][JB];
#pragma unroll
      for (int t = 0; t < IPT; ++t)
#pragma unroll
        for (int q = 0; q < JB; ++q) a[t][q] = v[q][ic[t]];
#pragma unroll
      for (int t = 0; t < IPT; ++t) {
        const cplx<T> b = wv[t];
#pragma unroll
        for (int q = 0; q < JB; ++q) {
          re[q] += (double)a[t][q].x * b.x + (double)a[t][q].y * b.y;
          im[q] += (double)a[t][q].x * b.y - (double)a[t][q].y * b.x;
        }
      }
#pragma unroll
      for (int q = 0; q < JB; ++q) {
        for (int off = 32; off >= 1; off >>= 1) {
          re[q] += __shfl_xor(re[q], off);
          im[q] += __shfl_xor(im[q], off);
        }
      }
      // lane q < JB (j0 + q < k) adds sum q to this wavefront's slot; every lane holds every sum
      double mr = re[0], mi = im[0];
#pragma unroll
      for (int q = 1; q < JB; ++q)
        if (lane == q) {
          mr = re[q];
          mi = im[q];
        }
      if (lane < JB && j0 + lane < k) sacc[wave][j0 + lane] = sacc[wave][j0 + lane] + mk<double>(mr, mi);
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < k; j += RED_THREADS) {
    c128 a = sacc[0][j];
#pragma unroll
    for (int r = 1; r < NW; ++r) a = a + sacc[r][j];
    partial[(int64_t)blockIdx.x * k + j] = a;
  }
}

// w += sgn * sum_j V_j h_j  (h in FP64 on the device); with partial != nullptr also one partial of
// ||w||^2 (of the updated w) per workgroup
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_gemv_acc(const cplx<T>* V, int64_t ldv, int k, const c128* h, double sgn,
                                                          cplx<T>* w, int64_t n, c128* partial = nullptr) {
  __shared__ cplx<T> hs[512];
  for (int j = threadIdx.x; j < k; j += RED_THREADS) hs[j] = ccast<T>(sgn * h[j]);
  __syncthreads();
  double re[1] = {0.0}, im[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_THREADS) {
    cplx<T> acc = w[i];
    for (int j = 0; j < k; ++j) acc = acc + V[(int64_t)j * ldv + i] * hs[j];
    w[i] = acc;
    re[0] += (double)acc.x * acc.x + (double)acc.y * acc.y;
  }
  if (partial) {
    c128 tot[1];
    block_sum<1>(re, im, tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot[0];
  }
}

// The middle two passes of CGS2 in one read of the basis: w1 = w + sgn V h (h in FP64 on the device, cast
// like k_gemv_acc), stored over w, and partial[blk*k + j] = sum_i conj(V_j[i]) w1[i] (the layout of
// k_mdot).  A workgroup stages a tile of GE elements of all k vectors in LDS (read once from memory, row
// pitch GE + 1), forms w1 of the tile (RED_THREADS / GE lanes per element over interleaved vectors, added
// in a fixed order), then takes the dot products from LDS, two threads per vector, summed in FP64 over the
// workgroup's tiles.  The next tile's loads (KB vectors' worth per lane, k <= KB, indices clamped; KB in
// 32 / 64 / 96 / 128 bounds the registers they take) are issued before the current tile is multiplied and
// land in LDS after it.
extern __shared__ __align__(16) unsigned char vs_gm_smem[];

template <typename T>
struct GmCfg {
  static constexpr int GE = 32;  // elements per tile (64 in FP32 measured 3.7 % slower: more prefetch registers)
  static constexpr int KMAX = RED_THREADS / 2;
  // dynamic LDS of a launch over k vectors (the tile and w1)
  static constexpr size_t lds(int k) { return sizeof(cplx<T>) * ((size_t)k * (GE + 1) + GE + RED_THREADS); }
};

template <typename T, int KB>
__global__ __launch_bounds__(RED_THREADS) void k_gemv_mdot(const cplx<T>* __restrict__ V, int64_t ldv, int k,
                                                           const c128* __restrict__ h, double sgn, cplx<T>* w,
                                                           int64_t n, c128* partial) {
  using C = GmCfg<T>;
  constexpr int GE = C::GE, GP = GE + 1, HALFE = GE / 2;
  constexpr int PER = (KB * GE + RED_THREADS - 1) / RED_THREADS;  // staged values per lane and tile
  static_assert(KB <= C::KMAX, "two threads per vector");
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(vs_gm_smem);  // [k][GP]
  cplx<T>* sw = tile + (size_t)k * GP;                       // w1 of the tile
  cplx<T>* sp = sw + GE;                                     // partial sums of w1, [RED_THREADS / GE][GE]
  __shared__ cplx<T> hs[KB];
  const int t = threadIdx.x;
  for (int j = t; j < k; j += RED_THREADS) hs[j] = ccast<T>(sgn * h[j]);
  const int jd = t >> 1, part = t & 1;  // dot phase: vector jd, elements part * GE/2 .. + GE/2
  double are = 0.0, aim = 0.0;
  const int64_t ntiles = (n + GE - 1) / GE;
  // staging: GE divides RED_THREADS, so lane t always takes element t % GE of a tile, of vectors
  // t / GE + q * (RED_THREADS / GE) (those past k clamped to k - 1)
  constexpr int JS = RED_THREADS / GE;
  const int es = t % GE, js = t / GE;
  // the prefetched values as separate real and imaginary arrays (an array of complex structs carried
  // across the loop is kept in scratch memory by the compiler in FP64)
  T pvr[PER], pvi[PER];
  cplx<T> pw = mk<T>(0, 0);
  int64_t tl = blockIdx.x;
  if (tl < ntiles) {
    const cplx<T>* vb = V + min(tl * GE + es, n - 1);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const cplx<T> x = vb[(int64_t)min(js + q * JS, k - 1) * ldv];
      pvr[q] = x.x;
      pvi[q] = x.y;
    }
    if (t < GE) pw = w[min(tl * GE + t, n - 1)];
  }
  for (; tl < ntiles; tl += gridDim.x) {
    const int64_t e0 = tl * GE;
    __syncthreads();  // the previous tile's dot phase has read the LDS
    // rows past k - 1 were loaded from row k - 1 and are stored there again (the same value), so that the
    // stores need no branch (a branch per store moves the prefetched values to scratch memory in FP64)
#pragma unroll
    for (int q = 0; q < PER; ++q) tile[min(js + q * JS, k - 1) * GP + es] = mk<T>(pvr[q], pvi[q]);
    const cplx<T> wv = pw;
    __syncthreads();
    const int64_t tn = tl + gridDim.x;
    if (tn < ntiles) {
      const cplx<T>* vb = V + min(tn * GE + es, n - 1);
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const cplx<T> x = vb[(int64_t)min(js + q * JS, k - 1) * ldv];
        pvr[q] = x.x;
        pvi[q] = x.y;
      }
      if (t < GE) pw = w[min(tn * GE + t, n - 1)];
    }
    {
      // w1: the RED_THREADS / GE lanes of element es take the vectors j = js (mod JS) in order; their sums
      // are added in a fixed order
      cplx<T> acc = mk<T>(0, 0);
      for (int j = js; j < k; j += JS) acc = acc + tile[j * GP + es] * hs[j];
      sp[js * GE + es] = acc;
    }
    __syncthreads();
    if (t < GE) {
      cplx<T> acc = wv;
#pragma unroll
      for (int q = 0; q < JS; ++q) acc = acc + sp[q * GE + t];
      const bool ok = e0 + t < n;
      if (ok) w[e0 + t] = acc;
      sw[t] = ok ? acc : mk<T>(0, 0);
    }
    __syncthreads();
    if (jd < k) {
      const cplx<T>* row = tile + jd * GP + part * HALFE;
      const cplx<T>* wr = sw + part * HALFE;
      // four independent partial sums per lane (the additions of one chain wait on each other)
      double pr[4] = {0, 0, 0, 0}, pi[4] = {0, 0, 0, 0};
#pragma unroll
      for (int e = 0; e < HALFE; ++e) {
        const cplx<T> a = row[e], b = wr[e];
        pr[e & 3] += (double)a.x * b.x + (double)a.y * b.y;
        pi[e & 3] += (double)a.x * b.y - (double)a.y * b.x;
      }
      are += (pr[0] + pr[1]) + (pr[2] + pr[3]);
      aim += (pi[0] + pi[1]) + (pi[2] + pi[3]);
    }
  }
  // the two halves of each vector are neighbouring lanes
  are += __shfl_xor(are, 1);
  aim += __shfl_xor(aim, 1);
  if (part == 0 && jd < k) partial[(int64_t)blockIdx.x * k + jd] = mk<double>(are, aim);
}

// w *= 1 / sqrt(nrm2[0].x) into y (Arnoldi normalisation with the norm left on the device)
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_scale_inv_norm(const cplx<T>* w, cplx<T>* y, int64_t n,
                                                                const c128* nrm2) {
  const double nn = sqrt(nrm2[0].x);
  const T s = (T)(1.0 / nn);
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_THREADS)
    y[i] = s * w[i];
}

template <typename To, typename Ti>
__global__ void k_cvt(const cplx<Ti>* __restrict__ in, cplx<To>* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = ccast<To>(in[i]);
}

// ------------------------------------------------------------------------------------------------
// host-side vector engine
// ------------------------------------------------------------------------------------------------
template <typename T>
Engine<T>::Engine(vs_ctx* c, int64_t n_) : ctx(c), n(n_), grid(red_grid(n_)) {
  ctx->red.ensure(sizeof(c128) * 1024 * 512);
  ctx->red_out.ensure(sizeof(c128) * 512);
  if (!ctx->pinned) VS_HIP(hipHostMalloc(&ctx->pinned, sizeof(c128) * VS_PINNED_C128, hipHostMallocDefault));
  if (!ctx->pinned_up) VS_HIP(hipHostMalloc(&ctx->pinned_up, sizeof(c128) * VS_PINNED_C128, hipHostMallocDefault));
}

template <typename T>
void Engine<T>::fetch(int K, c128* out) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(reduce_grid(K)), dim3(RED_THREADS), 0, ctx->stream, ctx->red.as<c128>(),
                     grid, K, ctx->red_out.as<c128>());
  VS_HIP(hipGetLastError());
  VS_HIP(hipMemcpyAsync(ctx->pinned, ctx->red_out.p, sizeof(c128) * K, hipMemcpyDeviceToHost, ctx->stream));
  VS_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(out, ctx->pinned, sizeof(c128) * K);
}

template <typename T>
void Engine<T>::dots(int K, const cplx<T>* const* a, const cplx<T>* const* b, c128* out) {
  DotArgs<T> args{};
  for (int k = 0; k < K; ++k) {
    args.a[k] = a[k];
    args.b[k] = b[k];
  }
  c128* part = ctx->red.as<c128>();
  switch (K) {
    case 1: hipLaunchKernelGGL((k_dots<T, 1>), dim3(grid), dim3(RED_THREADS), 0, ctx->stream, args, n, part); break;
    case 2: hipLaunchKernelGGL((k_dots<T, 2>), dim3(grid), dim3(RED_THREADS), 0, ctx->stream, args, n, part); break;
    case 3: hipLaunchKernelGGL((k_dots<T, 3>), dim3(grid), dim3(RED_THREADS), 0, ctx->stream, args, n, part); break;
    case 4: hipLaunchKernelGGL((k_dots<T, 4>), dim3(grid), dim3(RED_THREADS), 0, ctx->stream, args, n, part); break;
    default: throw Error{VS_ERR_INTERNAL, "dots: K out of range"};
  }
  VS_HIP(hipGetLastError());
  fetch(K, out);
}

template <typename T>
double Engine<T>::update(int op, const cplx<T>* x0, const cplx<T>* x1, const cplx<T>* x2, const cplx<T>* x3,
                         cplx<T>* y, cplx<T>* z, c128 al, c128 be, bool want_norm) {
  UpdArgs<T> a{x0, x1, x2, x3, y, z, ccast<T>(al), ccast<T>(be)};
  hipLaunchKernelGGL((k_update<T>), dim3(grid), dim3(RED_THREADS), 0, ctx->stream, op, a, n, want_norm ? 1 : 0,
                     ctx->red.as<c128>());
  VS_HIP(hipGetLastError());
  if (!want_norm) return 0.0;
  c128 r;
  fetch(1, &r);
  return std::sqrt(r.x);
}

template <typename T>
double Engine<T>::norm(const cplx<T>* x) {
  c128 r;
  const cplx<T>* a[1] = {x};
  dots(1, a, a, &r);
  return std::sqrt(r.x);
}

template <typename T>
void Engine<T>::copy(const cplx<T>* x, cplx<T>* y) {
  VS_HIP(hipMemcpyAsync(y, x, sizeof(cplx<T>) * n, hipMemcpyDeviceToDevice, ctx->stream));
}

template <typename T>
void Engine<T>::mdot(const cplx<T>* V, int64_t ldv, int k, const cplx<T>* w, c128* out) {
  VS_CHECK(k <= 512, VS_ERR_UNSUPPORTED, "Krylov basis larger than 512 vectors");
  hipLaunchKernelGGL((k_mdot<T>), dim3(grid), dim3(RED_THREADS), 0, ctx->stream, V, ldv, k, w, n, ctx->red.as<c128>());
  VS_HIP(hipGetLastError());
  fetch(k, out);
}

template <typename T>
void Engine<T>::mdot_dev(const cplx<T>* V, int64_t ldv, int k, const cplx<T>* w, c128* out_dev) {
  VS_CHECK(k <= 512, VS_ERR_UNSUPPORTED, "Krylov basis larger than 512 vectors");
  hipLaunchKernelGGL((k_mdot<T>), dim3(grid), dim3(RED_THREADS), 0, ctx->stream, V, ldv, k, w, n, ctx->red.as<c128>());
  hipLaunchKernelGGL(k_reduce_partials, dim3(reduce_grid(k)), dim3(RED_THREADS), 0, ctx->stream, ctx->red.as<c128>(),
                     grid, k, out_dev);
  VS_HIP(hipGetLastError());
}

template <typename T>
bool Engine<T>::gemv_mdot_dev(const cplx<T>* V, int64_t ldv, int k, const c128* h_dev, double sgn, cplx<T>* w,
                              c128* out_dev) {
  static const bool on = [] {
    const char* e = std::getenv("VS_GMDOT");
    return !(e && e[0] == '0');
  }();
  if (!on || k > GmCfg<T>::KMAX) return false;
  static const bool attr = [] {
    const int l = (int)GmCfg<T>::lds(GmCfg<T>::KMAX);
    return hipFuncSetAttribute((const void*)k_gemv_mdot<T, 32>, hipFuncAttributeMaxDynamicSharedMemorySize, l) ==
               hipSuccess &&
           hipFuncSetAttribute((const void*)k_gemv_mdot<T, 64>, hipFuncAttributeMaxDynamicSharedMemorySize, l) ==
               hipSuccess &&
           hipFuncSetAttribute((const void*)k_gemv_mdot<T, 96>, hipFuncAttributeMaxDynamicSharedMemorySize, l) ==
               hipSuccess &&
           hipFuncSetAttribute((const void*)k_gemv_mdot<T, 128>, hipFuncAttributeMaxDynamicSharedMemorySize, l) ==
               hipSuccess;
  }();
  VS_CHECK(attr, VS_ERR_HIP, "cannot raise the LDS limit of the fused Gram-Schmidt pass");
  // workgroups: up to six resident per CU, each over several tiles; the partials (gm * k) stay within the
  // reduction buffer (1024 x 512)
  const int64_t ntiles = (n + GmCfg<T>::GE - 1) / GmCfg<T>::GE;
  const int gm = (int)std::max<int64_t>(1, std::min<int64_t>({ntiles, 1536, (int64_t)1024 * 512 / k}));
  const size_t lds = GmCfg<T>::lds(k);
  c128* part = ctx->red.as<c128>();
  if (k <= 32)
    hipLaunchKernelGGL((k_gemv_mdot<T, 32>), dim3(gm), dim3(RED_THREADS), lds, ctx->stream, V, ldv, k, h_dev, sgn, w,
                       n, part);
  else if (k <= 64)
    hipLaunchKernelGGL((k_gemv_mdot<T, 64>), dim3(gm), dim3(RED_THREADS), lds, ctx->stream, V, ldv, k, h_dev, sgn, w,
                       n, part);
  else if (k <= 96)
    hipLaunchKernelGGL((k_gemv_mdot<T, 96>), dim3(gm), dim3(RED_THREADS), lds, ctx->stream, V, ldv, k, h_dev, sgn, w,
                       n, part);
  else
    hipLaunchKernelGGL((k_gemv_mdot<T, 128>), dim3(gm), dim3(RED_THREADS), lds, ctx->stream, V, ldv, k, h_dev, sgn,
                       w, n, part);
  hipLaunchKernelGGL(k_reduce_partials, dim3(reduce_grid(k)), dim3(RED_THREADS), 0, ctx->stream, part, gm, k, out_dev);
  VS_HIP(hipGetLastError());
  return true;
}

template <typename T>
void Engine<T>::gemv_acc_dev(const cplx<T>* V, int64_t ldv, int k, const c128* h_dev, double sgn, cplx<T>* w) {
  hipLaunchKernelGGL((k_gemv_acc<T>), dim3(grid), dim3(RED_THREADS), 0, ctx->stream, V, ldv, k, h_dev, sgn, w, n);
  VS_HIP(hipGetLastError());
}

template <typename T>
void Engine<T>::gemv_acc_norm2_dev(const cplx<T>* V, int64_t ldv, int k, const c128* h_dev, double sgn, cplx<T>* w,
                                   c128* nrm2_dev) {
  hipLaunchKernelGGL((k_gemv_acc<T>), dim3(grid), dim3(RED_THREADS), 0, ctx->stream, V, ldv, k, h_dev, sgn, w, n,
                     ctx->red.as<c128>());
  hipLaunchKernelGGL(k_reduce_partials, dim3(reduce_grid(1)), dim3(RED_THREADS), 0, ctx->stream, ctx->red.as<c128>(),
                     grid, 1, nrm2_dev);
  VS_HIP(hipGetLastError());
}

template <typename T>
void Engine<T>::norm2_dev(const cplx<T>* x, c128* out_dev) {
  DotArgs<T> a{};
  a.a[0] = x;
  a.b[0] = x;
  hipLaunchKernelGGL((k_dots<T, 1>), dim3(grid), dim3(RED_THREADS), 0, ctx->stream, a, n, ctx->red.as<c128>());
  hipLaunchKernelGGL(k_reduce_partials, dim3(reduce_grid(1)), dim3(RED_THREADS), 0, ctx->stream, ctx->red.as<c128>(),
                     grid, 1, out_dev);
  VS_HIP(hipGetLastError());
}

template <typename T>
void Engine<T>::scale_inv_norm_dev(const cplx<T>* w, cplx<T>* y, const c128* nrm2_dev) {
  hipLaunchKernelGGL((k_scale_inv_norm<T>), dim3(grid), dim3(RED_THREADS), 0, ctx->stream, w, y, n, nrm2_dev);
  VS_HIP(hipGetLastError());
}

template <typename T>
void Engine<T>::read(const c128* dev, int K, c128* host) {
  VS_CHECK(K <= VS_PINNED_C128, VS_ERR_INTERNAL, "read-back larger than the staging buffer");
  VS_HIP(hipMemcpyAsync(ctx->pinned, dev, sizeof(c128) * K, hipMemcpyDeviceToHost, ctx->stream));
  VS_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(host, ctx->pinned, sizeof(c128) * K);
}

template <typename T>
void Engine<T>::gemv_acc(const cplx<T>* V, int64_t ldv, int k, const c128* h_host, double sgn, cplx<T>* w) {
  VS_CHECK(k <= 512, VS_ERR_UNSUPPORTED, "Krylov basis larger than 512 vectors");
  // the staging buffer may still feed the previous upload: drain the stream before overwriting it
  VS_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(ctx->pinned_up, h_host, sizeof(c128) * k);
  VS_HIP(hipMemcpyAsync(ctx->red_out.p, ctx->pinned_up, sizeof(c128) * k, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL((k_gemv_acc<T>), dim3(grid), dim3(RED_THREADS), 0, ctx->stream, V, ldv, k, ctx->red_out.as<c128>(),
                     sgn, w, n);
  VS_HIP(hipGetLastError());
}


// Y(:, j) (+)= sum_i X(:, i) P(i, j) for up to 8 outputs j at once (P, q x kc column-major, in LDS):
// one pass over the q basis vectors per chunk of outputs (deflation-space updates of GMRES-DR)
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_combine(const cplx<T>* __restrict__ X, int64_t ldx, int q,
                                                         const c128* __restrict__ P, int kc, cplx<T>* Y, int64_t ldy,
                                                         int64_t n, int accumulate) {
  extern __shared__ __align__(16) unsigned char vs_comb_smem[];
  c128* Ps = reinterpret_cast<c128*>(vs_comb_smem);
  for (int e = threadIdx.x; e < q * kc; e += blockDim.x) Ps[e] = P[e];
  __syncthreads();
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    c128 acc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = mk<double>(0, 0);
    for (int i = 0; i < q; ++i) {
      const cplx<T> xv = X[e + (int64_t)i * ldx];
      const c128 x = mk<double>((double)xv.x, (double)xv.y);
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if (c < kc) acc[c] = acc[c] + x * Ps[i + q * c];
    }
#pragma unroll
    for (int c = 0; c < 8; ++c)
      if (c < kc) {
        cplx<T>* yp = Y + e + (int64_t)c * ldy;
        c128 v = acc[c];
        if (accumulate) v = v + mk<double>((double)yp->x, (double)yp->y);
        *yp = ccast<T>(v);
      }
  }
}

template <typename T>
void Engine<T>::combine(const cplx<T>* X, int64_t ldx, int q, const c128* P_host, int k, cplx<T>* Y, int64_t ldy,
                        bool accumulate) {
  VS_CHECK(q >= 1 && q <= 512 && k >= 0, VS_ERR_UNSUPPORTED, "basis combination of more than 512 vectors");
  if (k == 0) return;
  DevBuf dP(sizeof(c128) * (size_t)q * 8);
  std::vector<c128> chunk((size_t)q * 8);
  for (int j0 = 0; j0 < k; j0 += 8) {
    const int kc = std::min(8, k - j0);
    for (int c = 0; c < kc; ++c)
      for (int i = 0; i < q; ++i) chunk[i + (size_t)q * c] = P_host[i + (size_t)q * (j0 + c)];
    VS_HIP(hipMemcpyAsync(dP.p, chunk.data(), sizeof(c128) * (size_t)q * kc, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL((k_combine<T>), dim3(grid), dim3(RED_THREADS), sizeof(c128) * (size_t)q * kc, ctx->stream, X,
                       ldx, q, dP.as<c128>(), kc, Y + (int64_t)j0 * ldy, ldy, n, accumulate ? 1 : 0);
    VS_HIP(hipGetLastError());
    VS_HIP(hipStreamSynchronize(ctx->stream));  // the host chunk and dP are reused by the next chunk
  }
}

// ------------------------------------------------------------------------------------------------
// device-resident BiCGStab scalars: every scalar of an iteration is formed on the device, so a whole
// iteration runs without host round trips (the host reads a copy once per iteration, behind the next one)
// ------------------------------------------------------------------------------------------------
// BI_RHO: the batched kernels form rho in a stage of its own
enum BiStage : int { BI_INIT = 0, BI_ALPHA = 1, BI_HALF = 2, BI_OMEGA = 3, BI_FULL = 4, BI_RHO = 5 };

// MATLAB bicgstab's scalar logic of one stage, on the reduced sums tot[]
__device__ void bi_logic(int stage, const c128* tot, BiScal& S) {
  switch (stage) {
    case BI_INIT:
      S.rho_n = tot[0];
      return;
    case BI_ALPHA: {
      if (S.flag) return;
      // the start of the iteration: rho, beta and their breakdown checks
      c128 rho, beta;
      const int brk = bi_rho_beta(S, rho, beta);
      S.rho = rho;
      if (brk) { S.flag = 4; S.brk = brk; return; }
      S.beta = beta;
      const c128 rtv = tot[0];
      S.nph = sqrt(tot[1].x);
      S.nx = sqrt(tot[2].x);
      if (S.has_pc && !isfinite(S.nph)) { S.flag = 2; S.brk = 3; return; }
      if (cz(rtv) || cinf(rtv)) { S.flag = 4; S.brk = 3; return; }
      S.alpha = cdiv(S.rho, rtv);
      if (cinf(S.alpha)) { S.flag = 4; S.brk = 4; }
      return;
    }
    case BI_HALF:
      S.ns = sqrt(tot[0].x);
      return;
    case BI_OMEGA: {
      if (S.flag) return;
      const c128 tt = tot[0];
      S.nsh = sqrt(tot[2].x);
      S.nxh = sqrt(tot[3].x);
      if (S.has_pc && !isfinite(S.nsh)) { S.flag = 2; S.brk = 5; return; }
      if (cz(tt) || cinf(tt)) { S.flag = 4; S.brk = 5; return; }
      S.omega = cdiv(tot[1], tt);
      if (cinf(S.omega)) { S.flag = 4; S.brk = 6; }
      return;
    }
    default:  // BI_FULL
      S.nr = sqrt(tot[0].x);
      if (!S.flag) {
        S.rho1 = S.rho;
        S.first = 0;
        S.rho_n = tot[1];
      }
      return;
  }
}

// One wavefront per reduced quantity (K of `partial`, plus ||s|| of the half step when `part_half` is
// given): lane l sums blocks l, l + 64, ... in order, then a fixed shuffle tree; thread 0 then applies
// the scalar logic (the half step's first when it is folded in here).
__global__ __launch_bounds__(320) void k_bi_scalar(const c128* partial, int nblk, int K, const c128* part_half,
                                                   int stage, BiScal* sc) {
  __shared__ c128 tot[5];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const bool half = wave == K;
  if (wave < K || (half && part_half)) {
    const c128* pp = half ? part_half : partial;
    const int kk = half ? 0 : wave, KK = half ? 1 : K;
    double a = 0, b = 0;
    lane_sum(pp, nblk, KK, kk, lane, a, b);
    for (int off = 32; off >= 1; off >>= 1) {
      a += __shfl_xor(a, off);
      b += __shfl_xor(b, off);
    }
    if (lane == 0) tot[wave] = mk<double>(a, b);
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  if (part_half) bi_logic(BI_HALF, tot + K, *sc);
  bi_logic(stage, tot, *sc);
}

// K inner products; partial[blk * K + k]
template <typename T, int K>
__global__ __launch_bounds__(RED_THREADS) void k_bi_dots(DotArgs<T> args, int64_t n, c128* partial) {
  double re[K], im[K];
#pragma unroll
  for (int k = 0; k < K; ++k) re[k] = im[k] = 0.0;
  constexpr int U = K <= 2 ? 4 : 2;  // elements per thread whose loads are issued together
  for (int64_t i0 = (int64_t)blockIdx.x * RED_THREADS * U + threadIdx.x; i0 < n;
       i0 += (int64_t)gridDim.x * RED_THREADS * U) {
    cplx<T> a[U][K], b[U][K];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = min(i0 + u * RED_THREADS, n - 1);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        a[u][k] = args.a[k][i];
        b[u][k] = args.b[k][i];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (i0 + u * RED_THREADS >= n) break;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        re[k] += (double)a[u][k].x * b[u][k].x + (double)a[u][k].y * b[u][k].y;
        im[k] += (double)a[u][k].x * b[u][k].y - (double)a[u][k].y * b[u][k].x;
      }
    }
  }
  c128 tot[K];
  block_sum<K>(re, im, tot);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) partial[blockIdx.x * K + k] = tot[k];
}

// p = r (first iteration) or p = r + beta (p0 - omega v0); beta and its breakdown checks are evaluated by
// every workgroup from the same scalars (the alpha stage records them)
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_bi_p(const cplx<T>* r, const cplx<T>* v0, const cplx<T>* p0,
                                                      cplx<T>* p, int64_t n, const BiScal* sc) {
  const BiScal& S = *sc;
  if (S.flag) return;
  c128 rho, beta;
  if (bi_rho_beta(S, rho, beta)) return;
  const bool first = S.first != 0;
  const cplx<T> be = ccast<T>(beta), om = ccast<T>(S.omega);
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_THREADS)
    p[i] = first ? r[i] : r[i] + be * (p0[i] - om * v0[i]);
}

// y = x0 + c x1 ; z = x2 - c x3 with c = alpha (half step) or omega (full step); partials of ||z||^2 and,
// for the full step, of conj(rt) . z (the next iteration's rho): partial[blk * K + k], K = 1 + full
// partial_y (nullable): partials of ||y||^2 as well (the fused solve's ||x|| for the next alpha stage)
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_bi_axpy2(const cplx<T>* x0, const cplx<T>* x1, const cplx<T>* x2,
                                                          const cplx<T>* x3, cplx<T>* y, cplx<T>* z, const cplx<T>* rt,
                                                          int64_t n, const BiScal* sc, int full, c128* partial,
                                                          c128* partial_y) {
  double re[3] = {0.0, 0.0, 0.0}, im[3] = {0.0, 0.0, 0.0};
  if (!sc->flag) {
    const cplx<T> c = ccast<T>(full ? sc->omega : sc->alpha);
    constexpr int U = 4;  // elements per thread whose loads are issued together
    for (int64_t i0 = (int64_t)blockIdx.x * RED_THREADS * U + threadIdx.x; i0 < n;
         i0 += (int64_t)gridDim.x * RED_THREADS * U) {
      cplx<T> q0[U], q1[U], q2[U], q3[U], qr[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t i = min(i0 + u * RED_THREADS, n - 1);
        q0[u] = x0[i];
        q1[u] = x1[i];
        q2[u] = x2[i];
        q3[u] = x3[i];
        if (full) qr[u] = rt[i];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t i = i0 + u * RED_THREADS;
        if (i >= n) break;
        const cplx<T> yv = q0[u] + c * q1[u];
        y[i] = yv;
        const cplx<T> o = q2[u] - c * q3[u];
        z[i] = o;
        re[0] += (double)o.x * o.x + (double)o.y * o.y;
        if (full) {
          const cplx<T> a = qr[u];
          re[1] += (double)a.x * o.x + (double)a.y * o.y;
          im[1] += (double)a.x * o.y - (double)a.y * o.x;
        }
        if (partial_y) re[2] += (double)yv.x * yv.x + (double)yv.y * yv.y;
      }
    }
  }
  c128 tot[3];
  block_sum<3>(re, im, tot);
  if (threadIdx.x == 0) {
    partial[blockIdx.x * (1 + full)] = tot[0];
    if (full) partial[blockIdx.x * 2 + 1] = tot[1];
    if (partial_y) partial_y[blockIdx.x] = tot[2];
  }
}

// Reduced sums from several partial buffers (one wavefront per sum: lane l adds partials l, l + 64, ... in
// order, then a fixed shuffle tree), then the scalar logic of `stage`; half >= 0 names the sum that is the
// half step's ||s||^2, applied first (folded into the omega stage).
struct BiSrc {
  const c128* p[6];
  int nblk[6], K[6], k[6];
  int n, half;
};
__global__ __launch_bounds__(384) void k_bi_scalar_src(BiSrc src, int stage, BiScal* sc) {
  __shared__ c128 tot[6];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (wave < src.n) {
    const c128* pp = src.p[wave];
    const int KK = src.K[wave], kk = src.k[wave], nb = src.nblk[wave];
    double a = 0, b = 0;
    lane_sum(pp, nb, KK, kk, lane, a, b);
    for (int off = 32; off >= 1; off >>= 1) {
      a += __shfl_xor(a, off);
      b += __shfl_xor(b, off);
    }
    if (lane == 0) tot[wave] = mk<double>(a, b);
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  if (src.half >= 0) bi_logic(BI_HALF, tot + src.half, *sc);
  bi_logic(stage, tot, *sc);
}

// partial buffers of the stages: the half step's ||s|| partials live apart so that they can be reduced
// together with the omega stage's
static c128* part_main(vs_ctx* ctx) { return ctx->red.template as<c128>(); }
static c128* part_half(vs_ctx* ctx) { return ctx->red.template as<c128>() + 4 * 1024; }

static void launch_scalar(vs_ctx* ctx, const c128* part, int nblk, int K, const c128* ph, int stage, BiScal* sc) {
  hipLaunchKernelGGL(k_bi_scalar, dim3(1), dim3(64 * (K + (ph ? 1 : 0))), 0, ctx->stream, part, nblk, K, ph, stage,
                     sc);
}

__global__ void k_bi_set(BiScal* sc, BiScal h, int flag_only) {
  if (flag_only)
    sc->flag = 0;
  else
    *sc = h;
}

void bi_set_scalars(vs_ctx* ctx, BiScal* sc, const BiScal& h) {
  hipLaunchKernelGGL(k_bi_set, dim3(1), dim3(1), 0, ctx->stream, sc, h, 0);
  VS_HIP(hipGetLastError());
}

void bi_clear_flag(vs_ctx* ctx, BiScal* sc) {
  hipLaunchKernelGGL(k_bi_set, dim3(1), dim3(1), 0, ctx->stream, sc, BiScal{}, 1);
  VS_HIP(hipGetLastError());
}

template <typename T>
void bi_stage_init(Engine<T>& E, const cplx<T>* rt, const cplx<T>* r, BiScal* sc) {
  DotArgs<T> a{};
  a.a[0] = rt;
  a.b[0] = r;
  hipLaunchKernelGGL((k_bi_dots<T, 1>), dim3(E.grid), dim3(RED_THREADS), 0, E.ctx->stream, a, E.n, part_main(E.ctx));
  launch_scalar(E.ctx, part_main(E.ctx), E.grid, 1, nullptr, (int)BI_INIT, sc);
  VS_HIP(hipGetLastError());
}

// start of a BiCGStab solve from x0 = 0: x = 0, r = rt = b (and the minimum-residual iterate zeroed), with
// per-workgroup partials of conj(rt) . r = ||b||^2 and (px nullable) zero partials of ||x||^2
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_bi_start(const cplx<T>* __restrict__ b, cplx<T>* __restrict__ x,
                                                          cplx<T>* __restrict__ r, cplx<T>* __restrict__ rt,
                                                          cplx<T>* __restrict__ xmin, int64_t n, c128* part,
                                                          c128* px) {
  double re[1] = {0.0}, im[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_THREADS) {
    const cplx<T> v = b[i];
    x[i] = mk<T>(0, 0);
    r[i] = v;
    rt[i] = v;
    if (xmin) xmin[i] = mk<T>(0, 0);
    re[0] += (double)v.x * v.x + (double)v.y * v.y;
  }
  c128 tot[1];
  block_sum<1>(re, im, tot);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = tot[0];
    if (px) px[blockIdx.x] = mk<double>(0, 0);
  }
}

template <typename T>
void bi_start_dev(Engine<T>& E, const cplx<T>* b, cplx<T>* x, cplx<T>* r, cplx<T>* rt, cplx<T>* xmin, BiParts* P) {
  hipLaunchKernelGGL((k_bi_start<T>), dim3(E.grid), dim3(RED_THREADS), 0, E.ctx->stream, b, x, r, rt, xmin, E.n,
                     part_main(E.ctx), P ? P->x : (c128*)nullptr);
  if (P) P->n_x = E.grid;
  VS_HIP(hipGetLastError());
}

// *sc = h, then rho_n from the start pass's partials (the reduction of k_bi_scalar, one wavefront)
__global__ __launch_bounds__(64) void k_bi_set_init(const c128* partial, int nblk, BiScal h, BiScal* sc) {
  __shared__ c128 tot[1];
  const int lane = threadIdx.x & 63;
  double a = 0, b = 0;
  lane_sum(partial, nblk, 1, 0, lane, a, b);
  for (int off = 32; off >= 1; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
  }
  if (lane == 0) {
    tot[0] = mk<double>(a, b);
    BiScal S = h;
    bi_logic(BI_INIT, tot, S);
    *sc = S;
  }
}

template <typename T>
void bi_set_init_scalars(Engine<T>& E, BiScal* sc, const BiScal& h) {
  hipLaunchKernelGGL(k_bi_set_init, dim3(1), dim3(64), 0, E.ctx->stream, part_main(E.ctx), E.grid, h, sc);
  VS_HIP(hipGetLastError());
}

template void bi_start_dev<float>(Engine<float>&, const c64*, c64*, c64*, c64*, c64*, BiParts*);
template void bi_start_dev<double>(Engine<double>&, const c128*, c128*, c128*, c128*, c128*, BiParts*);
template void bi_set_init_scalars<float>(Engine<float>&, BiScal*, const BiScal&);
template void bi_set_init_scalars<double>(Engine<double>&, BiScal*, const BiScal&);

template <typename T>
void bi_stage_rho(Engine<T>& E, const cplx<T>* r, const cplx<T>* v0, const cplx<T>* p0, cplx<T>* p, BiScal* sc) {
  hipLaunchKernelGGL((k_bi_p<T>), dim3(E.grid), dim3(RED_THREADS), 0, E.ctx->stream, r, v0, p0, p, E.n, sc);
  VS_HIP(hipGetLastError());
}

template <typename T>
void bi_stage_alpha_half(Engine<T>& E, const cplx<T>* rt, const cplx<T>* v, const cplx<T>* ph, const cplx<T>* x,
                         const cplx<T>* r, cplx<T>* xh, cplx<T>* s, BiScal* sc, bool finish_half) {
  DotArgs<T> a{};
  a.a[0] = rt;
  a.b[0] = v;
  a.a[1] = ph;
  a.b[1] = ph;
  a.a[2] = x;
  a.b[2] = x;
  vs_ctx* ctx = E.ctx;
  hipLaunchKernelGGL((k_bi_dots<T, 3>), dim3(E.grid), dim3(RED_THREADS), 0, ctx->stream, a, E.n, part_main(ctx));
  launch_scalar(ctx, part_main(ctx), E.grid, 3, nullptr, (int)BI_ALPHA, sc);
  hipLaunchKernelGGL((k_bi_axpy2<T>), dim3(E.grid), dim3(RED_THREADS), 0, ctx->stream, x, ph, r, v, xh, s,
                     (const cplx<T>*)nullptr, E.n, sc, 0, part_half(ctx), (c128*)nullptr);
  if (finish_half) launch_scalar(ctx, part_half(ctx), E.grid, 1, nullptr, (int)BI_HALF, sc);
  VS_HIP(hipGetLastError());
}

template <typename T>
void bi_stage_omega_full(Engine<T>& E, const cplx<T>* t, const cplx<T>* s, const cplx<T>* sh, const cplx<T>* xh,
                         const cplx<T>* rt, cplx<T>* x, cplx<T>* r, BiScal* sc, bool fold_half) {
  DotArgs<T> a{};
  a.a[0] = t;
  a.b[0] = t;
  a.a[1] = t;
  a.b[1] = s;
  a.a[2] = sh;
  a.b[2] = sh;
  a.a[3] = xh;
  a.b[3] = xh;
  vs_ctx* ctx = E.ctx;
  hipLaunchKernelGGL((k_bi_dots<T, 4>), dim3(E.grid), dim3(RED_THREADS), 0, ctx->stream, a, E.n, part_main(ctx));
  launch_scalar(ctx, part_main(ctx), E.grid, 4, fold_half ? part_half(ctx) : nullptr, (int)BI_OMEGA, sc);
  hipLaunchKernelGGL((k_bi_axpy2<T>), dim3(E.grid), dim3(RED_THREADS), 0, ctx->stream, xh, sh, s, t, x, r, rt, E.n,
                     sc, 1, part_main(ctx), (c128*)nullptr);
  launch_scalar(ctx, part_main(ctx), E.grid, 2, nullptr, (int)BI_FULL, sc);
  VS_HIP(hipGetLastError());
}

// The fused halves' partial buffers: six regions of the context's reduction area past the stages' own
// (each region holds up to `cap` partials of two values).
template <typename T>
void bi_parts_take(Engine<T>& E, BiParts& P) {
  constexpr int64_t base = 16384, cap = 32768;
  static_assert(base + 6 * 2 * cap <= 1024 * 512, "fused partials exceed the reduction area");
  c128* r = part_main(E.ctx);
  P.ph = r + base;
  P.v = P.ph + 2 * cap;
  P.sx = P.v + 2 * cap;
  P.sh = P.sx + 2 * cap;
  P.t = P.sh + 2 * cap;
  P.x = P.t + 2 * cap;
  P.cap = (int)cap;
}

template <typename T>
void bi_stage_xnorm(Engine<T>& E, const cplx<T>* x, BiParts& P) {
  DotArgs<T> a{};
  a.a[0] = x;
  a.b[0] = x;
  hipLaunchKernelGGL((k_bi_dots<T, 1>), dim3(E.grid), dim3(RED_THREADS), 0, E.ctx->stream, a, E.n, P.x);
  P.n_x = E.grid;
  VS_HIP(hipGetLastError());
}

static void src_add(BiSrc& s, const c128* p, int nblk, int K, int k) {
  s.p[s.n] = p;
  s.nblk[s.n] = nblk;
  s.K[s.n] = K;
  s.k[s.n] = k;
  ++s.n;
}

template <typename T>
void bi_stage_alpha_fused(Engine<T>& E, BiScal* sc, const BiParts& P) {
  BiSrc s{};
  s.half = -1;
  src_add(s, P.v, P.n_v, 1, 0);
  src_add(s, P.ph, P.n_ph, 1, 0);
  src_add(s, P.x, P.n_x, 1, 0);
  hipLaunchKernelGGL(k_bi_scalar_src, dim3(1), dim3(64 * s.n), 0, E.ctx->stream, s, (int)BI_ALPHA, sc);
  VS_HIP(hipGetLastError());
}

template <typename T>
void bi_stage_omega_full_fused(Engine<T>& E, const cplx<T>* t, const cplx<T>* s, const cplx<T>* sh, const cplx<T>* xh,
                               const cplx<T>* rt, cplx<T>* x, cplx<T>* r, BiScal* sc, BiParts& P) {
  BiSrc q{};
  src_add(q, P.t, P.n_t, 2, 0);    // ||t||^2
  src_add(q, P.t, P.n_t, 2, 1);    // conj(t) . s
  src_add(q, P.sh, P.n_sh, 1, 0);  // ||sh||^2
  src_add(q, P.sx, P.n_sx, 2, 1);  // ||xh||^2
  src_add(q, P.sx, P.n_sx, 2, 0);  // ||s||^2 of the half step
  q.half = 4;
  vs_ctx* ctx = E.ctx;
  hipLaunchKernelGGL(k_bi_scalar_src, dim3(1), dim3(64 * q.n), 0, ctx->stream, q, (int)BI_OMEGA, sc);
  hipLaunchKernelGGL((k_bi_axpy2<T>), dim3(E.grid), dim3(RED_THREADS), 0, ctx->stream, xh, sh, s, t, x, r, rt, E.n,
                     sc, 1, part_main(ctx), P.x);
  P.n_x = E.grid;
  launch_scalar(ctx, part_main(ctx), E.grid, 2, nullptr, (int)BI_FULL, sc);
  VS_HIP(hipGetLastError());
}

#define VS_BI_INST(T)                                                                                             \
  template void bi_stage_init<T>(Engine<T>&, const cplx<T>*, const cplx<T>*, BiScal*);                            \
  template void bi_stage_rho<T>(Engine<T>&, const cplx<T>*, const cplx<T>*, const cplx<T>*, cplx<T>*, BiScal*);   \
  template void bi_stage_alpha_half<T>(Engine<T>&, const cplx<T>*, const cplx<T>*, const cplx<T>*, const cplx<T>*, \
                                       const cplx<T>*, cplx<T>*, cplx<T>*, BiScal*, bool);                       \
  template void bi_stage_omega_full<T>(Engine<T>&, const cplx<T>*, const cplx<T>*, const cplx<T>*, const cplx<T>*, \
                                       const cplx<T>*, cplx<T>*, cplx<T>*, BiScal*, bool);                       \
  template void bi_parts_take<T>(Engine<T>&, BiParts&);                                                           \
  template void bi_stage_xnorm<T>(Engine<T>&, const cplx<T>*, BiParts&);                                          \
  template void bi_stage_alpha_fused<T>(Engine<T>&, BiScal*, const BiParts&);                                     \
  template void bi_stage_omega_full_fused<T>(Engine<T>&, const cplx<T>*, const cplx<T>*, const cplx<T>*,          \
                                             const cplx<T>*, const cplx<T>*, cplx<T>*, cplx<T>*, BiScal*, BiParts&);
VS_BI_INST(float)
VS_BI_INST(double)
#undef VS_BI_INST


// ------------------------------------------------------------------------------------------------
// batched BiCGStab kernels: nrhs independent systems sharing every launch (blockIdx.y = right-hand side)
// ------------------------------------------------------------------------------------------------
template <typename T, int K>
__global__ __launch_bounds__(RED_THREADS) void k_dots_b(DotArgs<T> args, int64_t n, c128* partial) {
  const int64_t off = (int64_t)blockIdx.y * n;
  double re[K], im[K];
#pragma unroll
  for (int k = 0; k < K; ++k) re[k] = im[k] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_THREADS) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const cplx<T> a = args.a[k][off + i], b = args.b[k][off + i];
      re[k] += (double)a.x * b.x + (double)a.y * b.y;
      im[k] += (double)a.x * b.y - (double)a.y * b.x;
    }
  }
  c128 tot[K];
  block_sum<K>(re, im, tot);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) partial[((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * K + k] = tot[k];
}

// one workgroup per right-hand side
__global__ __launch_bounds__(RED_THREADS) void k_bi_scalar_b(const c128* partial, int nblk, int K, int stage,
                                                             BiScal* scs) {
  __shared__ double sre[RED_THREADS], sim[RED_THREADS];
  __shared__ c128 tot[4];
  const int rhs = blockIdx.x;
  const c128* pp = partial + (int64_t)rhs * nblk * K;
  for (int k = 0; k < K; ++k) {
    double a = 0, b = 0;
    for (int i = threadIdx.x; i < nblk; i += RED_THREADS) {
      a += pp[i * K + k].x;
      b += pp[i * K + k].y;
    }
    sre[threadIdx.x] = a;
    sim[threadIdx.x] = b;
    __syncthreads();
    for (int st = RED_THREADS / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st) {
        sre[threadIdx.x] += sre[threadIdx.x + st];
        sim[threadIdx.x] += sim[threadIdx.x + st];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) tot[k] = mk<double>(sre[0], sim[0]);
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  BiScal& S = scs[rhs];
  if (!S.active) return;
  switch (stage) {
    case BI_RHO:
      if (S.flag) return;
      S.rho = tot[0];
      if (cz(S.rho) || cinf(S.rho)) { S.flag = 4; S.brk = 1; return; }
      if (!S.first) {
        S.beta = cdiv(S.rho, S.rho1) * cdiv(S.alpha, S.omega);
        if (cz(S.beta) || !cfin(S.beta)) { S.flag = 4; S.brk = 2; }
      }
      return;
    case BI_ALPHA: {
      if (S.flag) return;
      const c128 rtv = tot[0];
      if (cz(rtv) || cinf(rtv)) { S.flag = 4; S.brk = 3; return; }
      S.alpha = cdiv(S.rho, rtv);
      if (cinf(S.alpha)) { S.flag = 4; S.brk = 4; }
      return;
    }
    case BI_HALF:
      S.ns = sqrt(tot[0].x);
      return;
    case BI_OMEGA: {
      if (S.flag) return;
      const c128 tt = tot[0];
      if (cz(tt) || cinf(tt)) { S.flag = 4; S.brk = 5; return; }
      S.omega = cdiv(tot[1], tt);
      if (cinf(S.omega)) { S.flag = 4; S.brk = 6; }
      return;
    }
    default:
      S.nr = sqrt(tot[0].x);
      if (!S.flag) {
        S.rho1 = S.rho;
        S.first = 0;
      }
      return;
  }
}

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_bi_p_b(const cplx<T>* r, const cplx<T>* v, cplx<T>* p, int64_t n,
                                                        const BiScal* scs) {
  const BiScal& S = scs[blockIdx.y];
  if (!S.active || S.flag) return;
  const int64_t off = (int64_t)blockIdx.y * n;
  const bool first = S.first != 0;
  const cplx<T> be = ccast<T>(S.beta), om = ccast<T>(S.omega);
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_THREADS)
    p[off + i] = first ? r[off + i] : r[off + i] + be * (p[off + i] - om * v[off + i]);
}

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_bi_axpy2_b(const cplx<T>* x0, const cplx<T>* x1, const cplx<T>* x2,
                                                            const cplx<T>* x3, cplx<T>* y, cplx<T>* z, int64_t n,
                                                            const BiScal* scs, int use_omega, c128* partial) {
  const BiScal& S = scs[blockIdx.y];
  const int64_t off = (int64_t)blockIdx.y * n;
  double re[1] = {0.0}, im[1] = {0.0};
  if (S.active && !S.flag) {
    const cplx<T> c = ccast<T>(use_omega ? S.omega : S.alpha);
    for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_THREADS) {
      y[off + i] = x0[off + i] + c * x1[off + i];
      const cplx<T> o = x2[off + i] - c * x3[off + i];
      z[off + i] = o;
      re[0] += (double)o.x * o.x + (double)o.y * o.y;
    }
  }
  c128 tot[1];
  block_sum<1>(re, im, tot);
  if (threadIdx.x == 0) partial[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = tot[0];
}

template <typename T>
void bib_stage_rho(Engine<T>& E, int nrhs, const cplx<T>* rt, const cplx<T>* r, const cplx<T>* v, cplx<T>* p,
                   BiScal* sc, c128* part) {
  DotArgs<T> a{};
  a.a[0] = rt;
  a.b[0] = r;
  hipLaunchKernelGGL((k_dots_b<T, 1>), dim3(E.grid, nrhs), dim3(RED_THREADS), 0, E.ctx->stream, a, E.n, part);
  hipLaunchKernelGGL(k_bi_scalar_b, dim3(nrhs), dim3(RED_THREADS), 0, E.ctx->stream, part, E.grid, 1, (int)BI_RHO, sc);
  hipLaunchKernelGGL((k_bi_p_b<T>), dim3(E.grid, nrhs), dim3(RED_THREADS), 0, E.ctx->stream, r, v, p, E.n, sc);
  VS_HIP(hipGetLastError());
}

template <typename T>
void bib_stage_alpha_half(Engine<T>& E, int nrhs, const cplx<T>* rt, const cplx<T>* v, const cplx<T>* ph,
                          const cplx<T>* x, const cplx<T>* r, cplx<T>* xh, cplx<T>* s, BiScal* sc, c128* part) {
  DotArgs<T> a{};
  a.a[0] = rt;
  a.b[0] = v;
  hipLaunchKernelGGL((k_dots_b<T, 1>), dim3(E.grid, nrhs), dim3(RED_THREADS), 0, E.ctx->stream, a, E.n, part);
  hipLaunchKernelGGL(k_bi_scalar_b, dim3(nrhs), dim3(RED_THREADS), 0, E.ctx->stream, part, E.grid, 1, (int)BI_ALPHA,
                     sc);
  hipLaunchKernelGGL((k_bi_axpy2_b<T>), dim3(E.grid, nrhs), dim3(RED_THREADS), 0, E.ctx->stream, x, ph, r, v, xh, s,
                     E.n, sc, 0, part);
  hipLaunchKernelGGL(k_bi_scalar_b, dim3(nrhs), dim3(RED_THREADS), 0, E.ctx->stream, part, E.grid, 1, (int)BI_HALF, sc);
  VS_HIP(hipGetLastError());
}

template <typename T>
void bib_stage_omega_full(Engine<T>& E, int nrhs, const cplx<T>* t, const cplx<T>* s, const cplx<T>* sh,
                          const cplx<T>* xh, cplx<T>* x, cplx<T>* r, BiScal* sc, c128* part) {
  DotArgs<T> a{};
  a.a[0] = t;
  a.b[0] = t;
  a.a[1] = t;
  a.b[1] = s;
  hipLaunchKernelGGL((k_dots_b<T, 2>), dim3(E.grid, nrhs), dim3(RED_THREADS), 0, E.ctx->stream, a, E.n, part);
  hipLaunchKernelGGL(k_bi_scalar_b, dim3(nrhs), dim3(RED_THREADS), 0, E.ctx->stream, part, E.grid, 2, (int)BI_OMEGA,
                     sc);
  hipLaunchKernelGGL((k_bi_axpy2_b<T>), dim3(E.grid, nrhs), dim3(RED_THREADS), 0, E.ctx->stream, xh, sh, s, t, x, r,
                     E.n, sc, 1, part);
  hipLaunchKernelGGL(k_bi_scalar_b, dim3(nrhs), dim3(RED_THREADS), 0, E.ctx->stream, part, E.grid, 1, (int)BI_FULL, sc);
  VS_HIP(hipGetLastError());
}

// ||x_i||^2 of nrhs vectors into out_dev[i] (deterministic)
template <typename T>
void batch_norm2(Engine<T>& E, int nrhs, const cplx<T>* x, c128* part, c128* out_dev) {
  DotArgs<T> a{};
  a.a[0] = x;
  a.b[0] = x;
  hipLaunchKernelGGL((k_dots_b<T, 1>), dim3(E.grid, nrhs), dim3(RED_THREADS), 0, E.ctx->stream, a, E.n, part);
  for (int i = 0; i < nrhs; ++i)
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(RED_THREADS), 0, E.ctx->stream, part + (int64_t)i * E.grid,
                       E.grid, 1, out_dev + i);
  VS_HIP(hipGetLastError());
}

#define VS_BIB_INST(T)                                                                                             \
  template void bib_stage_rho<T>(Engine<T>&, int, const cplx<T>*, const cplx<T>*, const cplx<T>*, cplx<T>*, BiScal*, \
                                 c128*);                                                                           \
  template void bib_stage_alpha_half<T>(Engine<T>&, int, const cplx<T>*, const cplx<T>*, const cplx<T>*,           \
                                        const cplx<T>*, const cplx<T>*, cplx<T>*, cplx<T>*, BiScal*, c128*);       \
  template void bib_stage_omega_full<T>(Engine<T>&, int, const cplx<T>*, const cplx<T>*, const cplx<T>*,           \
                                        const cplx<T>*, cplx<T>*, cplx<T>*, BiScal*, c128*);                       \
  template void batch_norm2<T>(Engine<T>&, int, const cplx<T>*, c128*, c128*);
VS_BIB_INST(float)
VS_BIB_INST(double)
#undef VS_BIB_INST


// per-right-hand-side scaling y_i = c_i x_i and y_i += c_i x_i (c on the device, FP64)
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_scale_b(const cplx<T>* x, cplx<T>* y, int64_t n, const double* c,
                                                         int accumulate) {
  const int64_t off = (int64_t)blockIdx.y * n;
  const T s = (T)c[blockIdx.y];
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_THREADS)
    y[off + i] = accumulate ? y[off + i] + s * x[off + i] : s * x[off + i];
}

template <typename T>
void batch_scale(Engine<T>& E, int nrhs, const cplx<T>* x, cplx<T>* y, const double* c_dev, bool accumulate) {
  hipLaunchKernelGGL((k_scale_b<T>), dim3(E.grid, nrhs), dim3(RED_THREADS), 0, E.ctx->stream, x, y, E.n, c_dev,
                     accumulate ? 1 : 0);
  VS_HIP(hipGetLastError());
}
template void batch_scale<float>(Engine<float>&, int, const c64*, c64*, const double*, bool);
template void batch_scale<double>(Engine<double>&, int, const c128*, c128*, const double*, bool);

// y = x - y over nrhs * n values (r = b - A x for a whole batch)
template <typename T>
__global__ void k_sub_all(const cplx<T>* x, cplx<T>* y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = x[i] - y[i];
}
template <typename T>
void batch_sub(Engine<T>& E, int nrhs, const cplx<T>* x, cplx<T>* y) {
  hipLaunchKernelGGL((k_sub_all<T>), dim3(2048), dim3(256), 0, E.ctx->stream, x, y, E.n * nrhs);
  VS_HIP(hipGetLastError());
}
template void batch_sub<float>(Engine<float>&, int, const c64*, c64*);
template void batch_sub<double>(Engine<double>&, int, const c128*, c128*);

template struct Engine<float>;
template struct Engine<double>;

// out = (FP32) (s * in), in FP64 (the refinement's first scaled residual, r = b)
__global__ __launch_bounds__(256) void k_scale_to32(const c128* __restrict__ in, c64* __restrict__ out, int64_t n,
                                                    double s) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const c128 v = in[i];
    out[i] = mk<float>((float)(s * v.x), (float)(s * v.y));
  }
}

// x = a d (first) or x += a d, d in FP32 (the refinement's update with the inner correction)
__global__ __launch_bounds__(256) void k_axpy_from32(const c64* __restrict__ d, c128* __restrict__ x, int64_t n,
                                                     double a, int first) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const c64 v = d[i];
    const c128 dv = mk<double>(a * (double)v.x, a * (double)v.y);
    x[i] = first ? dv : x[i] + dv;
  }
}

void scale_to32_dev(vs_ctx* ctx, const c128* in, c64* out, int64_t n, double s) {
  hipLaunchKernelGGL(k_scale_to32, dim3(1024), dim3(256), 0, ctx->stream, in, out, n, s);
  VS_HIP(hipGetLastError());
}

void axpy_from32_dev(vs_ctx* ctx, const c64* d, c128* x, int64_t n, double a, bool first) {
  hipLaunchKernelGGL(k_axpy_from32, dim3(1024), dim3(256), 0, ctx->stream, d, x, n, a, first ? 1 : 0);
  VS_HIP(hipGetLastError());
}

void convert_dev(vs_ctx* ctx, const void* in, int pin, void* out, int pout, int64_t n) {
  if (pin == pout) {
    VS_HIP(hipMemcpyAsync(out, in, (pin == VS_FP64 ? sizeof(c128) : sizeof(c64)) * n, hipMemcpyDeviceToDevice,
                          ctx->stream));
    return;
  }
  if (pin == VS_FP64)
    hipLaunchKernelGGL((k_cvt<float, double>), dim3(1024), dim3(256), 0, ctx->stream, (const c128*)in, (c64*)out, n);
  else
    hipLaunchKernelGGL((k_cvt<double, float>), dim3(1024), dim3(256), 0, ctx->stream, (const c64*)in, (c128*)out, n);
  VS_HIP(hipGetLastError());
}

}  // namespace vs
