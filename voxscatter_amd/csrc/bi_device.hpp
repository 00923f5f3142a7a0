// Device-side scalar helpers of BiCGStab shared by the solver kernels and the fused FFT passes.
#pragma once

#include "solver.hpp"

namespace vs {

__device__ __forceinline__ bool cz(c128 a) { return a.x == 0.0 && a.y == 0.0; }
__device__ __forceinline__ bool cinf(c128 a) { return isinf(a.x) || isinf(a.y); }
__device__ __forceinline__ bool cfin(c128 a) { return isfinite(a.x) && isfinite(a.y); }

// rho of this iteration (conj(rt) . r, formed when r was) and beta; brk = 1 / 2 on a breakdown
__device__ __forceinline__ int bi_rho_beta(const BiScal& S, c128& rho, c128& beta) {
  rho = S.rho_n;
  beta = mk<double>(0, 0);
  if (cz(rho) || cinf(rho)) return 1;
  if (!S.first) {
    beta = cdiv(rho, S.rho1) * cdiv(S.alpha, S.omega);
    if (cz(beta) || !cfin(beta)) return 2;
  }
  return 0;
}

// Sums of K complex doubles held per thread over a workgroup of TH threads (a multiple of 64), in a fixed
// order (shuffle tree per wavefront, then wavefronts in order); thread 0 writes out[0 .. K-1].  Every
// thread of the workgroup must call it.
template <int TH, int K>
__device__ __forceinline__ void block_partials(double (&re)[K], double (&im)[K], c128* out) {
  __shared__ double sre[K][TH / 64], sim[K][TH / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double a = re[k], b = im[k];
    for (int off = 32; off >= 1; off >>= 1) {
      a += __shfl_xor(a, off);
      b += __shfl_xor(b, off);
    }
    if (lane == 0) {
      sre[k][wave] = a;
      sim[k][wave] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double a = 0, b = 0;
      for (int w = 0; w < TH / 64; ++w) {
        a += sre[k][w];
        b += sim[k][w];
      }
      out[k] = mk<double>(a, b);
    }
  }
}

}  // namespace vs
