// Mixed-radix Stockham FFT on a tile of lines held in LDS (device side).
//
// A tile holds `nl` lines of length n; element (pos, line) lives at buf[pos * ld + line], so at a
// fixed position the lines are consecutive and a wavefront walking over lines hits consecutive banks.
// Each stage of radix R reads the R inputs of every butterfly into registers, synchronises, and writes
// the R outputs in place (the autosort Stockham index map of Govindaraju et al., SC'08):
//   inputs  j + q n/R            (q < R), twiddled by w_{Ns R}^{(j mod Ns) q}
//   outputs (j/Ns) Ns R + (j mod Ns) + r Ns
// Radices 2, 3, 4, 5, 8 have register butterflies; any other factor uses a direct DFT stage.
// All transforms are unnormalised; `inv` selects the +i exponent.
#pragma once

#include "vs_internal.hpp"

namespace vs {

template <typename T>
__device__ __forceinline__ cplx<T> mul_mi(cplx<T> a, bool inv) {  // a * (-i) forward, a * (+i) inverse
  return inv ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}

template <typename T, int R>
struct Bfly;

template <typename T>
struct Bfly<T, 2> {
  static __device__ __forceinline__ void run(cplx<T>* a, bool) {
    cplx<T> t = a[0];
    a[0] = t + a[1];
    a[1] = t - a[1];
  }
};

template <typename T>
struct Bfly<T, 4> {
  static __device__ __forceinline__ void run(cplx<T>* a, bool inv) {
    cplx<T> t0 = a[0] + a[2], t1 = a[0] - a[2], t2 = a[1] + a[3];
    cplx<T> t3 = mul_mi(a[1] - a[3], inv);
    a[0] = t0 + t2;
    a[2] = t0 - t2;
    a[1] = t1 + t3;
    a[3] = t1 - t3;
  }
};

template <typename T>
struct Bfly<T, 3> {
  static __device__ __forceinline__ void run(cplx<T>* a, bool inv) {
    const T h = (T)0.86602540378443864676372317075294;  // sqrt(3)/2
    cplx<T> s = a[1] + a[2], d = a[1] - a[2];
    cplx<T> m = a[0] - (T)0.5 * s;
    cplx<T> rd = mul_mi(h * d, inv);  // -i h d (forward)
    a[0] = a[0] + s;
    a[1] = m + rd;
    a[2] = m - rd;
  }
};

template <typename T>
struct Bfly<T, 5> {
  static __device__ __forceinline__ void run(cplx<T>* a, bool inv) {
    const T c1 = (T)0.30901699437494742410229341718282;   // cos(2pi/5)
    const T c2 = (T)-0.80901699437494742410229341718282;  // cos(4pi/5)
    const T s1 = (T)0.95105651629515357211643933337938;   // sin(2pi/5)
    const T s2 = (T)0.58778525229247312916870595463907;   // sin(4pi/5)
    cplx<T> S1 = a[1] + a[4], D1 = a[1] - a[4], S2 = a[2] + a[3], D2 = a[2] - a[3];
    cplx<T> m1 = a[0] + c1 * S1 + c2 * S2;
    cplx<T> m2 = a[0] + c2 * S1 + c1 * S2;
    cplx<T> r1 = mul_mi(s1 * D1 + s2 * D2, inv);
    cplx<T> r2 = mul_mi(s2 * D1 - s1 * D2, inv);
    a[0] = a[0] + S1 + S2;
    a[1] = m1 + r1;
    a[4] = m1 - r1;
    a[2] = m2 + r2;
    a[3] = m2 - r2;
  }
};

template <typename T>
struct Bfly<T, 8> {
  static __device__ __forceinline__ void run(cplx<T>* a, bool inv) {
    const T q = (T)0.70710678118654752440084436210485;
    cplx<T> u[4], v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      u[k] = a[k] + a[k + 4];
      v[k] = a[k] - a[k + 4];
    }
    // v[k] *= w8^k (forward w8 = (1-i)/sqrt2; inverse conj)
    if (!inv) {
      v[1] = mk<T>(q * (v[1].x + v[1].y), q * (v[1].y - v[1].x));
      v[2] = mk<T>(v[2].y, -v[2].x);
      v[3] = mk<T>(q * (v[3].y - v[3].x), -q * (v[3].x + v[3].y));
    } else {
      v[1] = mk<T>(q * (v[1].x - v[1].y), q * (v[1].x + v[1].y));
      v[2] = mk<T>(-v[2].y, v[2].x);
      v[3] = mk<T>(-q * (v[3].x + v[3].y), q * (v[3].x - v[3].y));
    }
    Bfly<T, 4>::run(u, inv);
    Bfly<T, 4>::run(v, inv);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[2 * k] = u[k];
      a[2 * k + 1] = v[k];
    }
  }
};

template <typename T>
__device__ __forceinline__ cplx<T> twiddle(const cplx<T>* tw, int e, bool inv) {
  cplx<T> w = tw[e];
  if (inv) w.y = -w.y;
  return w;
}

// one radix-R stage over the whole tile (all threads of the workgroup take part)
template <typename T, int R>
__device__ __forceinline__ void fft_stage(cplx<T>* buf, int ld, int nl, int n, int Ns, const cplx<T>* tw,
                                          bool inv) {
  constexpr int KM = (FFT_EMAX + R - 1) / R;
  const int nb = n / R;
  const int total = nb * nl;
  const int tstride = n / (Ns * R);
  const int tid = threadIdx.x;
  cplx<T> v[KM][R];
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const int b = tid + k * FFT_THREADS;
    if (b < total) {
      const int line = b % nl, j = b / nl;
      const int kk = j % Ns;
      v[k][0] = buf[j * ld + line];
#pragma unroll
      for (int r = 1; r < R; ++r) v[k][r] = buf[(j + r * nb) * ld + line] * twiddle(tw, kk * tstride * r, inv);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KM; ++k) {
    const int b = tid + k * FFT_THREADS;
    if (b < total) {
      const int line = b % nl, j = b / nl;
      const int kk = j % Ns;
      Bfly<T, R>::run(v[k], inv);
      const int idxD = (j / Ns) * Ns * R + kk;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(idxD + r * Ns) * ld + line] = v[k][r];
    }
  }
  __syncthreads();
}

// direct-DFT stage for any radix R (used for prime factors above 5)
template <typename T>
__device__ __noinline__ void fft_stage_generic(cplx<T>* buf, int ld, int nl, int n, int Ns, int R,
                                               const cplx<T>* tw, bool inv) {
  const int nb = n / R;
  const int total = n * nl;
  const int tid = threadIdx.x;
  const int tstep = n / (Ns * R), rstep = n / R;
  cplx<T> out[FFT_EMAX];
#pragma unroll
  for (int e = 0; e < FFT_EMAX; ++e) {
    const int o = tid + e * FFT_THREADS;
    if (o < total) {
      const int line = o % nl, pos = o / nl;
      const int kk = pos % Ns, t = pos / Ns;
      const int r = t % R, jq = t / R;
      const int j = jq * Ns + kk;
      const int m = kk * tstep + r * rstep;
      cplx<T> acc = mk<T>(0, 0);
      int ex = 0;
      for (int q = 0; q < R; ++q) {
        acc = acc + buf[(j + q * nb) * ld + line] * twiddle(tw, ex, inv);
        ex += m;
        if (ex >= n) ex -= n;
      }
      out[e] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < FFT_EMAX; ++e) {
    const int o = tid + e * FFT_THREADS;
    if (o < total) {
      const int line = o % nl, pos = o / nl;
      buf[pos * ld + line] = out[e];
    }
  }
  __syncthreads();
}

// full transform of the tile; callers synchronise before (tile loaded) -- this returns synchronised
template <typename T>
__device__ __forceinline__ void lds_fft(cplx<T>* buf, int ld, int nl, const FftDev<T>& p, bool inv) {
  int Ns = 1;
  for (int s = 0; s < p.nst; ++s) {
    const int R = p.rad[s];
    switch (R) {
      case 2: fft_stage<T, 2>(buf, ld, nl, p.n, Ns, p.tw, inv); break;
      case 3: fft_stage<T, 3>(buf, ld, nl, p.n, Ns, p.tw, inv); break;
      case 4: fft_stage<T, 4>(buf, ld, nl, p.n, Ns, p.tw, inv); break;
      case 5: fft_stage<T, 5>(buf, ld, nl, p.n, Ns, p.tw, inv); break;
      case 8: fft_stage<T, 8>(buf, ld, nl, p.n, Ns, p.tw, inv); break;
      default: fft_stage_generic<T>(buf, ld, nl, p.n, Ns, R, p.tw, inv); break;
    }
    Ns *= R;
  }
}

}  // namespace vs
