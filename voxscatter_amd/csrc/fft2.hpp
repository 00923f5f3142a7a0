// Compile-time two-pass FFT of one line held by a group of threads (device side).
//
// A length-N line (N = R1 * R2, R1, R2 <= 16) is transformed by the four-step factorisation
//   n = n1 + R1 n2,  k = k2 + R2 k1
//   X[k2 + R2 k1] = sum_n1 w_R1^{n1 k1} w_N^{n1 k2} sum_n2 x[n1 + R1 n2] w_R2^{n2 k2}
// Step 1: thread n1 (< R1) loads x[n1 + R1 n2] (consecutive threads, consecutive addresses), does a
//         register DFT of size R2, applies the twiddles w_N^{n1 k2} and writes Y[n1][k2] to LDS.
// Step 2: thread k2 (< R2) reads Y[.][k2], does a register DFT of size R1 and owns X[k2 + R2 k1].
// One LDS exchange per transform; all index arithmetic is compile-time.  Register butterflies of size
// 6, 10, 12, 16 are themselves four-step compositions of the 2/3/4/5/8 butterflies.
#pragma once

#include "fft_device.hpp"

namespace vs {

// twiddle constant w_R^m = exp(-2 pi i m / R) (conjugated for the inverse); R in {6, 10, 12, 16}
template <typename T, int R>
__device__ __forceinline__ cplx<T> wconst(int m, bool inv) {
  // cos / sin of 2 pi m / R for the composite radices, m < R
  constexpr double c16[16] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                              0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                              -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
                              0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
  constexpr double s16[16] = {0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
                              1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                              0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                              -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173};
  constexpr double c12[12] = {1.0, 0.86602540378443864676, 0.5, 0.0, -0.5, -0.86602540378443864676,
                              -1.0, -0.86602540378443864676, -0.5, 0.0, 0.5, 0.86602540378443864676};
  constexpr double s12[12] = {0.0, 0.5, 0.86602540378443864676, 1.0, 0.86602540378443864676, 0.5,
                              0.0, -0.5, -0.86602540378443864676, -1.0, -0.86602540378443864676, -0.5};
  constexpr double c10[10] = {1.0, 0.80901699437494742410, 0.30901699437494742410, -0.30901699437494742410,
                              -0.80901699437494742410, -1.0, -0.80901699437494742410, -0.30901699437494742410,
                              0.30901699437494742410, 0.80901699437494742410};
  constexpr double s10[10] = {0.0, 0.58778525229247312917, 0.95105651629515357212, 0.95105651629515357212,
                              0.58778525229247312917, 0.0, -0.58778525229247312917, -0.95105651629515357212,
                              -0.95105651629515357212, -0.58778525229247312917};
  constexpr double c6[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
  constexpr double s6[6] = {0.0, 0.86602540378443864676, 0.86602540378443864676, 0.0,
                            -0.86602540378443864676, -0.86602540378443864676};
  double c, s;
  if (R == 16) { c = c16[m]; s = s16[m]; }
  else if (R == 12) { c = c12[m]; s = s12[m]; }
  else if (R == 10) { c = c10[m]; s = s10[m]; }
  else { c = c6[m]; s = s6[m]; }
  return mk<T>((T)c, inv ? (T)s : (T)-s);
}

template <typename T, int A, int B>
struct BflyComp {
  static __device__ __forceinline__ void run(cplx<T>* v, bool inv) {
    constexpr int R = A * B;
    cplx<T> y[A][B];
#pragma unroll
    for (int a = 0; a < A; ++a) {
      cplx<T> t[B];
#pragma unroll
      for (int b = 0; b < B; ++b) t[b] = v[a + A * b];
      Bfly<T, B>::run(t, inv);
#pragma unroll
      for (int kb = 0; kb < B; ++kb) y[a][kb] = (a == 0 || kb == 0) ? t[kb] : t[kb] * wconst<T, R>((a * kb) % R, inv);
    }
#pragma unroll
    for (int kb = 0; kb < B; ++kb) {
      cplx<T> t[A];
#pragma unroll
      for (int a = 0; a < A; ++a) t[a] = y[a][kb];
      Bfly<T, A>::run(t, inv);
#pragma unroll
      for (int ka = 0; ka < A; ++ka) v[kb + B * ka] = t[ka];
    }
  }
};

template <typename T> struct Bfly<T, 1> { static __device__ __forceinline__ void run(cplx<T>*, bool) {} };
template <typename T> struct Bfly<T, 6> : BflyComp<T, 3, 2> {};
template <typename T> struct Bfly<T, 10> : BflyComp<T, 5, 2> {};
template <typename T> struct Bfly<T, 12> : BflyComp<T, 4, 3> {};
template <typename T> struct Bfly<T, 16> : BflyComp<T, 4, 4> {};

// radix split of the supported lengths (R1 >= R2, both in {1..6, 8, 10, 12, 16})
template <int N> struct Split { static constexpr int R1 = 0, R2 = 0; };
#define VS_SPLIT(N_, A_, B_) template <> struct Split<N_> { static constexpr int R1 = A_, R2 = B_; };
VS_SPLIT(2, 2, 1) VS_SPLIT(4, 2, 2) VS_SPLIT(6, 3, 2) VS_SPLIT(8, 4, 2) VS_SPLIT(10, 5, 2) VS_SPLIT(12, 4, 3)
VS_SPLIT(16, 4, 4) VS_SPLIT(20, 5, 4) VS_SPLIT(24, 6, 4) VS_SPLIT(32, 8, 4) VS_SPLIT(40, 8, 5) VS_SPLIT(48, 8, 6)
VS_SPLIT(64, 8, 8) VS_SPLIT(80, 10, 8) VS_SPLIT(96, 12, 8) VS_SPLIT(128, 16, 8) VS_SPLIT(160, 16, 10)
VS_SPLIT(192, 16, 12) VS_SPLIT(256, 16, 16)
#undef VS_SPLIT

// the lengths with a compile-time kernel (everything else uses the runtime-plan kernels)
#define VS_FFT2_SIZES(X) X(2) X(4) X(6) X(8) X(10) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) X(80) \
  X(96) X(128) X(160) X(192) X(256)

constexpr int cmax(int a, int b) { return a > b ? a : b; }

// Step 1 of the line transform for thread n1 < R1: DFT_R2 of v (already loaded: v[n2] = x[n1 + R1 n2]),
// twiddle by w_N^{n1 k2} from the table tw (length N).
template <typename T, int N, bool INV>
__device__ __forceinline__ void fft2_step1(cplx<T>* v, int n1, const cplx<T>* __restrict__ tw) {
  constexpr int R2 = Split<N>::R2;
  Bfly<T, R2>::run(v, INV);
#pragma unroll
  for (int k2 = 1; k2 < R2; ++k2) {
    cplx<T> w = tw[n1 * k2];
    if (INV) w.y = -w.y;
    v[k2] = v[k2] * w;
  }
}

// Step 2 for thread k2 < R2: u[n1] = Y[n1][k2] -> DFT_R1 -> u[k1] = X[k2 + R2 k1]
template <typename T, int N, bool INV>
__device__ __forceinline__ void fft2_step2(cplx<T>* u) {
  Bfly<T, Split<N>::R1>::run(u, INV);
}

}  // namespace vs
