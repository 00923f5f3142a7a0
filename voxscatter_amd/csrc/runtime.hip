// Runtime pieces: error state, device buffers, FFT plans, contexts.
#include "vs_internal.hpp"

#include <cstring>

namespace vs {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

static thread_local int g_forbid_alloc = 0;
NoAllocScope::NoAllocScope() { ++g_forbid_alloc; }
NoAllocScope::~NoAllocScope() { --g_forbid_alloc; }

void DevBuf::alloc(size_t n) {
  bytes = 0;
  p = nullptr;
  if (n == 0) return;
  if (g_forbid_alloc)
    throw Error{VS_ERR_ARG, "this call may not allocate device memory, and it would need " + std::to_string(n) +
                                " bytes (an FFT plan or buffer the operator has not built yet)"};
  hipError_t e = hipMalloc(&p, n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
    throw Error{VS_ERR_OOM, "hipMalloc of " + std::to_string(n) + " bytes failed: " + hipGetErrorString(e)};
  }
  bytes = n;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

// Radix factorisation: as many radix-8 stages as possible, then 4/2, 3, 5, then any remaining prime
// factor as a direct-DFT stage.
std::vector<int> fft_factor(int n) {
  std::vector<int> f;
  int m = n, p2 = 0;
  while (m % 2 == 0) { m /= 2; ++p2; }
  int n8 = p2 / 3, rem = p2 % 3;
  if (rem == 1 && n8 > 0) { --n8; f.push_back(4); f.push_back(4); }
  else if (rem == 1) f.push_back(2);
  else if (rem == 2) f.push_back(4);
  for (int i = 0; i < n8; ++i) f.push_back(8);
  while (m % 3 == 0) { f.push_back(3); m /= 3; }
  while (m % 5 == 0) { f.push_back(5); m /= 5; }
  for (int q = 7; m > 1; q += 2) {
    while (m % q == 0) { f.push_back(q); m /= q; }
  }
  return f;
}

template <>
FftDev<double> FftPlanHost::dev<double>() const {
  FftDev<double> d{};
  d.n = n;
  d.nst = (int)radices.size();
  for (int i = 0; i < d.nst; ++i) d.rad[i] = radices[i];
  d.tw = tw64.as<cplx<double>>();
  return d;
}

template <>
FftDev<float> FftPlanHost::dev<float>() const {
  FftDev<float> d{};
  d.n = n;
  d.nst = (int)radices.size();
  for (int i = 0; i < d.nst; ++i) d.rad[i] = radices[i];
  d.tw = tw32.as<cplx<float>>();
  return d;
}

}  // namespace vs

using namespace vs;

const vs::FftPlanHost& vs_ctx::plan(int n) {
  auto it = plans.find(n);
  if (it != plans.end()) return *it->second;
  VS_CHECK(n >= 1 && n <= FFT_THREADS * FFT_EMAX, VS_ERR_UNSUPPORTED,
           "FFT length " + std::to_string(n) + " outside 1.." + std::to_string(FFT_THREADS * FFT_EMAX));
  auto p = std::make_unique<FftPlanHost>();
  p->n = n;
  p->radices = fft_factor(n);
  VS_CHECK((int)p->radices.size() <= FFT_MAX_STAGES, VS_ERR_UNSUPPORTED, "too many FFT stages");
  // twiddles exp(-2 pi i k/n) in extended precision, reduced to the first octant by symmetry so that
  // w^k and w^(n-k) are exact conjugates
  std::vector<cplx<double>> t64(n);
  std::vector<cplx<float>> t32(n);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int k = 0; k < n; ++k) {
    const bool upper = 2 * k > n;               // w^k = conj(w^(n-k))
    const int kk = upper ? n - k : k;
    const long double a = two_pi * (long double)kk / (long double)n;
    const long double c = cosl(a);
    const long double s = upper ? sinl(a) : -sinl(a);
    t64[k] = mk<double>((double)c, (double)s);
    t32[k] = mk<float>((float)c, (float)s);
  }
  DeviceGuard g(device);
  p->tw64.alloc(sizeof(cplx<double>) * n);
  p->tw32.alloc(sizeof(cplx<float>) * n);
  VS_HIP(hipMemcpy(p->tw64.p, t64.data(), sizeof(cplx<double>) * n, hipMemcpyHostToDevice));
  VS_HIP(hipMemcpy(p->tw32.p, t32.data(), sizeof(cplx<float>) * n, hipMemcpyHostToDevice));
  auto& ref = *p;
  plans[n] = std::move(p);
  return ref;
}
