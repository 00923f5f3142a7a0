// Internal declarations of the voxscatter HIP library (gfx950).
//
// Conventions shared by every translation unit:
//  * complex data are interleaved (re, im) pairs of float or double (cplx<T>);
//  * volumes are MATLAB-ordered: x fastest, then y, then z, then component;
//  * the Krylov unknown vector is the compressed [Jx(idx); Jy(idx); Jz(idx)] (main_DDA.m:72), addressed
//    through map3: map3[g] = position of global unknown g = c*L*M*N + voxel in the compressed vector,
//    or -1 for air.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/voxscatter.h"

namespace vs {

// ------------------------------------------------------------------------------------------------
// complex arithmetic
// ------------------------------------------------------------------------------------------------
template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T x, y;
};

template <typename T> __host__ __device__ __forceinline__ cplx<T> mk(T a, T b) { cplx<T> r; r.x = a; r.y = b; return r; }
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a) { return mk<T>(-a.x, -a.y); }
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, cplx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __host__ __device__ __forceinline__ cplx<T> operator*(T s, cplx<T> a) { return mk<T>(s * a.x, s * a.y); }
template <typename T> __host__ __device__ __forceinline__ cplx<T> cconj(cplx<T> a) { return mk<T>(a.x, -a.y); }
// acc + a * b as four fused multiply-adds (acc + (a * b) written with the operators rounds the product
// first and costs a multiply, a fused multiply-add and an add per component)
template <typename T> __host__ __device__ __forceinline__ cplx<T> cfma(cplx<T> a, cplx<T> b, cplx<T> acc) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
  return acc;
}
// a * conj(b)
template <typename T> __host__ __device__ __forceinline__ cplx<T> mulconj(cplx<T> a, cplx<T> b) {
  return mk<T>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T> __host__ __device__ __forceinline__ T abs2(cplx<T> a) { return a.x * a.x + a.y * a.y; }
template <typename T> __host__ __device__ __forceinline__ cplx<T> cdiv(cplx<T> a, cplx<T> b) {
  // Smith's algorithm: no overflow for moderate ranges, exact for real divisors
  if (fabs(b.x) >= fabs(b.y)) {
    T r = b.y / b.x, den = b.x + b.y * r;
    return mk<T>((a.x + a.y * r) / den, (a.y - a.x * r) / den);
  } else {
    T r = b.x / b.y, den = b.x * r + b.y;
    return mk<T>((a.x * r + a.y) / den, (a.y * r - a.x) / den);
  }
}
template <typename To, typename Ti> __host__ __device__ __forceinline__ cplx<To> ccast(cplx<Ti> a) {
  return mk<To>((To)a.x, (To)a.y);
}

using c64 = cplx<float>;
using c128 = cplx<double>;

// ------------------------------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------------------------------
void set_error(const std::string& msg);
const char* last_error();

struct Error {
  int code;
  std::string msg;
};

#define VS_HIP(call)                                                                            \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      throw ::vs::Error{VS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)};         \
  } while (0)

#define VS_CHECK(cond, code, msg)                                                               \
  do {                                                                                          \
    if (!(cond)) throw ::vs::Error{(code), (msg)};                                              \
  } while (0)

// run body, convert exceptions into a status code + thread-local message
template <class F>
int guarded(F&& f) {
  try {
    f();
    return VS_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return VS_ERR_OOM;
  } catch (const std::exception& e) {
    set_error(e.what());
    return VS_ERR_INTERNAL;
  }
}

// ------------------------------------------------------------------------------------------------
// device memory
// ------------------------------------------------------------------------------------------------
// While a NoAllocScope lives on a thread, DevBuf allocations on that thread throw VS_ERR_ARG instead of
// calling hipMalloc (the guard of the entry points that must not allocate).
struct NoAllocScope {
  NoAllocScope();
  ~NoAllocScope();
  NoAllocScope(const NoAllocScope&) = delete;
  NoAllocScope& operator=(const NoAllocScope&) = delete;
};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n);
  void ensure(size_t n) { if (n > bytes) { release(); alloc(n); } }
  void release();
  template <typename U> U* as() const { return reinterpret_cast<U*>(p); }
};

// ------------------------------------------------------------------------------------------------
// FFT plans: radix factorisation + twiddle table exp(-2 pi i k / n), k < n
// ------------------------------------------------------------------------------------------------
constexpr int FFT_MAX_STAGES = 16;
constexpr int FFT_THREADS = 256;   // threads per workgroup of every FFT kernel
constexpr int FFT_EMAX = 16;       // complex elements per thread held in LDS tile

template <typename T>
struct FftDev {
  int n;
  int nst;
  int rad[FFT_MAX_STAGES];
  const cplx<T>* tw;
};

struct FftPlanHost {
  int n = 0;
  std::vector<int> radices;
  DevBuf tw64, tw32;
  template <typename T> FftDev<T> dev() const;
};

// ------------------------------------------------------------------------------------------------
// context
// ------------------------------------------------------------------------------------------------
}  // namespace vs

constexpr int VS_PINNED_C128 = 2048;  // complex doubles in each pinned staging buffer

struct vs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // Lifetime: the caller's reference (released by vs_ctx_destroy) plus one per live operator,
  // preconditioner, recycle handle and vs_dev_alloc buffer.  The stream and the context's memory are
  // freed when the last reference goes, so handles may be destroyed in any order.
  std::atomic<int> refs{1};
  std::mutex dev_mu;
  std::unordered_set<void*> dev_ptrs;  // live vs_dev_alloc buffers
  std::map<int, std::unique_ptr<vs::FftPlanHost>> plans;
  vs::DevBuf red;        // reduction partials
  vs::DevBuf red_out;    // reduced scalars
  void* pinned = nullptr;     // host staging for reduced scalars (device -> host)
  void* pinned_up = nullptr;  // host staging for small coefficient uploads (host -> device)
  // Reusable device workspaces of the solvers.  Solves on one context run one after another, so each
  // slot keeps its memory between calls instead of allocating per solve (hipFree also synchronises).
  enum WsSlot : int { WS_KRYLOV = 0, WS_KRYLOV_AUX, WS_REFINE, WS_REFINE_AUX, WS_NSLOTS };
  vs::DevBuf ws[WS_NSLOTS];
  void* workspace(int slot, size_t bytes) {
    ws[slot].ensure(bytes);
    return ws[slot].p;
  }
  const vs::FftPlanHost& plan(int n);
};

namespace vs {
// context reference counting (see vs_ctx::refs)
void ctx_retain(vs_ctx* ctx);
void ctx_release(vs_ctx* ctx);

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

std::vector<int> fft_factor(int n);
}  // namespace vs
