// extern "C" boundary: argument checks, host<->device staging, status codes.
#include <algorithm>
#include <atomic>
#include <complex>
#include <cstring>

#include "handles.hpp"
#include "solver.hpp"

#include <cstdlib>

namespace vs {
void op_build_dda(vs_op* op, const double* r_host, double ko, int nearby);
void op_build_fN(vs_op* op);
void op_compress_tensor(vs_op* op);
void build_circ_L(vs_ctx* ctx, const c128* toep, c128* cL, int L, int M, int N);
void build_circ_M(vs_ctx* ctx, const c128* cL, c128* F, int L, int M, int N);
template <typename T>
int matvec_variant(const vs_op* op);
void galerkin_singular(vs_ctx* ctx, double ko, double dx, c128 out[5]);
void galerkin_singular_directfn(double ko, double dx, int np, c128 out[5]);
void galerkin_toeplitz(vs_ctx* ctx, int L, int M, int N, double dx, double ko, const c128 sing[5], c128* G);
void assemble2(vs_ctx* ctx, const c128* F, int L, int M, int N, int Lb, int Mb, c128 a, const c128* w, c128 ws,
               c128* out);
void assemble1(vs_ctx* ctx, const c128* cL, int L, int M, int N, c128 a, const c128* w, c128 ws, c128* out);
void invert_blocks_dev(vs_ctx* ctx, c128* blocks, int64_t nblk, int64_t n);
void pc_store_inverse(vs_pc* pc, const c128* inv);
}  // namespace vs

using namespace vs;

// ------------------------------------------------------------------------------------------------
// Scratch of vs_mvp_dev_noalloc: a fixed device array, one copy per GPU, held until the process ends.
// In host code `scratch` names a host-side stand-in, never the device memory: its device address is
// resolved once per device by hipGetSymbolAddress (with that device current, when the first context on
// it is created) and only that address is handed to kernels.
// ------------------------------------------------------------------------------------------------
static constexpr size_t VS_NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static constexpr int VS_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[VS_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[VS_MAX_DEVICES] = {};

static void resolve_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  VS_CHECK(device >= 0 && device < VS_MAX_DEVICES, VS_ERR_UNSUPPORTED, "device index beyond the scratch table");
  if (g_scratch_ptr[device]) return;
  void* ptr = nullptr;
  VS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  VS_CHECK(ptr != nullptr, VS_ERR_HIP, "hipGetSymbolAddress returned no address for the matvec scratch");
  g_scratch_ptr[device] = ptr;
}

void vs::ctx_retain(vs_ctx* ctx) { ctx->refs.fetch_add(1, std::memory_order_relaxed); }

void vs::ctx_release(vs_ctx* ctx) {
  if (ctx->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  ctx->plans.clear();
  for (auto& w : ctx->ws) w.release();
  ctx->red.release();
  ctx->red_out.release();
  if (ctx->pinned) hipHostFree(ctx->pinned);
  if (ctx->pinned_up) hipHostFree(ctx->pinned_up);
  hipStreamDestroy(ctx->stream);
  delete ctx;
}

namespace {

void check_dims(int L, int M, int N) {
  VS_CHECK(L >= 2 && M >= 1 && N >= 1, VS_ERR_ARG, "grid dimensions must satisfy L >= 2, M >= 1, N >= 1");
  VS_CHECK(2 * L <= 2048 && 2 * M <= 2048 && 6 * N <= 2048, VS_ERR_UNSUPPORTED,
           "grid dimension too large for the single-pass LDS FFT (need 2L, 2M <= 2048 and 6N <= 2048)");
}

void h2d(vs_ctx* ctx, void* dst, const void* src, size_t bytes) {
  VS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
}
void d2h(vs_ctx* ctx, void* dst, const void* src, size_t bytes) {
  VS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  VS_HIP(hipStreamSynchronize(ctx->stream));
}

void set_coef(vs_op* op, Coef& c, bool scalar, c128 s, const std::vector<c128>* v) {
  c.scalar = scalar;
  c.s = s;
  if (!scalar) {
    const size_t V = (size_t)op->V;
    c.v64.ensure(sizeof(c128) * V);
    c.v32.ensure(sizeof(c64) * V);
    std::vector<c64> f(V);
    for (size_t i = 0; i < V; ++i) f[i] = ccast<float>((*v)[i]);
    h2d(op->ctx, c.v64.p, v->data(), sizeof(c128) * V);
    h2d(op->ctx, c.v32.p, f.data(), sizeof(c64) * V);
    VS_HIP(hipStreamSynchronize(op->ctx->stream));
  }
}

std::vector<int> build_map3(const int64_t* idx, int64_t nidx, int64_t V3) {
  VS_CHECK(nidx >= 1 && nidx <= V3, VS_ERR_ARG, "idx must hold between 1 and 3*L*M*N unknowns");
  VS_CHECK(V3 < (int64_t)INT32_MAX, VS_ERR_UNSUPPORTED, "volume too large for 32-bit unknown map");
  std::vector<int> map(V3, -1);
  for (int64_t i = 0; i < nidx; ++i) {
    const int64_t g = idx[i];
    VS_CHECK(g >= 0 && g < V3, VS_ERR_ARG, "idx entry out of range [0, 3*L*M*N)");
    VS_CHECK(map[g] < 0, VS_ERR_ARG, "idx has duplicate entries");
    map[g] = (int)i;
  }
  return map;
}

template <typename T>
LinOp<T> make_linop(vs_op* op, vs_pc* pc) {
  const int prec = sizeof(T) == 8 ? VS_FP64 : VS_FP32;
  LinOp<T> L;
  L.A = [op, prec](const cplx<T>* x, cplx<T>* y) { mvp_dev(op, x, y, 0, prec, 1); };
  if (pc) L.M = [pc, prec](const cplx<T>* x, cplx<T>* y) { pc_apply_dev(pc, x, y, prec, 1); };
  if (pc && bi_fused_available<T>(op, pc)) {
    L.F1 = [op, pc](const cplx<T>* r, const cplx<T>* p0, const cplx<T>* v0, cplx<T>* p, cplx<T>* ph, cplx<T>* v,
                    const cplx<T>* rt, const BiScal* sc, BiParts& P) {
      bi_fused_h1<T>(op, pc, r, p0, v0, p, ph, v, rt, sc, P);
    };
    L.F2 = [op, pc](const cplx<T>* x, const cplx<T>* ph, const cplx<T>* r, const cplx<T>* v, cplx<T>* xh, cplx<T>* s,
                    cplx<T>* sh, cplx<T>* t, const BiScal* sc, BiParts& P) {
      bi_fused_h2<T>(op, pc, x, ph, r, v, xh, s, sh, t, sc, P);
    };
  }
  L.Ab = [op, prec](const cplx<T>* x, cplx<T>* y, int k) { mvp_dev(op, x, y, 0, prec, k); };
  if (pc) L.Mb = [pc, prec](const cplx<T>* x, cplx<T>* y, int k) { pc_apply_dev(pc, x, y, prec, k); };
  L.Abd = [op, prec](const cplx<T>* x, cplx<T>* y, int k, const c128* d) { mvp_dev_diag(op, x, y, 0, prec, k, d); };
  return L;
}

void check_solver_args(vs_op* op, vs_pc* pc, int prec) {
  VS_CHECK(op != nullptr, VS_ERR_ARG, "null operator");
  VS_CHECK(op->dense || op->map3.p != nullptr, VS_ERR_ARG, "operator domain not set");
  if (pc) {
    VS_CHECK(pc->ctx == op->ctx, VS_ERR_ARG, "operator and preconditioner belong to different contexts");
    VS_CHECK(pc->nunk == op->nunk, VS_ERR_ARG, "preconditioner and operator unknown counts differ");
    VS_CHECK((pc->precisions & prec) != 0, VS_ERR_ARG, "preconditioner lacks the solver precision");
  }
  VS_CHECK((op->precisions & prec) != 0, VS_ERR_ARG, "operator lacks the solver precision");
}

}  // namespace

extern "C" {

const char* vs_last_error(void) { return vs::last_error(); }

int vs_version(void) { return 10000; }

int vs_device_count(int* n) {
  return guarded([&] {
    VS_CHECK(n != nullptr, VS_ERR_ARG, "null pointer");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      c = 0;
    }
    *n = c;
  });
}

int vs_ctx_create(int device, vs_ctx** out) {
  return guarded([&] {
    VS_CHECK(out != nullptr, VS_ERR_ARG, "null output pointer");
    *out = nullptr;
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) (void)hipGetLastError();
    VS_CHECK(e == hipSuccess && c > 0, VS_ERR_NODEV, "no HIP device available");
    VS_CHECK(device >= 0 && device < c, VS_ERR_NODEV, "device index out of range");
    VS_HIP(hipSetDevice(device));
    // the Krylov loops wait on the stream once per iteration: spin instead of yielding so the wait ends
    // as soon as the work does (a device whose runtime is already active keeps its flags)
    const char* sp = std::getenv("VS_SPIN");
    if (!(sp && sp[0] == '0') && hipSetDeviceFlags(hipDeviceScheduleSpin) != hipSuccess) (void)hipGetLastError();
    resolve_scratch(device);  // with `device` current
    auto ctx = std::make_unique<vs_ctx>();
    ctx->device = device;
    VS_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    *out = ctx.release();
  });
}

void vs_ctx_destroy(vs_ctx* ctx) {
  if (!ctx) return;
  ctx_release(ctx);  // the stream and memory go with the last handle made on this context
}

int vs_ctx_synchronize(vs_ctx* ctx) {
  return guarded([&] {
    VS_CHECK(ctx != nullptr, VS_ERR_ARG, "null context");
    VS_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int vs_dev_alloc(vs_ctx* ctx, int64_t bytes, void** ptr) {
  return guarded([&] {
    VS_CHECK(ctx && ptr && bytes >= 0, VS_ERR_ARG, "bad arguments");
    DeviceGuard g(ctx->device);
    *ptr = nullptr;
    hipError_t e = hipMalloc(ptr, (size_t)std::max<int64_t>(bytes, 1));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      *ptr = nullptr;
      throw Error{VS_ERR_OOM, "hipMalloc failed: " + std::string(hipGetErrorString(e))};
    }
    {
      std::lock_guard<std::mutex> lk(ctx->dev_mu);
      ctx->dev_ptrs.insert(*ptr);
    }
    ctx_retain(ctx);
  });
}

int vs_dev_free(vs_ctx* ctx, void* ptr) {
  return guarded([&] {
    VS_CHECK(ctx != nullptr, VS_ERR_ARG, "null context");
    if (!ptr) return;
    {
      std::lock_guard<std::mutex> lk(ctx->dev_mu);
      VS_CHECK(ctx->dev_ptrs.erase(ptr) == 1, VS_ERR_ARG, "pointer was not allocated by vs_dev_alloc on this context");
    }
    {
      DeviceGuard g(ctx->device);
      hipError_t e1 = hipStreamSynchronize(ctx->stream);
      hipError_t e2 = hipFree(ptr);
      ctx_release(ctx);
      VS_HIP(e1);
      VS_HIP(e2);
    }
  });
}

int vs_dev_copy(vs_ctx* ctx, void* dst, const void* src, int64_t bytes, int kind) {
  return guarded([&] {
    VS_CHECK(ctx && dst && src && bytes >= 0, VS_ERR_ARG, "bad arguments");
    VS_CHECK(kind >= 0 && kind <= 2, VS_ERR_ARG, "kind must be 0 (h2d), 1 (d2h) or 2 (d2d)");
    DeviceGuard g(ctx->device);
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost
                                                                           : hipMemcpyDeviceToDevice);
    VS_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, k, ctx->stream));
    VS_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int vs_alpha_ldr(const double refInd[2], double ko, double d, const double kvec[3], const double Eo[3],
                 double alpha_out[2]) {
  return guarded([&] {
    VS_CHECK(refInd && kvec && Eo && alpha_out, VS_ERR_ARG, "null pointer");
    using cd = std::complex<double>;
    const double b1 = -1.8915316, b2 = 0.1648469, b3 = -1.7700004;
    const double nk = std::sqrt(kvec[0] * kvec[0] + kvec[1] * kvec[1] + kvec[2] * kvec[2]);
    const double ne = std::sqrt(Eo[0] * Eo[0] + Eo[1] * Eo[1] + Eo[2] * Eo[2]);
    VS_CHECK(nk > 0 && ne > 0, VS_ERR_ARG, "kvec and Eo must be nonzero");
    double S = 0;
    for (int j = 0; j < 3; ++j) {
      const double t = (kvec[j] / nk) * (Eo[j] / ne);
      S += t * t;
    }
    const cd m(refInd[0], refInd[1]);
    const cd msqr = m * m;
    const double dcube = d * d * d;
    const cd aCM = 3.0 / (4.0 * M_PI) * (msqr - 1.0) / (msqr + 2.0);
    const cd aL = aCM / (1.0 + aCM * ((b1 + msqr * b2 + msqr * b3 * S) * (ko * d) * (ko * d) -
                                      2.0 / 3.0 * cd(0, 1) * ko * ko * ko * dcube));
    alpha_out[0] = aL.real();
    alpha_out[1] = aL.imag();
  });
}

int vs_op_create_dda(vs_ctx* ctx, int L, int M, int N, const double* r, double ko, int nearby_quad, int precisions,
                     vs_op** out) {
  return guarded([&] {
    VS_CHECK(ctx && r && out, VS_ERR_ARG, "null pointer");
    *out = nullptr;
    check_dims(L, M, N);
    VS_CHECK((precisions & (VS_FP64 | VS_FP32)) != 0, VS_ERR_ARG, "precisions must include VS_FP64 or VS_FP32");
    DeviceGuard g(ctx->device);
    auto op = std::make_unique<vs_op>();
    op->ctx = ctx;
    op->L = L;
    op->M = M;
    op->N = N;
    op->V = (int64_t)L * M * N;
    op->P = 8 * op->V;
    op->precisions = precisions;
    op_build_dda(op.get(), r, ko, nearby_quad);
    ctx_retain(ctx);
    *out = op.release();
  });
}

int vs_galerkin_singular(vs_ctx* ctx, double ko, double dx, double out[10]) {
  return guarded([&] {
    VS_CHECK(ctx && out, VS_ERR_ARG, "null pointer");
    VS_CHECK(dx > 0 && ko >= 0, VS_ERR_ARG, "dx must be positive and ko non-negative");
    DeviceGuard g(ctx->device);
    c128 v[5];
    galerkin_singular(ctx, ko, dx, v);
    for (int i = 0; i < 5; ++i) {
      out[2 * i] = v[i].x;
      out[2 * i + 1] = v[i].y;
    }
  });
}

int vs_op_create_galerkin(vs_ctx* ctx, int L, int M, int N, double dx, double ko, const double* singular,
                          int precisions, vs_op** out) {
  return guarded([&] {
    VS_CHECK(ctx && out, VS_ERR_ARG, "null pointer");
    *out = nullptr;
    check_dims(L, M, N);
    VS_CHECK(M >= 2 && N >= 2, VS_ERR_UNSUPPORTED, "Galerkin grids need L, M, N >= 2");
    VS_CHECK(dx > 0 && ko >= 0, VS_ERR_ARG, "dx must be positive and ko non-negative");
    VS_CHECK((precisions & (VS_FP64 | VS_FP32)) != 0, VS_ERR_ARG, "precisions must include VS_FP64 or VS_FP32");
    DeviceGuard g(ctx->device);
    auto op = std::make_unique<vs_op>();
    op->ctx = ctx;
    op->L = L;
    op->M = M;
    op->N = N;
    op->V = (int64_t)L * M * N;
    op->P = 8 * op->V;
    op->precisions = precisions;
    c128 sing[5];
    if (singular) {
      for (int i = 0; i < 5; ++i) sing[i] = mk<double>(singular[2 * i], singular[2 * i + 1]);
    } else {
      galerkin_singular_directfn(ko, dx, 20, sing);  // the reference's default (assembly_nop_const.m:16)
    }
    op->toep.ensure(sizeof(c128) * 6 * op->V);
    galerkin_toeplitz(ctx, L, M, N, dx, ko, sing, op->toep.as<c128>());
    op->has_toep = true;
    op_build_fN(op.get());
    VS_HIP(hipStreamSynchronize(ctx->stream));
    if (!(op->precisions & VS_FP64)) op->fN64.release();
    ctx_retain(ctx);
    *out = op.release();
  });
}

// The operator keeps one octant of fN and rebuilds the rest from the parity of each component (xy odd in
// x and y, xz in x and z, yz in y and z), which every circulant embedding of a dyadic Green's tensor has
// (circulant_nop_const.m, gperiodic_coeff_nop.m).  A caller's fN without that symmetry is refused rather
// than silently replaced: every mirrored entry must equal the signed entry it is rebuilt from to 1e-9
// of the largest magnitude of its component.
static void check_fN_parity(const double* fN, int L, int M, int N) {
  const int64_t L2 = 2 * L, M2 = 2 * M, N2 = 2 * N, P = L2 * M2 * N2;
  static const int sx[6] = {1, -1, -1, 1, 1, 1}, sy[6] = {1, -1, 1, 1, -1, 1}, sz[6] = {1, 1, -1, 1, -1, 1};
  auto at = [&](int p, int64_t kx, int64_t ky, int64_t kz) {
    const double* e = fN + 2 * (p * P + kx + L2 * (ky + M2 * kz));
    return std::complex<double>(e[0], e[1]);
  };
  for (int p = 0; p < 6; ++p) {
    double mx = 0;
    for (int64_t i = 0; i < P; ++i) mx = std::max(mx, std::abs(std::complex<double>(fN[2 * (p * P + i)], fN[2 * (p * P + i) + 1])));
    const double tol = 1e-9 * mx;
    double worst = 0;
    for (int64_t kz = 0; kz < N2; ++kz)
      for (int64_t ky = 0; ky < M2; ++ky)
        for (int64_t kx = 0; kx < L2; ++kx) {
          const std::complex<double> v = at(p, kx, ky, kz);
          if (kx > L) worst = std::max(worst, std::abs(v - (double)sx[p] * at(p, L2 - kx, ky, kz)));
          if (ky > M) worst = std::max(worst, std::abs(v - (double)sy[p] * at(p, kx, M2 - ky, kz)));
          if (kz > N) worst = std::max(worst, std::abs(v - (double)sz[p] * at(p, kx, ky, N2 - kz)));
        }
    VS_CHECK(worst <= tol, VS_ERR_ARG,
             "fN component " + std::to_string(p) +
                 " lacks the mirror parity of a circulant-embedded dyadic tensor (deviation " + std::to_string(worst) +
                 " against a largest entry of " + std::to_string(mx) + ")");
  }
}

int vs_op_create_fN(vs_ctx* ctx, int L, int M, int N, const double* fN, const double* Toep, int precisions,
                    vs_op** out) {
  return guarded([&] {
    VS_CHECK(ctx && fN && out, VS_ERR_ARG, "null pointer");
    *out = nullptr;
    check_dims(L, M, N);
    VS_CHECK((precisions & (VS_FP64 | VS_FP32)) != 0, VS_ERR_ARG, "precisions must include VS_FP64 or VS_FP32");
    check_fN_parity(fN, L, M, N);
    DeviceGuard g(ctx->device);
    auto op = std::make_unique<vs_op>();
    op->ctx = ctx;
    op->L = L;
    op->M = M;
    op->N = N;
    op->V = (int64_t)L * M * N;
    op->P = 8 * op->V;
    op->precisions = precisions;
    op->fN64.ensure(sizeof(c128) * 6 * op->P);
    h2d(ctx, op->fN64.p, fN, sizeof(c128) * 6 * op->P);
    if (Toep) {
      op->toep.ensure(sizeof(c128) * 6 * op->V);
      h2d(ctx, op->toep.p, Toep, sizeof(c128) * 6 * op->V);
      op->has_toep = true;
    }
    op_compress_tensor(op.get());
    VS_HIP(hipStreamSynchronize(ctx->stream));
    if (!(precisions & VS_FP64)) op->fN64.release();
    ctx_retain(ctx);
    *out = op.release();
  });
}

void vs_op_destroy(vs_op* op) {
  if (!op) return;
  vs_ctx* ctx = op->ctx;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  delete op;
  ctx_release(ctx);
}

int vs_op_get_toeplitz(vs_op* op, double* Toep) {
  return guarded([&] {
    VS_CHECK(op && Toep, VS_ERR_ARG, "null pointer");
    VS_CHECK(op->has_toep, VS_ERR_ARG, "operator has no Toeplitz tensor");
    DeviceGuard g(op->ctx->device);
    d2h(op->ctx, Toep, op->toep.p, sizeof(c128) * 6 * op->V);
  });
}

int vs_op_get_fN(vs_op* op, double* fN) {
  return guarded([&] {
    VS_CHECK(op && fN, VS_ERR_ARG, "null pointer");
    VS_CHECK(op->fN64.p != nullptr, VS_ERR_ARG, "operator holds no FP64 tensor");
    DeviceGuard g(op->ctx->device);
    d2h(op->ctx, fN, op->fN64.p, sizeof(c128) * 6 * op->P);
  });
}

int vs_op_set_domain(vs_op* op, const int64_t* idx, int64_t nidx, const double* Gram, int64_t gram_len,
                     const double* Mr, int64_t mr_len, const double* Mc, int64_t mc_len) {
  return guarded([&] {
    VS_CHECK(op && idx && Gram && Mr && Mc, VS_ERR_ARG, "null pointer");
    VS_CHECK(!op->dense, VS_ERR_ARG, "a dense operator has no unknown set or diagonal");
    const int64_t V = op->V;
    auto okl = [&](int64_t l) { return l == 1 || l == V; };
    VS_CHECK(okl(gram_len) && okl(mr_len) && okl(mc_len), VS_ERR_ARG, "Gram/Mr/Mc must have length 1 or L*M*N");
    DeviceGuard g(op->ctx->device);
    std::vector<int> map = build_map3(idx, nidx, 3 * V);
    op->map3.ensure(sizeof(int) * 3 * V);
    h2d(op->ctx, op->map3.p, map.data(), sizeof(int) * 3 * V);
    op->nunk = nidx;
    op->idx_host.assign(idx, idx + nidx);
    static std::atomic<uint64_t> next_dom{1};
    op->dom_id = next_dom++;
    auto at = [](const double* p, int64_t len, int64_t v) {
      return len == 1 ? mk<double>(p[0], p[1]) : mk<double>(p[2 * v], p[2 * v + 1]);
    };
    const bool sa = gram_len == 1 && mr_len == 1;
    if (sa) {
      const c128 G = at(Gram, 1, 0), R = at(Mr, 1, 0);
      set_coef(op, op->a, true, G * R, nullptr);
      set_coef(op, op->at, true, G * cconj(R), nullptr);
    } else {
      std::vector<c128> va(V), vt(V);
      for (int64_t v = 0; v < V; ++v) {
        const c128 G = at(Gram, gram_len, v), R = at(Mr, mr_len, v);
        va[v] = G * R;
        vt[v] = G * cconj(R);
      }
      set_coef(op, op->a, false, mk<double>(0, 0), &va);
      set_coef(op, op->at, false, mk<double>(0, 0), &vt);
    }
    if (mc_len == 1) {
      set_coef(op, op->b, true, at(Mc, 1, 0), nullptr);
    } else {
      std::vector<c128> vb(V);
      for (int64_t v = 0; v < V; ++v) vb[v] = at(Mc, mc_len, v);
      set_coef(op, op->b, false, mk<double>(0, 0), &vb);
    }
    op_autotune(op);  // first domain: pick the fastest matvec variant per precision
    VS_HIP(hipStreamSynchronize(op->ctx->stream));
  });
}

int64_t vs_op_nunknowns(const vs_op* op) { return op ? op->nunk : -1; }

int vs_set_fused_krylov(int on) {
  return guarded([&] { bi_fused_set(on != 0); });
}

int vs_fused_krylov_available(vs_op* op, vs_pc* pc, int precision, int* on) {
  return guarded([&] {
    VS_CHECK(op && pc && on, VS_ERR_ARG, "null pointer");
    VS_CHECK(precision == VS_FP64 || precision == VS_FP32, VS_ERR_ARG, "precision must be VS_FP64 or VS_FP32");
    *on = (precision == VS_FP64 ? bi_fused_available<double>(op, pc) : bi_fused_available<float>(op, pc)) ? 1 : 0;
  });
}

int vs_op_matvec_variant(const vs_op* op, int precision, int* variant) {
  return guarded([&] {
    VS_CHECK(op && variant, VS_ERR_ARG, "null pointer");
    VS_CHECK(!op->dense, VS_ERR_ARG, "a dense operator has no FFT matvec variant");
    VS_CHECK(precision == VS_FP64 || precision == VS_FP32, VS_ERR_ARG, "precision must be VS_FP64 or VS_FP32");
    *variant = precision == VS_FP64 ? matvec_variant<double>(op) : matvec_variant<float>(op);
  });
}

int vs_mvp(vs_op* op, const double* x, double* y, int transp) {
  return guarded([&] {
    VS_CHECK(op && x && y, VS_ERR_ARG, "null pointer");
    VS_CHECK(op->dense ? op->a64.p != nullptr : op->fc64.p != nullptr, VS_ERR_ARG,
             "vs_mvp needs an FP64 operator (precisions & VS_FP64)");
    VS_CHECK(op->dense || op->map3.p != nullptr, VS_ERR_ARG, "operator domain not set");
    DeviceGuard g(op->ctx->device);
    const size_t bytes = sizeof(c128) * (size_t)op->nunk;
    op->hx.ensure(std::max<size_t>(bytes, 16));
    op->hy.ensure(std::max<size_t>(bytes, 16));
    h2d(op->ctx, op->hx.p, x, bytes);
    mvp_dev(op, op->hx.p, op->hy.p, transp ? 1 : 0, VS_FP64, 1);
    d2h(op->ctx, y, op->hy.p, bytes);
  });
}

int vs_mvp_dev(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs) {
  return guarded([&] {
    VS_CHECK(op && x && y, VS_ERR_ARG, "null pointer");
    DeviceGuard g(op->ctx->device);
    mvp_dev(op, x, y, transp ? 1 : 0, precision, nrhs);
  });
}

int vs_mvp_dev_diag(vs_op* op, const void* x, void* y, const void* diag, int transp, int precision, int nrhs) {
  return guarded([&] {
    VS_CHECK(op && x && y && diag, VS_ERR_ARG, "null pointer");
    DeviceGuard g(op->ctx->device);
    mvp_dev_diag(op, x, y, transp ? 1 : 0, precision, nrhs, reinterpret_cast<const c128*>(diag));
  });
}

int vs_mvp_dev_noalloc(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs) {
  return guarded([&] {
    VS_CHECK(op && x && y, VS_ERR_ARG, "null pointer");
    const int dev = op->ctx->device;
    void* sp = (dev >= 0 && dev < VS_MAX_DEVICES) ? g_scratch_ptr[dev] : nullptr;
    VS_CHECK(sp != nullptr, VS_ERR_INTERNAL, "matvec scratch address not resolved for this device");
    DeviceGuard g(dev);
    NoAllocScope no_alloc;  // any allocation on this path becomes a VS_ERR_ARG error instead
    mvp_dev_scratch(op, x, y, transp ? 1 : 0, precision, nrhs, sp, VS_NOALLOC_SCRATCH_BYTES);
  });
}

int vs_circ_1_level(vs_ctx* ctx, int L, int M, int N, const double* Toep, double* circ_L, double* blocks) {
  return guarded([&] {
    VS_CHECK(ctx && Toep && circ_L, VS_ERR_ARG, "null pointer");
    check_dims(L, M, N);
    DeviceGuard g(ctx->device);
    const int64_t V = (int64_t)L * M * N;
    DevBuf t(sizeof(c128) * 6 * V), cL(sizeof(c128) * 6 * V);
    h2d(ctx, t.p, Toep, sizeof(c128) * 6 * V);
    build_circ_L(ctx, t.as<c128>(), cL.as<c128>(), L, M, N);
    d2h(ctx, circ_L, cL.p, sizeof(c128) * 6 * V);
    if (blocks) {
      const int64_t n = 3 * (int64_t)M * N;
      DevBuf B(sizeof(c128) * L * n * n);
      assemble1(ctx, cL.as<c128>(), L, M, N, mk<double>(0, 0), nullptr, mk<double>(-1, 0), B.as<c128>());
      d2h(ctx, blocks, B.p, sizeof(c128) * L * n * n);
    }
  });
}

int vs_circ_2_level(vs_ctx* ctx, int L, int M, int N, const double* circ_L, double* circ_M, double* blocks) {
  return guarded([&] {
    VS_CHECK(ctx && circ_L, VS_ERR_ARG, "null pointer");
    check_dims(L, M, N);
    DeviceGuard g(ctx->device);
    const int64_t V = (int64_t)L * M * N;
    DevBuf cL(sizeof(c128) * 6 * V), F(sizeof(c128) * 6 * V);
    h2d(ctx, cL.p, circ_L, sizeof(c128) * 6 * V);
    build_circ_M(ctx, cL.as<c128>(), F.as<c128>(), L, M, N);
    if (blocks) {
      const int64_t n = 3 * (int64_t)N;
      DevBuf B(sizeof(c128) * V / N * n * n);
      assemble2(ctx, F.as<c128>(), L, M, N, L, M, mk<double>(0, 0), nullptr, mk<double>(-1, 0), B.as<c128>());
      d2h(ctx, blocks, B.p, sizeof(c128) * (size_t)L * M * n * n);
    }
    if (circ_M) {
      d2h(ctx, circ_M, F.p, sizeof(c128) * 6 * V);
      // circ_2_level.m:44 stores the xz component negated
      c128* cm = reinterpret_cast<c128*>(circ_M) + 2 * V;
      for (int64_t i = 0; i < V; ++i) cm[i] = -cm[i];
    }
  });
}

int vs_invert_blocks(vs_ctx* ctx, int64_t nblk, int n, const double* blocks, const double a[2], const double* w,
                     const double wscalar[2], double* inv_out) {
  return guarded([&] {
    VS_CHECK(ctx && blocks && a && inv_out, VS_ERR_ARG, "null pointer");
    VS_CHECK(w || wscalar, VS_ERR_ARG, "either w or wscalar is required");
    VS_CHECK(nblk >= 1 && n >= 1, VS_ERR_ARG, "empty block set");
    DeviceGuard g(ctx->device);
    const size_t nn = (size_t)n * n;
    // form a I - diag(w) B on the host side of the copy (cheap), invert on the device
    std::vector<c128> A(nblk * nn);
    const c128* Bh = reinterpret_cast<const c128*>(blocks);
    const c128 av = mk<double>(a[0], a[1]);
    for (int64_t b = 0; b < nblk; ++b)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const c128 wi = w ? mk<double>(w[2 * i], w[2 * i + 1]) : mk<double>(wscalar[0], wscalar[1]);
          c128 v = -(wi * Bh[b * nn + i + (size_t)n * j]);
          if (i == j) v = v + av;
          A[b * nn + i + (size_t)n * j] = v;
        }
    DevBuf D(sizeof(c128) * nblk * nn);
    h2d(ctx, D.p, A.data(), sizeof(c128) * nblk * nn);
    invert_blocks_dev(ctx, D.as<c128>(), nblk, n);
    d2h(ctx, inv_out, D.p, sizeof(c128) * nblk * nn);
  });
}

int vs_pc_create_circ(vs_op* op, int levels, const double a[2], const double* w, const double wscalar[2],
                      int precisions, vs_pc** out) {
  return guarded([&] {
    VS_CHECK(op && a && out, VS_ERR_ARG, "null pointer");
    VS_CHECK(w || wscalar, VS_ERR_ARG, "either w or wscalar is required");
    VS_CHECK(levels == 1 || levels == 2, VS_ERR_ARG, "levels must be 1 or 2");
    VS_CHECK(op->has_toep, VS_ERR_ARG, "operator has no Toeplitz tensor");
    VS_CHECK(op->map3.p != nullptr, VS_ERR_ARG, "operator domain not set");
    VS_CHECK((precisions & (VS_FP64 | VS_FP32)) != 0, VS_ERR_ARG, "precisions must include VS_FP64 or VS_FP32");
    *out = nullptr;
    vs_ctx* ctx = op->ctx;
    DeviceGuard g(ctx->device);
    const int L = op->L, M = op->M, N = op->N;
    const int64_t V = op->V;
    auto pc = std::make_unique<vs_pc>();
    pc->ctx = ctx;
    pc->levels = levels;
    pc->L = L;
    pc->M = M;
    pc->N = N;
    pc->V = V;
    pc->precisions = precisions;
    pc->nunk = op->nunk;
    pc->dom_id = op->dom_id;
    pc->map3.ensure(sizeof(int) * 3 * V);
    VS_HIP(hipMemcpyAsync(pc->map3.p, op->map3.p, sizeof(int) * 3 * V, hipMemcpyDeviceToDevice, ctx->stream));
    const int64_t nb = levels == 2 ? 3LL * N : 3LL * M * N;
    // 2-level: only the blocks kx <= L/2, ky <= M/2 are built, inverted and kept (mirror symmetry)
    pc->mirror = levels == 2;
    pc->Lb = levels == 2 ? L / 2 + 1 : L;
    pc->Mb = levels == 2 ? M / 2 + 1 : 1;
    const int64_t nblk = levels == 2 ? (int64_t)pc->Lb * pc->Mb : L;
    VS_CHECK(nb <= INT32_MAX, VS_ERR_UNSUPPORTED, "block too large");
    const double bytes = (double)nblk * nb * nb * 16.0;
    VS_CHECK(bytes < 64e9, VS_ERR_OOM, "preconditioner blocks exceed the 64 GB build limit");
    VS_CHECK(levels == 2 || nb <= 4096, VS_ERR_UNSUPPORTED, "1-level blocks limited to 3MN <= 4096");
    pc->nb = (int)nb;
    DevBuf cL(sizeof(c128) * 6 * V);
    build_circ_L(ctx, op->toep.as<c128>(), cL.as<c128>(), L, M, N);
    DevBuf wdev;
    if (w) {
      wdev.ensure(sizeof(c128) * nb);
      h2d(ctx, wdev.p, w, sizeof(c128) * nb);
    }
    const c128 av = mk<double>(a[0], a[1]);
    const c128 ws = wscalar ? mk<double>(wscalar[0], wscalar[1]) : mk<double>(0, 0);
    DevBuf B(sizeof(c128) * nblk * nb * nb);
    if (levels == 2) {
      DevBuf F(sizeof(c128) * 6 * V);
      build_circ_M(ctx, cL.as<c128>(), F.as<c128>(), L, M, N);
      assemble2(ctx, F.as<c128>(), L, M, N, pc->Lb, pc->Mb, av, w ? wdev.as<c128>() : nullptr, ws, B.as<c128>());
    } else {
      assemble1(ctx, cL.as<c128>(), L, M, N, av, w ? wdev.as<c128>() : nullptr, ws, B.as<c128>());
    }
    invert_blocks_dev(ctx, B.as<c128>(), nblk, nb);
    // a scalar weight keeps inv(a I - w C)^T = D inv(a I - w C) D (D = -1 on the z unknowns): pack it
    pc->sym = levels == 2 && w == nullptr;
    pc_store_inverse(pc.get(), B.as<c128>());
    VS_HIP(hipStreamSynchronize(ctx->stream));
    ctx_retain(ctx);
    *out = pc.release();
  });
}

int vs_pc_create_blocks(vs_ctx* ctx, int levels, int L, int M, int N, const double* inv_blocks, const int64_t* idx,
                        int64_t nidx, int precisions, vs_pc** out) {
  return guarded([&] {
    VS_CHECK(ctx && inv_blocks && idx && out, VS_ERR_ARG, "null pointer");
    VS_CHECK(levels == 1 || levels == 2, VS_ERR_ARG, "levels must be 1 or 2");
    check_dims(L, M, N);
    VS_CHECK((precisions & (VS_FP64 | VS_FP32)) != 0, VS_ERR_ARG, "precisions must include VS_FP64 or VS_FP32");
    *out = nullptr;
    DeviceGuard g(ctx->device);
    auto pc = std::make_unique<vs_pc>();
    pc->ctx = ctx;
    pc->levels = levels;
    pc->L = L;
    pc->M = M;
    pc->N = N;
    pc->V = (int64_t)L * M * N;
    pc->precisions = precisions;
    const int64_t nb = levels == 2 ? 3LL * N : 3LL * M * N;
    // caller-supplied blocks: the full set, no symmetry assumed
    pc->Lb = L;
    pc->Mb = levels == 2 ? M : 1;
    const int64_t nblk = levels == 2 ? (int64_t)L * M : L;
    VS_CHECK(levels == 2 || nb <= 4096, VS_ERR_UNSUPPORTED, "1-level blocks limited to 3MN <= 4096");
    pc->nb = (int)nb;
    std::vector<int> map = build_map3(idx, nidx, 3 * pc->V);
    pc->map3.ensure(sizeof(int) * 3 * pc->V);
    h2d(ctx, pc->map3.p, map.data(), sizeof(int) * 3 * pc->V);
    pc->nunk = nidx;
    DevBuf B(sizeof(c128) * nblk * nb * nb);
    h2d(ctx, B.p, inv_blocks, sizeof(c128) * nblk * nb * nb);
    pc_store_inverse(pc.get(), B.as<c128>());
    VS_HIP(hipStreamSynchronize(ctx->stream));
    ctx_retain(ctx);
    *out = pc.release();
  });
}

void vs_pc_destroy(vs_pc* pc) {
  if (!pc) return;
  vs_ctx* ctx = pc->ctx;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  delete pc;
  ctx_release(ctx);
}

int vs_pc_apply(vs_pc* pc, const double* x, double* y) {
  return guarded([&] {
    VS_CHECK(pc && x && y, VS_ERR_ARG, "null pointer");
    VS_CHECK(pc->precisions & VS_FP64, VS_ERR_ARG, "vs_pc_apply needs FP64 blocks");
    DeviceGuard g(pc->ctx->device);
    const size_t bytes = sizeof(c128) * (size_t)pc->nunk;
    pc->hx.ensure(std::max<size_t>(bytes, 16));
    pc->hy.ensure(std::max<size_t>(bytes, 16));
    h2d(pc->ctx, pc->hx.p, x, bytes);
    pc_apply_dev(pc, pc->hx.p, pc->hy.p, VS_FP64, 1);
    d2h(pc->ctx, y, pc->hy.p, bytes);
  });
}

int vs_pc_apply_dev(vs_pc* pc, const void* x, void* y, int precision, int nrhs) {
  return guarded([&] {
    VS_CHECK(pc && x && y, VS_ERR_ARG, "null pointer");
    DeviceGuard g(pc->ctx->device);
    pc_apply_dev(pc, x, y, precision, nrhs);
  });
}

int vs_bicgstab(vs_op* op, vs_pc* pc, const double* b, double* x, double tol, int maxit, const double* x0, int* flag,
                double* relres, double* iter, double* resvec, int* nres) {
  return guarded([&] {
    VS_CHECK(op && b && x && flag && relres && iter && resvec && nres, VS_ERR_ARG, "null pointer");
    VS_CHECK(maxit >= 1, VS_ERR_ARG, "maxit must be >= 1");
    check_solver_args(op, pc, VS_FP64);
    vs_ctx* ctx = op->ctx;
    DeviceGuard g(ctx->device);
    const int64_t n = op->nunk;
    DevBuf db(sizeof(c128) * n), dx(sizeof(c128) * n);
    h2d(ctx, db.p, b, sizeof(c128) * n);
    if (x0) h2d(ctx, dx.p, x0, sizeof(c128) * n);
    Engine<double> E(ctx, n);
    SolveResult R = solve_bicgstab<double>(E, make_linop<double>(op, pc), db.as<c128>(), dx.as<c128>(), x0 != nullptr,
                                           tol, maxit);
    d2h(ctx, x, dx.p, sizeof(c128) * n);
    *flag = R.flag;
    *relres = R.relres;
    *iter = R.iter;
    *nres = (int)R.resvec.size();
    std::copy(R.resvec.begin(), R.resvec.end(), resvec);
  });
}

int vs_bicgstab_dev(vs_op* op, vs_pc* pc, const void* b, void* x, double tol, int maxit, int* flag, double* relres,
                    double* iter, int* nres) {
  return guarded([&] {
    VS_CHECK(op && b && x && flag && relres && iter && nres, VS_ERR_ARG, "null pointer");
    check_solver_args(op, pc, VS_FP64);
    DeviceGuard g(op->ctx->device);
    Engine<double> E(op->ctx, op->nunk);
    SolveResult R = solve_bicgstab<double>(E, make_linop<double>(op, pc), (const c128*)b, (c128*)x, false, tol, maxit);
    *flag = R.flag;
    *relres = R.relres;
    *iter = R.iter;
    *nres = (int)R.resvec.size();
  });
}

int vs_gmres(vs_op* op, vs_pc* pc, const double* b, double* x, int restart, double tol, int maxit, const double* x0,
             int* flag, double* relres, int iter[2], double* resvec, int* nres) {
  return guarded([&] {
    VS_CHECK(op && b && x && flag && relres && iter && resvec && nres, VS_ERR_ARG, "null pointer");
    VS_CHECK(maxit >= 1 && restart >= 1, VS_ERR_ARG, "maxit and restart must be >= 1");
    check_solver_args(op, pc, VS_FP64);
    vs_ctx* ctx = op->ctx;
    DeviceGuard g(ctx->device);
    const int64_t n = op->nunk;
    DevBuf db(sizeof(c128) * n), dx(sizeof(c128) * n);
    h2d(ctx, db.p, b, sizeof(c128) * n);
    if (x0) h2d(ctx, dx.p, x0, sizeof(c128) * n);
    Engine<double> E(ctx, n);
    SolveResult R = solve_gmres<double>(E, make_linop<double>(op, pc), db.as<c128>(), dx.as<c128>(), x0 != nullptr,
                                        restart, tol, maxit);
    d2h(ctx, x, dx.p, sizeof(c128) * n);
    *flag = R.flag;
    *relres = R.relres;
    iter[0] = R.outer;
    iter[1] = R.inner;
    *nres = (int)R.resvec.size();
    std::copy(R.resvec.begin(), R.resvec.end(), resvec);
  });
}

int vs_pgmres(vs_op* op, vs_pc* pc, int side, const double* b, double* x, int restart, double tol, int maxit,
              const double* x0, int* flag, double* relres, int iter[2], double* resvec, int* nres) {
  return guarded([&] {
    VS_CHECK(op && b && x && flag && relres && iter && resvec && nres, VS_ERR_ARG, "null pointer");
    VS_CHECK(maxit >= 1 && restart >= 1, VS_ERR_ARG, "maxit and restart must be >= 1");
    VS_CHECK(side == 0 || side == 1, VS_ERR_ARG, "side must be 0 (L1) or 1 (R1)");
    check_solver_args(op, pc, VS_FP64);
    vs_ctx* ctx = op->ctx;
    DeviceGuard g(ctx->device);
    const int64_t n = op->nunk;
    DevBuf db(sizeof(c128) * n), dx(sizeof(c128) * n);
    h2d(ctx, db.p, b, sizeof(c128) * n);
    if (x0) h2d(ctx, dx.p, x0, sizeof(c128) * n);
    Engine<double> E(ctx, n);
    SolveResult R = solve_pgmres<double>(E, make_linop<double>(op, pc), side, db.as<c128>(), dx.as<c128>(),
                                         x0 != nullptr, restart, tol, maxit);
    d2h(ctx, x, dx.p, sizeof(c128) * n);
    *flag = R.flag;
    *relres = R.relres;
    iter[0] = R.inner;
    iter[1] = R.outer;
    *nres = (int)R.resvec.size();
    std::copy(R.resvec.begin(), R.resvec.end(), resvec);
  });
}

int vs_pgmresdr(vs_op* op, vs_pc* pc, int side, const double* b, double* x, int restart, double tol, int maxit,
                int ritz, const double* x0, int* flag, double* relres, int iter[2], double* resvec, int* nres) {
  return guarded([&] {
    VS_CHECK(op && b && x && flag && relres && iter && resvec && nres, VS_ERR_ARG, "null pointer");
    VS_CHECK(maxit >= 0 && restart >= 2, VS_ERR_ARG, "maxit must be >= 0 and restart >= 2");
    VS_CHECK(side == 0 || side == 1, VS_ERR_ARG, "side must be 0 (L1) or 1 (R1)");
    check_solver_args(op, pc, VS_FP64);
    vs_ctx* ctx = op->ctx;
    DeviceGuard g(ctx->device);
    const int64_t n = op->nunk;
    DevBuf db(sizeof(c128) * n), dx(sizeof(c128) * n);
    h2d(ctx, db.p, b, sizeof(c128) * n);
    if (x0) h2d(ctx, dx.p, x0, sizeof(c128) * n);
    Engine<double> E(ctx, n);
    SolveResult R = solve_pgmresdr<double>(E, make_linop<double>(op, pc), side, db.as<c128>(), dx.as<c128>(),
                                           x0 != nullptr, restart, tol, maxit, ritz);
    d2h(ctx, x, dx.p, sizeof(c128) * n);
    *flag = R.flag;
    *relres = R.relres;
    iter[0] = R.inner;
    iter[1] = R.outer;
    *nres = (int)R.resvec.size();
    std::copy(R.resvec.begin(), R.resvec.end(), resvec);
  });
}

struct vs_recycle {
  vs_ctx* ctx = nullptr;
  vs::Recycle r;
};

int vs_recycle_create(vs_ctx* ctx, vs_recycle** out) {
  return guarded([&] {
    VS_CHECK(ctx && out, VS_ERR_ARG, "null pointer");
    auto rec = std::make_unique<vs_recycle>();
    rec->ctx = ctx;
    ctx_retain(ctx);
    *out = rec.release();
  });
}

void vs_recycle_destroy(vs_recycle* rec) {
  if (!rec) return;
  vs_ctx* ctx = rec->ctx;
  DeviceGuard g(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete rec;
  ctx_release(ctx);
}

int vs_recycle_size(const vs_recycle* rec) { return rec ? rec->r.k : 0; }

int vs_recycle_clear(vs_recycle* rec) {
  return guarded([&] {
    VS_CHECK(rec, VS_ERR_ARG, "null pointer");
    rec->r.k = 0;
  });
}

int vs_pgmresdr_rec(vs_op* op, vs_pc* pc, int side, const double* b, double* x, int restart, double tol, int maxit,
                    int ritz, const double* x0, vs_recycle* rec, int* flag, double* relres, int iter[2],
                    double* resvec, int* nres) {
  return guarded([&] {
    VS_CHECK(op && b && x && flag && relres && iter && resvec && nres, VS_ERR_ARG, "null pointer");
    VS_CHECK(maxit >= 0 && restart >= 2, VS_ERR_ARG, "maxit must be >= 0 and restart >= 2");
    VS_CHECK(side == 0 || side == 1, VS_ERR_ARG, "side must be 0 (L1) or 1 (R1)");
    VS_CHECK(!rec || rec->ctx == op->ctx, VS_ERR_ARG, "recycled space belongs to another context");
    check_solver_args(op, pc, VS_FP64);
    vs_ctx* ctx = op->ctx;
    DeviceGuard g(ctx->device);
    const int64_t n = op->nunk;
    DevBuf db(sizeof(c128) * n), dx(sizeof(c128) * n);
    h2d(ctx, db.p, b, sizeof(c128) * n);
    if (x0) h2d(ctx, dx.p, x0, sizeof(c128) * n);
    Engine<double> E(ctx, n);
    SolveResult R = solve_pgmresdr<double>(E, make_linop<double>(op, pc), side, db.as<c128>(), dx.as<c128>(),
                                           x0 != nullptr, restart, tol, maxit, ritz, rec ? &rec->r : nullptr);
    d2h(ctx, x, dx.p, sizeof(c128) * n);
    *flag = R.flag;
    *relres = R.relres;
    iter[0] = R.inner;
    iter[1] = R.outer;
    *nres = (int)R.resvec.size();
    std::copy(R.resvec.begin(), R.resvec.end(), resvec);
  });
}

int vs_gmres_dev(vs_op* op, vs_pc* pc, const void* b, void* x, int restart, double tol, int maxit, int* flag,
                 double* relres, int iter[2], int* nres) {
  return guarded([&] {
    VS_CHECK(op && b && x && flag && relres && iter && nres, VS_ERR_ARG, "null pointer");
    VS_CHECK(maxit >= 1 && restart >= 1, VS_ERR_ARG, "maxit and restart must be >= 1");
    check_solver_args(op, pc, VS_FP64);
    DeviceGuard g(op->ctx->device);
    Engine<double> E(op->ctx, op->nunk);
    SolveResult R = solve_gmres<double>(E, make_linop<double>(op, pc), (const c128*)b, (c128*)x, false, restart, tol,
                                        maxit);
    *flag = R.flag;
    *relres = R.relres;
    iter[0] = R.outer;
    iter[1] = R.inner;
    *nres = (int)R.resvec.size();
  });
}

}  // extern "C"

namespace {
SolveResult refine_common(vs_op* op, vs_pc* pc, const c128* b, c128* x, bool has_x0, double tol, int max_outer,
                          int inner_solver, double inner_tol, int inner_maxit, int restart, int* inner_total) {
  VS_CHECK(op != nullptr, VS_ERR_ARG, "null operator");
  VS_CHECK((op->precisions & (VS_FP64 | VS_FP32)) == (VS_FP64 | VS_FP32), VS_ERR_ARG,
           "refinement needs an operator holding both FP64 and FP32 tensors");
  check_solver_args(op, pc, VS_FP32);
  VS_CHECK(inner_solver == 0 || inner_solver == 1, VS_ERR_ARG, "inner_solver must be 0 (bicgstab) or 1 (gmres)");
  VS_CHECK(max_outer >= 0 && inner_maxit >= 1, VS_ERR_ARG, "bad iteration limits");
  Engine<double> E64(op->ctx, op->nunk);
  Engine<float> E32(op->ctx, op->nunk);
  LinOp<double> A64 = make_linop<double>(op, nullptr);
  LinOp<float> A32 = make_linop<float>(op, pc);
  return solve_refine(E64, E32, A64, A32, b, x, has_x0, tol, max_outer, inner_solver, inner_tol, inner_maxit,
                      restart, inner_total);
}
}  // namespace

extern "C" {

int vs_refine(vs_op* op, vs_pc* pc, const double* b, double* x, double tol, int max_outer, int inner_solver,
              double inner_tol, int inner_maxit, int restart, const double* x0, int* flag, double* relres,
              int* outer, int* inner_total, double* resvec_outer, int* nres) {
  return guarded([&] {
    VS_CHECK(op && b && x && flag && relres && outer && inner_total && resvec_outer && nres, VS_ERR_ARG,
             "null pointer");
    vs_ctx* ctx = op->ctx;
    DeviceGuard g(ctx->device);
    const int64_t n = op->nunk;
    DevBuf db(sizeof(c128) * n), dx(sizeof(c128) * n);
    h2d(ctx, db.p, b, sizeof(c128) * n);
    if (x0) h2d(ctx, dx.p, x0, sizeof(c128) * n);
    SolveResult R = refine_common(op, pc, db.as<c128>(), dx.as<c128>(), x0 != nullptr, tol, max_outer, inner_solver,
                                  inner_tol, inner_maxit, restart, inner_total);
    d2h(ctx, x, dx.p, sizeof(c128) * n);
    *flag = R.flag;
    *relres = R.relres;
    *outer = R.outer;
    *nres = (int)R.resvec.size();
    std::copy(R.resvec.begin(), R.resvec.end(), resvec_outer);
  });
}

int vs_refine_dev(vs_op* op, vs_pc* pc, const void* b, void* x, double tol, int max_outer, int inner_solver,
                  double inner_tol, int inner_maxit, int restart, int* flag, double* relres, int* outer,
                  int* inner_total) {
  return guarded([&] {
    VS_CHECK(op && b && x && flag && relres && outer && inner_total, VS_ERR_ARG, "null pointer");
    DeviceGuard g(op->ctx->device);
    SolveResult R = refine_common(op, pc, (const c128*)b, (c128*)x, false, tol, max_outer, inner_solver, inner_tol,
                                  inner_maxit, restart, inner_total);
    *flag = R.flag;
    *relres = R.relres;
    *outer = R.outer;
  });
}

int vs_bicgstab_batch_diag_dev(vs_op* op, vs_pc* pc, int nrhs, const void* B, void* X, const double* diag,
                               int precision, double tol, int maxit, int* flags, double* relres, double* iters) {
  return guarded([&] {
    VS_CHECK(op && B && X && flags && relres && iters && nrhs >= 1 && maxit >= 1, VS_ERR_ARG, "bad arguments");
    VS_CHECK(precision == VS_FP64 || precision == VS_FP32, VS_ERR_ARG, "precision must be VS_FP64 or VS_FP32");
    check_solver_args(op, pc, precision);
    DeviceGuard g(op->ctx->device);
    const c128* d = reinterpret_cast<const c128*>(diag);
    BatchResult R;
    if (precision == VS_FP64) {
      Engine<double> E(op->ctx, op->nunk);
      R = solve_bicgstab_batch<double>(E, make_linop<double>(op, pc), nrhs, (const c128*)B, (c128*)X, tol, maxit, d);
    } else {
      Engine<float> E(op->ctx, op->nunk);
      R = solve_bicgstab_batch<float>(E, make_linop<float>(op, pc), nrhs, (const c64*)B, (c64*)X, tol, maxit, d);
    }
    for (int i = 0; i < nrhs; ++i) {
      flags[i] = R.flag[i];
      relres[i] = R.relres[i];
      iters[i] = R.iter[i];
    }
  });
}

int vs_bicgstab_batch_dev(vs_op* op, vs_pc* pc, int nrhs, const void* B, void* X, int precision, double tol,
                          int maxit, int* flags, double* relres, double* iters) {
  return vs_bicgstab_batch_diag_dev(op, pc, nrhs, B, X, nullptr, precision, tol, maxit, flags, relres, iters);
}

int vs_refine_batch_diag_dev(vs_op* op, vs_pc* pc, int nrhs, const void* B, void* X, const double* diag, double tol,
                             int max_outer, double inner_tol, int inner_maxit, int* flags, double* relres, int* outer,
                             int* inner_total) {
  return guarded([&] {
    VS_CHECK(op && B && X && flags && relres && outer && inner_total && nrhs >= 1, VS_ERR_ARG, "bad arguments");
    VS_CHECK((op->precisions & (VS_FP64 | VS_FP32)) == (VS_FP64 | VS_FP32), VS_ERR_ARG,
             "refinement needs an operator holding both FP64 and FP32 tensors");
    check_solver_args(op, pc, VS_FP32);
    DeviceGuard g(op->ctx->device);
    Engine<double> E64(op->ctx, op->nunk);
    Engine<float> E32(op->ctx, op->nunk);
    BatchResult R = solve_refine_batch(E64, E32, make_linop<double>(op, nullptr), make_linop<float>(op, pc), nrhs,
                                       (const c128*)B, (c128*)X, tol, max_outer, inner_tol, inner_maxit, outer,
                                       inner_total, reinterpret_cast<const c128*>(diag));
    for (int i = 0; i < nrhs; ++i) {
      flags[i] = R.flag[i];
      relres[i] = R.relres[i];
    }
  });
}

int vs_refine_batch_dev(vs_op* op, vs_pc* pc, int nrhs, const void* B, void* X, double tol, int max_outer,
                        double inner_tol, int inner_maxit, int* flags, double* relres, int* outer, int* inner_total) {
  return vs_refine_batch_diag_dev(op, pc, nrhs, B, X, nullptr, tol, max_outer, inner_tol, inner_maxit, flags, relres,
                                  outer, inner_total);
}

}  // extern "C"

namespace {
template <class F>
double time_on_stream(vs_ctx* ctx, int nrep, F&& body) {
  hipEvent_t e0, e1;
  VS_HIP(hipEventCreate(&e0));
  VS_HIP(hipEventCreate(&e1));
  VS_HIP(hipEventRecord(e0, ctx->stream));
  for (int i = 0; i < nrep; ++i) body();
  VS_HIP(hipEventRecord(e1, ctx->stream));
  VS_HIP(hipEventSynchronize(e1));
  float ms = 0;
  VS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / nrep;
}
}  // namespace

extern "C" {

int vs_bench_mvp(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs, int nrep, double* ms) {
  return guarded([&] {
    VS_CHECK(op && x && y && ms && nrep >= 1, VS_ERR_ARG, "bad arguments");
    DeviceGuard g(op->ctx->device);
    mvp_dev(op, x, y, transp, precision, nrhs);  // workspace allocation outside the timed region
    *ms = time_on_stream(op->ctx, nrep, [&] { mvp_dev(op, x, y, transp, precision, nrhs); });
  });
}

int vs_bench_pc(vs_pc* pc, const void* x, void* y, int precision, int nrhs, int nrep, double* ms) {
  return guarded([&] {
    VS_CHECK(pc && x && y && ms && nrep >= 1, VS_ERR_ARG, "bad arguments");
    DeviceGuard g(pc->ctx->device);
    pc_apply_dev(pc, x, y, precision, nrhs);
    *ms = time_on_stream(pc->ctx, nrep, [&] { pc_apply_dev(pc, x, y, precision, nrhs); });
  });
}

int vs_bench_pc_blocks(vs_pc* pc, int precision, int nrhs, int nrep, double* ms) {
  return guarded([&] {
    VS_CHECK(pc && ms && nrep >= 1 && nrhs >= 1, VS_ERR_ARG, "bad arguments");
    DeviceGuard g(pc->ctx->device);
    pc_blocks_dev(pc, precision, nrhs);
    *ms = time_on_stream(pc->ctx, nrep, [&] { pc_blocks_dev(pc, precision, nrhs); });
  });
}

}  // extern "C"
