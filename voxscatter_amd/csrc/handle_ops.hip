// Operator "function handles" beyond the FFT operator (iterapp.m:41-69: A, M1 may be a matrix or a
// handle): a dense matrix operand, and a preconditioner that is the matvec of another operator (the
// material preconditioner of JVIE_Solver_const.m:232-234, mv_AN_const with 1./Mr and 1./Mr - 1).
#include "handles.hpp"
#include "solver.hpp"

namespace vs {

// y(:, r) = A x(:, r) for a column-major n x n A: one lane per row, the columns streamed in order
template <typename T>
__global__ __launch_bounds__(256) void k_dense_mv(const cplx<T>* __restrict__ A, const cplx<T>* __restrict__ x,
                                                  cplx<T>* __restrict__ y, int64_t n) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const cplx<T>* xr = x + (int64_t)blockIdx.y * n;
  if (r >= n) return;
  cplx<T> acc = mk<T>(0, 0);
  for (int64_t c = 0; c < n; ++c) acc = acc + A[r + n * c] * xr[c];
  y[(int64_t)blockIdx.y * n + r] = acc;
}

// y(:, r) = A^H x(:, r): one wavefront per column of A (contiguous), shuffle reduction
template <typename T>
__global__ __launch_bounds__(256) void k_dense_mvh(const cplx<T>* __restrict__ A, const cplx<T>* __restrict__ x,
                                                   cplx<T>* __restrict__ y, int64_t n) {
  const int lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const cplx<T>* xr = x + (int64_t)blockIdx.y * n;
  if (c >= n) return;
  T re = 0, im = 0;
  for (int64_t r = lane; r < n; r += 64) {
    const cplx<T> a = A[r + n * c], v = xr[r];
    re += a.x * v.x + a.y * v.y;  // conj(a) v
    im += a.x * v.y - a.y * v.x;
  }
  for (int off = 32; off >= 1; off >>= 1) {
    re += __shfl_xor(re, off);
    im += __shfl_xor(im, off);
  }
  if (lane == 0) y[(int64_t)blockIdx.y * n + c] = mk<T>(re, im);
}

template <typename T>
static void dense_mv(vs_op* op, const cplx<T>* x, cplx<T>* y, int transp, int nrhs) {
  const cplx<T>* A = sizeof(T) == 8 ? op->a64.as<cplx<T>>() : op->a32.as<cplx<T>>();
  VS_CHECK(A != nullptr, VS_ERR_ARG, "dense operator holds no matrix in the requested precision");
  const int64_t n = op->nunk;
  if (!transp) {
    hipLaunchKernelGGL((k_dense_mv<T>), dim3((unsigned)((n + 255) / 256), (unsigned)nrhs), dim3(256), 0,
                       op->ctx->stream, A, x, y, n);
  } else {
    hipLaunchKernelGGL((k_dense_mvh<T>), dim3((unsigned)((n + 3) / 4), (unsigned)nrhs), dim3(256), 0,
                       op->ctx->stream, A, x, y, n);
  }
  VS_HIP(hipGetLastError());
}

void dense_mvp_dev(vs_op* op, const void* x, void* y, int transp, int precision, int nrhs) {
  VS_CHECK(nrhs >= 1, VS_ERR_ARG, "nrhs must be >= 1");
  VS_CHECK(x != y, VS_ERR_ARG, "matvec input and output must differ");
  if (precision == VS_FP64)
    dense_mv<double>(op, (const c128*)x, (c128*)y, transp, nrhs);
  else if (precision == VS_FP32)
    dense_mv<float>(op, (const c64*)x, (c64*)y, transp, nrhs);
  else
    throw Error{VS_ERR_ARG, "precision must be VS_FP64 or VS_FP32"};
}

template <typename To>
__global__ void k_cast_dense(const c128* __restrict__ in, cplx<To>* __restrict__ out, int64_t tot) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < tot) out[e] = ccast<To>(in[e]);
}

}  // namespace vs

using namespace vs;

extern "C" {

int vs_op_create_dense(vs_ctx* ctx, int64_t n, const double* A, int precisions, vs_op** out) {
  return guarded([&] {
    VS_CHECK(ctx && A && out, VS_ERR_ARG, "null pointer");
    *out = nullptr;
    VS_CHECK(n >= 1 && n <= (int64_t)1 << 20, VS_ERR_ARG, "dense operator size must be between 1 and 2^20");
    VS_CHECK((precisions & (VS_FP64 | VS_FP32)) != 0, VS_ERR_ARG, "precisions must include VS_FP64 or VS_FP32");
    DeviceGuard g(ctx->device);
    auto op = std::make_unique<vs_op>();
    op->ctx = ctx;
    op->dense = true;
    op->nunk = n;
    op->precisions = precisions;
    const size_t tot = (size_t)n * (size_t)n;
    DevBuf a(sizeof(c128) * tot);
    VS_HIP(hipMemcpyAsync(a.p, A, sizeof(c128) * tot, hipMemcpyHostToDevice, ctx->stream));
    const unsigned grid = (unsigned)((tot + 255) / 256);
    if (precisions & VS_FP32) {
      op->a32.ensure(sizeof(c64) * tot);
      hipLaunchKernelGGL(k_cast_dense<float>, dim3(grid), dim3(256), 0, ctx->stream, a.as<c128>(), op->a32.as<c64>(),
                         (int64_t)tot);
      VS_HIP(hipGetLastError());
    }
    VS_HIP(hipStreamSynchronize(ctx->stream));
    if (precisions & VS_FP64) op->a64 = std::move(a);
    ctx_retain(ctx);
    *out = op.release();
  });
}

int vs_pc_create_op(vs_op* mop, vs_pc** out) {
  return guarded([&] {
    VS_CHECK(mop && out, VS_ERR_ARG, "null pointer");
    *out = nullptr;
    VS_CHECK(mop->nunk > 0 && (mop->dense || mop->map3.p != nullptr), VS_ERR_ARG, "operator domain not set");
    auto pc = std::make_unique<vs_pc>();
    pc->ctx = mop->ctx;
    pc->mop = mop;
    pc->levels = 0;
    pc->nunk = mop->nunk;
    pc->precisions = mop->precisions;
    ctx_retain(pc->ctx);
    *out = pc.release();
  });
}

}  // extern "C"
