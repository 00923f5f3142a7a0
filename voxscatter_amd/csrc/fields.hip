// Far field and cross sections of a DDA polarisation (no counterpart in the reference, which lists them
// as future work; Draine & Flatau conventions, Gaussian units):
//   dipoles        p_j = P_j dV
//   far field      F(n) = k^2 sum_j (I - n n) p_j exp(-i k n . r_j)      (E_sca ~ F exp(ikr)/r)
//   extinction     C_ext = 4 pi k / |E0|^2 sum_j Im(Einc_j^* . p_j)
//   absorption     C_abs = 4 pi k / |E0|^2 sum_j [ -Im(1/alpha_j) |P_j|^2 dV - (2/3) k^3 |p_j|^2 ]
// where alpha_j is the volume-normalised polarisability (the diagonal of the DDA system is 1/alpha_j).
#include "handles.hpp"

#include <cmath>

namespace vs {

// one wavefront per direction; lanes stride over the voxels
__global__ __launch_bounds__(256) void k_farfield(const double* __restrict__ r, const c128* __restrict__ P,
                                                  int64_t nvox, double dV, double k, const double* __restrict__ dirs,
                                                  int64_t ndir, c128* __restrict__ F) {
  const int lane = threadIdx.x & 63;
  const int64_t d = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (d >= ndir) return;
  const double nx = dirs[3 * d], ny = dirs[3 * d + 1], nz = dirs[3 * d + 2];
  double ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;  // sum p_j exp(-i k n.r_j), real / imag
  for (int64_t j = lane; j < nvox; j += 64) {
    const double ph = -k * (nx * r[3 * j] + ny * r[3 * j + 1] + nz * r[3 * j + 2]);
    double s, c;
    sincos(ph, &s, &c);
    const c128 px = P[j], py = P[nvox + j], pz = P[2 * nvox + j];
    ax += px.x * c - px.y * s;
    bx += px.x * s + px.y * c;
    ay += py.x * c - py.y * s;
    by += py.x * s + py.y * c;
    az += pz.x * c - pz.y * s;
    bz += pz.x * s + pz.y * c;
  }
  for (int off = 32; off >= 1; off >>= 1) {
    ax += __shfl_xor(ax, off);
    bx += __shfl_xor(bx, off);
    ay += __shfl_xor(ay, off);
    by += __shfl_xor(by, off);
    az += __shfl_xor(az, off);
    bz += __shfl_xor(bz, off);
  }
  if (lane == 0) {
    // (I - n n) applied to the summed moment
    const double dr = nx * ax + ny * ay + nz * az, di = nx * bx + ny * by + nz * bz;
    const double s = k * k * dV;
    F[3 * d] = mk<double>(s * (ax - nx * dr), s * (bx - nx * di));
    F[3 * d + 1] = mk<double>(s * (ay - ny * dr), s * (by - ny * di));
    F[3 * d + 2] = mk<double>(s * (az - nz * dr), s * (bz - nz * di));
  }
}

// per-workgroup partial sums of Im(Einc^* . P) and of -Im(1/alpha_j)|P_j|^2, |P_j|^2
__global__ __launch_bounds__(256) void k_xsec_partial(const c128* __restrict__ P, const c128* __restrict__ Einc,
                                                      const c128* __restrict__ ainv, int ainv_scalar, int64_t nvox,
                                                      double* __restrict__ part) {
  __shared__ double sh[3][256];
  double e = 0, a = 0, q = 0;
  for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < nvox; j += (int64_t)gridDim.x * 256) {
    const c128 ai = ainv_scalar ? ainv[0] : ainv[j];
    double p2 = 0;
    for (int c = 0; c < 3; ++c) {
      const c128 p = P[c * nvox + j], E = Einc[c * nvox + j];
      e += E.x * p.y - E.y * p.x;  // Im(conj(E) p)
      p2 += p.x * p.x + p.y * p.y;
    }
    a += -ai.y * p2;
    q += p2;
  }
  sh[0][threadIdx.x] = e;
  sh[1][threadIdx.x] = a;
  sh[2][threadIdx.x] = q;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) part[3 * blockIdx.x + c] = sh[c][0];
}

// g_j = -dainv_j * sum_c lambda_{c,j} P_{c,j}: derivative of a holomorphic objective w^T P with respect to
// eps_j, for the adjoint solution lambda of A^T lambda = w (A^T = A for the DDA system)
__global__ void k_grad_eps(const c128* __restrict__ P, const c128* __restrict__ lam, const c128* __restrict__ dainv,
                           int64_t nvox, c128* __restrict__ g) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nvox) return;
  c128 acc = mk<double>(0, 0);
  for (int c = 0; c < 3; ++c) acc = acc + lam[c * nvox + j] * P[c * nvox + j];
  g[j] = -(dainv[j] * acc);
}

}  // namespace vs

using namespace vs;

// plane waves E_p exp(i k khat_d . r_v) at the occupied voxels (PlaneWavePointWise.m:7-19) for ndir
// directions, npol polarisations each: B[(npol d + p) 3 nvox + c nvox + v] = e[d][p][c] exp(i k khat_d . r_v);
// one thread per (voxel, direction), the phase in FP64
__global__ void k_plane_waves(const double* __restrict__ r, int64_t nvox, const double* __restrict__ khat,
                              const double* __restrict__ pol, int ndir, int npol, double k, c128* __restrict__ B) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int d = blockIdx.y;
  if (v >= nvox) return;
  const double ph = k * (khat[3 * d] * r[3 * v] + khat[3 * d + 1] * r[3 * v + 1] + khat[3 * d + 2] * r[3 * v + 2]);
  double sn, cs;
  sincos(ph, &sn, &cs);
  for (int p = 0; p < npol; ++p) {
    const double* e = pol + 3 * (npol * d + p);
    c128* row = B + (int64_t)(npol * d + p) * 3 * nvox;
    for (int c = 0; c < 3; ++c) row[c * nvox + v] = mk<double>(e[c] * cs, e[c] * sn);
  }
}

extern "C" {

int vs_plane_wave_batch_dev(vs_ctx* ctx, int64_t nvox, const double* r, int ndir, const double* khat, int npol,
                            const double* pol, double k, void* B) {
  return guarded([&] {
    VS_CHECK(ctx && r && khat && pol && B && nvox >= 1 && ndir >= 1 && npol >= 1 && ndir <= 65535, VS_ERR_ARG,
             "bad arguments");
    DeviceGuard g(ctx->device);
    DevBuf dr(sizeof(double) * 3 * nvox), dk(sizeof(double) * 3 * ndir), dp(sizeof(double) * 3 * ndir * npol);
    VS_HIP(hipMemcpyAsync(dr.p, r, sizeof(double) * 3 * nvox, hipMemcpyHostToDevice, ctx->stream));
    VS_HIP(hipMemcpyAsync(dk.p, khat, sizeof(double) * 3 * ndir, hipMemcpyHostToDevice, ctx->stream));
    VS_HIP(hipMemcpyAsync(dp.p, pol, sizeof(double) * 3 * ndir * npol, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_plane_waves, dim3((unsigned)((nvox + 255) / 256), (unsigned)ndir), dim3(256), 0,
                       ctx->stream, dr.as<double>(), nvox, dk.as<double>(), dp.as<double>(), ndir, npol, k,
                       reinterpret_cast<c128*>(B));
    VS_HIP(hipGetLastError());
    VS_HIP(hipStreamSynchronize(ctx->stream));  // the staging buffers are released on return
  });
}

int vs_dda_farfield(vs_ctx* ctx, int64_t nvox, const double* r, const double* P, double dV, double k, int64_t ndir,
                    const double* dirs, double* F) {
  return guarded([&] {
    VS_CHECK(ctx && r && P && dirs && F && nvox >= 1 && ndir >= 1, VS_ERR_ARG, "bad arguments");
    DeviceGuard g(ctx->device);
    DevBuf dr(sizeof(double) * 3 * nvox), dP(sizeof(c128) * 3 * nvox), dd(sizeof(double) * 3 * ndir),
        dF(sizeof(c128) * 3 * ndir);
    VS_HIP(hipMemcpyAsync(dr.p, r, sizeof(double) * 3 * nvox, hipMemcpyHostToDevice, ctx->stream));
    VS_HIP(hipMemcpyAsync(dP.p, P, sizeof(c128) * 3 * nvox, hipMemcpyHostToDevice, ctx->stream));
    VS_HIP(hipMemcpyAsync(dd.p, dirs, sizeof(double) * 3 * ndir, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_farfield, dim3((unsigned)((ndir + 3) / 4)), dim3(256), 0, ctx->stream, dr.as<double>(),
                       dP.as<c128>(), nvox, dV, k, dd.as<double>(), ndir, dF.as<c128>());
    VS_HIP(hipGetLastError());
    VS_HIP(hipMemcpyAsync(F, dF.p, sizeof(c128) * 3 * ndir, hipMemcpyDeviceToHost, ctx->stream));
    VS_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int vs_dda_cross_sections(vs_ctx* ctx, int64_t nvox, const double* P, const double* Einc, const double* alpha_inv,
                          int64_t ainv_len, double dV, double k, double E0sq, double out[3]) {
  return guarded([&] {
    VS_CHECK(ctx && P && Einc && alpha_inv && out && nvox >= 1 && E0sq > 0, VS_ERR_ARG, "bad arguments");
    VS_CHECK(ainv_len == 1 || ainv_len == nvox, VS_ERR_ARG, "alpha_inv must hold 1 or nvox values");
    DeviceGuard g(ctx->device);
    const int grid = (int)std::min<int64_t>(1024, (nvox + 255) / 256);
    DevBuf dP(sizeof(c128) * 3 * nvox), dE(sizeof(c128) * 3 * nvox), da(sizeof(c128) * ainv_len),
        part(sizeof(double) * 3 * grid);
    VS_HIP(hipMemcpyAsync(dP.p, P, sizeof(c128) * 3 * nvox, hipMemcpyHostToDevice, ctx->stream));
    VS_HIP(hipMemcpyAsync(dE.p, Einc, sizeof(c128) * 3 * nvox, hipMemcpyHostToDevice, ctx->stream));
    VS_HIP(hipMemcpyAsync(da.p, alpha_inv, sizeof(c128) * ainv_len, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_xsec_partial, dim3(grid), dim3(256), 0, ctx->stream, dP.as<c128>(), dE.as<c128>(),
                       da.as<c128>(), ainv_len == 1 ? 1 : 0, nvox, part.as<double>());
    VS_HIP(hipGetLastError());
    std::vector<double> h(3 * grid);
    VS_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(double) * 3 * grid, hipMemcpyDeviceToHost, ctx->stream));
    VS_HIP(hipStreamSynchronize(ctx->stream));
    double e = 0, a = 0, q = 0;
    for (int b = 0; b < grid; ++b) {
      e += h[3 * b];
      a += h[3 * b + 1];
      q += h[3 * b + 2];
    }
    const double pre = 4.0 * M_PI * k / E0sq;
    const double cext = pre * e * dV;
    const double cabs = pre * (a * dV - (2.0 / 3.0) * k * k * k * q * dV * dV);
    out[0] = cext;
    out[1] = cabs;
    out[2] = cext - cabs;
  });
}

int vs_dda_grad_eps(vs_ctx* ctx, int64_t nvox, const double* P, const double* lam, const double* dainv, double* g) {
  return guarded([&] {
    VS_CHECK(ctx && P && lam && dainv && g && nvox >= 1, VS_ERR_ARG, "bad arguments");
    DeviceGuard gd(ctx->device);
    DevBuf dP(sizeof(c128) * 3 * nvox), dl(sizeof(c128) * 3 * nvox), da(sizeof(c128) * nvox), dg(sizeof(c128) * nvox);
    VS_HIP(hipMemcpyAsync(dP.p, P, sizeof(c128) * 3 * nvox, hipMemcpyHostToDevice, ctx->stream));
    VS_HIP(hipMemcpyAsync(dl.p, lam, sizeof(c128) * 3 * nvox, hipMemcpyHostToDevice, ctx->stream));
    VS_HIP(hipMemcpyAsync(da.p, dainv, sizeof(c128) * nvox, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_grad_eps, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, ctx->stream, dP.as<c128>(),
                       dl.as<c128>(), da.as<c128>(), nvox, dg.as<c128>());
    VS_HIP(hipGetLastError());
    VS_HIP(hipMemcpyAsync(g, dg.p, sizeof(c128) * nvox, hipMemcpyDeviceToHost, ctx->stream));
    VS_HIP(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
