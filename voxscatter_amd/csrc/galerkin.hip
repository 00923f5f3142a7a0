// Galerkin (piecewise-constant JVIE) N-operator Toeplitz tensor (assembly_nop_const.m:1-174):
//   J_mn(m)(i, j) = int_{face i of cell m} int_{face j of cell 0} exp(-i k R) / R  dS dS'
// over the 6 faces of the cubes (cube_faces.m), with 4-point Gauss rules per face dimension for far
// cells, 10 for cells within 5 of the origin cell and 20 for the 2 x 2 x 2 nearest cells, where the
// (weakly) singular face pairs are replaced by separately computed integrals (nearby_correct.m).  The
// six tensor components are G(kk) = 1/(4 pi) sum_ij J(i, j) P(i, j, kk), P(i, j, kk) the projection
// (n_i x e_a) . (n_j x e_b) of vector_projection.m.
//
// The singular face-pair integrals (coincident, edge-adjacent and vertex-adjacent squares, coplanar or
// orthogonal) are evaluated by this library's own method: for an observation point r the inner integral
// over a square, in polar coordinates about the projection of r on the square's plane (height h), is
//   int exp(-ikR)/R dA = sum_edges sign(t_e) int du G(|t_e| cosh u) / cosh u,
//   G(rho) = (exp(-ik h) - exp(-ik sqrt(rho^2 + h^2))) / (ik),
// with t_e the signed distance to the edge line and s = |t_e| sinh u the coordinate along the edge (the
// radial integral is done in closed form, the sinh substitution removes the near-singularity when r
// approaches an edge line); the outer integral runs over a tensor Gauss rule graded geometrically
// towards the edges of the observation square.  The static self term of a square of side a reproduces
// a^3 (4/3 (1 - sqrt 2) + 4 ln(1 + sqrt 2)).
#include "handles.hpp"

#include <cmath>
#include <vector>

namespace vs {

struct Sq {  // square o + s a + t b, s, t in [0, 1]
  double o[3], a[3], b[3];
};

__device__ __forceinline__ double d3(const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; }

// (1 - exp(-ix)) / (ix), stable for small x
__device__ __forceinline__ c128 onem_exp_over(double x) {
  if (fabs(x) < 1e-3) return mk<double>(1 - x * x / 6 + x * x * x * x / 120, -x / 2 + x * x * x / 24);
  double s, c;
  sincos(x, &s, &c);
  // (1 - c + i s) / (i x) = (s - i (1 - c)) / x
  return mk<double>(s / x, -(1 - c) / x);
}

__device__ c128 sq_inner(const Sq& Q, const double* r, double k, const double* gx, const double* gw, int ng) {
  double n[3] = {Q.a[1] * Q.b[2] - Q.a[2] * Q.b[1], Q.a[2] * Q.b[0] - Q.a[0] * Q.b[2], Q.a[0] * Q.b[1] - Q.a[1] * Q.b[0]};
  const double nn = sqrt(d3(n, n));
  for (int i = 0; i < 3; ++i) n[i] /= nn;
  double ro[3] = {r[0] - Q.o[0], r[1] - Q.o[1], r[2] - Q.o[2]};
  const double hs = d3(ro, n), h = fabs(hs);
  const double P[3] = {r[0] - hs * n[0], r[1] - hs * n[1], r[2] - hs * n[2]};
  double v[4][3];
  for (int i = 0; i < 3; ++i) {
    v[0][i] = Q.o[i];
    v[1][i] = Q.o[i] + Q.a[i];
    v[2][i] = Q.o[i] + Q.a[i] + Q.b[i];
    v[3][i] = Q.o[i] + Q.b[i];
  }
  double eh, ch;
  sincos(-k * h, &eh, &ch);
  const c128 ekh = mk<double>(ch, eh);  // exp(-ik h)
  c128 sum = mk<double>(0, 0);
  for (int e = 0; e < 4; ++e) {
    const double* A = v[e];
    const double* B = v[(e + 1) & 3];
    double td[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    const double len = sqrt(d3(td, td));
    for (int i = 0; i < 3; ++i) td[i] /= len;
    const double nu[3] = {td[1] * n[2] - td[2] * n[1], td[2] * n[0] - td[0] * n[2], td[0] * n[1] - td[1] * n[0]};
    const double AP[3] = {A[0] - P[0], A[1] - P[1], A[2] - P[2]};
    const double BP[3] = {B[0] - P[0], B[1] - P[1], B[2] - P[2]};
    const double t = d3(AP, nu), at = fabs(t);
    if (at < 1e-14 * len) continue;  // projection on the edge line: zero-area sector
    const double uA = asinh(d3(AP, td) / at), uB = asinh(d3(BP, td) / at);
    const double hu = 0.5 * (uB - uA), cu = 0.5 * (uB + uA);
    c128 acc = mk<double>(0, 0);
    for (int i = 0; i < ng; ++i) {
      const double u = cu + hu * gx[i];
      const double chu = cosh(u);
      const double rho = at * chu;
      const double R = sqrt(rho * rho + h * h);
      const double dr = rho * rho / (R + h);  // R - h
      const c128 g = ekh * (dr * onem_exp_over(k * dr));
      acc = acc + (gw[i] / chu) * g;
    }
    sum = sum + ((t > 0 ? hu : -hu) * acc);
  }
  return sum;
}

// one thread per (outer point i, j); blockIdx.y = integral; per-workgroup partial sums
__global__ __launch_bounds__(256) void k_singular(const Sq* __restrict__ obs, const Sq* __restrict__ src, double k,
                                                  const double* __restrict__ ox, const double* __restrict__ ow, int no,
                                                  const double* __restrict__ gx, const double* __restrict__ gw, int ng,
                                                  c128* __restrict__ part) {
  __shared__ double sre[256], sim[256];
  const int q = blockIdx.y;
  const Sq S1 = obs[q], S2 = src[q];
  const int64_t tot = (int64_t)no * no;
  double ar = 0, ai = 0;
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < tot; p += (int64_t)gridDim.x * 256) {
    const int i = (int)(p / no), j = (int)(p % no);
    const double r[3] = {S1.o[0] + ox[i] * S1.a[0] + ox[j] * S1.b[0], S1.o[1] + ox[i] * S1.a[1] + ox[j] * S1.b[1],
                         S1.o[2] + ox[i] * S1.a[2] + ox[j] * S1.b[2]};
    const c128 f = sq_inner(S2, r, k, gx, gw, ng);
    const double w = ow[i] * ow[j];
    ar += w * f.x;
    ai += w * f.y;
  }
  sre[threadIdx.x] = ar;
  sim[threadIdx.x] = ai;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sre[threadIdx.x] += sre[threadIdx.x + s];
      sim[threadIdx.x] += sim[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)q * gridDim.x + blockIdx.x] = mk<double>(sre[0], sim[0]);
}

// ------------------------------------------------------------------------------------------------
// regular face-face cubature
// ------------------------------------------------------------------------------------------------
// face f of the unit cube: corner c0 and edge vectors ea (c1 - c0), eb (c3 - c0), in units of dx
// (cube_faces.m: -x, +x, -y, +y, -z, +z)
__constant__ double c_face[6][3][3] = {
    {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}}, {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, {{0, 0, 0}, {1, 0, 0}, {0, 0, 1}},
    {{0, 1, 0}, {1, 0, 0}, {0, 0, 1}}, {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}}, {{0, 0, 1}, {1, 0, 0}, {0, 1, 0}}};
// the 21 face pairs computed; the other 15 follow by symmetry (cubature_nop_const.cpp)
__constant__ int c_pair_obs[21] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 5};
__constant__ int c_pair_src[21] = {0, 1, 2, 3, 4, 5, 0, 2, 3, 4, 5, 2, 3, 4, 5, 2, 4, 5, 4, 5, 4};

// sum_{p, q} w_p w_q exp(-ik R)/R between face fo of the cell at offset m and face fs of cell 0, in
// units where dx = 1 (k scaled by dx); u, v, w: the rule on [0,1]^2 (weights summing to 1)
__device__ c128 face_pair(int fo, int fs, const double* m, double kd, const double* u, const double* v,
                          const double* w, int np) {
  c128 acc = mk<double>(0, 0);
  for (int p = 0; p < np; ++p) {
    double rp[3];
    for (int d = 0; d < 3; ++d) rp[d] = c_face[fo][0][d] + c_face[fo][1][d] * u[p] + c_face[fo][2][d] * v[p] + m[d];
    double ar = 0, ai = 0;
    for (int q = 0; q < np; ++q) {
      double dd[3];
      for (int d = 0; d < 3; ++d) dd[d] = c_face[fs][0][d] + c_face[fs][1][d] * u[q] + c_face[fs][2][d] * v[q] - rp[d];
      const double R = sqrt(d3(dd, dd));
      double s, c;
      sincos(kd * R, &s, &c);
      const double wr = w[q] / R;
      ar += wr * c;
      ai -= wr * s;
    }
    acc = acc + w[p] * mk<double>(ar, ai);
  }
  return acc;
}

// 6 x 6 symmetric fill of the 21 computed pairs (cubature_nop_const.cpp:104-125)
__host__ __device__ inline void fill_sym(c128 J[6][6]) {
  J[1][1] = J[0][0];
  J[2][0] = J[1][3];
  J[2][1] = J[0][3];
  J[3][0] = J[1][2];
  J[3][1] = J[0][2];
  J[3][3] = J[2][2];
  J[4][0] = J[1][5];
  J[4][1] = J[0][5];
  J[4][2] = J[3][5];
  J[4][3] = J[2][5];
  J[5][0] = J[1][4];
  J[5][1] = J[0][4];
  J[5][2] = J[3][4];
  J[5][3] = J[2][4];
  J[5][5] = J[4][4];
}

// projection weights P(i, j, kk) (vector_projection.m), kk = xx, xy, xz, yy, yz, zz
struct ProjTable {
  double p[6][6][6];
};

// far cells: one thread per cell outside the medium box, all 21 pairs with the far rule; writes the
// six tensor components
__global__ __launch_bounds__(128) void k_gal_far(int L, int M, int N, int nmL, int nmM, int nmN, double kd,
                                                 const double* __restrict__ rule, int np, double scale, ProjTable P,
                                                 c128* __restrict__ G) {
  const int64_t V = (int64_t)L * M * N;
  const int64_t vox = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (vox >= V) return;
  const int mx = (int)(vox % L), my = (int)((vox / L) % M), mz = (int)(vox / ((int64_t)L * M));
  if (mx < nmL && my < nmM && mz < nmN) return;  // medium / near box: separate kernel
  const double m[3] = {(double)mx, (double)my, (double)mz};
  const double* u = rule;
  const double* v = rule + np;
  const double* w = rule + 2 * np;
  c128 J[6][6];
  for (int i = 0; i < 21; ++i) J[c_pair_obs[i]][c_pair_src[i]] = face_pair(c_pair_obs[i], c_pair_src[i], m, kd, u, v, w, np);
  fill_sym(J);
  for (int kk = 0; kk < 6; ++kk) {
    c128 g = mk<double>(0, 0);
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j)
        if (P.p[i][j][kk] != 0.0) g = g + P.p[i][j][kk] * J[i][j];
    G[kk * V + vox] = scale * g;
  }
}

// medium / near box: one workgroup per (box cell, pair); threads split the observation points
__global__ __launch_bounds__(256) void k_gal_box(int nmL, int nmM, int nmN, int nnL, int nnM, int nnN, double kd,
                                                 const double* __restrict__ rule_med, int np_med,
                                                 const double* __restrict__ rule_near, int np_near,
                                                 c128* __restrict__ Jbox) {
  __shared__ double sre[256], sim[256];
  const int cell = blockIdx.x / 21, pair = blockIdx.x % 21;
  const int mx = cell % nmL, my = (cell / nmL) % nmM, mz = cell / (nmL * nmM);
  const bool nearc = mx < nnL && my < nnM && mz < nnN;
  const double* rule = nearc ? rule_near : rule_med;
  const int np = nearc ? np_near : np_med;
  const double* u = rule;
  const double* v = rule + np;
  const double* w = rule + 2 * np;
  const int fo = c_pair_obs[pair], fs = c_pair_src[pair];
  const double m[3] = {(double)mx, (double)my, (double)mz};
  double ar = 0, ai = 0;
  for (int p = threadIdx.x; p < np; p += 256) {
    double rp[3];
    for (int d = 0; d < 3; ++d) rp[d] = c_face[fo][0][d] + c_face[fo][1][d] * u[p] + c_face[fo][2][d] * v[p] + m[d];
    double br = 0, bi = 0;
    for (int q = 0; q < np; ++q) {
      double dd[3];
      for (int d = 0; d < 3; ++d) dd[d] = c_face[fs][0][d] + c_face[fs][1][d] * u[q] + c_face[fs][2][d] * v[q] - rp[d];
      const double R = sqrt(d3(dd, dd));
      double s, c;
      sincos(kd * R, &s, &c);
      const double wr = w[q] / R;
      br += wr * c;
      bi -= wr * s;
    }
    ar += w[p] * br;
    ai += w[p] * bi;
  }
  sre[threadIdx.x] = ar;
  sim[threadIdx.x] = ai;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sre[threadIdx.x] += sre[threadIdx.x + s];
      sim[threadIdx.x] += sim[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) Jbox[blockIdx.x] = mk<double>(sre[0], sim[0]);
}

// ------------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------------
static void h2d(vs_ctx* ctx, void* dst, const void* src, size_t bytes) {
  VS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  VS_HIP(hipStreamSynchronize(ctx->stream));  // the sources are host temporaries
}
static void d2h(vs_ctx* ctx, void* dst, const void* src, size_t bytes) {
  VS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  VS_HIP(hipStreamSynchronize(ctx->stream));
}

static ProjTable projection_table() {
  // face normals -x, +x, -y, +y, -z, +z; components (a, b) = xx, xy, xz, yy, yz, zz
  const double nrm[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  const int ca[6] = {0, 0, 0, 1, 1, 2}, cb[6] = {0, 1, 2, 1, 2, 2};
  auto cross = [](const double* x, const double* y, double* z) {
    z[0] = x[1] * y[2] - x[2] * y[1];
    z[1] = x[2] * y[0] - x[0] * y[2];
    z[2] = x[0] * y[1] - x[1] * y[0];
  };
  ProjTable P;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j)
      for (int kk = 0; kk < 6; ++kk) {
        double ea[3] = {0, 0, 0}, eb[3] = {0, 0, 0}, x[3], y[3];
        ea[ca[kk]] = 1;
        eb[cb[kk]] = 1;
        cross(nrm[i], ea, x);
        cross(nrm[j], eb, y);
        P.p[i][j][kk] = x[0] * y[0] + x[1] * y[1] + x[2] * y[2];
      }
  return P;
}

// 2-D rule on [0,1]^2 from n-point Gauss-Legendre (gauss_2d.m): u = (x_I + 1)/2, v = (x_J + 1)/2,
// w = w_I w_J / 4; stored as [u(0..n^2), v(...), w(...)]
static std::vector<double> rule2d(int n) {
  std::vector<double> x, w;
  gauss_legendre(n, x, w);
  std::vector<double> r(3 * n * n);
  for (int I = 0; I < n; ++I)
    for (int J = 0; J < n; ++J) {
      const int c = I * n + J;
      r[c] = (x[I] + 1) / 2;
      r[n * n + c] = (x[J] + 1) / 2;
      r[2 * n * n + c] = 0.25 * w[I] * w[J];
    }
  return r;
}

// singular integrals ST, EAc, EAo, VAc, VAo for squares of side dx (singular_*_const.m geometry)
void galerkin_singular(vs_ctx* ctx, double ko, double dx, c128 out[5]) {
  const double X[3] = {dx, 0, 0}, Y[3] = {0, dx, 0}, Z[3] = {0, 0, dx};
  auto sq = [](const double* o, const double* a, const double* b) {
    Sq s;
    for (int i = 0; i < 3; ++i) {
      s.o[i] = o[i];
      s.a[i] = a[i];
      s.b[i] = b[i];
    }
    return s;
  };
  const double O[3] = {0, 0, 0}, Ox[3] = {dx, 0, 0}, Oxy[3] = {dx, dx, 0};
  Sq obs[5], src[5];
  for (int q = 0; q < 5; ++q) obs[q] = sq(O, X, Y);
  src[0] = sq(O, X, Y);     // coincident
  src[1] = sq(Ox, X, Y);    // edge-adjacent, coplanar
  src[2] = sq(Ox, Z, Y);    // edge-adjacent, orthogonal (plane x = dx, sharing the edge x = dx)
  src[3] = sq(Oxy, X, Y);   // vertex-adjacent, coplanar
  src[4] = sq(Oxy, Y, Z);   // vertex-adjacent, orthogonal (plane x = dx, sharing the vertex (dx, dx, 0))
  // outer rule: 16-point Gauss on each of 2 x 9 intervals graded geometrically (ratio 0.15) to both ends
  std::vector<double> gx, gw;
  gauss_legendre(16, gx, gw);
  std::vector<double> br{0};
  for (int l = 8; l >= 1; --l) br.push_back(0.5 * std::pow(0.15, l));
  br.push_back(0.5);
  std::vector<double> ox, ow;
  for (size_t i = 0; i + 1 < br.size(); ++i)
    for (int j = 0; j < 16; ++j) {
      ox.push_back(0.5 * (br[i] + br[i + 1]) + 0.5 * (br[i + 1] - br[i]) * gx[j]);
      ow.push_back(0.5 * (br[i + 1] - br[i]) * gw[j]);
    }
  const size_t half = ox.size();
  for (size_t i = 0; i < half; ++i) {
    ox.push_back(1 - ox[i]);
    ow.push_back(ow[i]);
  }
  std::vector<double> ix, iw;
  gauss_legendre(32, ix, iw);
  const int no = (int)ox.size(), ng = 32;
  const int nblk = 256;
  DevBuf dobs(sizeof(obs)), dsrc(sizeof(src)), dox(sizeof(double) * no), dow(sizeof(double) * no),
      dgx(sizeof(double) * ng), dgw(sizeof(double) * ng), part(sizeof(c128) * 5 * nblk);
  h2d(ctx, dobs.p, obs, sizeof(obs));
  h2d(ctx, dsrc.p, src, sizeof(src));
  h2d(ctx, dox.p, ox.data(), sizeof(double) * no);
  h2d(ctx, dow.p, ow.data(), sizeof(double) * no);
  h2d(ctx, dgx.p, ix.data(), sizeof(double) * ng);
  h2d(ctx, dgw.p, iw.data(), sizeof(double) * ng);
  hipLaunchKernelGGL(k_singular, dim3(nblk, 5), dim3(256), 0, ctx->stream, dobs.as<Sq>(), dsrc.as<Sq>(), ko,
                     dox.as<double>(), dow.as<double>(), no, dgx.as<double>(), dgw.as<double>(), ng, part.as<c128>());
  VS_HIP(hipGetLastError());
  std::vector<c128> h(5 * nblk);
  d2h(ctx, h.data(), part.p, sizeof(c128) * 5 * nblk);
  const double area = dx * dx;
  for (int q = 0; q < 5; ++q) {
    double re = 0, im = 0;
    for (int b = 0; b < nblk; ++b) {
      re += h[q * nblk + b].x;
      im += h[q * nblk + b].y;
    }
    out[q] = mk<double>(area * re, area * im);
  }
}

// nearby_correct.m: in the 2 x 2 x 2 box the face pairs that touch take a singular integral instead of
// the cubature: code 1 ST (the squares coincide), 2 EAc / 3 EAo (they share an edge, coplanar /
// orthogonal), 4 VAc / 5 VAo (they share a vertex only); keep = 1 for disjoint squares, which keep the
// cubature.  The table is derived here from the face geometry (observation face on the cell at offset
// (mx, my, mz), source face on cell 0, both closed axis-aligned unit squares); it equals the matrices
// written out in nearby_correct.m.
static const int h_face[6][3][3] = {
    {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}}, {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, {{0, 0, 0}, {1, 0, 0}, {0, 0, 1}},
    {{0, 1, 0}, {1, 0, 0}, {0, 0, 1}}, {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}}, {{0, 0, 1}, {1, 0, 0}, {0, 1, 0}}};

static void nearby_table(int mx, int my, int mz, int code[6][6], int keep[6][6]) {
  const int off[3] = {mx, my, mz};
  for (int fo = 0; fo < 6; ++fo)
    for (int fs = 0; fs < 6; ++fs) {
      int dim = 0, nrm_o = 0, nrm_s = 0;
      bool empty = false;
      for (int d = 0; d < 3; ++d) {
        const int ext_o = h_face[fo][1][d] + h_face[fo][2][d], ext_s = h_face[fs][1][d] + h_face[fs][2][d];
        if (!ext_o) nrm_o = d;
        if (!ext_s) nrm_s = d;
        const int lo = std::max(h_face[fo][0][d] + off[d], h_face[fs][0][d]);
        const int hi = std::min(h_face[fo][0][d] + off[d] + ext_o, h_face[fs][0][d] + ext_s);
        if (lo > hi) empty = true;
        if (hi > lo) ++dim;
      }
      keep[fo][fs] = empty ? 1 : 0;
      code[fo][fs] = empty ? 0 : (dim == 2 ? 1 : dim == 1 ? (nrm_o == nrm_s ? 2 : 3) : (nrm_o == nrm_s ? 4 : 5));
    }
}

// Toeplitz tensor (L x M x N x 6, FP64, device) of the Galerkin N operator
void galerkin_toeplitz(vs_ctx* ctx, int L, int M, int N, double dx, double ko, const c128 sing[5], c128* G) {
  const int64_t V = (int64_t)L * M * N;
  const double kd = ko * dx;
  // in units of dx the face-face sums carry dx^3: dx^4 of area Jacobians and 1/dx of 1/R
  const double scale = dx * dx * dx / (4.0 * M_PI);
  const ProjTable P = projection_table();
  // medium box (assembly_nop_const.m:21-29) and near box (:96-146)
  const int nmL = std::min(5, L), nmM = std::min(5, M), nmN = std::min(5, N);
  const int nnL = 2, nnM = 2, nnN = 2;  // Galerkin grids have L, M, N >= 2
  std::vector<double> rf = rule2d(4), rm = rule2d(10), rn = rule2d(20);
  DevBuf drf(sizeof(double) * rf.size()), drm(sizeof(double) * rm.size()), drn(sizeof(double) * rn.size());
  h2d(ctx, drf.p, rf.data(), sizeof(double) * rf.size());
  h2d(ctx, drm.p, rm.data(), sizeof(double) * rm.size());
  h2d(ctx, drn.p, rn.data(), sizeof(double) * rn.size());
  hipLaunchKernelGGL(k_gal_far, dim3((unsigned)((V + 127) / 128)), dim3(128), 0, ctx->stream, L, M, N, nmL, nmM, nmN,
                     kd, drf.as<double>(), 16, scale, P, G);
  VS_HIP(hipGetLastError());
  const int nbox = nmL * nmM * nmN;
  DevBuf jb(sizeof(c128) * nbox * 21);
  hipLaunchKernelGGL(k_gal_box, dim3((unsigned)(nbox * 21)), dim3(256), 0, ctx->stream, nmL, nmM, nmN, nnL, nnM, nnN,
                     kd, drm.as<double>(), 100, drn.as<double>(), 400, jb.as<c128>());
  VS_HIP(hipGetLastError());
  std::vector<c128> hj(nbox * 21);
  d2h(ctx, hj.data(), jb.p, sizeof(c128) * nbox * 21);
  static const int po[21] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 5};
  static const int ps[21] = {0, 1, 2, 3, 4, 5, 0, 2, 3, 4, 5, 2, 3, 4, 5, 2, 4, 5, 4, 5, 4};
  const double dx3 = dx * dx * dx;
  std::vector<c128> gbox(6 * nbox);
  for (int cell = 0; cell < nbox; ++cell) {
    const int mx = cell % nmL, my = (cell / nmL) % nmM, mz = cell / (nmL * nmM);
    c128 J[6][6];
    for (int i = 0; i < 21; ++i) J[po[i]][ps[i]] = dx3 * hj[cell * 21 + i];
    fill_sym(J);
    if (mx < nnL && my < nnM && mz < nnN) {
      int code[6][6], keep[6][6];
      nearby_table(mx, my, mz, code, keep);
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j)
          if (!keep[i][j]) J[i][j] = code[i][j] ? sing[code[i][j] - 1] : mk<double>(0, 0);
    }
    for (int kk = 0; kk < 6; ++kk) {
      c128 g = mk<double>(0, 0);
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) g = g + P.p[i][j][kk] * J[i][j];
      gbox[kk * nbox + cell] = (1.0 / (4.0 * M_PI)) * g;
    }
  }
  // scatter the box values into G
  for (int kk = 0; kk < 6; ++kk)
    for (int mz = 0; mz < nmN; ++mz)
      for (int my = 0; my < nmM; ++my) {
        const int cell0 = nmL * (my + nmM * mz);
        const int64_t v0 = (int64_t)L * (my + (int64_t)M * mz);
        VS_HIP(hipMemcpyAsync(G + kk * V + v0, gbox.data() + kk * nbox + cell0, sizeof(c128) * nmL,
                              hipMemcpyHostToDevice, ctx->stream));
      }
  VS_HIP(hipStreamSynchronize(ctx->stream));
}

}  // namespace vs

extern "C" int vs_galerkin_near_table(int* code, int* keep) {
  if (!code || !keep) return VS_ERR_ARG;
  for (int cell = 0; cell < 8; ++cell) {
    int c[6][6], k[6][6];
    vs::nearby_table(cell & 1, (cell >> 1) & 1, cell >> 2, c, k);
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) {
        code[36 * cell + 6 * i + j] = c[i][j];
        keep[36 * cell + 6 * i + j] = k[i][j];
      }
  }
  return VS_OK;
}
