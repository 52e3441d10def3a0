// Quenched SU(3) gauge-field generation on the device: Cabibbo-Marinari
// heat-bath over the SU(2) subgroups (0,1), (0,2), (1,2) with the
// Kennedy-Pendleton generator, Wilson plaquette action
//   S = beta sum_P (1 - Re tr U_P / 3).
// The reference's heat-bath is SU(2) only (HBActions_SUN.cpp:85-92 aborts
// for NC != 2), so this has no reference counterpart; it produces the
// beta = 5.8 / 6.0 ensembles the solve path is measured on.
//
// Links U_mu(x) of one direction and one parity do not appear in each
// other's staples, so a (mu, parity) pass updates all of them at once.
// Random numbers are counter-based (SplitMix64 finaliser keyed by seed,
// sweep, mu, lexicographic site, subgroup and attempt), identical to the
// CPU oracle's, so a sweep is independent of the thread schedule.
#include "ii_ops.h"
#include "ii_stencil.h"

namespace ii {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double uniform(uint64_t key, uint64_t counter) {
  uint64_t z = mix64(key + counter * 0xD1B54A32D192ED03ULL);
  return (double)((z >> 11) + 1) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ void ldU(const double2* __restrict__ u, const Geom& g, int par, int mu,
                                    int hs, double2 U[3][3]) {
  const double2* b = u + (size_t)((par * 4 + mu) * 9) * g.stride + hs;
#pragma unroll
  for (int k = 0; k < 9; ++k) U[k / 3][k % 3] = b[(size_t)k * g.stride];
}

// C = A B, A B^dag, A^dag B
__device__ __forceinline__ void mul_nn(double2 C[3][3], const double2 A[3][3], const double2 B[3][3]) {
  double2 T[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 s = make_double2(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) cmac(s, A[i][k], B[k][j]);
      T[i][j] = s;
    }
#pragma unroll
  for (int i = 0; i < 9; ++i) C[i / 3][i % 3] = T[i / 3][i % 3];
}
__device__ __forceinline__ void mul_nd(double2 C[3][3], const double2 A[3][3], const double2 B[3][3]) {
  double2 T[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 s = make_double2(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) cmacc(s, B[j][k], A[i][k]);
      T[i][j] = s;
    }
#pragma unroll
  for (int i = 0; i < 9; ++i) C[i / 3][i % 3] = T[i / 3][i % 3];
}
__device__ __forceinline__ void mul_dn(double2 C[3][3], const double2 A[3][3], const double2 B[3][3]) {
  double2 T[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 s = make_double2(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) cmacc(s, A[k][i], B[k][j]);
      T[i][j] = s;
    }
#pragma unroll
  for (int i = 0; i < 9; ++i) C[i / 3][i % 3] = T[i / 3][i % 3];
}

// Kennedy-Pendleton SU(2) deviate with density ~ exp(alpha X0); leaves the
// identity if no candidate is accepted within the attempt budget.
__device__ void kp_su2(double alpha, uint64_t key, uint64_t base, double X[4]) {
  const double twopi = 6.283185307179586476925;
  X[0] = 1.0; X[1] = X[2] = X[3] = 0.0;
  for (int att = 0; att < 64; ++att) {
    uint64_t c = base + (uint64_t)att * 8;
    double r1 = uniform(key, c), r2 = uniform(key, c + 1);
    double r3 = uniform(key, c + 2), r4 = uniform(key, c + 3);
    double cs = cos(twopi * r2);
    double lam2 = -(log(r1) + cs * cs * log(r3)) / (2.0 * alpha);
    if (lam2 > 1.0 || r4 * r4 > 1.0 - lam2) continue;
    double x0 = 1.0 - 2.0 * lam2;
    double rad = sqrt(fmax(0.0, 1.0 - x0 * x0));
    double ct = 2.0 * uniform(key, c + 4) - 1.0;
    double st = sqrt(fmax(0.0, 1.0 - ct * ct));
    double ph = twopi * uniform(key, c + 5);
    X[0] = x0;
    X[1] = rad * st * cos(ph);
    X[2] = rad * st * sin(ph);
    X[3] = rad * ct;
    return;
  }
}

__device__ void su3_project(double2 M[3][3]) {
  double n0 = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) n0 += M[0][c].x * M[0][c].x + M[0][c].y * M[0][c].y;
  n0 = 1.0 / sqrt(n0);
#pragma unroll
  for (int c = 0; c < 3; ++c) { M[0][c].x *= n0; M[0][c].y *= n0; }
  double pr = 0.0, pi = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    pr += M[0][c].x * M[1][c].x + M[0][c].y * M[1][c].y;
    pi += M[0][c].x * M[1][c].y - M[0][c].y * M[1][c].x;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    M[1][c].x -= pr * M[0][c].x - pi * M[0][c].y;
    M[1][c].y -= pr * M[0][c].y + pi * M[0][c].x;
  }
  double n1 = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) n1 += M[1][c].x * M[1][c].x + M[1][c].y * M[1][c].y;
  n1 = 1.0 / sqrt(n1);
#pragma unroll
  for (int c = 0; c < 3; ++c) { M[1][c].x *= n1; M[1][c].y *= n1; }
  r12_third_row(M, false);
}

__global__ void __launch_bounds__(kBlock)
k_heatbath(double2* __restrict__ u, Geom g, int mu, int par, double beta, uint64_t seed,
           int sweep) {
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= g.Vh) return;
  const SiteCoord c = decode_site(g, hs, par);
  const int op = 1 - par;
  const int sp = neighbour(g, c, hs, mu, true);  // x+mu, parity op
  double2 A[3][3];
#pragma unroll
  for (int i = 0; i < 9; ++i) A[i / 3][i % 3] = make_double2(0.0, 0.0);
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == mu) continue;
    double2 X[3][3], Y[3][3];
    // upper: U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
    const int snp = neighbour(g, c, hs, nu, true);
    ldU(u, g, op, nu, sp, X);
    ldU(u, g, op, mu, snp, Y);
    mul_nd(X, X, Y);
    ldU(u, g, par, nu, hs, Y);
    mul_nd(X, X, Y);
#pragma unroll
    for (int i = 0; i < 9; ++i) A[i / 3][i % 3] = cadd(A[i / 3][i % 3], X[i / 3][i % 3]);
    // lower: U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu); x+mu-nu has parity par
    const int snm = neighbour(g, c, hs, nu, false);
    const SiteCoord cp = decode_site(g, sp, op);
    const int spm = neighbour(g, cp, sp, nu, false);
    ldU(u, g, par, nu, spm, X);
    ldU(u, g, op, mu, snm, Y);
    double2 Z[3][3];
    // X^dag Y^dag = (Y X)^dag
    mul_nn(Z, Y, X);
    double2 W[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) W[i][j] = make_double2(Z[j][i].x, -Z[j][i].y);
    ldU(u, g, op, nu, snm, Y);
    mul_nn(W, W, Y);
#pragma unroll
    for (int i = 0; i < 9; ++i) A[i / 3][i % 3] = cadd(A[i / 3][i % 3], W[i / 3][i % 3]);
  }
  double2 U[3][3], W[3][3];
  ldU(u, g, par, mu, hs, U);
  mul_nn(W, U, A);
  const int lex = 2 * hs + (c.x & 1);
  const uint64_t stream = ((uint64_t)sweep * 4 + (uint64_t)mu) * (uint64_t)g.V + (uint64_t)lex;
  const uint64_t key = mix64(seed ^ mix64(stream));
  const int sub[3][2] = {{0, 1}, {0, 2}, {1, 2}};
#pragma unroll 1
  for (int sg = 0; sg < 3; ++sg) {
    const int i = sub[sg][0], j = sub[sg][1];
    double a0 = 0.5 * (W[i][i].x + W[j][j].x);
    double a1 = 0.5 * (W[i][j].y + W[j][i].y);
    double a2 = 0.5 * (W[i][j].x - W[j][i].x);
    double a3 = 0.5 * (W[i][i].y - W[j][j].y);
    double k = sqrt(a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3);
    if (k < 1e-300) continue;
    double v0 = a0 / k, v1 = a1 / k, v2 = a2 / k, v3 = a3 / k;
    double X[4];
    kp_su2(2.0 * beta * k / 3.0, key, (uint64_t)sg * 1024, X);
    // R = X v^dag as 2x2 complex matrices, q -> [[q0+iq3, q2+iq1], [-q2+iq1, q0-iq3]]
    double2 Xm[2][2] = {{make_double2(X[0], X[3]), make_double2(X[2], X[1])},
                        {make_double2(-X[2], X[1]), make_double2(X[0], -X[3])}};
    double2 Vm[2][2] = {{make_double2(v0, v3), make_double2(v2, v1)},
                        {make_double2(-v2, v1), make_double2(v0, -v3)}};
    double2 R[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        double2 s = make_double2(0.0, 0.0);
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) cmacc(s, Vm[b][cc], Xm[a][cc]);
        R[a][b] = s;
      }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        double2 xv = q ? W[i][cc] : U[i][cc];
        double2 yv = q ? W[j][cc] : U[j][cc];
        double2 ni = make_double2(0.0, 0.0), nj = make_double2(0.0, 0.0);
        cmac(ni, R[0][0], xv); cmac(ni, R[0][1], yv);
        cmac(nj, R[1][0], xv); cmac(nj, R[1][1], yv);
        if (q) { W[i][cc] = ni; W[j][cc] = nj; }
        else { U[i][cc] = ni; U[j][cc] = nj; }
      }
    }
  }
  su3_project(U);
  double2* b = u + (size_t)((par * 4 + mu) * 9) * g.stride + hs;
#pragma unroll
  for (int k = 0; k < 9; ++k) b[(size_t)k * g.stride] = U[k / 3][k % 3];
}

__global__ void __launch_bounds__(kBlock) k_unit_gauge(double2* __restrict__ u, Geom g) {
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  const int pm = blockIdx.y;
  if (hs >= g.Vh) return;
  double2* b = u + (size_t)(pm * 9) * g.stride + hs;
#pragma unroll
  for (int k = 0; k < 9; ++k)
    b[(size_t)k * g.stride] = make_double2((k == 0 || k == 4 || k == 8) ? 1.0 : 0.0, 0.0);
}

// Antiperiodic time boundary on the device copy: U_t -> -U_t on t = Lt-1
// (antiPeriodicBC.hpp:25-37).
__global__ void __launch_bounds__(kBlock) k_apbc(double2* __restrict__ u, Geom g) {
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  const int par = blockIdx.y;
  if (hs >= g.Vh) return;
  const SiteCoord c = decode_site(g, hs, par);
  if (c.t != g.L[3] - 1) return;
  double2* b = u + (size_t)((par * 4 + 3) * 9) * g.stride + hs;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    double2 v = b[(size_t)k * g.stride];
    b[(size_t)k * g.stride] = make_double2(-v.x, -v.y);
  }
}

static void drop_lowprec(ii_context* ctx) {
  for (int i = 0; i < ii_context::kGaugeCopies; ++i)
    if (ctx->gauge_lp[i]) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipFree(ctx->gauge_lp[i]);
      ctx->gauge_lp[i] = nullptr;
    }
  ctx->r12_ok = false;
  ctx->tbc_sign = 1;
}

}  // namespace ii

using namespace ii;

extern "C" {

ii_status ii_gauge_unit(ii_context* ctx) {
  if (!ctx) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  hipLaunchKernelGGL(k_unit_gauge, dim3(blocks_for(g.Vh), 8), dim3(kBlock), 0, ctx->stream,
                     ctx->gauge64, g);
  II_HIP(ctx, hipGetLastError());
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  drop_lowprec(ctx);
  ctx->gauge_valid = true;
  return II_OK;
}

ii_status ii_gauge_heatbath(ii_context* ctx, double beta, unsigned long long seed,
                            int first_sweep, int nsweeps) {
  if (!ctx || beta <= 0.0 || nsweeps < 0 || first_sweep < 0) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  for (int sw = first_sweep; sw < first_sweep + nsweeps; ++sw)
    for (int mu = 0; mu < 4; ++mu)
      for (int par = 0; par < 2; ++par) {
        hipLaunchKernelGGL(k_heatbath, dim3(blocks_for(g.Vh)), dim3(kBlock), 0, ctx->stream,
                           ctx->gauge64, g, mu, par, beta, (uint64_t)seed, sw);
        II_HIP(ctx, hipGetLastError());
      }
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  drop_lowprec(ctx);
  return II_OK;
}

ii_status ii_gauge_apply_apbc(ii_context* ctx) {
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  hipLaunchKernelGGL(k_apbc, dim3(blocks_for(g.Vh), 2), dim3(kBlock), 0, ctx->stream,
                     ctx->gauge64, g);
  II_HIP(ctx, hipGetLastError());
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  drop_lowprec(ctx);
  return II_OK;
}

}  // extern "C"
