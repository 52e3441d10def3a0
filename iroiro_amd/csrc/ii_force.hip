// Molecular-dynamics force of the Wilson operator on the device.
//
// Dirac_Wilson::mkfrc (dirac_wilson.cpp:62-92) adds, for every output site
// x of an operator and direction mu,
//   F_mu(x)_{ab} += sum_s [U_mu(x) (1 - gamma_mu) psi(x+mu)]_{a,s} conj(chi_{b,s}(x))
// for a pair (psi, chi); md_force_p uses (eta, zeta) and md_force_m
// (g5 zeta, g5 eta) (:41-50).  Because (1 - gamma_mu) = 2 P_mu with P_mu a
// rank-2 projector, sum_s [(1-gamma) psi]_s conj(chi_s) equals
// sum_{j=1,2} w_j conj(v_j), where w = the two projected colour vectors of
// psi that the hop forms (with the factor 2 of the time direction) and v
// the same projection of chi (without it).  So one link costs
//   M = sum_pairs sum_j w_j (x) conj(v_j)      (two 3x3 outer products a pair)
//   F = scale * U_mu(x) M,
// instead of reconstructing four spin components and summing 12 products
// per matrix element.  A block covers 64 half sites of one output parity
// and the four directions, one wavefront per direction: the four waves read
// the same own-site spinors chi (shared through the CU's L1) and each thread
// holds one link's worth of state.  Each output parity holds at most two
// pairs.
//
// Dirac_Wilson_EvenOdd::md_force (dirac_wilson_EvenOdd.hpp:94-116) with
// eta_o = Doe eta and zeta_o = Deo^dag zeta (odd fields) becomes
//   even x:  (eta_o, zeta) + (g5 zeta_o, g5 eta)       [Deo p, Deo m]
//   odd  x:  (g5 zeta, g5 eta_o) + (eta, zeta_o)       [Doe m, Doe p]
// with scale +kappa; the full-index operator's md_force is
//   (eta, zeta) + (g5 zeta, g5 eta) on both parities with scale -kappa.
//
// Memory traffic per output site with two pairs: 4 links (576 B) read,
// 4 force matrices (576 B) written, two neighbour-parity spinors and two
// own-parity spinors (4 x 192 B) read once each when the caches hold the
// neighbour planes: 1920 B; the kernel is HBM-bound.
#include "ii_ops.h"
#include "ii_stencil.h"

namespace ii {

// The pairs of each output parity: psi lives on the neighbour parity, chi
// on the output parity.  Bit (2*parity + pair) of g5psi / g5chi applies
// gamma5 to that field on load.
struct ForcePairs {
  const double2* psi00; const double2* psi01;  // output parity 0, pairs 0/1
  const double2* psi10; const double2* psi11;  // output parity 1
  const double2* chi00; const double2* chi01;
  const double2* chi10; const double2* chi11;
  unsigned g5psi, g5chi;
  int npairs;
};

// M += w (x) conj(v) over the two projected components of one pair.
template <int MU, bool G5P, bool G5C>
__device__ __forceinline__ void force_pair(double2 M[3][3], const double2* psi,
                                           const double2* chi, int stride, int nb, int hs) {
  double2 w1[3], w2[3], v1[3], v2[3];
  load_project<MU, true, G5P>(SpF64{psi, stride}, nb, w1, w2);
  load_project<MU, true, G5C>(SpF64{chi, stride}, hs, v1, v2);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      cmacc(M[a][b], v1[b], w1[a]);
      cmacc(M[a][b], v2[b], w2[a]);
    }
}

template <int MU>
__device__ __forceinline__ void force_pair_rt(double2 M[3][3], const double2* psi,
                                              const double2* chi, int stride, int nb, int hs,
                                              bool g5p, bool g5c) {
  if (g5p) {
    if (g5c) force_pair<MU, true, true>(M, psi, chi, stride, nb, hs);
    else force_pair<MU, true, false>(M, psi, chi, stride, nb, hs);
  } else {
    if (g5c) force_pair<MU, false, true>(M, psi, chi, stride, nb, hs);
    else force_pair<MU, false, false>(M, psi, chi, stride, nb, hs);
  }
}

// (m - m^dag)/2 - i Im tr(m)/3 (SUNmatUtils::anti_hermite_traceless,
// sunMatUtils.cpp:50-60; SUNmatrix::anti_hermite, sunMat.hpp:139-157).
__device__ __forceinline__ void traceless_antihermite(double2 F[3][3]) {
  const double imtr = (F[0][0].y + F[1][1].y + F[2][2].y) / 3.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int b = a + 1; b < 3; ++b) {
      const double re = 0.5 * (F[a][b].x - F[b][a].x);
      const double im = 0.5 * (F[a][b].y + F[b][a].y);
      F[a][b] = make_double2(re, im);
      F[b][a] = make_double2(-re, im);
    }
    F[a][a] = make_double2(0.0, F[a][a].y - imtr);
  }
}

template <int MU>
__device__ __forceinline__ void force_dir(const Geom& g, const ForcePairs& fp,
                                          const double2* __restrict__ U, double2* __restrict__ F,
                                          double scale, int par, int hs, const SiteCoord& c,
                                          bool accumulate, bool ta) {
  const int nb = neighbour(g, c, hs, MU, true);
  double2 M[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) M[a][b] = make_double2(0.0, 0.0);
  const unsigned b0 = 2u * (unsigned)par;
  force_pair_rt<MU>(M, par ? fp.psi10 : fp.psi00, par ? fp.chi10 : fp.chi00, g.stride, nb, hs,
                    (fp.g5psi >> b0) & 1u, (fp.g5chi >> b0) & 1u);
  if (fp.npairs > 1)
    force_pair_rt<MU>(M, par ? fp.psi11 : fp.psi01, par ? fp.chi11 : fp.chi01, g.stride, nb, hs,
                      (fp.g5psi >> (b0 + 1)) & 1u, (fp.g5chi >> (b0 + 1)) & 1u);
  // both projections carry the factor 2 of the time direction; the force
  // carries it once
  const double s = MU == 3 ? 0.5 * scale : scale;
  // one descriptor per (parity, direction) block of U and F: its base in 64 bits
  const size_t lblk = (size_t)((par * 4 + MU) * 9) * (size_t)g.stride;
  const double2* Ub = U + lblk;
  double2* Fb = F + lblk;
  double2 Ul[3][3];
#pragma unroll
  for (int k = 0; k < 9; ++k)
    Ul[k / 3][k % 3] = ld_at<double2, kNT>(Ub, (uint32_t)(k * g.stride), (uint32_t)hs);
  double2 Fl[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double2 acc = make_double2(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 3; ++k) cmac(acc, Ul[a][k], M[k][b]);
      Fl[a][b] = cscale(acc, s);
    }
  if (accumulate) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const double2 o = ld_at<double2>(Fb, (uint32_t)(k * g.stride), (uint32_t)hs);
      Fl[k / 3][k % 3].x += o.x;
      Fl[k / 3][k % 3].y += o.y;
    }
  }
  if (ta) traceless_antihermite(Fl);
#pragma unroll
  for (int k = 0; k < 9; ++k)
    st_at<double2, kNT>(Fb, (uint32_t)(k * g.stride), (uint32_t)hs, Fl[k / 3][k % 3]);
}

// blockDim.x = 256: lanes 0..63 of wave w handle direction mu = w for the
// block's 64 half sites.
__global__ void __launch_bounds__(256)
k_md_force(Geom g, ForcePairs fp, const double2* __restrict__ U, double2* __restrict__ F,
           double scale, int par0, int accumulate, int ta) {
  const int mu = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int hs = blockIdx.x * 64 + ((int)threadIdx.x & 63);
  const int par = par0 + (int)blockIdx.y;
  if (hs >= g.Vh) return;
  const SiteCoord c = decode_site(g, hs, par);
  const bool acc = accumulate != 0, pr = ta != 0;
  switch (mu) {
    case 0: force_dir<0>(g, fp, U, F, scale, par, hs, c, acc, pr); break;
    case 1: force_dir<1>(g, fp, U, F, scale, par, hs, c, acc, pr); break;
    case 2: force_dir<2>(g, fp, U, F, scale, par, hs, c, acc, pr); break;
    default: force_dir<3>(g, fp, U, F, scale, par, hs, c, acc, pr); break;
  }
}

// Per-link projection on an array of 3x3 matrices in any layout that keeps
// the 18 doubles of a matrix together (the host Format_G layout); in and out
// may be the same array.
__global__ void __launch_bounds__(kBlock)
k_traceless_antihermite(const double* in, double* out, long nlinks) {
  const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlinks) return;
  double2 F[3][3];
#pragma unroll
  for (int k = 0; k < 9; ++k) F[k / 3][k % 3] = make_double2(in[18 * l + 2 * k], in[18 * l + 2 * k + 1]);
  traceless_antihermite(F);
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    out[18 * l + 2 * k] = F[k / 3][k % 3].x;
    out[18 * l + 2 * k + 1] = F[k / 3][k % 3].y;
  }
}

static ii_status launch_force(ii_context* ctx, const ForcePairs& fp, double2* F, double scale,
                              int par0, int npar, bool accumulate, bool ta) {
  const Geom& g = ctx->g;
  hipLaunchKernelGGL(k_md_force, dim3(blocks_for(g.Vh, 64), npar), dim3(256), 0, ctx->stream, g,
                     fp, (const double2*)ctx->gauge64, F, scale, par0, accumulate ? 1 : 0,
                     ta ? 1 : 0);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

// Workspace slots of this file: 16-19 action fields, 22 the device force,
// 23 a second host staging area, 24/25 odd temporaries, 26/27 the input
// fields (half or full), 28 the host-layout force.
enum { kWsPhi = 16, kWsEta = 17, kWsZeta = 18, kWsTmp = 19, kWsF = 22, kWsStage = 23,
       kWsEtaO = 24, kWsZetaO = 25, kWsIn0 = 26, kWsIn1 = 27, kWsOut = 28 };

static ii_status ws_half(ii_context* ctx, int slot, int nhalf, double2** p) {
  void* v = nullptr;
  II_CHECK(workspace(ctx, slot, half_elems(ctx->g) * nhalf * sizeof(double2), &v));
  *p = (double2*)v;
  return II_OK;
}
static ii_status ws_force(ii_context* ctx, double2** p) {
  void* v = nullptr;
  II_CHECK(workspace(ctx, kWsF, (size_t)72 * ctx->g.stride * sizeof(double2), &v));
  *p = (double2*)v;
  return II_OK;
}

ii_status md_force_eo64(ii_context* ctx, double kappa, const double2* eta, const double2* zeta,
                        double2* F, bool ta, double weight, bool accumulate) {
  double2 *eta_o, *zeta_o;
  II_CHECK(ws_half(ctx, kWsEtaO, 1, &eta_o));
  II_CHECK(ws_half(ctx, kWsZetaO, 1, &zeta_o));
  HopSpec h;
  h.out_par = 1;
  h.a = -kappa;
  II_CHECK(hop64(ctx, h, eta, eta_o, nullptr, nullptr));   // Doe eta
  h.g5in = h.g5out = true;
  II_CHECK(hop64(ctx, h, zeta, zeta_o, nullptr, nullptr));  // Deo^dag zeta = g5 Doe g5 zeta
  ForcePairs fp{};
  fp.psi00 = eta_o;  fp.chi00 = zeta;    // Deo.md_force_p(eta_o, zeta)
  fp.psi01 = zeta_o; fp.chi01 = eta;     // Deo.md_force_m(eta, zeta_o): (g5 zeta_o, g5 eta)
  fp.psi10 = zeta;   fp.chi10 = eta_o;   // Doe.md_force_m(eta_o, zeta): (g5 zeta, g5 eta_o)
  fp.psi11 = eta;    fp.chi11 = zeta_o;  // Doe.md_force_p(eta, zeta_o)
  fp.g5psi = (1u << 1) | (1u << 2);
  fp.g5chi = (1u << 1) | (1u << 2);
  fp.npairs = 2;
  return launch_force(ctx, fp, F, kappa * weight, 0, 2, accumulate, ta);
}

// Full-index pairs on fields split into even/odd halves.
static ForcePairs full_pairs(const double2* ee, const double2* eo, const double2* ze,
                             const double2* zo, int which) {
  ForcePairs fp{};
  // which: 0 = md_force_p, 1 = md_force_m, 2 = both
  if (which == 1) {
    fp.psi00 = zo; fp.chi00 = ee;
    fp.psi10 = ze; fp.chi10 = eo;
    fp.g5psi = 1u | (1u << 2);
    fp.g5chi = 1u | (1u << 2);
    fp.npairs = 1;
    return fp;
  }
  fp.psi00 = eo; fp.chi00 = ze;
  fp.psi10 = ee; fp.chi10 = zo;
  fp.psi01 = zo; fp.chi01 = ee;
  fp.psi11 = ze; fp.chi11 = eo;
  fp.g5psi = (1u << 1) | (1u << 3);
  fp.g5chi = (1u << 1) | (1u << 3);
  fp.npairs = which == 2 ? 2 : 1;
  return fp;
}

static ii_status upload_to(ii_context* ctx, int slot, const double* host, size_t n, double** dev) {
  void* p = nullptr;
  II_CHECK(workspace(ctx, slot, n * sizeof(double), &p));
  II_HIP(ctx, hipMemcpyAsync(p, host, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  *dev = (double*)p;
  return II_OK;
}

static ii_status download_force(ii_context* ctx, const double2* F, double* host) {
  const size_t n = (size_t)72 * ctx->g.V;
  void* p = nullptr;
  II_CHECK(workspace(ctx, kWsOut, n * sizeof(double), &p));
  II_CHECK(unpack_gauge_field(ctx, F, (double*)p));
  II_HIP(ctx, hipMemcpyAsync(host, p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return II_OK;
}

// Two host spinor arrays (half or full) into device fields of slots
// kWsIn0 / kWsIn1.
static ii_status upload_pair(ii_context* ctx, const double* eta, const double* zeta, bool full,
                             double2** de, double2** dz) {
  const Geom& g = ctx->g;
  const size_t n = (size_t)24 * (full ? g.V : g.Vh);
  double *s0, *s1;
  II_CHECK(upload_to(ctx, 0, eta, n, &s0));
  II_CHECK(upload_to(ctx, kWsStage, zeta, n, &s1));
  II_CHECK(ws_half(ctx, kWsIn0, full ? 2 : 1, de));
  II_CHECK(ws_half(ctx, kWsIn1, full ? 2 : 1, dz));
  const size_t he = half_elems(g);
  if (full) {
    II_CHECK(pack_full_spinor(ctx, s0, *de, *de + he));
    II_CHECK(pack_full_spinor(ctx, s1, *dz, *dz + he));
  } else {
    II_CHECK(pack_half_spinor(ctx, s0, *de));
    II_CHECK(pack_half_spinor(ctx, s1, *dz));
  }
  return II_OK;
}

// Deterministic fill of the Vh sites of a half field with values in
// [-1, 1) (benchmark inputs; padding sites stay untouched).
__global__ void __launch_bounds__(kBlock)
k_fill_bench(double2* __restrict__ f, Geom g, uint32_t seed) {
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= g.Vh) return;
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    uint32_t h = (uint32_t)hs * 2654435761u ^ (seed + 97u * (uint32_t)k);
    h ^= h >> 15; h *= 2246822519u; h ^= h >> 13; h *= 3266489917u; h ^= h >> 16;
    const double re = (double)(h & 0xFFFF) / 32768.0 - 1.0;
    const double im = (double)(h >> 16) / 32768.0 - 1.0;
    f[(size_t)k * g.stride + hs] = make_double2(re, im);
  }
}

}  // namespace ii

using namespace ii;

extern "C" {

ii_status ii_dirac_md_force(ii_context* ctx, double kappa, double* fce, const double* eta,
                            const double* zeta) {
  if (!fce || !eta || !zeta) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  double2 *de, *dz, *F;
  II_CHECK(upload_pair(ctx, eta, zeta, false, &de, &dz));
  II_CHECK(ws_force(ctx, &F));
  II_CHECK(md_force_eo64(ctx, kappa, de, dz, F, false));
  return download_force(ctx, F, fce);
}

ii_status ii_dirac_full_md_force(ii_context* ctx, double kappa, double* fce, const double* eta,
                                 const double* zeta) {
  if (!fce || !eta || !zeta) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  double2 *de, *dz, *F;
  II_CHECK(upload_pair(ctx, eta, zeta, true, &de, &dz));
  II_CHECK(ws_force(ctx, &F));
  const size_t he = half_elems(ctx->g);
  ForcePairs fp = full_pairs(de, de + he, dz, dz + he, 2);
  II_CHECK(launch_force(ctx, fp, F, -kappa, 0, 2, false, false));
  return download_force(ctx, F, fce);
}

ii_status ii_dirac_full_md_force_pm(ii_context* ctx, double* fce, const double* eta,
                                    const double* zeta, int minus) {
  if (!fce || !eta || !zeta || (minus != 0 && minus != 1)) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  double2 *de, *dz, *F;
  II_CHECK(upload_pair(ctx, eta, zeta, true, &de, &dz));
  II_CHECK(ws_force(ctx, &F));
  double* s = nullptr;
  II_CHECK(upload_to(ctx, kWsOut, fce, (size_t)72 * ctx->g.V, &s));
  II_CHECK(pack_gauge_field(ctx, s, F));
  const size_t he = half_elems(ctx->g);
  ForcePairs fp = full_pairs(de, de + he, dz, dz + he, minus);
  II_CHECK(launch_force(ctx, fp, F, 1.0, 0, 2, true, false));
  return download_force(ctx, F, fce);
}

ii_status ii_traceless_antihermite(ii_context* ctx, double* out, const double* in) {
  if (!out || !in) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  const long nl = 4L * ctx->g.V;
  double* d = nullptr;
  II_CHECK(upload_to(ctx, kWsStage, in, (size_t)18 * nl, &d));
  hipLaunchKernelGGL(k_traceless_antihermite, dim3(blocks_for(nl)), dim3(kBlock), 0, ctx->stream,
                     d, d, nl);
  II_HIP(ctx, hipGetLastError());
  II_HIP(ctx, hipMemcpyAsync(out, d, (size_t)18 * nl * sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return II_OK;
}

ii_status ii_bench_md_force(ii_context* ctx, int nrep, double* force_ms, double* total_ms) {
  if (!force_ms || !total_ms || nrep < 1) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  double2 *eta, *zeta, *eta_o, *zeta_o, *F;
  II_CHECK(ws_half(ctx, kWsIn0, 1, &eta));
  II_CHECK(ws_half(ctx, kWsIn1, 1, &zeta));
  II_CHECK(ws_force(ctx, &F));
  hipLaunchKernelGGL(k_fill_bench, dim3(blocks_for(g.Vh)), dim3(kBlock), 0, ctx->stream, eta, g,
                     11u);
  hipLaunchKernelGGL(k_fill_bench, dim3(blocks_for(g.Vh)), dim3(kBlock), 0, ctx->stream, zeta, g,
                     23u);
  II_HIP(ctx, hipGetLastError());
  const double kappa = 0.1355;
  for (int i = 0; i < 2; ++i) II_CHECK(md_force_eo64(ctx, kappa, eta, zeta, F, false));
  II_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int i = 0; i < nrep; ++i) II_CHECK(md_force_eo64(ctx, kappa, eta, zeta, F, false));
  II_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  II_HIP(ctx, hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  II_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  *total_ms = ms / nrep;
  // the force kernel alone, on the odd fields the last call left behind
  II_CHECK(ws_half(ctx, kWsEtaO, 1, &eta_o));
  II_CHECK(ws_half(ctx, kWsZetaO, 1, &zeta_o));
  ForcePairs fp{};
  fp.psi00 = eta_o;  fp.chi00 = zeta;
  fp.psi01 = zeta_o; fp.chi01 = eta;
  fp.psi10 = zeta;   fp.chi10 = eta_o;
  fp.psi11 = eta;    fp.chi11 = zeta_o;
  fp.g5psi = fp.g5chi = (1u << 1) | (1u << 2);
  fp.npairs = 2;
  II_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int i = 0; i < nrep; ++i) II_CHECK(launch_force(ctx, fp, F, kappa, 0, 2, false, false));
  II_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  II_HIP(ctx, hipEventSynchronize(ctx->ev1));
  II_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  *force_ms = ms / nrep;
  return II_OK;
}

ii_status ii_action_nf2_md_force(ii_context* ctx, double kappa, const double* phi, double prec,
                                 int maxiter, double* force, ii_solver_output* out) {
  if (!phi || !force || maxiter < 1) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  double* s = nullptr;
  double2 *b, *eta, *zeta, *t, *F;
  II_CHECK(upload_to(ctx, 0, phi, (size_t)24 * ctx->g.Vh, &s));
  II_CHECK(ws_half(ctx, kWsPhi, 1, &b));
  II_CHECK(ws_half(ctx, kWsEta, 1, &eta));
  II_CHECK(ws_half(ctx, kWsZeta, 1, &zeta));
  II_CHECK(ws_half(ctx, kWsTmp, 1, &t));
  II_CHECK(ws_force(ctx, &F));
  II_CHECK(pack_half_spinor(ctx, s, b));
  // eta = (Dhat^dag Dhat)^-1 phi (Action_Nf2::DdagD_inv, action_Nf2.cpp:32-39)
  ii_status st = cg_reference(ctx, kappa, eta, b, prec, maxiter, out);
  if (st != II_OK && st != II_ERR_NOT_CONVERGED) return st;
  const std::string msg = ctx->err;
  // zeta = Dhat eta; fce = md_force(eta, zeta); TracelessAntihermite(fce)
  II_CHECK(dhat64(ctx, kappa, eta, zeta, t));
  II_CHECK(md_force_eo64(ctx, kappa, eta, zeta, F, true));
  II_CHECK(download_force(ctx, F, force));
  if (st != II_OK) ctx->err = msg;
  return st;
}

}  // extern "C"
