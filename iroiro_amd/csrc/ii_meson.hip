// Meson two-point functions of 12-column propagators on the device:
// MesonCorrelator::calculate / calculate_mom / calculate_PSmom
// (lib/Measurements/FermionicM/meson_correlator.hpp:138-296) for all 18
// MesonTypes in one pass over the propagators.
//
// Per site x and type T with gamma tables G1, G2 (lib/Tools/gammaMatrices.cpp):
//   z_T(x) = sum_{s4,s2} gf(s4,s2) sum_{c1,c2}
//            conj(q1[c1 + 3 s1][s2, c2](x)) q2[c1 + 3 s4][s3, c2](x),
//   s1 = G2(s4).spn, s3 = G1(s2).spn, gf = Re(G2(s4).fac G1(s2).fac);
// calculate sums Re z over slices along a direction; calculate_mom sums
// z e^{i p.x} over time slices; calculate_PSmom sums Re z cos(px x)
// cos(py y) cos(pz z).
//
// k_meson_sites: one thread per site.  For each (c1, c2) it loads the 4x4
// spin blocks of both propagators (16 complex each; one block serves both
// when q1 == q2, the SAME instantiation) and
// updates the 18 accumulators; the type tables are constexpr, so every
// spin index and factor folds at compile time.  It writes one value (two
// for the momentum projection) per type and site.  k_meson_slices then sums
// each slice in a fixed order, so results are deterministic.
#include "ii_ops.h"
#include "ii_stencil.h"

#include <cmath>

namespace ii {

struct GammaTab {
  int spn[4];
  double fr[4], fi[4];
};

// gammaMatrices.cpp:13-170 (index order: Unit, Gamma1..5, Gamma1_5..4_5,
// Sigma12_5, 13_5, 14_5, 23_5, 24_5, 34_5)
constexpr GammaTab kGamma[16] = {
    {{0, 1, 2, 3}, {1, 1, 1, 1}, {0, 0, 0, 0}},        // Unit
    {{3, 2, 1, 0}, {0, 0, 0, 0}, {-1, -1, 1, 1}},      // Gamma1
    {{3, 2, 1, 0}, {-1, 1, 1, -1}, {0, 0, 0, 0}},      // Gamma2
    {{2, 3, 0, 1}, {0, 0, 0, 0}, {-1, 1, 1, -1}},      // Gamma3
    {{0, 1, 2, 3}, {1, 1, -1, -1}, {0, 0, 0, 0}},      // Gamma4
    {{2, 3, 0, 1}, {1, 1, 1, 1}, {0, 0, 0, 0}},        // Gamma5
    {{1, 0, 3, 2}, {0, 0, 0, 0}, {-1, -1, 1, 1}},      // Gamma1_5
    {{1, 0, 3, 2}, {-1, 1, 1, -1}, {0, 0, 0, 0}},      // Gamma2_5
    {{0, 1, 2, 3}, {0, 0, 0, 0}, {-1, 1, 1, -1}},      // Gamma3_5
    {{2, 3, 0, 1}, {1, 1, -1, -1}, {0, 0, 0, 0}},      // Gamma4_5
    {{2, 3, 0, 1}, {0, 0, 0, 0}, {1, -1, 1, -1}},      // Sigma12_5
    {{3, 2, 1, 0}, {1, -1, 1, -1}, {0, 0, 0, 0}},      // Sigma13_5
    {{1, 0, 2, 3}, {1, 1, 1, 1}, {0, 0, 0, 0}},        // Sigma14_5 (as in the reference)
    {{3, 2, 1, 0}, {0, 0, 0, 0}, {1, 1, 1, 1}},        // Sigma23_5
    {{1, 0, 3, 2}, {1, -1, 1, -1}, {0, 0, 0, 0}},      // Sigma24_5
    {{0, 1, 2, 3}, {0, 0, 0, 0}, {1, -1, 1, -1}},      // Sigma34_5
};

// MesonType (meson_correlator.hpp:20-24) -> (G1, G2) (:37-120)
constexpr int kMesonTypes = 18;
constexpr int kTypeG[kMesonTypes][2] = {
    {0, 0},   {5, 5},   {6, 6},   {7, 7},   {8, 8},   {9, 9},   {1, 1},   {2, 2},   {3, 3},
    {4, 4},   {10, 10}, {11, 11}, {12, 12}, {13, 13}, {14, 14}, {15, 15}, {4, 0},   {0, 4}};

constexpr double gamma_factor(int T, int s4, int s2) {
  return kGamma[kTypeG[T][1]].fr[s4] * kGamma[kTypeG[T][0]].fr[s2] -
         kGamma[kTypeG[T][1]].fi[s4] * kGamma[kTypeG[T][0]].fi[s2];
}

template <int T>
__device__ __forceinline__ void meson_type(double* re, double* im, const double2 (&a)[4][4],
                                           const double2 (&b)[4][4], bool want_im) {
#pragma unroll
  for (int s4 = 0; s4 < 4; ++s4) {
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2) {
      constexpr int g1 = kTypeG[T][0], g2 = kTypeG[T][1];
      const double gf = gamma_factor(T, s4, s2);
      if (gf == 0.0) continue;
      const int s1 = kGamma[g2].spn[s4], s3 = kGamma[g1].spn[s2];
      const double2 x = a[s1][s2], y = b[s4][s3];
      re[T] = fma(gf, fma(x.x, y.x, x.y * y.y), re[T]);
      if (want_im) im[T] = fma(gf, fma(x.x, y.y, -(x.y * y.x)), im[T]);
    }
  }
}

template <int T>
__device__ __forceinline__ void meson_all(double* re, double* im, const double2 (&a)[4][4],
                                          const double2 (&b)[4][4], bool want_im) {
  meson_type<T>(re, im, a, b, want_im);
  if constexpr (T + 1 < kMesonTypes) meson_all<T + 1>(re, im, a, b, want_im);
}

// mode 0: Re z; 1: z e^{i p.x} (re, im); 2: Re z cos cos cos.
// vals: [k][2][stride], k = type (mode 0, 2) or 2 type + re/im (mode 1).
template <int MODE, bool SAME>
__global__ void __launch_bounds__(kBlock)
k_meson_sites(Geom g, const double2* __restrict__ q1, const double2* __restrict__ q2,
              double px, double py, double pz, double* __restrict__ vals) {
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  const int par = blockIdx.y;
  if (hs >= g.Vh) return;
  const size_t he = (size_t)12 * g.stride;
  const size_t colstride = 2 * he;  // one full field per column: [par][12][stride]
  const double2* b1 = q1 + (size_t)par * he + hs;
  const double2* b2 = q2 + (size_t)par * he + hs;
  double re[kMesonTypes], im[kMesonTypes];
#pragma unroll
  for (int T = 0; T < kMesonTypes; ++T) re[T] = im[T] = 0.0;
  for (int c1 = 0; c1 < 3; ++c1) {
    for (int c2 = 0; c2 < 3; ++c2) {
      double2 a[4][4];
#pragma unroll
      for (int s1 = 0; s1 < 4; ++s1)
#pragma unroll
        for (int s = 0; s < 4; ++s)
          a[s1][s] = b1[(size_t)(c1 + 3 * s1) * colstride + (size_t)(3 * s + c2) * g.stride];
      if constexpr (SAME) {
        // q1 == q2: one block serves both operands (half the registers)
        meson_all<0>(re, im, a, a, MODE == 1);
      } else {
        double2 b[4][4];
#pragma unroll
        for (int s1 = 0; s1 < 4; ++s1)
#pragma unroll
          for (int s = 0; s < 4; ++s)
            b[s1][s] = b2[(size_t)(c1 + 3 * s1) * colstride + (size_t)(3 * s + c2) * g.stride];
        meson_all<0>(re, im, a, b, MODE == 1);
      }
    }
  }
  const size_t vstride = 2 * (size_t)g.stride;
  const size_t o = (size_t)par * g.stride + hs;
  if (MODE == 0) {
#pragma unroll
    for (int T = 0; T < kMesonTypes; ++T) vals[T * vstride + o] = re[T];
    return;
  }
  const SiteCoord c = decode_site(g, hs, par);
  if (MODE == 1) {
    const double ph = px * c.x + py * c.y + pz * c.z;
    const double cs = cos(ph), sn = sin(ph);
#pragma unroll
    for (int T = 0; T < kMesonTypes; ++T) {
      vals[(2 * T) * vstride + o] = re[T] * cs - im[T] * sn;
      vals[(2 * T + 1) * vstride + o] = re[T] * sn + im[T] * cs;
    }
  } else {
    const double w = cos(px * c.x) * cos(py * c.y) * cos(pz * c.z);
#pragma unroll
    for (int T = 0; T < kMesonTypes; ++T) vals[T * vstride + o] = re[T] * w;
  }
}

// out[k * L + slice] = sum over the sites with coordinate `slice` in
// direction dir, in a fixed order
__global__ void __launch_bounds__(kBlock)
k_meson_slices(Geom g, const double* __restrict__ vals, int dir, double* __restrict__ out) {
  const int k = blockIdx.x, slice = blockIdx.y;
  const int L = g.L[dir];
  const int nsl = g.V / L;
  int dims[3], m = 0;
  for (int d = 0; d < 4; ++d)
    if (d != dir) dims[m++] = g.L[d];
  double acc = 0.0;
  for (int j = threadIdx.x; j < nsl; j += blockDim.x) {
    int r = j, cc[4];
    int o[3];
    o[0] = r % dims[0]; r /= dims[0];
    o[1] = r % dims[1];
    o[2] = r / dims[1];
    m = 0;
    for (int d = 0; d < 4; ++d) cc[d] = (d == dir) ? slice : o[m++];
    const int site = cc[0] + g.L[0] * (cc[1] + g.L[1] * (cc[2] + g.L[2] * cc[3]));
    const int par = (cc[0] + cc[1] + cc[2] + cc[3]) & 1;
    acc += vals[(size_t)k * 2 * g.stride + (size_t)par * g.stride + (site >> 1)];
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) out[(size_t)k * L + slice] = acc;
}

}  // namespace ii

using namespace ii;

extern "C" {

ii_status ii_meson_correlators(ii_context* ctx, const ii_field* q1, const ii_field* q2, int mode,
                               int direction, const double* mom, double* out) {
  if (!ctx || !q1 || !q2 || !out || mode < 0 || mode > 2 || direction < 0 || direction > 3)
    return II_ERR_BAD_ARG;
  if (q1->ctx != ctx || q2->ctx != ctx || q1->subset != II_FULL || q2->subset != II_FULL ||
      q1->nrhs != 12 || q2->nrhs != 12)
    return set_error(ctx, II_ERR_BAD_ARG, "propagators must be full fields of 12 columns");
  if (mode != 0 && direction != 3)
    return set_error(ctx, II_ERR_BAD_ARG, "momentum projections are over time slices");
  if (mode != 0 && !mom) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  const int nk = mode == 1 ? 2 * kMesonTypes : kMesonTypes;
  void* wv = nullptr;
  II_CHECK(workspace(ctx, 48, (size_t)nk * 2 * g.stride * sizeof(double), &wv));
  const int L = g.L[direction];
  void* wo = nullptr;
  II_CHECK(workspace(ctx, 49, (size_t)nk * L * sizeof(double), &wo));
  const double px = mom ? mom[0] : 0.0, py = mom ? mom[1] : 0.0, pz = mom ? mom[2] : 0.0;
  const bool same = q1->data == q2->data;
  const dim3 grid(blocks_for(g.Vh), 2);
#define II_MESON(M, S)                                                                       \
  hipLaunchKernelGGL((k_meson_sites<M, S>), grid, dim3(kBlock), 0, ctx->stream, g, q1->data, \
                     q2->data, px, py, pz, (double*)wv)
  if (mode == 0) {
    if (same) II_MESON(0, true); else II_MESON(0, false);
  } else if (mode == 1) {
    if (same) II_MESON(1, true); else II_MESON(1, false);
  } else {
    if (same) II_MESON(2, true); else II_MESON(2, false);
  }
#undef II_MESON
  II_HIP(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_meson_slices, dim3(nk, L), dim3(kBlock), 0, ctx->stream, g,
                     (const double*)wv, direction, (double*)wo);
  II_HIP(ctx, hipGetLastError());
  std::vector<double> h((size_t)nk * L);
  II_HIP(ctx, hipMemcpyAsync(h.data(), wo, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (mode == 1) {
    // calculate_mom's layout: per type, (re, im) interleaved per time slice
    for (int T = 0; T < kMesonTypes; ++T)
      for (int t = 0; t < L; ++t) {
        out[(size_t)T * 2 * L + 2 * t] = h[(size_t)(2 * T) * L + t];
        out[(size_t)T * 2 * L + 2 * t + 1] = h[(size_t)(2 * T + 1) * L + t];
      }
  } else {
    for (size_t i = 0; i < h.size(); ++i) out[i] = h[i];
  }
  return II_OK;
}

}  // extern "C"
