// The Wilson hopping stencil for gfx950, shared by all storage precisions.
//
// One thread computes one output site (of one right-hand side):
//   sum_mu (1 - gamma_mu) U_mu(x) psi(x+mu) + (1 + gamma_mu) U_mu(x-mu)^dag psi(x-mu)
// in the Dirac basis of lib/Dirac_ops/wilsonLikeUtils.cpp:9-49, i.e. the
// operator H of D = 1 - kappa H (lib/Dirac_ops/dirac_wilson.cpp:107-118).
// Spin projection/reconstruction follow the tables of
// lib/Dirac_ops/dirac_wilson_improved.code (x+ :5-78, y+ :80-160,
// z+ :162-243, t+ :245-329, x- :331-412, y- :414-492, z- :494-572,
// t- :574-648); the summation order x+,x-,y+,y-,z+,z-,t+,t- is the
// reference's (mult_offdiag, dirac_wilson.cpp:107-113).
//
// gamma5 is folded into the loads (G5IN) and stores (G5OUT) as the spin
// permutation s -> s^2 (wilsonLikeUtils.cpp:42-49), so D^dag = g5 D g5
// costs no extra pass.  Storage is abstracted by accessor types so that the
// same stencil runs on FP64, FP32 and FP16/BF16-with-norm fields and on
// r18/r12 links.
#pragma once

#include "ii_internal.h"

namespace ii {

// ---------------------------------------------------------------- accessors
// FP64 half spinor, [12][stride] double2.
struct SpF64 {
  using R = double;
  using C = double2;
  const double2* __restrict__ p;
  int stride;
  __device__ __forceinline__ C ld(int site, int comp) const { return p[comp * stride + site]; }
  __device__ __forceinline__ R scale(int) const { return 1.0; }
  __device__ __forceinline__ SpF64 rhs(int r) const { return SpF64{p + (size_t)r * 12 * stride, stride}; }
};
struct StF64 {
  using R = double;
  using C = double2;
  double2* __restrict__ p;
  int stride;
  __device__ __forceinline__ void st(int site, const C v[12]) const {
#pragma unroll
    for (int k = 0; k < 12; ++k) p[k * stride + site] = v[k];
  }
  __device__ __forceinline__ StF64 rhs(int r) const { return StF64{p + (size_t)r * 12 * stride, stride}; }
};

// FP32 half spinor, [12][stride] float2.
struct SpF32 {
  using R = float;
  using C = float2;
  const float2* __restrict__ p;
  int stride;
  __device__ __forceinline__ C ld(int site, int comp) const { return p[comp * stride + site]; }
  __device__ __forceinline__ R scale(int) const { return 1.0f; }
  __device__ __forceinline__ SpF32 rhs(int r) const { return SpF32{p + (size_t)r * 12 * stride, stride}; }
};
struct StF32 {
  using R = float;
  using C = float2;
  float2* __restrict__ p;
  int stride;
  __device__ __forceinline__ void st(int site, const C v[12]) const {
#pragma unroll
    for (int k = 0; k < 12; ++k) p[k * stride + site] = v[k];
  }
  __device__ __forceinline__ StF32 rhs(int r) const { return StF32{p + (size_t)r * 12 * stride, stride}; }
};

// 16-bit half spinor with a per-site FP32 norm: value = stored * norm[site],
// stored components in [-1, 1].  T = __half2 or __hip_bfloat162.
template <typename T> __device__ __forceinline__ float2 lo_to_f2(T v);
template <> __device__ __forceinline__ float2 lo_to_f2<__half2>(__half2 v) { return __half22float2(v); }
template <> __device__ __forceinline__ float2 lo_to_f2<__hip_bfloat162>(__hip_bfloat162 v) {
  return make_float2(__bfloat162float(v.x), __bfloat162float(v.y));
}
template <typename T> __device__ __forceinline__ T f2_to_lo(float2 v);
template <> __device__ __forceinline__ __half2 f2_to_lo<__half2>(float2 v) { return __float22half2_rn(v); }
template <> __device__ __forceinline__ __hip_bfloat162 f2_to_lo<__hip_bfloat162>(float2 v) {
  __hip_bfloat162 r;
  r.x = __float2bfloat16(v.x);
  r.y = __float2bfloat16(v.y);
  return r;
}

template <typename T>
struct SpLo {
  using R = float;
  using C = float2;
  const T* __restrict__ p;
  const float* __restrict__ nrm;
  int stride;
  __device__ __forceinline__ C ld(int site, int comp) const { return lo_to_f2<T>(p[comp * stride + site]); }
  __device__ __forceinline__ R scale(int site) const { return nrm[site]; }
  __device__ __forceinline__ SpLo rhs(int r) const {
    return SpLo{p + (size_t)r * 12 * stride, nrm + (size_t)r * stride, stride};
  }
};
template <typename T>
struct StLo {
  using R = float;
  using C = float2;
  T* __restrict__ p;
  float* __restrict__ nrm;
  int stride;
  __device__ __forceinline__ void st(int site, const C v[12]) const {
    float m = 0.0f;
#pragma unroll
    for (int k = 0; k < 12; ++k) m = fmaxf(m, fmaxf(fabsf(v[k].x), fabsf(v[k].y)));
    float inv = m > 0.0f ? 1.0f / m : 0.0f;
    nrm[site] = m;
#pragma unroll
    for (int k = 0; k < 12; ++k) p[k * stride + site] = f2_to_lo<T>(make_float2(v[k].x * inv, v[k].y * inv));
  }
  __device__ __forceinline__ StLo rhs(int r) const {
    return StLo{p + (size_t)r * 12 * stride, nrm + (size_t)r * stride, stride};
  }
};

// Links.  ld() returns the full 3x3 matrix U[row][col].
struct GaugeF64 {
  using R = double;
  using C = double2;
  const double2* __restrict__ p;  // [2][4][9][stride]
  int stride;
  __device__ __forceinline__ void ld(int par, int mu, int site, bool /*tflip*/, C U[3][3]) const {
    const double2* b = p + (size_t)((par * 4 + mu) * 9) * stride + site;
#pragma unroll
    for (int k = 0; k < 9; ++k) U[k / 3][k % 3] = b[(size_t)k * stride];
  }
};

// Reconstruct-12: rows 0 and 1 stored, row 2 = conj(row0 x row1) * sign,
// exact for SU(3) (unitary rows, det 1).  The antiperiodic time boundary
// makes the last-slice t-links -SU(3); their sign is restored by tflip.
template <typename C>
__device__ __forceinline__ void r12_third_row(C U[3][3], bool flip) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int k = (j + 1) % 3, l = (j + 2) % 3;
    C a = U[0][k], b = U[1][l], c = U[0][l], d = U[1][k];
    auto re = a.x * b.x - a.y * b.y - (c.x * d.x - c.y * d.y);
    auto im = a.x * b.y + a.y * b.x - (c.x * d.y + c.y * d.x);
    U[2][j].x = flip ? -re : re;
    U[2][j].y = flip ? im : -im;
  }
}

struct GaugeF32r12 {
  using R = float;
  using C = float2;
  const float2* __restrict__ p;  // [2][4][6][stride]
  int stride;
  __device__ __forceinline__ void ld(int par, int mu, int site, bool tflip, C U[3][3]) const {
    const float2* b = p + (size_t)((par * 4 + mu) * 6) * stride + site;
#pragma unroll
    for (int k = 0; k < 6; ++k) U[k / 3][k % 3] = b[(size_t)k * stride];
    r12_third_row(U, tflip);
  }
};

template <typename T>
struct GaugeLoR12 {
  using R = float;
  using C = float2;
  const T* __restrict__ p;  // [2][4][6][stride]
  int stride;
  __device__ __forceinline__ void ld(int par, int mu, int site, bool tflip, C U[3][3]) const {
    const T* b = p + (size_t)((par * 4 + mu) * 6) * stride + site;
#pragma unroll
    for (int k = 0; k < 6; ++k) U[k / 3][k % 3] = lo_to_f2<T>(b[(size_t)k * stride]);
    r12_third_row(U, tflip);
  }
};

// Ends one direction of the stencil: the accumulators pass through an empty
// asm statement (so every product of this direction is computed here, not
// sunk to the final store), and the memory clobber keeps the next
// direction's loads after it.  Without this the compiler keeps all eight
// directions' spinors and links live at once (about 500 registers in FP64,
// one wave per SIMD).
template <typename C>
__device__ __forceinline__ void pin_acc(C acc[4][3]) {
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < 3; ++c) asm volatile("" : "+v"(acc[s][c].x), "+v"(acc[s][c].y) : : "memory");
}
#define II_DIR_BARRIER() pin_acc(acc)

// ---------------------------------------------------------------- stencil
template <int MU, bool FWD, typename C>
__device__ __forceinline__ void project(const C v[4][3], C w1[3], C w2[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (MU == 0) {
      if (FWD) { w1[c] = cadd_i(v[0][c], v[3][c]); w2[c] = cadd_i(v[1][c], v[2][c]); }
      else     { w1[c] = csub_i(v[0][c], v[3][c]); w2[c] = csub_i(v[1][c], v[2][c]); }
    } else if (MU == 1) {
      if (FWD) { w1[c] = cadd(v[0][c], v[3][c]); w2[c] = csub(v[1][c], v[2][c]); }
      else     { w1[c] = csub(v[0][c], v[3][c]); w2[c] = cadd(v[1][c], v[2][c]); }
    } else if (MU == 2) {
      if (FWD) { w1[c] = cadd_i(v[0][c], v[2][c]); w2[c] = csub_i(v[1][c], v[3][c]); }
      else     { w1[c] = csub_i(v[0][c], v[2][c]); w2[c] = cadd_i(v[1][c], v[3][c]); }
    } else {
      if (FWD) { w1[c] = cadd(v[2][c], v[2][c]); w2[c] = cadd(v[3][c], v[3][c]); }
      else     { w1[c] = cadd(v[0][c], v[0][c]); w2[c] = cadd(v[1][c], v[1][c]); }
    }
  }
}

template <int MU, bool FWD, typename C>
__device__ __forceinline__ void reconstruct(C acc[4][3], const C h1[3], const C h2[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (!(MU == 3 && FWD)) {
      acc[0][c] = cadd(acc[0][c], h1[c]);
      acc[1][c] = cadd(acc[1][c], h2[c]);
    }
    if (MU == 0) {
      if (FWD) { acc[2][c] = csub_i(acc[2][c], h2[c]); acc[3][c] = csub_i(acc[3][c], h1[c]); }
      else     { acc[2][c] = cadd_i(acc[2][c], h2[c]); acc[3][c] = cadd_i(acc[3][c], h1[c]); }
    } else if (MU == 1) {
      if (FWD) { acc[2][c] = csub(acc[2][c], h2[c]); acc[3][c] = cadd(acc[3][c], h1[c]); }
      else     { acc[2][c] = cadd(acc[2][c], h2[c]); acc[3][c] = csub(acc[3][c], h1[c]); }
    } else if (MU == 2) {
      if (FWD) { acc[2][c] = csub_i(acc[2][c], h1[c]); acc[3][c] = cadd_i(acc[3][c], h2[c]); }
      else     { acc[2][c] = cadd_i(acc[2][c], h1[c]); acc[3][c] = csub_i(acc[3][c], h2[c]); }
    } else if (FWD) {
      acc[2][c] = cadd(acc[2][c], h1[c]);
      acc[3][c] = cadd(acc[3][c], h2[c]);
    }
  }
}

// One direction: load neighbour spinor (only the spins the projector
// needs), project, multiply by U or U^dag, reconstruct into acc.
template <int MU, bool FWD, bool G5IN, typename SpIn, typename Gauge>
__device__ __forceinline__ void hop_dir(typename SpIn::C acc[4][3], const SpIn& in,
                                        const Gauge& gauge, int nb, int lpar, int lsite,
                                        bool tflip) {
  using C = typename SpIn::C;
  using R = typename SpIn::R;
  C v[4][3];
  // spins needed: t+ uses 2,3; t- uses 0,1; others all four
  constexpr int s_lo = (MU == 3) ? (FWD ? 2 : 0) : 0;
  constexpr int s_hi = (MU == 3) ? (FWD ? 4 : 2) : 4;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (s >= s_lo && s < s_hi) {
        const int sp = G5IN ? (s ^ 2) : s;
        v[s][c] = in.ld(nb, 3 * sp + c);
      } else {
        v[s][c] = cmk<C>(R(0), R(0));
      }
    }
  const R sc = in.scale(nb);
  C w1[3], w2[3];
  project<MU, FWD>(v, w1, w2);
  typename Gauge::C U[3][3];
  gauge.ld(lpar, MU, lsite, tflip, U);
  C h1[3], h2[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    h1[a] = cmk<C>(R(0), R(0));
    h2[a] = cmk<C>(R(0), R(0));
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      C u = cmk<C>((R)U[FWD ? a : b][FWD ? b : a].x, (R)U[FWD ? a : b][FWD ? b : a].y);
      if (FWD) { cmac(h1[a], u, w1[b]); cmac(h2[a], u, w2[b]); }
      else     { cmacc(h1[a], u, w1[b]); cmacc(h2[a], u, w2[b]); }
    }
    h1[a] = cscale(h1[a], sc);
    h2[a] = cscale(h2[a], sc);
  }
  reconstruct<MU, FWD>(acc, h1, h2);
}

// The full 8-direction hop at half site hs of parity out_par.
// tbc_flip: r12 links on the last t-slice carry a -1 (antiperiodic time BC).
template <bool G5IN, typename SpIn, typename Gauge>
__device__ __forceinline__ void wilson_hop_site(typename SpIn::C acc[4][3], const Geom& g,
                                                const SpIn& in, const Gauge& gauge, int hs,
                                                int out_par, bool tbc_flip) {
  using C = typename SpIn::C;
  using R = typename SpIn::R;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[s][c] = cmk<C>(R(0), R(0));
  const SiteCoord sc = decode_site(g, hs, out_par);
  const int op = 1 - out_par;
  const bool tf_fw = tbc_flip && sc.t == g.L[3] - 1;
  const bool tf_bw = tbc_flip && sc.t == 0;
  // The scheduling barriers keep the compiler from hoisting the loads of
  // all eight directions at once, which would need ~500 registers (FP64)
  // and drop occupancy to one wave per SIMD.
  hop_dir<0, true, G5IN>(acc, in, gauge, neighbour(g, sc, hs, 0, true), out_par, hs, false);
  II_DIR_BARRIER();
  { int nb = neighbour(g, sc, hs, 0, false); hop_dir<0, false, G5IN>(acc, in, gauge, nb, op, nb, false); }
  II_DIR_BARRIER();
  hop_dir<1, true, G5IN>(acc, in, gauge, neighbour(g, sc, hs, 1, true), out_par, hs, false);
  II_DIR_BARRIER();
  { int nb = neighbour(g, sc, hs, 1, false); hop_dir<1, false, G5IN>(acc, in, gauge, nb, op, nb, false); }
  II_DIR_BARRIER();
  hop_dir<2, true, G5IN>(acc, in, gauge, neighbour(g, sc, hs, 2, true), out_par, hs, false);
  II_DIR_BARRIER();
  { int nb = neighbour(g, sc, hs, 2, false); hop_dir<2, false, G5IN>(acc, in, gauge, nb, op, nb, false); }
  II_DIR_BARRIER();
  hop_dir<3, true, G5IN>(acc, in, gauge, neighbour(g, sc, hs, 3, true), out_par, hs, tf_fw);
  II_DIR_BARRIER();
  { int nb = neighbour(g, sc, hs, 3, false); hop_dir<3, false, G5IN>(acc, in, gauge, nb, op, nb, tf_bw); }
}

// XCD-aware block order: hardware dispatches block b to XCD b % 8; the
// remap gives each XCD a contiguous range of logical blocks so that
// neighbouring sites (and the rhs of one site tile) share an L2.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
  const int nx = 8;
  if (nblocks % nx) return b;
  return (b % nx) * (nblocks / nx) + b / nx;
}

}  // namespace ii
