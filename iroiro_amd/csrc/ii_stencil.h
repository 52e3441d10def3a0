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
//
// Addressing: every access is a raw buffer load/store with the array's
// descriptor in scalar registers, the component offset as the scalar
// offset and the site as a 32-bit per-lane offset; this keeps the per-load
// integer work to one shift (64-bit per-lane address arithmetic cost ~20%
// of the VALU instructions of the 16-bit multi-rhs hop).
#pragma once

#include "ii_internal.h"

#include <type_traits>

namespace ii {

// ---------------------------------------------------------------- addressing
// Raw buffer access: one descriptor per array (scalar registers), the
// component offset as the scalar offset and the site as a 32-bit lane
// offset.  Third descriptor word of the gfx9 family (raw, dword data).
constexpr int kBufWord3 = 0x00020000;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)0xFFFFFFFF,
                                           kBufWord3);
}

// element `off + idx` of base (off uniform, idx per lane); the bytes are
// moved with memcpy because __half2 / __hip_bfloat162 are not trivially
// copyable
using U32x2 = uint32_t __attribute__((ext_vector_type(2)));
using U32x4 = uint32_t __attribute__((ext_vector_type(4)));
// Cache policy of a buffer access (instruction bits of gfx950): 0 = default,
// kNT = non-temporal.  Links (each read once per hop) and stencil outputs
// (written once, read by the next kernel) are streamed with kNT so they do
// not evict the neighbour spinors the stencil re-reads from L2; measured on
// the FP64 hop at 32^3x64: 0.362 ms -> 0.289 ms.
constexpr int kNT = 2;
template <typename T, int CP = 0>
__device__ __forceinline__ T ld_at(const T* __restrict__ base, uint32_t off, uint32_t idx) {
  const uint32_t vo = idx * (uint32_t)sizeof(T), so = off * (uint32_t)sizeof(T);
  T r;
  if constexpr (sizeof(T) == 4) {
    uint32_t v = __builtin_amdgcn_raw_buffer_load_b32(rsrc(base), vo, so, CP);
    __builtin_memcpy((void*)&r, &v, 4);
  } else if constexpr (sizeof(T) == 8) {
    U32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rsrc(base), vo, so, CP);
    __builtin_memcpy((void*)&r, &v, 8);
  } else {
    static_assert(sizeof(T) == 16, "4, 8 or 16-byte elements");
    U32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rsrc(base), vo, so, CP);
    __builtin_memcpy((void*)&r, &v, 16);
  }
  return r;
}
template <typename T, int CP = 0>
__device__ __forceinline__ void st_at(T* __restrict__ base, uint32_t off, uint32_t idx, T v) {
  const uint32_t vo = idx * (uint32_t)sizeof(T), so = off * (uint32_t)sizeof(T);
  if constexpr (sizeof(T) == 4) {
    uint32_t w;
    __builtin_memcpy(&w, &v, 4);
    __builtin_amdgcn_raw_buffer_store_b32(w, rsrc(base), vo, so, CP);
  } else if constexpr (sizeof(T) == 8) {
    U32x2 w;
    __builtin_memcpy(&w, &v, 8);
    __builtin_amdgcn_raw_buffer_store_b64(w, rsrc(base), vo, so, CP);
  } else {
    U32x4 w;
    __builtin_memcpy(&w, &v, 16);
    __builtin_amdgcn_raw_buffer_store_b128(w, rsrc(base), vo, so, CP);
  }
}

// ---------------------------------------------------------------- accessors
// FP64 half spinor, [12][stride] double2.
struct SpF64 {
  using R = double;
  using C = double2;
  static constexpr bool kScaled = false;
  const double2* __restrict__ p;
  int stride;
  int rs = 0;  // elements between rhs (0: 12 * stride, back to back)
  __device__ __forceinline__ C ld(int site, int comp) const {
    return ld_at(p, (uint32_t)(comp * stride), (uint32_t)site);
  }
  __device__ __forceinline__ R scale(int) const { return 1.0; }
  __device__ __forceinline__ SpF64 rhs(int r) const {
    return SpF64{p + (size_t)r * (rs ? rs : 12 * stride), stride, rs};
  }
};
struct StF64 {
  using R = double;
  using C = double2;
  double2* __restrict__ p;
  int stride;
  __device__ __forceinline__ void st(int site, const C v[12]) const {
#pragma unroll
    for (int k = 0; k < 12; ++k) st_at<double2, kNT>(p, (uint32_t)(k * stride), (uint32_t)site, v[k]);
  }
  __device__ __forceinline__ StF64 rhs(int r) const { return StF64{p + (size_t)r * 12 * stride, stride}; }
};

// FP32 half spinor, [12][stride] float2.
struct SpF32 {
  using R = float;
  using C = float2;
  static constexpr bool kScaled = false;
  const float2* __restrict__ p;
  int stride;
  __device__ __forceinline__ C ld(int site, int comp) const {
    return ld_at(p, (uint32_t)(comp * stride), (uint32_t)site);
  }
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
    for (int k = 0; k < 12; ++k) st_at<float2, kNT>(p, (uint32_t)(k * stride), (uint32_t)site, v[k]);
  }
  __device__ __forceinline__ StF32 rhs(int r) const { return StF32{p + (size_t)r * 12 * stride, stride}; }
};

// 16-bit half spinor with a per-site FP32 norm: value = stored * norm[site],
// stored components in [-1, 1].  T = __half2 or __hip_bfloat162.
template <typename T> __device__ __forceinline__ float lo_re(T v);
template <typename T> __device__ __forceinline__ float lo_im(T v);
template <> __device__ __forceinline__ float lo_re<__half2>(__half2 v) { return __low2float(v); }
template <> __device__ __forceinline__ float lo_im<__half2>(__half2 v) { return __high2float(v); }
template <> __device__ __forceinline__ float lo_re<__hip_bfloat162>(__hip_bfloat162 v) {
  return __bfloat162float(v.x);
}
template <> __device__ __forceinline__ float lo_im<__hip_bfloat162>(__hip_bfloat162 v) {
  return __bfloat162float(v.y);
}
template <typename T> __device__ __forceinline__ float2 lo_to_f2(T v) {
  return make_float2(lo_re<T>(v), lo_im<T>(v));
}
template <typename T> __device__ __forceinline__ T f2_to_lo(float2 v);
template <> __device__ __forceinline__ __half2 f2_to_lo<__half2>(float2 v) { return __float22half2_rn(v); }
template <> __device__ __forceinline__ __hip_bfloat162 f2_to_lo<__hip_bfloat162>(float2 v) {
  __hip_bfloat162 r;
  r.x = __float2bfloat16(v.x);
  r.y = __float2bfloat16(v.y);
  return r;
}

// Layout of one rhs: the two spins of a pair side by side, [colour c][spin
// pair sp][site][2] of T: component comp = 3 s + c sits at 4-byte word
// ((2 c + s/2) stride + site) 2 + s%2, so one 8-byte load returns spins
// (2 sp, 2 sp + 1) of a colour.  Every spin projector of the stencil needs
// either both spins of a pair or (the time direction) one whole pair.
template <typename T>
struct LoPair {
  T lo, hi;  // spins 2 sp and 2 sp + 1
};
template <typename T>
struct SpLo {
  using R = float;
  using C = float2;
  using Raw = T;
  static constexpr bool kScaled = true;
  const T* __restrict__ p;
  const float* __restrict__ nrm;
  int stride;
  __device__ __forceinline__ T raw(int site, int comp) const {
    const int sp = (comp / 3) >> 1, h = (comp / 3) & 1, c = comp % 3;
    return ld_at(p, (uint32_t)((2 * c + sp) * stride * 2 + h), (uint32_t)site * 2u);
  }
  __device__ __forceinline__ LoPair<T> raw2(int site, int c, int sp) const {
    return ld_at(reinterpret_cast<const LoPair<T>*>(p), (uint32_t)((2 * c + sp) * stride),
                 (uint32_t)site);
  }
  __device__ __forceinline__ C ld(int site, int comp) const { return lo_to_f2<T>(raw(site, comp)); }
  __device__ __forceinline__ R scale(int site) const { return ld_at(nrm, 0u, (uint32_t)site); }
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
    // Sites below 1e-30 are stored as zero: near the front of a point
    // source's Krylov vectors the values underflow to FP32 denormals, whose
    // reciprocal overflows.  They are far below the 16-bit resolution of
    // any field the solver stores.
    if (!(m > 1e-30f)) m = 0.0f;
    const float inv = m > 0.0f ? 1.0f / m : 0.0f;
    st_at<float, kNT>(nrm, 0u, (uint32_t)site, m);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int sp = 0; sp < 2; ++sp) {
        const float2 a = v[3 * (2 * sp) + c], b = v[3 * (2 * sp + 1) + c];
        LoPair<T> pr;
        pr.lo = f2_to_lo<T>(make_float2(a.x * inv, a.y * inv));
        pr.hi = f2_to_lo<T>(make_float2(b.x * inv, b.y * inv));
        st_at<LoPair<T>, kNT>(reinterpret_cast<LoPair<T>*>(p), (uint32_t)((2 * c + sp) * stride),
                              (uint32_t)site, pr);
      }
  }
  __device__ __forceinline__ StLo rhs(int r) const {
    return StLo{p + (size_t)r * 12 * stride, nrm + (size_t)r * stride, stride};
  }
};

// Reconstruct-12: rows 0 and 1 stored, row 2 = conj(row0 x row1) * sign,
// exact for SU(3) (unitary rows, det 1).  The antiperiodic time boundary
// makes the last-slice t-links -SU(3); their sign is restored by flip.
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

// Links.  ld() returns the full 3x3 matrix U[row][col] of the link that
// starts at half site `site` of parity `par` in direction mu.
// Gauge fields are addressed with one descriptor per (parity, direction):
// the block's base is formed in 64 bits, so only one direction's components
// (9 or 6 x stride) ride in the 32-bit offsets.
struct GaugeF64 {
  using R = double;
  using C = double2;
  const double2* __restrict__ p;  // [2][4][9][stride]
  int stride;
  __device__ __forceinline__ void ld(int par, int mu, int site, bool /*tflip*/, C U[3][3]) const {
    const double2* b = p + (size_t)((par * 4 + mu) * 9) * (size_t)stride;
#pragma unroll
    for (int k = 0; k < 9; ++k)
      U[k / 3][k % 3] = ld_at<double2, kNT>(b, (uint32_t)(k * stride), (uint32_t)site);
  }
};

// The whole FP64 gauge field in one descriptor (72 components in the scalar
// offset): the single-rhs FP64 hop runs 3-4% faster this way than with one
// descriptor per block (0.282 against 0.293 ms at 32^3x64), so it is used
// wherever the field fits 4 GiB (V/2 up to about 3.7M sites).
struct GaugeF64Flat {
  using R = double;
  using C = double2;
  const double2* __restrict__ p;  // [2][4][9][stride]
  int stride;
  __device__ __forceinline__ void ld(int par, int mu, int site, bool /*tflip*/, C U[3][3]) const {
#pragma unroll
    for (int k = 0; k < 9; ++k)
      U[k / 3][k % 3] =
          ld_at<double2, kNT>(p, (uint32_t)(((par * 4 + mu) * 9 + k) * stride), (uint32_t)site);
  }
};

struct GaugeF32r12 {
  using R = float;
  using C = float2;
  const float2* __restrict__ p;  // [2][4][6][stride]
  int stride;
  __device__ __forceinline__ void ld(int par, int mu, int site, bool tflip, C U[3][3]) const {
    const float2* b = p + (size_t)((par * 4 + mu) * 6) * (size_t)stride;
#pragma unroll
    for (int k = 0; k < 6; ++k)
      U[k / 3][k % 3] = ld_at<float2, kNT>(b, (uint32_t)(k * stride), (uint32_t)site);
    r12_third_row(U, tflip);
  }
};

template <typename TT>
struct GaugeLoR12 {
  using R = float;
  using C = float2;
  const TT* __restrict__ p;  // [2][4][6][stride]
  int stride;
  __device__ __forceinline__ void ld(int par, int mu, int site, bool tflip, C U[3][3]) const {
    const TT* b = p + (size_t)((par * 4 + mu) * 6) * (size_t)stride;
#pragma unroll
    for (int k = 0; k < 6; ++k)
      U[k / 3][k % 3] = lo_to_f2<TT>(ld_at<TT, kNT>(b, (uint32_t)(k * stride), (uint32_t)site));
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

// ---------------------------------------------------------------- projection
// w1 = v[A1] + c1 v[B1], w2 = v[A2] + c2 v[B2], with c in {+1, -1, +i, -i}
// (codes 0..3), or w = 2 v[A] (code 4, the time direction, where
// 1 -/+ gamma_t = 2 diag(0,0,1,1) / 2 diag(1,1,0,0)).
template <int MU, bool FWD> struct Proj;
template <> struct Proj<0, true>  { static constexpr int A1 = 0, B1 = 3, C1 = 2, A2 = 1, B2 = 2, C2 = 2; };
template <> struct Proj<0, false> { static constexpr int A1 = 0, B1 = 3, C1 = 3, A2 = 1, B2 = 2, C2 = 3; };
template <> struct Proj<1, true>  { static constexpr int A1 = 0, B1 = 3, C1 = 0, A2 = 1, B2 = 2, C2 = 1; };
template <> struct Proj<1, false> { static constexpr int A1 = 0, B1 = 3, C1 = 1, A2 = 1, B2 = 2, C2 = 0; };
template <> struct Proj<2, true>  { static constexpr int A1 = 0, B1 = 2, C1 = 2, A2 = 1, B2 = 3, C2 = 3; };
template <> struct Proj<2, false> { static constexpr int A1 = 0, B1 = 2, C1 = 3, A2 = 1, B2 = 3, C2 = 2; };
template <> struct Proj<3, true>  { static constexpr int A1 = 2, B1 = 2, C1 = 4, A2 = 3, B2 = 3, C2 = 4; };
template <> struct Proj<3, false> { static constexpr int A1 = 0, B1 = 0, C1 = 4, A2 = 1, B2 = 1, C2 = 4; };

template <int CODE, typename R>
__device__ __forceinline__ void combine(R ax, R ay, R bx, R by, R n, R& wx, R& wy) {
  // (a + c b) * n, written as one product and one fused multiply-add per
  // real component so that 16-bit operands convert inside the FMA
  if (CODE == 0) { wx = fma(ax, n, bx * n); wy = fma(ay, n, by * n); }
  else if (CODE == 1) { wx = fma(ax, n, -(bx * n)); wy = fma(ay, n, -(by * n)); }
  else if (CODE == 2) { wx = fma(ax, n, -(by * n)); wy = fma(ay, n, bx * n); }
  else if (CODE == 3) { wx = fma(ax, n, by * n); wy = fma(ay, n, -(bx * n)); }
  else { wx = ax * (n + n); wy = ay * (n + n); }
}

// Loads the neighbour spinor components the projector needs and forms the
// two projected colour vectors (scaled by the per-site norm for 16-bit
// storage).
template <int MU, bool FWD, bool G5IN, typename SpIn>
__device__ __forceinline__ void load_project(const SpIn& in, int nb, typename SpIn::C w1[3],
                                             typename SpIn::C w2[3]) {
  using R = typename SpIn::R;
  using P = Proj<MU, FWD>;
  constexpr int a1 = G5IN ? (P::A1 ^ 2) : P::A1, b1 = G5IN ? (P::B1 ^ 2) : P::B1;
  constexpr int a2 = G5IN ? (P::A2 ^ 2) : P::A2, b2 = G5IN ? (P::B2 ^ 2) : P::B2;
  const R n = in.scale(nb);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    auto va = in.ld(nb, 3 * a1 + c);
    if (P::C1 == 4) {
      combine<4, R>(va.x, va.y, R(0), R(0), n, w1[c].x, w1[c].y);
    } else {
      auto vb = in.ld(nb, 3 * b1 + c);
      combine<P::C1, R>(va.x, va.y, vb.x, vb.y, n, w1[c].x, w1[c].y);
    }
    auto vc = in.ld(nb, 3 * a2 + c);
    if (P::C2 == 4) {
      combine<4, R>(vc.x, vc.y, R(0), R(0), n, w2[c].x, w2[c].y);
    } else {
      auto vd = in.ld(nb, 3 * b2 + c);
      combine<P::C2, R>(vc.x, vc.y, vd.x, vd.y, n, w2[c].x, w2[c].y);
    }
  }
}

// 16-bit storage: the projection is formed from the stored (unscaled)
// components, a + c b, and the per-site norm (times 2 for the time
// direction) is applied once, inside the reconstruction's fused
// multiply-adds (reconstruct_scaled): 12 fewer multiplies per direction
// than scaling the projected vectors.
template <int CODE>
__device__ __forceinline__ void combine_u(float ax, float ay, float bx, float by, float& wx,
                                          float& wy) {
  if (CODE == 0) { wx = ax + bx; wy = ay + by; }
  else if (CODE == 1) { wx = ax - bx; wy = ay - by; }
  else if (CODE == 2) { wx = ax - by; wy = ay + bx; }
  else if (CODE == 3) { wx = ax + by; wy = ay - bx; }
  else { wx = ax; wy = ay; }
}
template <int MU, bool FWD, bool G5IN, typename SpIn>
__device__ __forceinline__ float load_project_u(const SpIn& in, int nb, float2 w1[3],
                                                float2 w2[3]) {
  using P = Proj<MU, FWD>;
  constexpr int a1 = G5IN ? (P::A1 ^ 2) : P::A1, b1 = G5IN ? (P::B1 ^ 2) : P::B1;
  constexpr int a2 = G5IN ? (P::A2 ^ 2) : P::A2, b2 = G5IN ? (P::B2 ^ 2) : P::B2;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    // the spins of colour c by pairs (one 8-byte load each); the time
    // directions need the one pair a1 / 2
    float2 v[4];
    if (P::C1 == 4) {
      const auto pr = in.raw2(nb, c, a1 >> 1);
      v[a1] = lo_to_f2(pr.lo);
      v[a1 + 1] = lo_to_f2(pr.hi);
      v[b1] = v[a1];  // unused by code 4
    } else {
      const auto p0 = in.raw2(nb, c, 0), p1 = in.raw2(nb, c, 1);
      v[0] = lo_to_f2(p0.lo);
      v[1] = lo_to_f2(p0.hi);
      v[2] = lo_to_f2(p1.lo);
      v[3] = lo_to_f2(p1.hi);
    }
    combine_u<P::C1>(v[a1].x, v[a1].y, v[b1].x, v[b1].y, w1[c].x, w1[c].y);
    combine_u<P::C2>(v[a2].x, v[a2].y, v[b2].x, v[b2].y, w2[c].x, w2[c].y);
  }
  const float n = in.scale(nb);
  return P::C1 == 4 ? n + n : n;
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

// acc += sc * (reconstruction of h1, h2), with the same spin tables
template <typename C>
__device__ __forceinline__ void fma_c(C& acc, float sc, C h) {
  acc.x = fmaf(sc, h.x, acc.x);
  acc.y = fmaf(sc, h.y, acc.y);
}
template <typename C>
__device__ __forceinline__ void fma_ci(C& acc, float sc, C h) {  // acc += sc * i h
  acc.x = fmaf(-sc, h.y, acc.x);
  acc.y = fmaf(sc, h.x, acc.y);
}
template <int MU, bool FWD, typename C>
__device__ __forceinline__ void reconstruct_scaled(C acc[4][3], const C h1[3], const C h2[3],
                                                   float sc) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (!(MU == 3 && FWD)) {
      fma_c(acc[0][c], sc, h1[c]);
      fma_c(acc[1][c], sc, h2[c]);
    }
    if (MU == 0) {
      if (FWD) { fma_ci(acc[2][c], -sc, h2[c]); fma_ci(acc[3][c], -sc, h1[c]); }
      else     { fma_ci(acc[2][c], sc, h2[c]);  fma_ci(acc[3][c], sc, h1[c]); }
    } else if (MU == 1) {
      if (FWD) { fma_c(acc[2][c], -sc, h2[c]); fma_c(acc[3][c], sc, h1[c]); }
      else     { fma_c(acc[2][c], sc, h2[c]);  fma_c(acc[3][c], -sc, h1[c]); }
    } else if (MU == 2) {
      if (FWD) { fma_ci(acc[2][c], -sc, h1[c]); fma_ci(acc[3][c], sc, h2[c]); }
      else     { fma_ci(acc[2][c], sc, h1[c]);  fma_ci(acc[3][c], -sc, h2[c]); }
    } else if (FWD) {
      fma_c(acc[2][c], sc, h1[c]);
      fma_c(acc[3][c], sc, h2[c]);
    }
  }
}

// One direction: project the neighbour spinor, multiply by U (forward) or
// U^dag (backward), reconstruct into acc.  The link comes from getU(U)
// (global memory or LDS).
template <int MU, bool FWD, bool G5IN, typename UC, typename SpIn, typename GetU>
__device__ __forceinline__ void hop_dir_with(typename SpIn::C acc[4][3], const SpIn& in, int nb,
                                             GetU&& getU) {
  using C = typename SpIn::C;
  using R = typename SpIn::R;
  C w1[3], w2[3];
  float sc = 1.0f;
  if constexpr (SpIn::kScaled) sc = load_project_u<MU, FWD, G5IN>(in, nb, w1, w2);
  else load_project<MU, FWD, G5IN>(in, nb, w1, w2);
  UC U[3][3];
  getU(U);
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
  }
  if constexpr (SpIn::kScaled) reconstruct_scaled<MU, FWD>(acc, h1, h2, sc);
  else reconstruct<MU, FWD>(acc, h1, h2);
}

// ---------------------------------------------------------------- FP16 dot2
// The FP16 multi-rhs hop does its colour products with v_dot2_f32_f16 (two
// FP16 products summed into an FP32 accumulator): one complex multiply-add
// is two instructions instead of four FP32 FMAs.  The operands stay in
// FP16: the projected vectors a + c b are formed with packed FP16 adds from
// the stored (norm-scaled) components, |a + c b| <= 2, and the links are
// staged in LDS as FP16.
typedef _Float16 h2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ h2v as_h2v(__half2 v) { return __builtin_bit_cast(h2v, v); }

// One entry M_ab = mx + i my of the colour matrix a direction applies (U for
// a forward hop, U^dag for a backward one), staged so that for w = (wx, wy)
// Re(M_ab w) = dot2(re, w) and Im(M_ab w) = dot2(im, w):
// re = (mx, -my), im = (my, mx).
struct MDot2 {
  h2v re, im;
};
__device__ __forceinline__ MDot2 mdot2(float mx, float my) {
  MDot2 m;
  m.re = h2v{(_Float16)mx, (_Float16)(-my)};
  m.im = h2v{(_Float16)my, (_Float16)mx};
  return m;
}

// dot2(a, b) + 0 as the three-operand v_dot2_f32_f16 with an inline zero:
// left to itself the compiler starts every chain with a v_mov of 0 and the
// accumulating two-operand form (96 extra moves per site and rhs).
__device__ __forceinline__ float dot2_first(h2v a, h2v b) {
  float r;
  asm("v_dot2_f32_f16 %0, %1, %2, 0" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// a + c b for the projector codes of combine_u, in packed FP16 (each form
// is one v_pk_add_f16 with operand-select and negate modifiers)
template <int CODE>
__device__ __forceinline__ h2v combine_h(h2v a, h2v b) {
  if (CODE == 0) return a + b;
  else if (CODE == 1) return a - b;
  else if (CODE == 2) return a + h2v{-b.y, b.x};
  else if (CODE == 3) return a + h2v{b.y, -b.x};
  else return a;
}

// One direction of the FP16 stencil, split into its loads and its
// arithmetic so that the caller can issue the next direction's loads before
// this one's products (software pipelining).  getM(k) returns the staged
// entry k = 3a + b of M; projection / reconstruction tables as in
// hop_dir_with.
struct RawDir {
  __half2 v[4][3];  // [spin][colour]; the time directions load one spin pair
  float n;          // site norm
};
template <int MU, bool FWD, bool G5IN, typename SpIn>
__device__ __forceinline__ void load_dir_raw(const SpIn& in, int nb, RawDir& r) {
  using P = Proj<MU, FWD>;
  constexpr int a1 = G5IN ? (P::A1 ^ 2) : P::A1;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (P::C1 == 4) {
      // 1 -/+ gamma_t: spins a1, a1 + 1 = the pair a1 / 2
      const auto pr = in.raw2(nb, c, a1 >> 1);
      r.v[a1][c] = pr.lo;
      r.v[a1 + 1][c] = pr.hi;
    } else {
      const auto p0 = in.raw2(nb, c, 0), p1 = in.raw2(nb, c, 1);
      r.v[0][c] = p0.lo;
      r.v[1][c] = p0.hi;
      r.v[2][c] = p1.lo;
      r.v[3][c] = p1.hi;
    }
  }
  r.n = in.scale(nb);
}
// The input spinor of a lane's own tile site (the 64 sites of the tile in
// the input parity), kept in registers: a neighbour inside the tile is
// taken from the lane that holds it (ds_bpermute) instead of memory.  With
// 16-site rows every x neighbour and 3 of 4 y neighbours are inside.
struct RawSite {
  __half2 v[4][3];
  float n;
};
template <typename SpIn>
__device__ __forceinline__ void load_site_raw(const SpIn& in, int hs, RawSite& o) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const auto p0 = in.raw2(hs, c, 0), p1 = in.raw2(hs, c, 1);
    o.v[0][c] = p0.lo;
    o.v[1][c] = p0.hi;
    o.v[2][c] = p1.lo;
    o.v[3][c] = p1.hi;
  }
  o.n = in.scale(hs);
}
__device__ __forceinline__ __half2 shfl_h2(__half2 v, int lane) {
  return __builtin_bit_cast(__half2, __shfl(__builtin_bit_cast(int, v), lane));
}
// load_dir_raw for the neighbour nb, from the owning lane when nb is one of
// the tile's sites [hs0, hs0 + 64), else from memory
template <int MU, bool FWD, bool G5IN, typename SpIn>
__device__ __forceinline__ void load_dir_tile(const SpIn& in, int nb, int hs0, int lane,
                                              const RawSite& own, RawDir& r) {
  using P = Proj<MU, FWD>;
  constexpr int a1 = G5IN ? (P::A1 ^ 2) : P::A1;
  const int src = nb - hs0;
  const bool inside = (unsigned)src < 64u;
  if (__any(inside)) {
    const int sl = inside ? src : lane;  // every active lane takes part
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (P::C1 == 4) {
        r.v[a1][c] = shfl_h2(own.v[a1][c], sl);
        r.v[a1 + 1][c] = shfl_h2(own.v[a1 + 1][c], sl);
      } else {
#pragma unroll
        for (int sp = 0; sp < 4; ++sp) r.v[sp][c] = shfl_h2(own.v[sp][c], sl);
      }
    }
    r.n = __shfl(own.n, sl);
  }
  if (!__all(inside)) {
    if (!inside) load_dir_raw<MU, FWD, G5IN>(in, nb, r);
  }
}

template <int MU, bool FWD, bool G5IN, typename GetM>
__device__ __forceinline__ void hop_dir_dot2_raw(float2 acc[4][3], const RawDir& r, GetM&& getM) {
  using P = Proj<MU, FWD>;
  constexpr int a1 = G5IN ? (P::A1 ^ 2) : P::A1, b1 = G5IN ? (P::B1 ^ 2) : P::B1;
  constexpr int a2 = G5IN ? (P::A2 ^ 2) : P::A2, b2 = G5IN ? (P::B2 ^ 2) : P::B2;
  h2v w1[3], w2[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    w1[c] = combine_h<P::C1>(as_h2v(r.v[a1][c]), P::C1 == 4 ? h2v{} : as_h2v(r.v[b1][c]));
    w2[c] = combine_h<P::C2>(as_h2v(r.v[a2][c]), P::C2 == 4 ? h2v{} : as_h2v(r.v[b2][c]));
  }
  const float sc = P::C1 == 4 ? r.n + r.n : r.n;
  float2 h1[3], h2[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const MDot2 m0 = getM(3 * a);
    float r1 = dot2_first(m0.re, w1[0]), i1 = dot2_first(m0.im, w1[0]);
    float r2 = dot2_first(m0.re, w2[0]), i2 = dot2_first(m0.im, w2[0]);
#pragma unroll
    for (int b = 1; b < 3; ++b) {
      const MDot2 m = getM(3 * a + b);
      r1 = __builtin_amdgcn_fdot2(m.re, w1[b], r1, false);
      i1 = __builtin_amdgcn_fdot2(m.im, w1[b], i1, false);
      r2 = __builtin_amdgcn_fdot2(m.re, w2[b], r2, false);
      i2 = __builtin_amdgcn_fdot2(m.im, w2[b], i2, false);
    }
    h1[a] = make_float2(r1, i1);
    h2[a] = make_float2(r2, i2);
  }
  reconstruct_scaled<MU, FWD>(acc, h1, h2, sc);
}

template <int MU, bool FWD, bool G5IN, typename SpIn, typename Gauge>
__device__ __forceinline__ void hop_dir(typename SpIn::C acc[4][3], const SpIn& in,
                                        const Gauge& gauge, int nb, int lpar, int lsite,
                                        bool tflip) {
  hop_dir_with<MU, FWD, G5IN, typename Gauge::C>(
      acc, in, nb, [&](typename Gauge::C U[3][3]) { gauge.ld(lpar, MU, lsite, tflip, U); });
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
  // directions between register barriers: FP64 keeps four directions'
  // loads in flight (256 VGPRs, 0.293 -> 0.281 ms at 32^3x64); the 32- and
  // 16-bit hops pin every direction
  constexpr int kDirGroup = std::is_same<R, double>::value ? 4 : 1;
  hop_dir<0, true, G5IN>(acc, in, gauge, neighbour(g, sc, hs, 0, true), out_par, hs, false);
  if constexpr (1 % kDirGroup == 0) II_DIR_BARRIER();
  { int nb = neighbour(g, sc, hs, 0, false); hop_dir<0, false, G5IN>(acc, in, gauge, nb, op, nb, false); }
  if constexpr (2 % kDirGroup == 0) II_DIR_BARRIER();
  hop_dir<1, true, G5IN>(acc, in, gauge, neighbour(g, sc, hs, 1, true), out_par, hs, false);
  if constexpr (3 % kDirGroup == 0) II_DIR_BARRIER();
  { int nb = neighbour(g, sc, hs, 1, false); hop_dir<1, false, G5IN>(acc, in, gauge, nb, op, nb, false); }
  if constexpr (4 % kDirGroup == 0) II_DIR_BARRIER();
  hop_dir<2, true, G5IN>(acc, in, gauge, neighbour(g, sc, hs, 2, true), out_par, hs, false);
  if constexpr (5 % kDirGroup == 0) II_DIR_BARRIER();
  { int nb = neighbour(g, sc, hs, 2, false); hop_dir<2, false, G5IN>(acc, in, gauge, nb, op, nb, false); }
  if constexpr (6 % kDirGroup == 0) II_DIR_BARRIER();
  hop_dir<3, true, G5IN>(acc, in, gauge, neighbour(g, sc, hs, 3, true), out_par, hs, tf_fw);
  if constexpr (7 % kDirGroup == 0) II_DIR_BARRIER();
  { int nb = neighbour(g, sc, hs, 3, false); hop_dir<3, false, G5IN>(acc, in, gauge, nb, op, nb, tf_bw); }
}

// XCD-aware block order: hardware dispatches block b to XCD b % 8; the
// remap gives each XCD a contiguous range of logical blocks so that
// neighbouring sites (and the rhs of one site tile) share an L2.
__host__ __device__ __forceinline__ int xcd_remap(int b, int nblocks) {
  const int nx = 8;
  if (nblocks % nx) return b;
  return (b % nx) * (nblocks / nx) + b / nx;
}

// First half site of the logical block that physical block b handles.
__host__ __device__ __forceinline__ int hop_block_base(const Geom& g, const HopOrder& o, int b,
                                                       int nblk, int bs) {
  if (o.mode == 1) {
    const int x = b & 7, j = b >> 3;
    const int bb = j % o.bpp;
    int r = j / o.bpp;
    const int tt = r % o.W;
    r /= o.W;
    const int z = r % g.L[2];
    const int strip = r / g.L[2];
    const int t = x * (g.L[3] >> 3) + strip * o.W + tt;
    return ((z + g.L[2] * t) * o.bpp + bb) * bs;
  }
  return xcd_remap(b, nblk) * bs;
}

}  // namespace ii
