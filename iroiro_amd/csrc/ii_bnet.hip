// The learned preconditioner M applied to many right-hand sides at once.
// No counterpart in the reference (its solvers are unpreconditioned:
// lib/Solver/solver_CG.cpp, multiShiftSolver_CG.cpp).
//
// M r = P_e N(r on even sites, 0 on odd sites), N the 4-layer linear network
// (widths 1 -> 8 -> 8 -> 8 -> 1) of ii_learned.hip:
//   L_l[o](x) = sum_{d=0..8} sum_i W_l[d][o][i] (T_d L_{l-1}[i])(x),
// T_0 = 1, T_{2mu+1} psi(x) = U_mu(x) psi(x+mu), T_{2mu+2} psi(x) =
// U_mu(x-mu)^dag psi(x-mu).  The spin/channel matrices W commute with the
// colour/site transports T, and the input vanishes on odd sites, so
//   L1(even) = W1_0 r,
//   L1(odd)  = sum_{d>0} W1_d T_d r,
//   L2(odd)  = sum_{d>0} (W2_0 W1_d + W2_d W1_0) T_d r,
//   L2(even) = (W2_0 W1_0) r + sum_{d>0} W2_d T_d L1(odd),
//   L3(odd)  = W3_0 L2(odd) + sum_{d>0} W3_d T_d L2(even),
//   z = L4(even) = (W4_0 W3_0) L2(even)
//                  + sum_{d>0} T_d [(W4_0 W3_d) L2(odd) + W4_d L3(odd)].
// L1(even) and L3(even) are never formed, and only two of the four stages
// transport 8 channels, each on one parity: half the transports of the
// layer-by-layer evaluation.  Four kernels, each over 16-site tiles of one
// parity and every rhs:
//   k_bnet_first  odd:  16 channels [L1(odd) | L2(odd)] from 8 transports of r
//   k_bnet_mid<0> even: L2(even), 8 transported channels of L1(odd) + r
//   k_bnet_mid<1> odd:  L3(odd) over L1(odd)'s slot, 8 transported channels
//                       of L2(even) + L2(odd)
//   k_bnet_last   even: z, mixing [L3(odd) | L2(odd)] at the neighbour first
//                       (16 -> 1 channel), then one transport per path
// Channel/spin mixing runs on v_mfma_f32_16x16x32_bf16 (bf16 weights and
// activations, FP32 accumulation) as in ii_learned.hip: B operand k = 8 i +
// 2 t + re/im (input channel i, spin t), n = site; A the real embedding
// [[Re X, -Im X], [Im X, Re X]] of the (composite) channel-spin matrix.
// Transports run on v_dot2_f32_bf16: the link matrices are stored per site
// and path as pairs P = (Re M, -Im M), Q = (Im M, Re M) of bf16, so
// Re(M v) = sum_b dot2(P_ab, v_b) and Im(M v) = sum_b dot2(Q_ab, v_b) for
// v_b = (Re, Im) bf16.
//
// Activations are bfloat16, one dword (re, im) per complex number, in
// records [rhs][site][channel][colour][spin]: the 16 bytes of one channel
// and colour are exactly one B operand (spins 0..3, re/im).  The input r and
// the output z are FP32 half fields [rhs][12 = 3 spin + colour][stride].
#include "ii_ops.h"
#include "ii_stencil.h"

#include <algorithm>
#include <complex>
#include <cstring>
#include <vector>

namespace ii {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kHop = 8;          // transported paths: 2 mu (forward), 2 mu + 1 (backward)
constexpr int kLinkWords = 20;   // P_ab (9), Q_ab (9), 2 words of padding: 80 bytes
constexpr int kBThreads = 256;   // 4 waves; a block runs its tiles for every rhs
constexpr int kOddCh = 16;       // odd-site record: [L1 or L3 (0..7) | L2 (8..15)]
constexpr int kEvenCh = 8;       // even-site record: L2(even)

constexpr int fr_mb(int cout) { return (8 * cout + 15) / 16; }
constexpr int fr_kb(int cin) { return (8 * cin + 31) / 32; }

// ------------------------------------------------------------ device helpers
__device__ __forceinline__ U32x4 ld128(const void* base, uint32_t soff, uint32_t voff) {
  return __builtin_amdgcn_raw_buffer_load_b128(rsrc(base), voff, soff, 0);
}
__device__ __forceinline__ void st64(void* base, uint32_t soff, uint32_t voff, U32x2 v) {
  __builtin_amdgcn_raw_buffer_store_b64(v, rsrc(base), voff, soff, 0);
}

__device__ __forceinline__ bf16x8 as_frag(U32x4 v) {
  bf16x8 r;
  __builtin_memcpy(&r, &v, 16);
  return r;
}
__device__ __forceinline__ bf16x2v as_b2(uint32_t v) {
  bf16x2v r;
  __builtin_memcpy(&r, &v, 4);
  return r;
}
__device__ __forceinline__ uint32_t pack_b2(float re, float im) {
  bf16x2v r;
  r[0] = (__bf16)re;
  r[1] = (__bf16)im;
  uint32_t w;
  __builtin_memcpy(&w, &r, 4);
  return w;
}
__device__ __forceinline__ float lo_f(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float hi_f(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

// B operands of M v: v[b] holds spins 0..3 of colour b as (re, im) bf16
// pairs; L the 5 x 16 bytes of the link record (P then Q).
__device__ __forceinline__ void transport(const U32x4 L[5], const U32x4 v[3], bf16x8 w[3]) {
  uint32_t W[20];
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int e = 0; e < 4; ++e) W[4 * k + e] = L[k][e];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float re[4], im[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      float r = 0.0f, i = 0.0f;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        r = __builtin_amdgcn_fdot2_f32_bf16(as_b2(W[3 * a + b]), as_b2(v[b][s]), r, false);
        i = __builtin_amdgcn_fdot2_f32_bf16(as_b2(W[9 + 3 * a + b]), as_b2(v[b][s]), i, false);
      }
      re[s] = r;
      im[s] = i;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      w[a][2 * s] = (__bf16)re[s];
      w[a][2 * s + 1] = (__bf16)im[s];
    }
  }
}

// r (FP32 half field of one rhs, [12][stride] float2) at site s as bf16
// pairs per colour
__device__ __forceinline__ void load_r_b2(const float2* __restrict__ r, int stride, int s,
                                          U32x4 v[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int sp = 0; sp < 4; ++sp) {
      const float2 x = ld_at(r, (uint32_t)((3 * sp + c) * stride), (uint32_t)s);
      v[c][sp] = pack_b2(x.x, x.y);
    }
}

struct BArgs {
  Geom g;
  int nrhs;
  int tiles;            // 16-site tiles of the output parity
  int tiles_per_block;
  const uint4* frag;    // hop-path A operands
  const uint4* frag_id; // identity-path A operands
  const uint32_t* links;  // link records of the output parity: [8][stride][20]
  const void* in;       // hop input (the other parity)
  const void* in_id;    // identity input (the output parity)
  void* out;
};

__device__ __forceinline__ uint32_t link_off(const Geom& g, int dh) {
  return (uint32_t)dh * (uint32_t)g.stride * (uint32_t)(kLinkWords * 4);
}

// ------------------------------------------------------------------ links
// Transport matrices of the output sites of parity blockIdx.y, path
// blockIdx.z: M = U_mu(x) (dh = 2 mu) or U_mu(x - mu)^dag (dh = 2 mu + 1),
// from the FP64 gauge field with its boundary phase.
__global__ void __launch_bounds__(kBlock)
k_bnet_links(Geom g, const double2* __restrict__ u, uint32_t* __restrict__ out) {
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= g.Vh) return;
  const int par = blockIdx.y, dh = blockIdx.z, mu = dh >> 1;
  const bool fw = (dh & 1) == 0;
  double2 M[3][3];
  if (fw) {
#pragma unroll
    for (int k = 0; k < 9; ++k)
      M[k / 3][k % 3] = u[((size_t)(par * 4 + mu) * 9 + k) * g.stride + hs];
  } else {
    const SiteCoord c = decode_site(g, hs, par);
    const int nb = neighbour(g, c, hs, mu, false);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const double2 v = u[((size_t)((1 - par) * 4 + mu) * 9 + 3 * a + b) * g.stride + nb];
        M[b][a] = make_double2(v.x, -v.y);
      }
  }
  uint32_t* o = out + ((size_t)(par * kHop + dh) * g.stride + hs) * kLinkWords;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const double2 m = M[k / 3][k % 3];
    o[k] = pack_b2((float)m.x, (float)-m.y);
    o[9 + k] = pack_b2((float)m.y, (float)m.x);
  }
  o[18] = 0u;
  o[19] = 0u;
}

// --------------------------------------------------------------- stage 1
// odd sites: [L1(odd) | L2(odd)] (16 channels) from the 8 transports of r.
// The B operand's channel index is the path: lane group q of k-block kb
// transports r along path 4 kb + q.
__global__ void __launch_bounds__(kBThreads) k_bnet_first(BArgs a) {
  constexpr int MB = fr_mb(kOddCh), KB = 2, NF = MB * KB;
  __shared__ uint4 sA[NF * 64];
  for (int i = threadIdx.x; i < NF * 64; i += kBThreads) sA[i] = a.frag[i];
  __syncthreads();
  const Geom& g = a.g;
  const int lane = threadIdx.x & 63, q = lane >> 4, n = lane & 15, wave = threadIdx.x >> 6;
  const int t0 = blockIdx.x * a.tiles_per_block, t1 = min(t0 + a.tiles_per_block, a.tiles);
  for (int tile = t0; tile < t1; ++tile) {
    const int site = tile * 16 + n;
    const bool valid = site < g.Vh;
    const int s = valid ? site : g.Vh - 1;
    const SiteCoord sc = decode_site(g, s, 1);
    int nbk[KB];
    U32x4 L[KB][5];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const int dh = 4 * kb + q;
      nbk[kb] = neighbour(g, sc, s, dh >> 1, (dh & 1) == 0);
#pragma unroll
      for (int k = 0; k < 5; ++k)
        L[kb][k] = ld128(a.links, 0u, link_off(g, dh) + (uint32_t)s * 80u + 16u * k);
    }
    for (int rr = wave; rr < a.nrhs; rr += kBThreads / 64) {
      const float2* r = (const float2*)a.in + (size_t)rr * 12 * g.stride;
      bf16x8 B[KB][3];
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        U32x4 v[3];
        load_r_b2(r, g.stride, nbk[kb], v);
        transport(L[kb], v, B[kb]);
      }
      f32x4 acc[3][MB];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) acc[c][mb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
          const bf16x8 A = as_frag(*reinterpret_cast<const U32x4*>(&sA[(mb * KB + kb) * 64 + lane]));
#pragma unroll
          for (int c = 0; c < 3; ++c)
            acc[c][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[kb][c], acc[c][mb], 0, 0, 0);
        }
      if (valid) {
        // C row 4q + j of block mb: channel 2 mb + q/2, spin 2 (q & 1) + j/2, re/im j & 1
        char* o = (char*)a.out + (size_t)rr * g.stride * (kOddCh * 48);
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
          const int ch = 2 * mb + (q >> 1);
          const uint32_t vo = ((uint32_t)s * kOddCh + ch) * 48u + 8u * (q & 1);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            U32x2 w;
            w[0] = pack_b2(acc[c][mb][0], acc[c][mb][1]);
            w[1] = pack_b2(acc[c][mb][2], acc[c][mb][3]);
            st64(o, 16u * c, vo, w);
          }
        }
      }
    }
  }
}

// --------------------------------------------------------------- stages 2, 3
// ID = 0: even sites, L2(even) = (W2_0 W1_0) r + sum_{d>0} W2_d T_d L1(odd)
//         (hop input: odd records, channels 0..7; identity input: r, FP32)
// ID = 1: odd sites, L3(odd) = W3_0 L2(odd) + sum_{d>0} W3_d T_d L2(even)
//         (hop input: even records; identity input and output: odd records,
//         reading channels 8..15 and writing channels 0..7 of the same site)
template <int ID>
__global__ void __launch_bounds__(kBThreads) k_bnet_mid(BArgs a) {
  constexpr int MB = 4, KB = 2;
  constexpr int KBI = ID == 0 ? 1 : 2;
  constexpr int NH = kHop * MB * KB, NI = MB * KBI;
  constexpr int INREC = ID == 0 ? kOddCh : kEvenCh;   // hop input record
  constexpr int OUTREC = ID == 0 ? kEvenCh : kOddCh;  // output record
  constexpr int PAR = ID == 0 ? 0 : 1;                // output parity
  __shared__ uint4 sA[(NH + NI) * 64];
  for (int i = threadIdx.x; i < NH * 64; i += kBThreads) sA[i] = a.frag[i];
  for (int i = threadIdx.x; i < NI * 64; i += kBThreads) sA[NH * 64 + i] = a.frag_id[i];
  __syncthreads();
  const Geom& g = a.g;
  const int lane = threadIdx.x & 63, q = lane >> 4, n = lane & 15, wave = threadIdx.x >> 6;
  const int t0 = blockIdx.x * a.tiles_per_block, t1 = min(t0 + a.tiles_per_block, a.tiles);
  for (int tile = t0; tile < t1; ++tile) {
    const int site = tile * 16 + n;
    const bool valid = site < g.Vh;
    const int s = valid ? site : g.Vh - 1;
    const SiteCoord sc = decode_site(g, s, PAR);
    for (int rr = wave; rr < a.nrhs; rr += kBThreads / 64) {
      f32x4 acc[3][MB];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) acc[c][mb] = f32x4{0.f, 0.f, 0.f, 0.f};
      // identity path
      {
        bf16x8 B[KBI][3];
        if constexpr (ID == 0) {
          const float2* r = (const float2*)a.in_id + (size_t)rr * 12 * g.stride;
          U32x4 v[3];
          load_r_b2(r, g.stride, s, v);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            if (q != 0) v[c] = U32x4{0u, 0u, 0u, 0u};
            B[0][c] = as_frag(v[c]);
          }
        } else {
          const char* in = (const char*)a.in_id + (size_t)rr * g.stride * (kOddCh * 48);
#pragma unroll
          for (int kb = 0; kb < KBI; ++kb) {
            const uint32_t vo = ((uint32_t)s * kOddCh + 8 + 4 * kb + q) * 48u;
#pragma unroll
            for (int c = 0; c < 3; ++c) B[kb][c] = as_frag(ld128(in, 16u * c, vo));
          }
        }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
          for (int kb = 0; kb < KBI; ++kb) {
            const bf16x8 A =
                as_frag(*reinterpret_cast<const U32x4*>(&sA[(NH + mb * KBI + kb) * 64 + lane]));
#pragma unroll
            for (int c = 0; c < 3; ++c)
              acc[c][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[kb][c], acc[c][mb], 0, 0, 0);
          }
      }
      // the 8 transported paths
      const char* in = (const char*)a.in + (size_t)rr * g.stride * (INREC * 48);
#pragma unroll 1
      for (int dh = 0; dh < kHop; ++dh) {
        const int nb = neighbour(g, sc, s, dh >> 1, (dh & 1) == 0);
        U32x4 L[5];
#pragma unroll
        for (int k = 0; k < 5; ++k)
          L[k] = ld128(a.links, 0u, link_off(g, dh) + (uint32_t)s * 80u + 16u * k);
        bf16x8 B[KB][3];
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
          const uint32_t vo = ((uint32_t)nb * INREC + 4 * kb + q) * 48u;
          U32x4 v[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) v[c] = ld128(in, 16u * c, vo);
          transport(L, v, B[kb]);
        }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
#pragma unroll
          for (int kb = 0; kb < KB; ++kb) {
            const bf16x8 A =
                as_frag(*reinterpret_cast<const U32x4*>(&sA[((dh * MB + mb) * KB + kb) * 64 + lane]));
#pragma unroll
            for (int c = 0; c < 3; ++c)
              acc[c][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[kb][c], acc[c][mb], 0, 0, 0);
          }
      }
      if (valid) {
        char* o = (char*)a.out + (size_t)rr * g.stride * (OUTREC * 48);
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
          const int ch = 2 * mb + (q >> 1);
          const uint32_t vo = ((uint32_t)s * OUTREC + ch) * 48u + 8u * (q & 1);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            U32x2 w;
            w[0] = pack_b2(acc[c][mb][0], acc[c][mb][1]);
            w[1] = pack_b2(acc[c][mb][2], acc[c][mb][3]);
            st64(o, 16u * c, vo, w);
          }
        }
      }
    }
  }
}

// --------------------------------------------------------------- stage 4
// even sites: z = (W4_0 W3_0) L2(even) + sum_{d>0} T_d [X_d (L3 | L2)(odd)],
// X_d = [W4_d | W4_0 W3_d] (16 -> 1 channel).  The mixing runs at the
// neighbour on the matrix cores (fresh accumulators per path), then lanes
// of groups 0 and 1 (C rows: spins 2q, 2q + 1, re/im) transport their two
// spins in FP32.  Output: FP32 half fields [rhs][12][stride].
__global__ void __launch_bounds__(kBThreads) k_bnet_last(BArgs a) {
  constexpr int KB = 4, KBI = 2;
  constexpr int NH = kHop * KB, NI = KBI;
  __shared__ uint4 sA[(NH + NI) * 64];
  for (int i = threadIdx.x; i < NH * 64; i += kBThreads) sA[i] = a.frag[i];
  for (int i = threadIdx.x; i < NI * 64; i += kBThreads) sA[NH * 64 + i] = a.frag_id[i];
  __syncthreads();
  const Geom& g = a.g;
  const int lane = threadIdx.x & 63, q = lane >> 4, n = lane & 15, wave = threadIdx.x >> 6;
  const int t0 = blockIdx.x * a.tiles_per_block, t1 = min(t0 + a.tiles_per_block, a.tiles);
  for (int tile = t0; tile < t1; ++tile) {
    const int site = tile * 16 + n;
    const bool valid = site < g.Vh;
    const int s = valid ? site : g.Vh - 1;
    const SiteCoord sc = decode_site(g, s, 0);
    for (int rr = wave; rr < a.nrhs; rr += kBThreads / 64) {
      float2 z[2][3];
      // identity path: (W4_0 W3_0) L2(even)(x)
      {
        const char* in = (const char*)a.in_id + (size_t)rr * g.stride * (kEvenCh * 48);
        f32x4 C[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) C[c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KBI; ++kb) {
          const bf16x8 A = as_frag(*reinterpret_cast<const U32x4*>(&sA[(NH + kb) * 64 + lane]));
          const uint32_t vo = ((uint32_t)s * kEvenCh + 4 * kb + q) * 48u;
#pragma unroll
          for (int c = 0; c < 3; ++c)
            C[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, as_frag(ld128(in, 16u * c, vo)), C[c],
                                                           0, 0, 0);
        }
#pragma unroll
        for (int sp = 0; sp < 2; ++sp)
#pragma unroll
          for (int c = 0; c < 3; ++c) z[sp][c] = make_float2(C[c][2 * sp], C[c][2 * sp + 1]);
      }
      const char* in = (const char*)a.in + (size_t)rr * g.stride * (kOddCh * 48);
#pragma unroll 1
      for (int dh = 0; dh < kHop; ++dh) {
        const int nb = neighbour(g, sc, s, dh >> 1, (dh & 1) == 0);
        f32x4 C[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) C[c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
          const bf16x8 A = as_frag(*reinterpret_cast<const U32x4*>(&sA[(dh * KB + kb) * 64 + lane]));
          const uint32_t vo = ((uint32_t)nb * kOddCh + 4 * kb + q) * 48u;
#pragma unroll
          for (int c = 0; c < 3; ++c)
            C[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, as_frag(ld128(in, 16u * c, vo)), C[c],
                                                           0, 0, 0);
        }
        if (q < 2) {
          U32x4 L[5];
#pragma unroll
          for (int k = 0; k < 5; ++k)
            L[k] = ld128(a.links, 0u, link_off(g, dh) + (uint32_t)s * 80u + 16u * k);
          uint32_t W[20];
#pragma unroll
          for (int k = 0; k < 5; ++k)
#pragma unroll
            for (int e = 0; e < 4; ++e) W[4 * k + e] = L[k][e];
#pragma unroll
          for (int ca = 0; ca < 3; ++ca)
#pragma unroll
            for (int cb = 0; cb < 3; ++cb) {
              // M_ab = (Re, Im) from P = (Re, -Im)
              const float mr = lo_f(W[3 * ca + cb]), mi = -hi_f(W[3 * ca + cb]);
#pragma unroll
              for (int sp = 0; sp < 2; ++sp) {
                const float vr = C[cb][2 * sp], vi = C[cb][2 * sp + 1];
                z[sp][ca].x = fmaf(mr, vr, z[sp][ca].x);
                z[sp][ca].x = fmaf(-mi, vi, z[sp][ca].x);
                z[sp][ca].y = fmaf(mr, vi, z[sp][ca].y);
                z[sp][ca].y = fmaf(mi, vr, z[sp][ca].y);
              }
            }
        }
      }
      if (valid && q < 2) {
        float2* o = (float2*)a.out + (size_t)rr * 12 * g.stride;
#pragma unroll
        for (int sp = 0; sp < 2; ++sp)
#pragma unroll
          for (int c = 0; c < 3; ++c)
            st_at(o, (uint32_t)((3 * (2 * q + sp) + c) * g.stride), (uint32_t)s, z[sp][c]);
      }
    }
  }
}

// ------------------------------------------------------------- host side
using Cd = std::complex<double>;

// A channel-spin mixing matrix X[path][o][i][s][t] (out channel o, in
// channel i, out spin s, in spin t)
struct Mix {
  int np = 1, cout = 1, cin = 1;
  std::vector<Cd> x;
  Mix(int np_, int cout_, int cin_) : np(np_), cout(cout_), cin(cin_),
                                     x((size_t)np_ * cout_ * cin_ * 16) {}
  Cd& at(int p, int o, int i, int s, int t) {
    return x[(((size_t)(p * cout + o) * cin + i) * 4 + s) * 4 + t];
  }
};

// spin matrix of the trained weights W_l[d][o][i]
struct Spin {
  Cd m[4][4];
};
Spin wmat(const std::vector<float2>& w, int cout, int cin, int d, int o, int i) {
  Spin r;
  for (int s = 0; s < 4; ++s)
    for (int t = 0; t < 4; ++t) {
      const float2 v = w[((size_t)(d * cout + o) * cin + i) * 16 + 4 * s + t];
      r.m[s][t] = Cd(v.x, v.y);
    }
  return r;
}
Spin mul(const Spin& a, const Spin& b) {
  Spin r;
  for (int s = 0; s < 4; ++s)
    for (int t = 0; t < 4; ++t) {
      Cd acc = 0.0;
      for (int u = 0; u < 4; ++u) acc += a.m[s][u] * b.m[u][t];
      r.m[s][t] = acc;
    }
  return r;
}
void add_to(Spin& a, const Spin& b) {
  for (int s = 0; s < 4; ++s)
    for (int t = 0; t < 4; ++t) a.m[s][t] += b.m[s][t];
}
void put(Mix& X, int p, int o, int i, const Spin& m) {
  for (int s = 0; s < 4; ++s)
    for (int t = 0; t < 4; ++t) X.at(p, o, i, s, t) = m.m[s][t];
}

uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);                     // round to nearest even
  return (uint16_t)(u >> 16);
}

// A operands of v_mfma_f32_16x16x32_bf16 for X: entry ((p MB + mb) KB + kb)
// 64 + lane holds A[16 mb + lane % 16][32 kb + 8 (lane / 16) + e], e = 0..7,
// A the real embedding: row 8 o + 2 s + ro, column 8 i + 2 t + ri.
void append_frags(const Mix& X, std::vector<uint4>& out) {
  const int MB = fr_mb(X.cout), KB = fr_kb(X.cin);
  for (int p = 0; p < X.np; ++p)
    for (int mb = 0; mb < MB; ++mb)
      for (int kb = 0; kb < KB; ++kb)
        for (int lane = 0; lane < 64; ++lane) {
          uint16_t h[8];
          const int row = 16 * mb + (lane & 15);
          for (int e = 0; e < 8; ++e) {
            const int k = 32 * kb + 8 * (lane >> 4) + e;
            double v = 0.0;
            if (row < 8 * X.cout && k < 8 * X.cin) {
              const int o = row >> 3, s = (row & 7) >> 1, ro = row & 1;
              const int i = k >> 3, t = (k & 7) >> 1, ri = k & 1;
              const Cd c = const_cast<Mix&>(X).at(p, o, i, s, t);
              v = ro == 0 ? (ri == 0 ? c.real() : -c.imag()) : (ri == 0 ? c.imag() : c.real());
            }
            h[e] = bf16_bits((float)v);
          }
          uint4 w;
          std::memcpy(&w, h, 16);
          out.push_back(w);
        }
}

struct BNet {
  int version = -1;            // weight version the fragments were built from
  uint4* frag = nullptr;
  size_t off[7] = {};          // first, mid0, mid0_id, mid1, mid1_id, last, last_id
  uint32_t* odd = nullptr;     // [rhs][stride][16][12] bf16 pairs
  uint32_t* even = nullptr;    // [rhs][stride][8][12]
  int act_rhs = 0;
};

BNet* bnet_of(ii_context* ctx) { return static_cast<BNet*>(ctx->bnet); }

ii_status bnet_get(ii_context* ctx, BNet** out) {
  if (!ctx->bnet) ctx->bnet = new BNet();
  *out = bnet_of(ctx);
  return II_OK;
}

// the bfloat16 transport matrices of the current gauge field
ii_status bnet_links(ii_context* ctx) {
  void*& L = ctx->gauge_lp[ii_context::kGaugePairs];
  if (L) return II_OK;
  const Geom& g = ctx->g;
  II_HIP(ctx, hipMalloc(&L, (size_t)2 * kHop * g.stride * kLinkWords * sizeof(uint32_t)));
  hipLaunchKernelGGL(k_bnet_links, dim3(blocks_for(g.Vh), 2, kHop), dim3(kBlock), 0, ctx->stream,
                     g, (const double2*)ctx->gauge64, (uint32_t*)L);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

// composite A operands of the four stages from the trained weights
ii_status bnet_frags(ii_context* ctx, BNet* B) {
  int ver = 0;
  II_CHECK(net_host_weights(ctx, &ver, nullptr));
  if (ver == B->version && B->frag) return II_OK;
  std::vector<std::vector<float2>> w;
  II_CHECK(net_host_weights(ctx, &ver, &w));
  auto W1 = [&](int d, int o) { return wmat(w[0], 8, 1, d, o, 0); };
  auto W2 = [&](int d, int o, int i) { return wmat(w[1], 8, 8, d, o, i); };
  auto W3 = [&](int d, int o, int i) { return wmat(w[2], 8, 8, d, o, i); };
  auto W4 = [&](int d, int i) { return wmat(w[3], 1, 8, d, 0, i); };
  std::vector<uint4> f;
  // stage 1: [L1(odd) | L2(odd)] from T_d r, d = dh + 1
  {
    Mix X(1, kOddCh, kHop);
    for (int dh = 0; dh < kHop; ++dh)
      for (int o = 0; o < 8; ++o) {
        put(X, 0, o, dh, W1(dh + 1, o));
        Spin e{};
        for (int i = 0; i < 8; ++i) {
          add_to(e, mul(W2(0, o, i), W1(dh + 1, i)));
          add_to(e, mul(W2(dh + 1, o, i), W1(0, i)));
        }
        put(X, 0, 8 + o, dh, e);
      }
    B->off[0] = f.size();
    append_frags(X, f);
  }
  // stage 2: W2_d on L1(odd); identity W2_0 W1_0 on r
  {
    Mix X(kHop, 8, 8), I(1, 8, 1);
    for (int dh = 0; dh < kHop; ++dh)
      for (int o = 0; o < 8; ++o)
        for (int i = 0; i < 8; ++i) put(X, dh, o, i, W2(dh + 1, o, i));
    for (int o = 0; o < 8; ++o) {
      Spin e{};
      for (int i = 0; i < 8; ++i) add_to(e, mul(W2(0, o, i), W1(0, i)));
      put(I, 0, o, 0, e);
    }
    B->off[1] = f.size();
    append_frags(X, f);
    B->off[2] = f.size();
    append_frags(I, f);
  }
  // stage 3: W3_d on L2(even); identity W3_0 on L2(odd)
  {
    Mix X(kHop, 8, 8), I(1, 8, 8);
    for (int dh = 0; dh < kHop; ++dh)
      for (int o = 0; o < 8; ++o)
        for (int i = 0; i < 8; ++i) put(X, dh, o, i, W3(dh + 1, o, i));
    for (int o = 0; o < 8; ++o)
      for (int i = 0; i < 8; ++i) put(I, 0, o, i, W3(0, o, i));
    B->off[3] = f.size();
    append_frags(X, f);
    B->off[4] = f.size();
    append_frags(I, f);
  }
  // stage 4: [W4_d | W4_0 W3_d] on (L3 | L2)(odd) at the neighbour; identity
  // W4_0 W3_0 on L2(even)
  {
    Mix X(kHop, 1, kOddCh), I(1, 1, 8);
    for (int dh = 0; dh < kHop; ++dh)
      for (int j = 0; j < 8; ++j) {
        put(X, dh, 0, j, W4(dh + 1, j));
        Spin e{};
        for (int k = 0; k < 8; ++k) add_to(e, mul(W4(0, k), W3(dh + 1, k, j)));
        put(X, dh, 0, 8 + j, e);
      }
    for (int i = 0; i < 8; ++i) {
      Spin e{};
      for (int k = 0; k < 8; ++k) add_to(e, mul(W4(0, k), W3(0, k, i)));
      put(I, 0, 0, i, e);
    }
    B->off[5] = f.size();
    append_frags(X, f);
    B->off[6] = f.size();
    append_frags(I, f);
  }
  if (!B->frag) II_HIP(ctx, hipMalloc(&B->frag, f.size() * sizeof(uint4)));
  II_HIP(ctx, hipMemcpy(B->frag, f.data(), f.size() * sizeof(uint4), hipMemcpyHostToDevice));
  B->version = ver;
  return II_OK;
}

ii_status bnet_acts(ii_context* ctx, BNet* B, int nrhs) {
  if (B->act_rhs >= nrhs) return II_OK;
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (B->odd) (void)hipFree(B->odd);
  if (B->even) (void)hipFree(B->even);
  B->odd = B->even = nullptr;
  B->act_rhs = 0;
  const size_t st = ctx->g.stride;
  II_HIP(ctx, hipMalloc(&B->odd, (size_t)nrhs * st * kOddCh * 48));
  II_HIP(ctx, hipMalloc(&B->even, (size_t)nrhs * st * kEvenCh * 48));
  B->act_rhs = nrhs;
  return II_OK;
}

// blocks of a stage: 2 per CU (the A operands of stages 2 and 3 take ~70 KB
// of LDS), each a contiguous run of tiles, so that the blocks a time slice
// apart (8 blocks, the same XCD under round-robin dispatch) read the same
// planes at about the same time
void bnet_grid(const Geom& g, BArgs& a, int* nblk) {
  a.tiles = (g.Vh + 15) / 16;
  int nb = std::min(512, a.tiles);
  a.tiles_per_block = (a.tiles + nb - 1) / nb;
  *nblk = (a.tiles + a.tiles_per_block - 1) / a.tiles_per_block;
}

}  // namespace

// links, weights and activation buffers for nrhs rhs: everything
// bnet_apply needs to launch without allocating (so that it can be captured)
ii_status bnet_prepare(ii_context* ctx, int nrhs) {
  BNet* B = nullptr;
  II_CHECK(bnet_get(ctx, &B));
  II_CHECK(bnet_links(ctx));
  II_CHECK(bnet_frags(ctx, B));
  return bnet_acts(ctx, B, nrhs);
}

// launches stage `stage` (0..3) of z = M r
ii_status bnet_stage(ii_context* ctx, int stage, int nrhs, const float2* r, float2* z) {
  BNet* B = bnet_of(ctx);
  const Geom& g = ctx->g;
  const uint32_t* links = (const uint32_t*)ctx->gauge_lp[ii_context::kGaugePairs];
  const size_t lpar = (size_t)kHop * g.stride * kLinkWords;
  BArgs a{};
  a.g = g;
  a.nrhs = nrhs;
  int nblk = 0;
  bnet_grid(g, a, &nblk);
  switch (stage) {
    case 0:  // odd: [L1 | L2](odd) from the transports of r
      a.frag = B->frag + B->off[0];
      a.links = links + lpar;
      a.in = r;
      a.out = B->odd;
      hipLaunchKernelGGL(k_bnet_first, dim3(nblk), dim3(kBThreads), 0, ctx->stream, a);
      break;
    case 1:  // even: L2(even)
      a.frag = B->frag + B->off[1];
      a.frag_id = B->frag + B->off[2];
      a.links = links;
      a.in = B->odd;
      a.in_id = r;
      a.out = B->even;
      hipLaunchKernelGGL(k_bnet_mid<0>, dim3(nblk), dim3(kBThreads), 0, ctx->stream, a);
      break;
    case 2:  // odd: L3 over L1's channels of the odd records
      a.frag = B->frag + B->off[3];
      a.frag_id = B->frag + B->off[4];
      a.links = links + lpar;
      a.in = B->even;
      a.in_id = B->odd;
      a.out = B->odd;
      hipLaunchKernelGGL(k_bnet_mid<1>, dim3(nblk), dim3(kBThreads), 0, ctx->stream, a);
      break;
    default:  // even: z
      a.frag = B->frag + B->off[5];
      a.frag_id = B->frag + B->off[6];
      a.links = links;
      a.in = B->odd;
      a.in_id = B->even;
      a.out = z;
      hipLaunchKernelGGL(k_bnet_last, dim3(nblk), dim3(kBThreads), 0, ctx->stream, a);
      break;
  }
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

// z = M r for nrhs FP32 even half fields [rhs][12][stride] on the device
ii_status bnet_apply(ii_context* ctx, int nrhs, const float2* r, float2* z) {
  II_CHECK(bnet_prepare(ctx, nrhs));
  for (int st = 0; st < 4; ++st) II_CHECK(bnet_stage(ctx, st, nrhs, r, z));
  return II_OK;
}

void bnet_release(ii_context* ctx) {
  BNet* B = bnet_of(ctx);
  if (!B) return;
  if (B->frag) (void)hipFree(B->frag);
  if (B->odd) (void)hipFree(B->odd);
  if (B->even) (void)hipFree(B->even);
  delete B;
  ctx->bnet = nullptr;
}

}  // namespace ii

using namespace ii;

extern "C" {

ii_status ii_net_apply_batched(ii_context* ctx, int nrhs, const double* r_even, double* z_even) {
  if (!ctx || !r_even || !z_even || nrhs < 1) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  const size_t nh = (size_t)24 * g.Vh, he = half_elems(g);
  void *p64 = nullptr, *p32 = nullptr, *pz = nullptr;
  II_CHECK(workspace(ctx, 80, nrhs * he * sizeof(double2), &p64));
  II_CHECK(workspace(ctx, 81, nrhs * he * sizeof(float2), &p32));
  II_CHECK(workspace(ctx, 82, nrhs * he * sizeof(float2), &pz));
  double2* f64 = (double2*)p64;
  for (int k = 0; k < nrhs; ++k) {
    double* d = nullptr;
    II_CHECK(stage_in(ctx, r_even + k * nh, nh, &d));
    II_CHECK(pack_half_spinor(ctx, d, f64 + k * he));
  }
  II_CHECK(convert_64_to_32(ctx, f64, (float2*)p32, nrhs));
  II_CHECK(bnet_apply(ctx, nrhs, (const float2*)p32, (float2*)pz));
  II_CHECK(convert_32_to_64(ctx, (const float2*)pz, f64, nrhs));
  for (int k = 0; k < nrhs; ++k) {
    double* d = nullptr;
    II_CHECK(stage_buf(ctx, nh, &d));
    II_CHECK(unpack_half_spinor(ctx, f64 + k * he, d));
    II_CHECK(stage_out(ctx, z_even + k * nh, d, nh));
  }
  return II_OK;
}

ii_status ii_bench_net_batched(ii_context* ctx, int nrhs, int nrep, double* stage_ms) {
  if (!ctx || !stage_ms || nrhs < 1 || nrep < 1) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const size_t he = half_elems(ctx->g);
  void *pr = nullptr, *pz = nullptr;
  II_CHECK(workspace(ctx, 81, nrhs * he * sizeof(float2), &pr));
  II_CHECK(workspace(ctx, 82, nrhs * he * sizeof(float2), &pz));
  // a defined input: zeros (the stage times do not depend on the values)
  II_HIP(ctx, hipMemsetAsync(pr, 0, nrhs * he * sizeof(float2), ctx->stream));
  II_CHECK(bnet_apply(ctx, nrhs, (const float2*)pr, (float2*)pz));
  for (int st = 0; st < 4; ++st) {
    II_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int i = 0; i < nrep; ++i)
      II_CHECK(bnet_stage(ctx, st, nrhs, (const float2*)pr, (float2*)pz));
    II_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    II_HIP(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    II_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    stage_ms[st] = ms / nrep;
  }
  return II_OK;
}

}  // extern "C"
