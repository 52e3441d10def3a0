// Learned gauge-covariant preconditioner for the even/odd Wilson solve,
// trained and applied on gfx950.  It has no counterpart in the reference,
// whose solvers (lib/Solver/solver_CG.cpp, multiShiftSolver_CG.cpp) are
// unpreconditioned.
//
// Network: 4 linear layers, widths 1 -> 8 -> 8 -> 8 -> 1 spinor channels on
// the full lattice.  Layer: out[o](x) = sum_d sum_i W[d][o][i] (T_d in[i])(x)
// over the 9 paths d (identity, U_mu(x) psi(x+mu), U_mu(x-mu)^dag psi(x-mu)),
// W a complex 4x4 spin matrix.  Field layout per channel is the full-field
// layout of the library, [parity][12][stride] float2, channels back to back.
//
// Layer kernel (forward and adjoint): one wavefront owns 16 sites of one
// parity.  Lane (n, q) = (lane & 15, lane >> 4) transports channels q and
// q + 4 of the path's neighbour of site n in FP32 (r12 FP32 links) and
// rounds them to bfloat16: these are the B operands of
// v_mfma_f32_16x16x32_bf16 (k = 8q + 2*spin + re/im of one channel, n =
// site), one per colour.  The A operands are the real embedding
// [[Re W, -Im W], [Im W, Re W]] of the channel-spin mixing matrix of the
// path, pre-arranged per lane (8 bf16 = 16 bytes) and staged in LDS once per
// block.  C rows are (out channel, spin, re/im), accumulated in FP32 over
// the 9 paths.  The adjoint layer is the same kernel with the paths
// mirrored and W -> W^dag (channel-transposed): T_{+mu}^dag = T_{-mu}.
//
// Weight gradient: VALU, gw[d][o][i] = sum_{x,c} g[o](x) conj(T_d in[i](x)),
// per-block partials in FP32 then a fixed-order FP64 sum.
#include "ii_ops.h"
#include "ii_stencil.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>

namespace ii {

namespace {

constexpr int kNL = 4;
constexpr int kWidth[kNL + 1] = {1, 8, 8, 8, 1};
constexpr int kPaths = 9;
constexpr int kLayerThreads = 256;
constexpr int kWgThreads = 256;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// MFMA tiling of a layer: 16 C rows per block, 32 k per step
__host__ __device__ constexpr int frag_mb(int cout) { return (8 * cout + 15) / 16; }
__host__ __device__ constexpr int frag_kb(int cin) { return (8 * cin + 31) / 32; }
__host__ __device__ constexpr int frag_count(int cin, int cout) {
  return kPaths * frag_mb(cout) * frag_kb(cin) * 64;
}
__host__ __device__ constexpr int mirror_path(int d) {
  return d == 0 ? 0 : (((d - 1) & 1) == 0 ? d + 1 : d - 1);
}

struct LayerParams {
  int cin = 0, cout = 0;
  size_t nw = 0;              // complex weights: 9*cout*cin*16
  float2* w = nullptr;        // FP32 master weights
  float2* gacc = nullptr;     // gradient accumulator
  float2* m = nullptr;        // Adam first moment
  float2* v = nullptr;        // Adam second moment
  uint4* frag_fw = nullptr;   // bf16 A operands, forward
  uint4* frag_adj = nullptr;  // bf16 A operands, adjoint
};

struct Net {
  LayerParams L[kNL];
  bool init = false;
  int version = 0;             // bumped whenever the weights change
  int adam_step = 0;
  // activations / gradients, FP32 full fields [ch][2][12][stride]
  float2* act[kNL + 1] = {};   // act[0] 1 ch, act[1..3] 8 ch, act[4] 1 ch
  float2* grad[2] = {};        // 8-ch backward ping-pong
  float2* gout = nullptr;      // 1-ch gradient at the output
  float2* partial = nullptr;   // weight-gradient partials
  size_t partial_len = 0;      // float2 entries
  double* dpart = nullptr;     // loss partials
  double* dhost = nullptr;     // pinned
  // host staging for the test-facing calls
  double* stage = nullptr;
  size_t stage_len = 0;
  // training pairs
  float2* pb = nullptr;        // [n][2][12][stride]: source on even sites, odd zero
  float2* py = nullptr;        // [n][12][stride]: solution on even sites
  std::vector<double> inv_ny;  // 1/|y_k|^2
  int npairs = 0;
  // FGCR workspace (FP64 half fields)
  double2* kry = nullptr;      // [2*nk][12][stride]: z_j then q_j
  int kry_n = 0;
  double2* fw[6] = {};         // x, r, b, t (odd temp), w, spare
};

Net* net_of(ii_context* ctx) { return static_cast<Net*>(ctx->learned); }

size_t chan_elems(const Geom& g) { return (size_t)24 * g.stride; }  // float2 per channel

template <typename T>
void dfree(T* p) {
  if (p) (void)hipFree((void*)p);
}

template <typename T>
ii_status dev_alloc(ii_context* ctx, T** p, size_t n) {
  if (*p) return II_OK;
  II_HIP(ctx, hipMalloc((void**)p, n * sizeof(T)));
  return II_OK;
}

__device__ __forceinline__ uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// ------------------------------------------------------------------ kernels
// A operands of the layer (adjoint = 0) or of its adjoint.  Entry
// ((d*MB + mb)*KB + kb)*64 + lane holds A[16mb + lane%16][32kb + 8(lane/16) + e],
// e = 0..7, A = real embedding of the path-d mixing matrix.
__global__ void k_net_frags(const float2* __restrict__ w, int cin, int cout, int adjoint,
                            uint4* __restrict__ frag) {
  const int fin = adjoint ? cout : cin, fout = adjoint ? cin : cout;
  const int MB = frag_mb(fout), KB = frag_kb(fin);
  const int total = kPaths * MB * KB * 64;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= total) return;
  const int lane = id & 63;
  int rest = id >> 6;
  const int kb = rest % KB;
  rest /= KB;
  const int mb = rest % MB;
  const int d = rest / MB;
  const int row = 16 * mb + (lane & 15);
  __bf16 vals[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int k = 32 * kb + 8 * (lane >> 4) + e;
    float val = 0.0f;
    if (row < 8 * fout && k < 8 * fin) {
      const int o = row >> 3, s = (row & 7) >> 1, ro = row & 1;
      const int i = k >> 3, t = (k & 7) >> 1, ri = k & 1;
      float2 c;
      if (!adjoint) {
        c = w[((size_t)(d * cout + o) * cin + i) * 16 + 4 * s + t];
      } else {
        // W'_d[o][i] = W_{mirror d}[i][o]^dag: spin-transposed, conjugated
        const float2 x = w[((size_t)(mirror_path(d) * cout + i) * cin + o) * 16 + 4 * t + s];
        c = make_float2(x.x, -x.y);
      }
      val = (ro == 0) ? (ri == 0 ? c.x : -c.y) : (ri == 0 ? c.y : c.x);
    }
    vals[e] = (__bf16)val;
  }
  uint4 r;
  __builtin_memcpy(&r, vals, 16);
  frag[id] = r;
}

struct LayerArgs {
  Geom g;
  const float2* in;    // [CIN][2][12][stride]
  float2* out;         // [COUT][2][12][stride]
  const uint4* frag;
  const float2* gauge; // r12 FP32, [2][4][6][stride]
  int flip;            // antiperiodic t-links carry -1
  int par0, npar;      // output parities
  int tiles_per_block;
};

// Transported 4-spinor of a channel field at one site: w = M v with M = 1,
// U or U^dag.  The buffer base is the whole activation array (uniform), the
// scalar offset `soff` must be wave-uniform (component blocks), and the
// per-lane element offset `loff` carries the channel and the site, so no
// load needs a per-lane descriptor or scalar offset.
__device__ __forceinline__ void net_transport(const float2* __restrict__ base, uint32_t soff,
                                              int stride, uint32_t loff, bool has_link, bool dag,
                                              const float2 U[3][3], float2 w[4][3]) {
  float2 v[4][3];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int c = 0; c < 3; ++c) v[t][c] = ld_at(base, soff + (uint32_t)((3 * t + c) * stride), loff);
  if (!has_link) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int c = 0; c < 3; ++c) w[t][c] = v[t][c];
    return;
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        if (dag) cmacc(acc, U[b][a], v[t][b]);
        else cmac(acc, U[a][b], v[t][b]);
      }
      w[t][a] = acc;
    }
}

// r12 FP32 link with every offset in the lane offset (for kernels whose
// threads may hold sites of different parities)
__device__ __forceinline__ void ld_link_lane(const float2* __restrict__ gp, int stride, int par,
                                             int mu, int site, bool flip, float2 U[3][3]) {
  const uint32_t loff = (uint32_t)((par * 4 + mu) * 6 * stride + site);
#pragma unroll
  for (int k = 0; k < 6; ++k) U[k / 3][k % 3] = ld_at(gp, (uint32_t)(k * stride), loff);
  r12_third_row(U, flip);
}

// Path d of half site s (parity par): neighbour half index (in parity
// par ^ (d > 0)) and link.
__device__ __forceinline__ int net_path(const Geom& g, const GaugeF32r12& gauge, int flip, int d,
                                        int par, int s, const SiteCoord& sc, float2 U[3][3],
                                        bool& dag) {
  dag = false;
  if (d == 0) return s;
  const int mu = (d - 1) >> 1;
  const bool fw = ((d - 1) & 1) == 0;
  const int nb = neighbour(g, sc, s, mu, fw);
  if (fw) {
    gauge.ld(par, mu, s, flip && mu == 3 && sc.t == g.L[3] - 1, U);
  } else {
    gauge.ld(1 - par, mu, nb, flip && mu == 3 && sc.t == 0, U);
    dag = true;
  }
  return nb;
}

template <int CIN, int COUT>
__global__ void __launch_bounds__(kLayerThreads) k_net_layer(LayerArgs a) {
  constexpr int MB = frag_mb(COUT), KB = frag_kb(CIN);
  constexpr int NF = frag_count(CIN, COUT);
  __shared__ uint4 sA[NF];
  for (int i = threadIdx.x; i < NF; i += kLayerThreads) sA[i] = a.frag[i];
  __syncthreads();
  const Geom& g = a.g;
  const int lane = threadIdx.x & 63, q = lane >> 4, n = lane & 15;
  const int wave = threadIdx.x >> 6;
  const int tpp = (g.Vh + 15) >> 4;
  const int ntiles = tpp * a.npar;
  const size_t cs = (size_t)24 * g.stride;
  const GaugeF32r12 gauge{a.gauge, g.stride};
  const int t_begin = blockIdx.x * a.tiles_per_block;
  const int t_end = min(t_begin + a.tiles_per_block, ntiles);
  for (int tile = t_begin + wave; tile < t_end; tile += kLayerThreads / 64) {
    // wave-uniform: keeps the component offsets of every access scalar
    const int par = __builtin_amdgcn_readfirstlane(a.par0 + tile / tpp);
    const int site = (tile % tpp) * 16 + n;
    const bool valid = site < g.Vh;
    const int s = valid ? site : g.Vh - 1;
    const SiteCoord sc = decode_site(g, s, par);
    f32x4 acc[3][MB];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[c][mb][r] = 0.0f;
#pragma unroll 1
    for (int d = 0; d < kPaths; ++d) {
      float2 U[3][3];
      bool dag;
      const int nb = net_path(g, gauge, a.flip, d, par, s, sc, U, dag);
      const uint32_t soff = (uint32_t)((d == 0 ? par : 1 - par) * 12 * g.stride);
      bf16x8 B[3][KB];
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        const int ch = 4 * kb + q;
        float2 w[4][3];
        if (ch >= CIN) {
#pragma unroll
          for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int c = 0; c < 3; ++c) w[t][c] = make_float2(0.f, 0.f);
        } else {
          net_transport(a.in, soff, g.stride, (uint32_t)(ch * cs + nb), d != 0, dag, U, w);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          bf16x8 b;
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            b[2 * t] = (__bf16)w[t][c].x;
            b[2 * t + 1] = (__bf16)w[t][c].y;
          }
          B[c][kb] = b;
        }
      }
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
          const uint4 araw = sA[((d * MB + mb) * KB + kb) * 64 + lane];
          bf16x8 A;
          __builtin_memcpy(&A, &araw, 16);
#pragma unroll
          for (int c = 0; c < 3; ++c)
            acc[c][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[c][kb], acc[c][mb], 0, 0, 0);
        }
    }
    if (valid) {
      // C row 4q + r of block mb = (channel 2mb + q/2, spin 2(q&1) + r/2, re/im r&1)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const int o = 2 * mb + (q >> 1);
        if (o < COUT) {
          // spin 2(q&1)+sp varies across lane groups: its offset goes with the lane
          const uint32_t loff = (uint32_t)(o * cs + 6 * (q & 1) * g.stride + site);
          const uint32_t pbase = (uint32_t)(par * 12 * g.stride);
#pragma unroll
          for (int sp = 0; sp < 2; ++sp)
#pragma unroll
            for (int c = 0; c < 3; ++c)
              st_at(a.out, pbase + (uint32_t)((3 * sp + c) * g.stride), loff,
                    make_float2(acc[c][mb][2 * sp], acc[c][mb][2 * sp + 1]));
        }
      }
    }
  }
}

// Weight gradient partials.  Block b covers global sites q in
// [b*spb, (b+1)*spb) of the 2*Vh sites (q = par*Vh + hs), in stages of 16
// sites; thread groups split the sites of a stage, each thread of a group
// owns a 2x2 tile of the (4*COUT) x (4*CIN) complex matrix of every path.
template <int CIN, int COUT>
__global__ void __launch_bounds__(kWgThreads)
k_net_wgrad(Geom g, const float2* __restrict__ in, const float2* __restrict__ gout,
            const float2* __restrict__ gaugep, int flip, int spb, float2* __restrict__ partial) {
  constexpr int ROWS = 4 * COUT, COLS = 4 * CIN;
  constexpr int OWN = ROWS * COLS / 4;
  constexpr int G = kWgThreads / OWN;
  constexpr int ST = 16;
  static_assert(OWN * G == kWgThreads, "tile owners must divide the block");
  __shared__ float2 sg[ST][3][ROWS];
  __shared__ float2 sta[ST][3][COLS];
  const int tid = threadIdx.x;
  const int owner = tid % OWN, grp = tid / OWN;
  const int r0 = 2 * (owner / (COLS / 2)), c0 = 2 * (owner % (COLS / 2));
  const size_t cs = (size_t)24 * g.stride;
  const GaugeF32r12 gauge{gaugep, g.stride};
  float2 acc[kPaths][2][2];
#pragma unroll
  for (int d = 0; d < kPaths; ++d)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[d][i][j] = make_float2(0.f, 0.f);
  const int nq = 2 * g.Vh;
  const int q0 = blockIdx.x * spb;
  const int q1 = min(q0 + spb, nq);
  for (int qb = q0; qb < q1; qb += ST) {
    for (int e = tid; e < ST * COUT * 12; e += kWgThreads) {
      const int x = e % ST, rest = e / ST, comp = rest % 12, o = rest / 12;
      const int q = qb + x;
      float2 v = make_float2(0.f, 0.f);
      if (q < q1) {
        const int par = q >= g.Vh, hs = q - par * g.Vh;
        v = ld_at(gout, 0u, (uint32_t)(o * cs + (par * 12 + comp) * g.stride + hs));
      }
      sg[x][comp % 3][4 * o + comp / 3] = v;
    }
#pragma unroll
    for (int d = 0; d < kPaths; ++d) {
      __syncthreads();
      for (int e = tid; e < ST * CIN; e += kWgThreads) {
        const int x = e % ST, i = e / ST;
        const int q = qb + x;
        float2 w[4][3];
        if (q < q1) {
          const int par = q >= g.Vh, hs = q - par * g.Vh;
          const SiteCoord sc = decode_site(g, hs, par);
          float2 U[3][3];
          bool dag = false;
          int nb = hs;
          if (d > 0) {
            const int mu = (d - 1) >> 1;
            const bool fw = ((d - 1) & 1) == 0;
            nb = neighbour(g, sc, hs, mu, fw);
            if (fw) {
              ld_link_lane(gaugep, g.stride, par, mu, hs, flip && mu == 3 && sc.t == g.L[3] - 1, U);
            } else {
              ld_link_lane(gaugep, g.stride, 1 - par, mu, nb, flip && mu == 3 && sc.t == 0, U);
              dag = true;
            }
          }
          const int npar = d == 0 ? par : 1 - par;
          net_transport(in, 0u, g.stride, (uint32_t)(i * cs + npar * 12 * g.stride + nb), d != 0,
                        dag, U, w);
        } else {
#pragma unroll
          for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int c = 0; c < 3; ++c) w[t][c] = make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int c = 0; c < 3; ++c) sta[x][c][4 * i + t] = w[t][c];
      }
      __syncthreads();
      for (int x = grp; x < ST; x += G)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float2 ga = sg[x][c][r0], gb = sg[x][c][r0 + 1];
          const float2 ta = sta[x][c][c0], tb = sta[x][c][c0 + 1];
          cmacc(acc[d][0][0], ta, ga);
          cmacc(acc[d][0][1], tb, ga);
          cmacc(acc[d][1][0], ta, gb);
          cmacc(acc[d][1][1], tb, gb);
        }
    }
    __syncthreads();
  }
  constexpr int NE = COUT * CIN * 16;
  float2* dst = partial + (size_t)(blockIdx.x * G + grp) * kPaths * NE;
#pragma unroll
  for (int d = 0; d < kPaths; ++d)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int row = r0 + i, col = c0 + j;
        const int o = row >> 2, s = row & 3, ci = col >> 2, t = col & 3;
        dst[(size_t)d * NE + (o * CIN + ci) * 16 + 4 * s + t] = acc[d][i][j];
      }
}

// gacc[e] += scale * sum_p partial[p][e], fixed order, FP64 sum.
__global__ void k_net_wreduce(const float2* __restrict__ partial, int np, int ne, float scale,
                              float2* __restrict__ gacc) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double sr = 0.0, si = 0.0;
  for (int p = 0; p < np; ++p) {
    const float2 v = partial[(size_t)p * ne + e];
    sr += v.x;
    si += v.y;
  }
  float2 a = gacc[e];
  a.x += (float)(scale * sr);
  a.y += (float)(scale * si);
  gacc[e] = a;
}

__global__ void k_net_adam(float* __restrict__ w, const float* __restrict__ gr,
                           float* __restrict__ m, float* __restrict__ v, int n, float lr, float b1,
                           float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = gr[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Even half field (FP64 [12][stride]) -> channel 0 of a full FP32 field,
// odd sites zero.
__global__ void __launch_bounds__(kBlock)
k_net_embed(Geom g, const double2* __restrict__ r, float2* __restrict__ a0) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= g.Vh) return;
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const double2 v = r[(size_t)k * g.stride + s];
    a0[(size_t)k * g.stride + s] = make_float2((float)v.x, (float)v.y);
    a0[(size_t)(12 + k) * g.stride + s] = make_float2(0.f, 0.f);
  }
}

// Even part of a 1-channel FP32 full field -> FP64 half field.
__global__ void __launch_bounds__(kBlock)
k_net_extract(Geom g, const float2* __restrict__ a, double2* __restrict__ z) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= g.Vh) return;
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const float2 v = a[(size_t)k * g.stride + s];
    z[(size_t)k * g.stride + s] = make_double2(v.x, v.y);
  }
}

// Loss of one pair: e = out_even - y; partial sums of |e|^2; gradient at the
// output 2*wscale*e on even sites, zero on odd sites.
__global__ void __launch_bounds__(kBlock)
k_net_loss(Geom g, const float2* __restrict__ out, const float2* __restrict__ y,
           float2* __restrict__ gout, float wscale, double* __restrict__ part) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double e2 = 0.0;
  if (s < g.Vh) {
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const float2 o = out[(size_t)k * g.stride + s];
      const float2 t = y[(size_t)k * g.stride + s];
      const float ex = o.x - t.x, ey = o.y - t.y;
      e2 += (double)ex * ex + (double)ey * ey;
      gout[(size_t)k * g.stride + s] = make_float2(2.0f * wscale * ex, 2.0f * wscale * ey);
      gout[(size_t)(12 + k) * g.stride + s] = make_float2(0.f, 0.f);
    }
  }
  e2 = block_sum(e2);
  if (threadIdx.x == 0) part[blockIdx.x] = e2;
}

// Lexicographic Format_F doubles [nch][V][24] <-> FP32 channel fields.
__global__ void __launch_bounds__(kBlock)
k_net_pack_lex(Geom g, const double* __restrict__ lex, float2* __restrict__ dst, int nch) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long)nch * g.V) return;
  const int ch = (int)(id / g.V), site = (int)(id % g.V);
  int r = site;
  const int x = r % g.L[0]; r /= g.L[0];
  const int y = r % g.L[1]; r /= g.L[1];
  const int z = r % g.L[2];
  const int t = r / g.L[2];
  const int par = (x + y + z + t) & 1, hs = site >> 1;
  const double* src = lex + ((size_t)ch * g.V + site) * 24;
  float2* d = dst + (size_t)ch * 24 * g.stride + (size_t)par * 12 * g.stride;
#pragma unroll
  for (int k = 0; k < 12; ++k)
    d[(size_t)k * g.stride + hs] = make_float2((float)src[2 * k], (float)src[2 * k + 1]);
}
__global__ void __launch_bounds__(kBlock)
k_net_unpack_lex(Geom g, const float2* __restrict__ src, double* __restrict__ lex, int nch) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long)nch * g.V) return;
  const int ch = (int)(id / g.V), site = (int)(id % g.V);
  int r = site;
  const int x = r % g.L[0]; r /= g.L[0];
  const int y = r % g.L[1]; r /= g.L[1];
  const int z = r % g.L[2];
  const int t = r / g.L[2];
  const int par = (x + y + z + t) & 1, hs = site >> 1;
  double* d = lex + ((size_t)ch * g.V + site) * 24;
  const float2* s = src + (size_t)ch * 24 * g.stride + (size_t)par * 12 * g.stride;
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const float2 v = s[(size_t)k * g.stride + hs];
    d[2 * k] = v.x;
    d[2 * k + 1] = v.y;
  }
}

// Gaussian even half field (unit variance per real component), FP64,
// counter-based: stream = (seed, pair), counter = component index.
__global__ void __launch_bounds__(kBlock)
k_net_gauss(Geom g, uint64_t seed, int pair, double2* __restrict__ dst) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= g.Vh) return;
  const uint64_t key = splitmix(seed ^ splitmix(0x6e657400ULL + (uint64_t)pair));
#pragma unroll 1
  for (int k = 0; k < 12; ++k) {
    double re[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const uint64_t c = ((uint64_t)s * 24 + 2 * k + p) * 2;
      const double u1 = (double)((splitmix(key + c * 0xD1B54A32D192ED03ULL) >> 11) + 1) *
                        (1.0 / 9007199254740992.0);
      const double u2 = (double)((splitmix(key + (c + 1) * 0xD1B54A32D192ED03ULL) >> 11) + 1) *
                        (1.0 / 9007199254740992.0);
      re[p] = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
    }
    dst[(size_t)k * g.stride + s] = make_double2(re[0], re[1]);
  }
}

// FP64 half field -> FP32 half field (k < 12 components), optional odd zero
// block after it (full channel layout).
__global__ void __launch_bounds__(kBlock)
k_net_to32(Geom g, const double2* __restrict__ src, float2* __restrict__ dst, int zero_odd) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= g.Vh) return;
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const double2 v = src[(size_t)k * g.stride + s];
    dst[(size_t)k * g.stride + s] = make_float2((float)v.x, (float)v.y);
    if (zero_odd) dst[(size_t)(12 + k) * g.stride + s] = make_float2(0.f, 0.f);
  }
}

// ---------------------------------------------------- complex FP64 BLAS
// partials of <a, b> = sum conj(a) b: part[2*blk] re, part[2*blk+1] im
__global__ void __launch_bounds__(kBlock)
k_cdot(Geom g, const double2* __restrict__ a, const double2* __restrict__ b,
       double* __restrict__ part) {
  double re = 0.0, im = 0.0;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < g.Vh; s += gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const double2 x = a[(size_t)k * g.stride + s], y = b[(size_t)k * g.stride + s];
      re += x.x * y.x + x.y * y.y;
      im += x.x * y.y - x.y * y.x;
    }
  }
  re = block_sum(re);
  im = block_sum(im);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = re;
    part[2 * blockIdx.x + 1] = im;
  }
}
__global__ void k_cdot_fin(const double* __restrict__ part, int nb, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double re = 0.0, im = 0.0;
    for (int i = 0; i < nb; ++i) {
      re += part[2 * i];
      im += part[2 * i + 1];
    }
    out[0] = re;
    out[1] = im;
  }
}
// y += (ar + i ai) x
__global__ void __launch_bounds__(kBlock)
k_caxpy(Geom g, double ar, double ai, const double2* __restrict__ x, double2* __restrict__ y) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= g.Vh) return;
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const double2 v = x[(size_t)k * g.stride + s];
    double2 w = y[(size_t)k * g.stride + s];
    w.x += ar * v.x - ai * v.y;
    w.y += ar * v.y + ai * v.x;
    y[(size_t)k * g.stride + s] = w;
  }
}

// ------------------------------------------------------------------ host
template <int CIN, int COUT>
ii_status launch_layer(ii_context* ctx, const float2* in, float2* out, const uint4* frag,
                       int par0, int npar) {
  const Geom& g = ctx->g;
  LayerArgs a;
  a.g = g;
  a.in = in;
  a.out = out;
  a.frag = frag;
  a.gauge = (const float2*)ctx->gauge_lp[II_PREC_SINGLE];
  a.flip = ctx->tbc_sign < 0;
  a.par0 = par0;
  a.npar = npar;
  const int ntiles = ((g.Vh + 15) / 16) * npar;
  // two blocks per CU (LDS of the 8->8 operands ~74 KB), each a contiguous
  // run of tiles
  int nblk = std::min(512, std::max(1, (ntiles + 3) / 4));
  a.tiles_per_block = (ntiles + nblk - 1) / nblk;
  nblk = (ntiles + a.tiles_per_block - 1) / a.tiles_per_block;
  hipLaunchKernelGGL((k_net_layer<CIN, COUT>), dim3(nblk), dim3(kLayerThreads), 0, ctx->stream,
                     a);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

// The layer kernels address a whole 8-channel activation array ([channel]
// [parity][12][stride] FP32 complex) through one buffer descriptor with
// 32-bit offsets, so they take lattices whose such array stays below 4 GiB
// (V/2 up to about 2.79M sites).  The batched network (ii_bnet.hip) has no
// such limit below the context's own.
ii_status layer_fits(ii_context* ctx) {
  if ((unsigned long long)8 * chan_elems(ctx->g) * sizeof(float2) >= (1ull << 32))
    return set_error(ctx, II_ERR_BAD_ARG,
                     "lattice too large for the single-rhs network layers (an 8-channel "
                     "activation array beyond 4 GiB); the batched network takes it");
  return II_OK;
}

// layer l forward (adj = false) or its adjoint, output parities [par0, par0+npar)
ii_status run_layer(ii_context* ctx, Net* N, int l, bool adj, const float2* in, float2* out,
                    int par0 = 0, int npar = 2) {
  II_CHECK(layer_fits(ctx));
  const LayerParams& P = N->L[l];
  const int cin = adj ? P.cout : P.cin, cout = adj ? P.cin : P.cout;
  const uint4* fr = adj ? P.frag_adj : P.frag_fw;
  if (cin == 1 && cout == 8) return launch_layer<1, 8>(ctx, in, out, fr, par0, npar);
  if (cin == 8 && cout == 8) return launch_layer<8, 8>(ctx, in, out, fr, par0, npar);
  if (cin == 8 && cout == 1) return launch_layer<8, 1>(ctx, in, out, fr, par0, npar);
  return set_error(ctx, II_ERR_STATE, "unsupported layer shape");
}

template <int CIN, int COUT>
ii_status launch_wgrad(ii_context* ctx, Net* N, const float2* in, const float2* gout,
                       float2* gacc, float scale) {
  const Geom& g = ctx->g;
  constexpr int G = kWgThreads / (4 * COUT * CIN);
  constexpr int NE = COUT * CIN * 16;
  const int nq = 2 * g.Vh;
  int nblk = std::min(1024, std::max(1, nq / 64));
  const int spb = ((nq + nblk - 1) / nblk + 15) / 16 * 16;
  nblk = (nq + spb - 1) / spb;
  const size_t need = (size_t)nblk * G * kPaths * NE;
  if (need > N->partial_len) {
    if (N->partial) dfree(N->partial);
    N->partial = nullptr;
    II_CHECK(dev_alloc(ctx, &N->partial, need));
    N->partial_len = need;
  }
  hipLaunchKernelGGL((k_net_wgrad<CIN, COUT>), dim3(nblk), dim3(kWgThreads), 0, ctx->stream, g,
                     in, gout, (const float2*)ctx->gauge_lp[II_PREC_SINGLE], ctx->tbc_sign < 0,
                     spb, N->partial);
  II_HIP(ctx, hipGetLastError());
  const int ne = kPaths * NE;
  hipLaunchKernelGGL(k_net_wreduce, dim3((ne + 255) / 256), dim3(256), 0, ctx->stream, N->partial,
                     nblk * G, ne, scale, gacc);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

ii_status run_wgrad(ii_context* ctx, Net* N, int l, const float2* in, const float2* gout,
                    float2* gacc, float scale) {
  II_CHECK(layer_fits(ctx));
  const LayerParams& P = N->L[l];
  if (P.cin == 1 && P.cout == 8) return launch_wgrad<1, 8>(ctx, N, in, gout, gacc, scale);
  if (P.cin == 8 && P.cout == 8) return launch_wgrad<8, 8>(ctx, N, in, gout, gacc, scale);
  if (P.cin == 8 && P.cout == 1) return launch_wgrad<8, 1>(ctx, N, in, gout, gacc, scale);
  return set_error(ctx, II_ERR_STATE, "unsupported layer shape");
}

ii_status build_frags(ii_context* ctx, Net* N) {
  N->version++;
  for (int l = 0; l < kNL; ++l) {
    LayerParams& P = N->L[l];
    const int nf = frag_count(P.cin, P.cout), na = frag_count(P.cout, P.cin);
    hipLaunchKernelGGL(k_net_frags, dim3((nf + 255) / 256), dim3(256), 0, ctx->stream, P.w, P.cin,
                       P.cout, 0, P.frag_fw);
    hipLaunchKernelGGL(k_net_frags, dim3((na + 255) / 256), dim3(256), 0, ctx->stream, P.w, P.cin,
                       P.cout, 1, P.frag_adj);
    II_HIP(ctx, hipGetLastError());
  }
  return II_OK;
}

ii_status net_get(ii_context* ctx, Net** out) {
  II_CHECK(require_gauge(ctx));
  if (!ctx->gauge_lp[II_PREC_SINGLE]) II_CHECK(build_lowprec_gauge(ctx, II_PREC_SINGLE));
  Net* N = net_of(ctx);
  if (!N) {
    N = new Net();
    ctx->learned = N;
    for (int l = 0; l < kNL; ++l) {
      LayerParams& P = N->L[l];
      P.cin = kWidth[l];
      P.cout = kWidth[l + 1];
      P.nw = (size_t)kPaths * P.cout * P.cin * 16;
      II_CHECK(dev_alloc(ctx, &P.w, P.nw));
      II_CHECK(dev_alloc(ctx, &P.gacc, P.nw));
      II_CHECK(dev_alloc(ctx, &P.m, P.nw));
      II_CHECK(dev_alloc(ctx, &P.v, P.nw));
      II_CHECK(dev_alloc(ctx, &P.frag_fw, (size_t)frag_count(P.cin, P.cout)));
      II_CHECK(dev_alloc(ctx, &P.frag_adj, (size_t)frag_count(P.cout, P.cin)));
      II_HIP(ctx, hipMemsetAsync(P.w, 0, P.nw * sizeof(float2), ctx->stream));
      II_HIP(ctx, hipMemsetAsync(P.m, 0, P.nw * sizeof(float2), ctx->stream));
      II_HIP(ctx, hipMemsetAsync(P.v, 0, P.nw * sizeof(float2), ctx->stream));
    }
    const Geom& g = ctx->g;
    const size_t ce = chan_elems(g);
    for (int l = 0; l <= kNL; ++l) II_CHECK(dev_alloc(ctx, &N->act[l], ce * kWidth[l]));
    II_CHECK(dev_alloc(ctx, &N->grad[0], ce * 8));
    II_CHECK(dev_alloc(ctx, &N->grad[1], ce * 8));
    II_CHECK(dev_alloc(ctx, &N->gout, ce));
    II_CHECK(dev_alloc(ctx, &N->dpart, (size_t)2 * kRedBlocksMax + 64));
    II_HIP(ctx, hipHostMalloc((void**)&N->dhost, 4096, hipHostMallocDefault));
  }
  *out = N;
  return II_OK;
}

ii_status stage(ii_context* ctx, Net* N, size_t n) {
  if (n > N->stage_len) {
    if (N->stage) dfree(N->stage);
    N->stage = nullptr;
    II_CHECK(dev_alloc(ctx, &N->stage, n));
    N->stage_len = n;
  }
  return II_OK;
}

// net forward on act[0] (already embedded); keeps every activation;
// last layer evaluated on even output sites only
ii_status forward(ii_context* ctx, Net* N) {
  for (int l = 0; l < kNL; ++l)
    II_CHECK(run_layer(ctx, N, l, false, N->act[l], N->act[l + 1], 0, l == kNL - 1 ? 1 : 2));
  return II_OK;
}

// gradient of the loss whose output gradient is in N->gout into every gacc
ii_status backward(ii_context* ctx, Net* N, float scale) {
  const float2* g = N->gout;
  int pp = 0;
  for (int l = kNL - 1; l >= 0; --l) {
    II_CHECK(run_wgrad(ctx, N, l, N->act[l], g, N->L[l].gacc, scale));
    if (l > 0) {
      float2* gn = N->grad[pp];
      II_CHECK(run_layer(ctx, N, l, true, g, gn));
      g = gn;
      pp ^= 1;
    }
  }
  return II_OK;
}

ii_status zero_grads(ii_context* ctx, Net* N) {
  for (int l = 0; l < kNL; ++l)
    II_HIP(ctx, hipMemsetAsync(N->L[l].gacc, 0, N->L[l].nw * sizeof(float2), ctx->stream));
  return II_OK;
}

ii_status adam_step(ii_context* ctx, Net* N, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  N->adam_step++;
  const float c1 = 1.0f - std::pow(b1, (float)N->adam_step);
  const float c2 = 1.0f - std::pow(b2, (float)N->adam_step);
  for (int l = 0; l < kNL; ++l) {
    const int n = (int)(2 * N->L[l].nw);
    hipLaunchKernelGGL(k_net_adam, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                       (float*)N->L[l].w, (const float*)N->L[l].gacc, (float*)N->L[l].m,
                       (float*)N->L[l].v, n, lr, b1, b2, eps, c1, c2);
    II_HIP(ctx, hipGetLastError());
  }
  return build_frags(ctx, N);
}

// loss of pair k (device pair storage) accumulated into the host value;
// leaves the output gradient (scaled by gscale) in N->gout
ii_status pair_loss(ii_context* ctx, Net* N, const float2* b_full, const float2* y_half,
                    double inv_ny, float gscale, double* loss) {
  const Geom& g = ctx->g;
  II_HIP(ctx, hipMemcpyAsync(N->act[0], b_full, chan_elems(g) * sizeof(float2),
                             hipMemcpyDeviceToDevice, ctx->stream));
  II_CHECK(forward(ctx, N));
  const int nb = blocks_for(g.Vh);
  if (nb > kRedBlocksMax * 2) return set_error(ctx, II_ERR_STATE, "lattice too large for loss");
  hipLaunchKernelGGL(k_net_loss, dim3(nb), dim3(kBlock), 0, ctx->stream, g, N->act[kNL], y_half,
                     N->gout, (float)(inv_ny * gscale), N->dpart);
  II_HIP(ctx, hipGetLastError());
  std::vector<double> hp(nb);
  II_HIP(ctx, hipMemcpyAsync(hp.data(), N->dpart, nb * sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double s = 0.0;
  for (int i = 0; i < nb; ++i) s += hp[i];
  *loss += s * inv_ny;
  return II_OK;
}

// Dirac gamma matrices of the reference basis (wilsonLikeUtils.cpp:9-40)
void gamma_mat(int mu, std::complex<double> G[4][4]) {
  using C = std::complex<double>;
  const C I(0, 1), O(0, 0), one(1, 0);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) G[a][b] = O;
  if (mu == 0) { G[0][3] = -I; G[1][2] = -I; G[2][1] = I; G[3][0] = I; }
  if (mu == 1) { G[0][3] = -one; G[1][2] = one; G[2][1] = one; G[3][0] = -one; }
  if (mu == 2) { G[0][2] = -I; G[1][3] = I; G[2][0] = I; G[3][1] = -I; }
  if (mu == 3) { G[0][0] = one; G[1][1] = one; G[2][2] = -one; G[3][3] = -one; }
}

// ------------------------------------------------- applying M on the device
// z = M r for FP64 even half fields on the device
ii_status apply_M(ii_context* ctx, Net* N, const double2* r, double2* z) {
  const Geom& g = ctx->g;
  hipLaunchKernelGGL(k_net_embed, dim3(blocks_for(g.Vh)), dim3(kBlock), 0, ctx->stream, g, r,
                     N->act[0]);
  II_HIP(ctx, hipGetLastError());
  II_CHECK(forward(ctx, N));
  hipLaunchKernelGGL(k_net_extract, dim3(blocks_for(g.Vh)), dim3(kBlock), 0, ctx->stream, g,
                     N->act[kNL], z);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

ii_status cdot(ii_context* ctx, Net* N, const double2* a, const double2* b, double out[2]) {
  const Geom& g = ctx->g;
  const int nb = std::min(kRedBlocksMax, blocks_for(g.Vh));
  hipLaunchKernelGGL(k_cdot, dim3(nb), dim3(kBlock), 0, ctx->stream, g, a, b, N->dpart);
  hipLaunchKernelGGL(k_cdot_fin, dim3(1), dim3(64), 0, ctx->stream, N->dpart, nb,
                     N->dpart + 2 * kRedBlocksMax);
  II_HIP(ctx, hipGetLastError());
  II_HIP(ctx, hipMemcpyAsync(N->dhost, N->dpart + 2 * kRedBlocksMax, 2 * sizeof(double),
                             hipMemcpyDeviceToHost, ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  out[0] = N->dhost[0];
  out[1] = N->dhost[1];
  return II_OK;
}

ii_status caxpy(ii_context* ctx, double ar, double ai, const double2* x, double2* y) {
  const Geom& g = ctx->g;
  hipLaunchKernelGGL(k_caxpy, dim3(blocks_for(g.Vh)), dim3(kBlock), 0, ctx->stream, g, ar, ai, x,
                     y);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

}  // namespace

// ------------------------------------------------------------ flexible GCR
// Right-preconditioned flexible GCR on Dhat x = b (FP64 vectors, FP64
// Dhat), Krylov directions z_j = M r_j, q_j = Dhat z_j orthogonalised by
// modified Gram-Schmidt; restart after nk directions with the true
// residual.  Returns II_ERR_NOT_CONVERGED when maxiter directions do not
// reach tol (the caller falls back to CG).
ii_status fgcr(ii_context* ctx, double kappa, double2* x, const double2* b, double tol,
               int maxiter, int nk, bool use_net, ii_solver_output* out) {
  Net* N = nullptr;
  II_CHECK(net_get(ctx, &N));
  if (use_net && !N->init) return set_error(ctx, II_ERR_STATE, "network not initialised");
  const Geom& g = ctx->g;
  const size_t he = half_elems(g);
  nk = std::max(1, std::min(nk, 64));
  if (N->kry_n < nk) {
    if (N->kry) dfree(N->kry);
    N->kry = nullptr;
    II_CHECK(dev_alloc(ctx, &N->kry, 2 * nk * he));
    N->kry_n = nk;
  }
  for (int i = 0; i < 3; ++i) II_CHECK(dev_alloc(ctx, &N->fw[i], he));
  double2 *r = N->fw[0], *t = N->fw[1], *w = N->fw[2];
  auto Z = [&](int j) { return N->kry + (size_t)j * he; };
  auto Q = [&](int j) { return N->kry + (size_t)(nk + j) * he; };
  double bb[2];
  II_CHECK(cdot(ctx, N, b, b, bb));
  const double bn = std::sqrt(bb[0]);
  auto t0 = std::chrono::steady_clock::now();
  // r = b - Dhat x
  II_CHECK(dhat64(ctx, kappa, x, w, t));
  II_CHECK(copy64(ctx, r, b));
  II_CHECK(axpby64(ctx, -1.0, w, 1.0, r));
  double rr[2];
  II_CHECK(cdot(ctx, N, r, r, rr));
  int its = 0, restarts = 0;
  std::vector<double> qq(nk);
  double best = std::sqrt(rr[0]);
  int stall = 0;
  while (std::sqrt(rr[0]) > tol * bn && its < maxiter) {
    int k = 0;
    for (; k < nk && its < maxiter; ++k, ++its) {
      if (use_net) II_CHECK(apply_M(ctx, N, r, Z(k)));
      else II_CHECK(copy64(ctx, Z(k), r));
      II_CHECK(dhat64(ctx, kappa, Z(k), Q(k), t));
      for (int j = 0; j < k; ++j) {
        double c[2];
        II_CHECK(cdot(ctx, N, Q(j), Q(k), c));
        const double br = c[0] / qq[j], bi = c[1] / qq[j];
        II_CHECK(caxpy(ctx, -br, -bi, Q(j), Q(k)));
        II_CHECK(caxpy(ctx, -br, -bi, Z(j), Z(k)));
      }
      double c[2];
      II_CHECK(cdot(ctx, N, Q(k), Q(k), c));
      qq[k] = c[0];
      if (!(qq[k] > 0.0) || !std::isfinite(qq[k]))
        return set_error(ctx, II_ERR_NAN, "flexible GCR: degenerate direction");
      II_CHECK(cdot(ctx, N, Q(k), r, c));
      const double ar = c[0] / qq[k], ai = c[1] / qq[k];
      II_CHECK(caxpy(ctx, ar, ai, Z(k), x));
      II_CHECK(caxpy(ctx, -ar, -ai, Q(k), r));
      II_CHECK(cdot(ctx, N, r, r, rr));
      if (std::sqrt(rr[0]) <= tol * bn) {
        ++k;
        ++its;
        break;
      }
    }
    // restart with the true residual
    II_CHECK(dhat64(ctx, kappa, x, w, t));
    II_CHECK(copy64(ctx, r, b));
    II_CHECK(axpby64(ctx, -1.0, w, 1.0, r));
    II_CHECK(cdot(ctx, N, r, r, rr));
    ++restarts;
    const double cur = std::sqrt(rr[0]);
    if (cur < 0.5 * best) {
      best = cur;
      stall = 0;
    } else if (++stall >= 3) {
      break;
    }
  }
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (out) {
    out->iterations = its;
    out->restarts = restarts;
    out->inner_iterations = 0;
    out->diff = std::sqrt(rr[0]) / bn;
    out->true_residual = std::sqrt(rr[0]) / bn;
    out->timing_ms = ms;
  }
  if (std::sqrt(rr[0]) > tol * bn)
    return set_error(ctx, II_ERR_NOT_CONVERGED, "flexible GCR did not reach the tolerance");
  return II_OK;
}

// Flexible GCR, then, if it stalls or runs out of directions, plain CG on
// Dhat^dag Dhat e = Dhat^dag r for the remaining defect (x += e), repeated
// with a tighter CG threshold until the FP64 true residual meets tol.
ii_status solve_even(ii_context* ctx, double kappa, double2* x, const double2* b, double tol,
                     int maxiter, int nk, bool use_net, ii_solver_output* out) {
  ii_solver_output o{};
  ii_status st = fgcr(ctx, kappa, x, b, tol, maxiter, nk, use_net, &o);
  if (st != II_ERR_NOT_CONVERGED) {
    if (out) *out = o;
    return st;
  }
  ctx->err.clear();
  Net* N = net_of(ctx);
  const Geom& g = ctx->g;
  const size_t he = half_elems(g);
  for (int i = 5; i < 6; ++i) II_CHECK(dev_alloc(ctx, &N->fw[i], 2 * he));
  double2 *r = N->fw[0], *t = N->fw[1], *w = N->fw[2], *c = N->fw[5], *e = N->fw[5] + he;
  double bb[2], rr[2], cc[2];
  II_CHECK(cdot(ctx, N, b, b, bb));
  int cg_its = 0;
  double shrink = 1e-2;
  for (int attempt = 0; attempt < 6; ++attempt) {
    II_CHECK(dhat64(ctx, kappa, x, w, t));
    II_CHECK(copy64(ctx, r, b));
    II_CHECK(axpby64(ctx, -1.0, w, 1.0, r));
    II_CHECK(cdot(ctx, N, r, r, rr));
    if (std::sqrt(rr[0]) <= tol * std::sqrt(bb[0])) break;
    II_CHECK(dhat_dag64(ctx, kappa, r, c, t));
    II_CHECK(cdot(ctx, N, c, c, cc));
    // squared criterion on |Dhat^dag r_new|^2 / |Dhat^dag r|^2
    const double prec = std::max(1e-30, shrink * tol * tol * bb[0] / rr[0]);
    ii_solver_output co{};
    ii_status cs = cg_reference(ctx, kappa, e, c, prec, 100000, &co);
    if (cs != II_OK && cs != II_ERR_NOT_CONVERGED) return cs;
    ctx->err.clear();
    cg_its += co.iterations;
    II_CHECK(axpby64(ctx, 1.0, e, 1.0, x));
    shrink *= 1e-2;
  }
  const bool ok = std::sqrt(rr[0]) <= tol * std::sqrt(bb[0]);
  o.inner_iterations = cg_its;
  o.true_residual = o.diff = std::sqrt(rr[0] / bb[0]);
  if (out) *out = o;
  if (!ok) return set_error(ctx, II_ERR_NOT_CONVERGED, "flexible GCR and CG fallback failed");
  return II_OK;
}

ii_status net_host_weights(ii_context* ctx, int* version,
                           std::vector<std::vector<float2>>* w) {
  Net* N = net_of(ctx);
  if (!N || !N->init) return set_error(ctx, II_ERR_STATE, "network not initialised");
  *version = N->version;
  if (!w) return II_OK;
  w->assign(kNL, {});
  for (int l = 0; l < kNL; ++l) {
    (*w)[l].resize(N->L[l].nw);
    II_HIP(ctx, hipMemcpyAsync((*w)[l].data(), N->L[l].w, N->L[l].nw * sizeof(float2),
                               hipMemcpyDeviceToHost, ctx->stream));
  }
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return II_OK;
}

void learned_release(ii_context* ctx) {
  Net* N = net_of(ctx);
  if (!N) return;
  for (int l = 0; l < kNL; ++l) {
    LayerParams& P = N->L[l];
    dfree(P.w);
    dfree(P.gacc);
    dfree(P.m);
    dfree(P.v);
    dfree(P.frag_fw);
    dfree(P.frag_adj);
  }
  for (int l = 0; l <= kNL; ++l) dfree(N->act[l]);
  dfree(N->grad[0]);
  dfree(N->grad[1]);
  dfree(N->gout);
  dfree(N->partial);
  dfree(N->dpart);
  if (N->dhost) (void)hipHostFree(N->dhost);
  dfree(N->stage);
  dfree(N->pb);
  dfree(N->py);
  dfree(N->kry);
  for (auto* p : N->fw) dfree(p);
  delete N;
  ctx->learned = nullptr;
}

ii_status learned_propagator(ii_context* ctx, double kappa, const ii_prop_params* p,
                             const ii_field* src, ii_field* sol, ii_solver_output* outs);

}  // namespace ii

using namespace ii;

extern "C" {

ii_status ii_net_init(ii_context* ctx, double kappa, double noise, unsigned long long seed) {
  if (!ctx) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  Net* N = nullptr;
  II_CHECK(net_get(ctx, &N));
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  using C = std::complex<double>;
  C G[4][4][4];
  for (int mu = 0; mu < 4; ++mu) gamma_mat(mu, G[mu]);
  for (int l = 0; l < kNL; ++l) {
    LayerParams& P = N->L[l];
    std::vector<float2> w(P.nw);
    auto W = [&](int d, int o, int i, int s, int t) -> float2& {
      return w[((size_t)(d * P.cout + o) * P.cin + i) * 16 + 4 * s + t];
    };
    // channel 0 carries the input, channel 1 the Horner sum:
    // ch1 <- ch0 + kappa H ch1, output <- ch0 + kappa H ch1
    auto set_hop = [&](int o, int i) {
      for (int mu = 0; mu < 4; ++mu)
        for (int s = 0; s < 4; ++s)
          for (int t = 0; t < 4; ++t) {
            const C id = (s == t) ? C(1, 0) : C(0, 0);
            const C fwv = kappa * (id - G[mu][s][t]);
            const C bwv = kappa * (id + G[mu][s][t]);
            W(2 * mu + 1, o, i, s, t) = make_float2((float)fwv.real(), (float)fwv.imag());
            W(2 * mu + 2, o, i, s, t) = make_float2((float)bwv.real(), (float)bwv.imag());
          }
    };
    auto set_id = [&](int o, int i) {
      for (int s = 0; s < 4; ++s) W(0, o, i, s, s) = make_float2(1.f, 0.f);
    };
    if (l == 0) {
      set_id(0, 0);
      set_id(1, 0);
      set_hop(1, 0);
    } else if (l < kNL - 1) {
      set_id(0, 0);
      set_id(1, 0);
      set_hop(1, 1);
    } else {
      set_id(0, 0);
      set_hop(0, 1);
    }
    const double sc = noise * kappa;
    for (auto& v : w) {
      v.x += (float)(sc * nd(rng));
      v.y += (float)(sc * nd(rng));
    }
    II_HIP(ctx, hipMemcpyAsync(P.w, w.data(), P.nw * sizeof(float2), hipMemcpyHostToDevice,
                               ctx->stream));
    II_HIP(ctx, hipMemsetAsync(P.m, 0, P.nw * sizeof(float2), ctx->stream));
    II_HIP(ctx, hipMemsetAsync(P.v, 0, P.nw * sizeof(float2), ctx->stream));
  }
  N->adam_step = 0;
  II_CHECK(build_frags(ctx, N));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  N->init = true;
  return II_OK;
}

ii_status ii_net_set_weights(ii_context* ctx, int layer, const float* w) {
  if (!ctx || !w || layer < 0 || layer >= kNL) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  Net* N = nullptr;
  II_CHECK(net_get(ctx, &N));
  LayerParams& P = N->L[layer];
  II_HIP(ctx, hipMemcpyAsync(P.w, w, P.nw * sizeof(float2), hipMemcpyHostToDevice, ctx->stream));
  II_CHECK(build_frags(ctx, N));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  N->init = true;
  return II_OK;
}

ii_status ii_net_get_weights(ii_context* ctx, int layer, float* w) {
  if (!ctx || !w || layer < 0 || layer >= kNL) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  Net* N = nullptr;
  II_CHECK(net_get(ctx, &N));
  LayerParams& P = N->L[layer];
  II_HIP(ctx, hipMemcpyAsync(w, P.w, P.nw * sizeof(float2), hipMemcpyDeviceToHost, ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return II_OK;
}

ii_status ii_net_layer_apply(ii_context* ctx, int layer, int adjoint, const double* in,
                             double* out) {
  if (!ctx || !in || !out || layer < 0 || layer >= kNL) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  Net* N = nullptr;
  II_CHECK(net_get(ctx, &N));
  const Geom& g = ctx->g;
  const LayerParams& P = N->L[layer];
  const int cin = adjoint ? P.cout : P.cin, cout = adjoint ? P.cin : P.cout;
  const size_t nin = (size_t)cin * 24 * g.V, nout = (size_t)cout * 24 * g.V;
  II_CHECK(stage(ctx, N, std::max(nin, nout)));
  // input channels in grad[0] (8-channel capacity), output in grad[1]
  II_HIP(ctx, hipMemcpyAsync(N->stage, in, nin * sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
  hipLaunchKernelGGL(k_net_pack_lex, dim3(blocks_for((long)cin * g.V)), dim3(kBlock), 0,
                     ctx->stream, g, N->stage, N->grad[0], cin);
  II_HIP(ctx, hipGetLastError());
  II_CHECK(run_layer(ctx, N, layer, adjoint != 0, N->grad[0], N->grad[1]));
  hipLaunchKernelGGL(k_net_unpack_lex, dim3(blocks_for((long)cout * g.V)), dim3(kBlock), 0,
                     ctx->stream, g, N->grad[1], N->stage, cout);
  II_HIP(ctx, hipGetLastError());
  II_HIP(ctx, hipMemcpyAsync(out, N->stage, nout * sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return II_OK;
}

ii_status ii_net_layer_wgrad(ii_context* ctx, int layer, const double* in, const double* grad_out,
                             float* gw) {
  if (!ctx || !in || !grad_out || !gw || layer < 0 || layer >= kNL) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  Net* N = nullptr;
  II_CHECK(net_get(ctx, &N));
  const Geom& g = ctx->g;
  LayerParams& P = N->L[layer];
  const size_t nin = (size_t)P.cin * 24 * g.V, nout = (size_t)P.cout * 24 * g.V;
  II_CHECK(stage(ctx, N, nin + nout));
  II_HIP(ctx, hipMemcpyAsync(N->stage, in, nin * sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
  II_HIP(ctx, hipMemcpyAsync(N->stage + nin, grad_out, nout * sizeof(double),
                             hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_net_pack_lex, dim3(blocks_for((long)P.cin * g.V)), dim3(kBlock), 0,
                     ctx->stream, g, N->stage, N->grad[0], P.cin);
  hipLaunchKernelGGL(k_net_pack_lex, dim3(blocks_for((long)P.cout * g.V)), dim3(kBlock), 0,
                     ctx->stream, g, N->stage + nin, N->grad[1], P.cout);
  II_HIP(ctx, hipGetLastError());
  II_HIP(ctx, hipMemsetAsync(P.gacc, 0, P.nw * sizeof(float2), ctx->stream));
  II_CHECK(run_wgrad(ctx, N, layer, N->grad[0], N->grad[1], P.gacc, 1.0f));
  II_HIP(ctx, hipMemcpyAsync(gw, P.gacc, P.nw * sizeof(float2), hipMemcpyDeviceToHost,
                             ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return II_OK;
}

ii_status ii_net_loss_grad(ii_context* ctx, int npairs, const double* b_even, const double* y_even,
                           double* loss, float* grads) {
  if (!ctx || npairs < 1 || !b_even || !y_even || !loss) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  Net* N = nullptr;
  II_CHECK(net_get(ctx, &N));
  if (!N->init) return set_error(ctx, II_ERR_STATE, "network not initialised");
  const Geom& g = ctx->g;
  const size_t nh = (size_t)24 * g.Vh;
  const size_t he = half_elems(g);
  II_CHECK(stage(ctx, N, 2 * nh));
  double2 *b64 = nullptr, *y64 = nullptr;
  II_CHECK(dev_alloc(ctx, &N->fw[3], he));
  II_CHECK(dev_alloc(ctx, &N->fw[4], he));
  b64 = N->fw[3];
  y64 = N->fw[4];
  float2 *bfull = nullptr, *yhalf = nullptr;
  II_CHECK(dev_alloc(ctx, &N->fw[5], he * 2));  // FP32 storage: full b (24*stride float2) + y
  bfull = (float2*)N->fw[5];
  yhalf = bfull + 24 * (size_t)g.stride;
  II_CHECK(zero_grads(ctx, N));
  double L = 0.0;
  for (int k = 0; k < npairs; ++k) {
    II_HIP(ctx, hipMemcpyAsync(N->stage, b_even + k * nh, nh * sizeof(double),
                               hipMemcpyHostToDevice, ctx->stream));
    II_HIP(ctx, hipMemcpyAsync(N->stage + nh, y_even + k * nh, nh * sizeof(double),
                               hipMemcpyHostToDevice, ctx->stream));
    II_CHECK(pack_half_spinor(ctx, N->stage, b64));
    II_CHECK(pack_half_spinor(ctx, N->stage + nh, y64));
    hipLaunchKernelGGL(k_net_to32, dim3(blocks_for(g.Vh)), dim3(kBlock), 0, ctx->stream, g, b64,
                       bfull, 1);
    hipLaunchKernelGGL(k_net_to32, dim3(blocks_for(g.Vh)), dim3(kBlock), 0, ctx->stream, g, y64,
                       yhalf, 0);
    II_HIP(ctx, hipGetLastError());
    double yy[2];
    II_CHECK(cdot(ctx, N, y64, y64, yy));
    II_CHECK(pair_loss(ctx, N, bfull, yhalf, 1.0 / yy[0], 1.0f, &L));
    II_CHECK(backward(ctx, N, 1.0f));
  }
  *loss = L;
  if (grads) {
    size_t off = 0;
    for (int l = 0; l < kNL; ++l) {
      II_HIP(ctx, hipMemcpyAsync(grads + 2 * off, N->L[l].gacc, N->L[l].nw * sizeof(float2),
                                 hipMemcpyDeviceToHost, ctx->stream));
      off += N->L[l].nw;
    }
  }
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return II_OK;
}

ii_status ii_net_apply(ii_context* ctx, const double* r_even, double* z_even) {
  if (!ctx || !r_even || !z_even) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  Net* N = nullptr;
  II_CHECK(net_get(ctx, &N));
  if (!N->init) return set_error(ctx, II_ERR_STATE, "network not initialised");
  const Geom& g = ctx->g;
  const size_t nh = (size_t)24 * g.Vh, he = half_elems(g);
  II_CHECK(stage(ctx, N, nh));
  II_CHECK(dev_alloc(ctx, &N->fw[3], he));
  II_CHECK(dev_alloc(ctx, &N->fw[4], he));
  II_HIP(ctx, hipMemcpyAsync(N->stage, r_even, nh * sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
  II_CHECK(pack_half_spinor(ctx, N->stage, N->fw[3]));
  II_CHECK(apply_M(ctx, N, N->fw[3], N->fw[4]));
  II_CHECK(unpack_half_spinor(ctx, N->fw[4], N->stage));
  II_HIP(ctx, hipMemcpyAsync(z_even, N->stage, nh * sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return II_OK;
}

ii_status ii_net_train(ii_context* ctx, double kappa, const ii_train_params* p,
                       ii_train_report* rep) {
  if (!ctx || !p || p->npairs < 1 || p->epochs < 0 || p->batch < 1) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  Net* N = nullptr;
  II_CHECK(net_get(ctx, &N));
  if (!p->keep_weights || !N->init) II_CHECK(ii_net_init(ctx, kappa, 0.05, p->seed));
  const Geom& g = ctx->g;
  const size_t he = half_elems(g), ce = chan_elems(g);
  auto t0 = std::chrono::steady_clock::now();
  // ---- pairs from the plain FP64 solver, 12 at a time
  if (N->npairs < p->npairs) {
    dfree(N->pb);
    dfree(N->py);
    N->pb = N->py = nullptr;
    II_CHECK(dev_alloc(ctx, &N->pb, ce * p->npairs));
    II_CHECK(dev_alloc(ctx, &N->py, (size_t)12 * g.stride * p->npairs));
  }
  N->npairs = p->npairs;
  N->inv_ny.assign(p->npairs, 0.0);
  {
    const int nbat = 12;
    ii_field src, sol;
    src.ctx = sol.ctx = ctx;
    src.subset = sol.subset = II_FULL;
    src.nrhs = sol.nrhs = nbat;
    II_HIP(ctx, hipMalloc(&src.data, (size_t)nbat * 2 * he * sizeof(double2)));
    ii_status st = II_OK;
    if (hipMalloc(&sol.data, (size_t)nbat * 2 * he * sizeof(double2)) != hipSuccess) {
      dfree(src.data);
      return set_error(ctx, II_ERR_DEVICE, "out of device memory for training pairs");
    }
    ii_prop_params pp{};
    pp.tol = p->pair_tol > 0 ? p->pair_tol : 1e-8;
    pp.maxiter = 100000;
    pp.precision = II_PREC_DOUBLE;
    pp.batched = 1;
    std::vector<ii_solver_output> outs(nbat);
    for (int k0 = 0; k0 < p->npairs && st == II_OK; k0 += nbat) {
      const int nr = std::min(nbat, p->npairs - k0);
      src.nrhs = sol.nrhs = nr;
      if (hipMemsetAsync(src.data, 0, (size_t)nr * 2 * he * sizeof(double2), ctx->stream) !=
          hipSuccess) {
        st = set_error(ctx, II_ERR_DEVICE, "memset failed");
        break;
      }
      for (int k = 0; k < nr; ++k)
        hipLaunchKernelGGL(k_net_gauss, dim3(blocks_for(g.Vh)), dim3(kBlock), 0, ctx->stream, g,
                           (uint64_t)p->seed, k0 + k, src.data + (size_t)k * 2 * he);
      st = propagator(ctx, kappa, &pp, &src, &sol, outs.data());
      if (st != II_OK)
        st = set_error(ctx, st, "training pairs " + std::to_string(k0) + ".." +
                                    std::to_string(k0 + nr - 1) + ": " + ctx->err);
      for (int k = 0; k < nr && st == II_OK; ++k) {
        const double2* b = src.data + (size_t)k * 2 * he;
        const double2* y = sol.data + (size_t)k * 2 * he;
        hipLaunchKernelGGL(k_net_to32, dim3(blocks_for(g.Vh)), dim3(kBlock), 0, ctx->stream, g,
                           b, N->pb + (size_t)(k0 + k) * ce, 1);
        hipLaunchKernelGGL(k_net_to32, dim3(blocks_for(g.Vh)), dim3(kBlock), 0, ctx->stream, g,
                           y, N->py + (size_t)(k0 + k) * 12 * g.stride, 0);
        double yy[2];
        st = cdot(ctx, N, y, y, yy);
        if (st == II_OK) N->inv_ny[k0 + k] = 1.0 / yy[0];
      }
    }
    dfree(src.data);
    dfree(sol.data);
    II_CHECK(st);
  }
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  auto t1 = std::chrono::steady_clock::now();
  // ---- Adam on mean_k |M b_k - y_k|^2/|y_k|^2 over minibatches
  std::mt19937_64 rng(p->seed ^ 0x7472616e);
  std::vector<int> order(p->npairs);
  for (int k = 0; k < p->npairs; ++k) order[k] = k;
  double first = -1.0, last = -1.0;
  int steps = 0;
  for (int ep = 0; ep < p->epochs; ++ep) {
    std::shuffle(order.begin(), order.end(), rng);
    double eloss = 0.0;
    for (int k0 = 0; k0 < p->npairs; k0 += p->batch) {
      const int nb = std::min(p->batch, p->npairs - k0);
      II_CHECK(zero_grads(ctx, N));
      for (int j = 0; j < nb; ++j) {
        const int k = order[k0 + j];
        II_CHECK(pair_loss(ctx, N, N->pb + (size_t)k * ce, N->py + (size_t)k * 12 * g.stride,
                           N->inv_ny[k], 1.0f / nb, &eloss));
        II_CHECK(backward(ctx, N, 1.0f));
      }
      II_CHECK(adam_step(ctx, N, (float)p->lr));
      ++steps;
    }
    eloss /= p->npairs;
    if (ep == 0) first = eloss;
    last = eloss;
  }
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  auto t2 = std::chrono::steady_clock::now();
  if (rep) {
    rep->loss_first = first;
    rep->loss_last = last;
    rep->pairs_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    rep->train_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    rep->steps = steps;
  }
  return II_OK;
}

ii_status ii_field_gaussian(ii_field* f, int rhs, unsigned long long seed, int stream) {
  if (!f || !f->ctx || rhs < 0 || rhs >= f->nrhs) return II_ERR_BAD_ARG;
  ii_context* ctx = f->ctx;
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  const size_t he = half_elems(g);
  const int nblk = f->subset == II_FULL ? 2 : 1;
  double2* base = f->data + (size_t)rhs * nblk * he;
  II_HIP(ctx, hipMemsetAsync(base, 0, nblk * he * sizeof(double2), ctx->stream));
  hipLaunchKernelGGL(k_net_gauss, dim3(blocks_for(g.Vh)), dim3(kBlock), 0, ctx->stream, g,
                     (uint64_t)seed, stream, base);
  II_HIP(ctx, hipGetLastError());
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return II_OK;
}

ii_status ii_net_save(ii_context* ctx, const char* path) {
  if (!ctx || !path) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  Net* N = nullptr;
  II_CHECK(net_get(ctx, &N));
  if (!N->init) return set_error(ctx, II_ERR_STATE, "network not initialised");
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_error(ctx, II_ERR_BAD_ARG, std::string("cannot write ") + path);
  const char magic[8] = {'I', 'I', 'N', 'E', 'T', '0', '0', '1'};
  bool ok = std::fwrite(magic, 1, 8, f) == 8;
  for (int l = 0; l < kNL && ok; ++l) {
    const LayerParams& P = N->L[l];
    std::vector<float2> w(P.nw);
    if (hipMemcpy(w.data(), P.w, P.nw * sizeof(float2), hipMemcpyDeviceToHost) != hipSuccess) {
      ok = false;
      break;
    }
    int32_t hdr[2] = {P.cin, P.cout};
    ok = std::fwrite(hdr, sizeof(hdr), 1, f) == 1 &&
         std::fwrite(w.data(), sizeof(float2), P.nw, f) == P.nw;
  }
  std::fclose(f);
  if (!ok) return set_error(ctx, II_ERR_DEVICE, std::string("failed writing ") + path);
  return II_OK;
}

ii_status ii_net_load(ii_context* ctx, const char* path) {
  if (!ctx || !path) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  Net* N = nullptr;
  II_CHECK(net_get(ctx, &N));
  FILE* f = std::fopen(path, "rb");
  if (!f) return set_error(ctx, II_ERR_BAD_ARG, std::string("cannot read ") + path);
  char magic[8];
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "IINET001", 8) == 0;
  std::vector<std::vector<float2>> ws(kNL);
  for (int l = 0; l < kNL && ok; ++l) {
    const LayerParams& P = N->L[l];
    int32_t hdr[2];
    ok = std::fread(hdr, sizeof(hdr), 1, f) == 1 && hdr[0] == P.cin && hdr[1] == P.cout;
    if (!ok) break;
    ws[l].resize(P.nw);
    ok = std::fread(ws[l].data(), sizeof(float2), P.nw, f) == P.nw;
  }
  std::fclose(f);
  if (!ok) return set_error(ctx, II_ERR_BAD_ARG, std::string("not a network file: ") + path);
  for (int l = 0; l < kNL; ++l)
    II_HIP(ctx, hipMemcpy(N->L[l].w, ws[l].data(), N->L[l].nw * sizeof(float2),
                          hipMemcpyHostToDevice));
  II_CHECK(build_frags(ctx, N));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  N->init = true;
  return II_OK;
}

ii_status ii_net_bench(ii_context* ctx, int nrep, double* layer_ms) {
  if (!ctx || !layer_ms || nrep < 1) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  Net* N = nullptr;
  II_CHECK(net_get(ctx, &N));
  if (!N->init) return set_error(ctx, II_ERR_STATE, "network not initialised");
  const Geom& g = ctx->g;
  const size_t ce = chan_elems(g);
  for (int l = 0; l <= kNL; ++l)
    II_HIP(ctx, hipMemsetAsync(N->act[l], 0, ce * kWidth[l] * sizeof(float2), ctx->stream));
  for (int l = 0; l < kNL; ++l) {
    const int npar = l == kNL - 1 ? 1 : 2;
    II_CHECK(run_layer(ctx, N, l, false, N->act[l], N->act[l + 1], 0, npar));
    II_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int r = 0; r < nrep; ++r)
      II_CHECK(run_layer(ctx, N, l, false, N->act[l], N->act[l + 1], 0, npar));
    II_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    II_HIP(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    II_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    layer_ms[l] = ms / nrep;
  }
  return II_OK;
}

ii_status ii_fgcr_solve(ii_context* ctx, double kappa, double* x_even, const double* b_even,
                        double tol, int maxiter, int nkrylov, int use_net,
                        ii_solver_output* out) {
  if (!ctx || !x_even || !b_even || tol <= 0 || maxiter < 1) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  Net* N = nullptr;
  II_CHECK(net_get(ctx, &N));
  const Geom& g = ctx->g;
  const size_t nh = (size_t)24 * g.Vh, he = half_elems(g);
  II_CHECK(stage(ctx, N, nh));
  II_CHECK(dev_alloc(ctx, &N->fw[3], he));
  II_CHECK(dev_alloc(ctx, &N->fw[4], he));
  double2 *b = N->fw[3], *x = N->fw[4];
  II_HIP(ctx, hipMemcpyAsync(N->stage, b_even, nh * sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
  II_CHECK(pack_half_spinor(ctx, N->stage, b));
  II_CHECK(zero64(ctx, x));
  ii_solver_output o{};
  ii_status st = solve_even(ctx, kappa, x, b, tol, maxiter, nkrylov, use_net != 0, &o);
  if (out) *out = o;
  II_CHECK(st);
  II_CHECK(unpack_half_spinor(ctx, x, N->stage));
  II_HIP(ctx, hipMemcpyAsync(x_even, N->stage, nh * sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (out) *out = o;
  return II_OK;
}

}  // extern "C"

namespace ii {
// Propagator columns through the flexible GCR with the trained network
// (Inverter_WilsonLike::invert structure, eoUtils.cpp:43-80):
// bhat = b_e + kappa H_eo b_o, Dhat y = bhat, x_o = b_o + kappa H_oe y.
// The Schur residual equals the full-D residual on the even sites and
// vanishes on the odd ones, so |bhat - Dhat y| <= tol |b| is the target.
ii_status learned_propagator(ii_context* ctx, double kappa, const ii_prop_params* p,
                             const ii_field* src, ii_field* sol, ii_solver_output* outs) {
  Net* N = nullptr;
  II_CHECK(net_get(ctx, &N));
  if (!N->init) return set_error(ctx, II_ERR_STATE, "no trained preconditioner on this context");
  const Geom& g = ctx->g;
  const size_t he = half_elems(g);
  for (int i = 3; i < 5; ++i) II_CHECK(dev_alloc(ctx, &N->fw[i], he));
  double2 *bh = N->fw[3], *y = N->fw[4];
  const double tol = p->tol > 0 ? p->tol : 1e-10;
  const int maxiter = p->maxiter > 0 ? p->maxiter : 100000;
  bool failed = false;
  for (int k = 0; k < src->nrhs; ++k) {
    const double2* be = src->data + (size_t)k * 2 * he;
    const double2* bo = be + he;
    double2* xe = sol->data + (size_t)k * 2 * he;
    double2* xo = xe + he;
    auto t0 = std::chrono::steady_clock::now();
    double nbe[2], nbo[2], nbh[2];
    II_CHECK(cdot(ctx, N, be, be, nbe));
    II_CHECK(cdot(ctx, N, bo, bo, nbo));
    HopSpec h;
    h.out_par = 0;
    h.a = kappa;
    h.xpay = true;
    II_CHECK(hop64(ctx, h, bo, bh, be, nullptr));
    II_CHECK(cdot(ctx, N, bh, bh, nbh));
    II_CHECK(zero64(ctx, y));
    const double tol_hat = tol * std::sqrt((nbe[0] + nbo[0]) / nbh[0]);
    ii_solver_output o{};
    ii_status st = solve_even(ctx, kappa, y, bh, tol_hat, maxiter, 16, true, &o);
    if (st == II_ERR_NOT_CONVERGED) {
      failed = true;
      ctx->err.clear();
    } else {
      II_CHECK(st);
    }
    h.out_par = 1;
    II_CHECK(hop64(ctx, h, y, xo, bo, nullptr));
    II_CHECK(copy64(ctx, xe, y));
    II_HIP(ctx, hipStreamSynchronize(ctx->stream));
    o.timing_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    o.diff = o.true_residual * std::sqrt(nbh[0] / (nbe[0] + nbo[0]));
    o.true_residual = -1.0;
    outs[k] = o;
  }
  if (failed)
    return set_error(ctx, II_ERR_NOT_CONVERGED, "learned propagator did not reach the tolerance");
  return II_OK;
}
}  // namespace ii
