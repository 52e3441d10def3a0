// Even/odd Wilson hop kernels for gfx950 and their host dispatch.
//
// k_hop computes, for every half site of parity out_par and every rhs,
//   out = a * [g5] H [g5] in  (+ x)          (XPAY adds the own-parity field x)
// and per-block partial sums for the CG scalars as the epilogue mode asks
// (HopEpilogue: Re<d, out>, |out|^2, or the residual update r -= alpha out
// with |r|^2).  With a = -kappa this is Dirac_Wilson::mult for the EO
// constructors (dirac_wilson.cpp:28-32, 107-113); with g5 folded on both
// sides it is mult_dag (:34-36); with XPAY and a = -kappa^2 on the second
// hop it is Dirac_Wilson_EvenOdd::mult/mult_dag
// (dirac_wilson_EvenOdd.hpp:79-92).
#include "ii_ops.h"
#include "ii_stencil.h"

#include <algorithm>
#include <cstdio>
#include <type_traits>
#include <cstdlib>
#include <vector>

namespace ii {

// The epilogue (HopEpilogue) of one site's output o of rhs-offset accessors
// dr (the dv field) and outr; the site's partial sums go to ds[0..2].  xs
// holds the XPAY term the kernel added (kEpiTS).  16- and 32-bit fields: the
// 24-term site sums in FP32, the sums over sites in FP64.
template <int EPI, typename R, typename C, typename SpX, typename StOut>
__device__ __forceinline__ void hop_epilogue(C o[12], const C* xs, int hs, const SpX& dr,
                                             const StOut& outr, double alpha, double ds[3]) {
  constexpr bool kF64 = std::is_same<R, double>::value;
  if constexpr (EPI == kEpiDot) {
    if constexpr (kF64) {
      const double sc = (double)dr.scale(hs);
      double sum = 0.0;
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        auto dvv = dr.ld(hs, k);
        sum += ((double)dvv.x * (double)o[k].x + (double)dvv.y * (double)o[k].y) * sc;
      }
      ds[0] = sum;
    } else {
      float sd = 0.0f;
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        auto dvv = dr.ld(hs, k);
        sd = fmaf((float)dvv.x, o[k].x, sd);
        sd = fmaf((float)dvv.y, o[k].y, sd);
      }
      ds[0] = (double)sd * (double)dr.scale(hs);
    }
  } else if constexpr (EPI == kEpiNorm) {
    R sd = R(0);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      sd = fma(o[k].x, o[k].x, sd);
      sd = fma(o[k].y, o[k].y, sd);
    }
    ds[0] = (double)sd;
  } else if constexpr (EPI == kEpiRUpd) {
    const R sc = (R)dr.scale(hs), al = (R)alpha;
    R sd = R(0);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      auto dvv = dr.ld(hs, k);
      o[k].x = fma(-al, o[k].x, (R)dvv.x * sc);
      o[k].y = fma(-al, o[k].y, (R)dvv.y * sc);
      sd = fma(o[k].x, o[k].x, sd);
      sd = fma(o[k].y, o[k].y, sd);
    }
    ds[0] = (double)sd;
  } else if constexpr (EPI == kEpiCDot) {
    // sum conj(d) o: re = d.x o.x + d.y o.y, im = d.x o.y - d.y o.x
    R re = R(0), im = R(0);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      auto dvv = dr.ld(hs, k);
      const R dx = (R)dvv.x, dy = (R)dvv.y;
      re = fma(dx, o[k].x, fma(dy, o[k].y, re));
      im = fma(dx, o[k].y, fma(-dy, o[k].x, im));
    }
    const double sc = (double)dr.scale(hs);
    ds[0] = (double)re * sc;
    ds[1] = (double)im * sc;
  } else if constexpr (EPI == kEpiTS) {
    // sum conj(o) x and |o|^2
    R re = R(0), im = R(0), nn = R(0);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      re = fma(o[k].x, xs[k].x, fma(o[k].y, xs[k].y, re));
      im = fma(o[k].x, xs[k].y, fma(-o[k].y, xs[k].x, im));
      nn = fma(o[k].x, o[k].x, fma(o[k].y, o[k].y, nn));
    }
    ds[0] = (double)re;
    ds[1] = (double)im;
    ds[2] = (double)nn;
  }
  outr.st(hs, o);
}

// BS threads per block (kBlock for the variants with partial sums, which are
// indexed by kBlock-site ranges; the autotuner picks it for the others).
template <bool G5IN, bool G5OUT, bool XPAY, int EPI, typename SpIn, typename Gauge,
          typename StOut, typename SpX, int BS = kBlock>
__global__ void __launch_bounds__(BS)
k_hop(Geom g, SpIn in, Gauge gauge, StOut out, SpX x, SpX dv, int out_par,
      typename SpIn::R a, bool tbc_flip, double* __restrict__ partials,
      const int* __restrict__ skip, int skip_stride, HopOrder order,
      const double* __restrict__ coef, int coef_stride) {
  static_assert(EPI == kEpiStore || BS == kBlock, "partial sums are per kBlock sites");
  using C = typename SpIn::C;
  const int nblk = gridDim.x;
  const int base = hop_block_base(g, order, blockIdx.x, nblk, BS);
  const int hs = base + threadIdx.x;
  const int rhs = blockIdx.y;
  if (skip && skip[rhs * skip_stride]) return;  // uniform per block
  double dsum[3] = {0.0, 0.0, 0.0};
  if (hs < g.Vh) {
    const SpIn inr = in.rhs(rhs);
    C acc[4][3];
    wilson_hop_site<G5IN>(acc, g, inr, gauge, hs, out_par, tbc_flip);
    C o[12];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int so = G5OUT ? (s ^ 2) : s;
#pragma unroll
      for (int c = 0; c < 3; ++c) o[3 * so + c] = cscale(acc[s][c], a);
    }
    C xsv[12];
    if (XPAY) {
      const SpX xr = x.rhs(rhs);
      const auto xs = xr.scale(hs);
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        auto xv = xr.ld(hs, k);
        xsv[k].x = (typename SpIn::R)(xv.x * xs);
        xsv[k].y = (typename SpIn::R)(xv.y * xs);
        o[k].x += xsv[k].x;
        o[k].y += xsv[k].y;
      }
    }
    hop_epilogue<EPI, typename SpIn::R>(o, xsv, hs, dv.rhs(rhs), out.rhs(rhs),
                                        EPI == kEpiRUpd ? coef[rhs * coef_stride] : 0.0, dsum);
  }
  constexpr int NS = epi_sums(EPI);
  const size_t plane = (size_t)gridDim.y * nblk;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    double s = block_sum(dsum[j]);
    // indexed by the site range, so sums do not depend on the traversal
    if (threadIdx.x == 0) partials[j * plane + (size_t)rhs * nblk + base / kBlock] = s;
  }
}

// Multi-rhs hop.  A block is TS sites x blockDim.y right-hand sides.  All
// threads first stage the tile's neighbour indices, time-boundary flags
// and its 8 links per site in LDS, so each link is read from HBM once per
// site for all rhs; then thread (tx, ty) applies the stencil to site tx of
// rhs ty (one wavefront per rhs for TS = 64).  Partial dot products are
// per (rhs, tile): partials[rhs * gridDim.x + blockIdx.x].
template <bool G5IN, bool G5OUT, bool XPAY, int EPI, typename SpIn, typename Gauge,
          typename StOut, typename SpX, int TS, int MAXT>
__global__ void __launch_bounds__(MAXT)
k_hop_mrhs(Geom g, SpIn in, Gauge gauge, StOut out, SpX x, SpX dv, int out_par,
           typename SpIn::R a, bool tbc_flip, double* __restrict__ partials,
           const int* __restrict__ skip, int skip_stride, int nrhs, HopOrder order,
           const double* __restrict__ coef, int coef_stride) {
  using C = typename SpIn::C;
  using UC = typename Gauge::C;
  // FP16 fields and links: colour products with v_dot2_f32_f16 on links
  // staged as MDot2 entries of U (forward) or U^dag (backward)
  constexpr bool kDot2 = std::is_same<SpIn, SpLo<__half2>>::value &&
                         std::is_same<Gauge, GaugeLoR12<__half2>>::value;
  using LC = typename std::conditional<kDot2, MDot2, UC>::type;
  __shared__ LC lnk[8 * 9 * TS];  // [direction][row*3+col][site], full matrices
  __shared__ int nbi[8][TS];      // neighbour half index per direction
  __shared__ int flg[TS];         // bit 0: t+ link flipped, bit 1: t- link flipped
  const int ntile = gridDim.x;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int nthr = blockDim.x * blockDim.y;
  const int tid = tx + ty * blockDim.x;
  const int hs0 = hop_block_base(g, order, blockIdx.x, ntile, TS);
  const int op = 1 - out_par;
  for (int i = tid; i < TS; i += nthr) {
    const int hs = hs0 + i;
    if (hs < g.Vh) {
      const SiteCoord c = decode_site(g, hs, out_par);
#pragma unroll
      for (int mu = 0; mu < 4; ++mu) {
        nbi[2 * mu][i] = neighbour(g, c, hs, mu, true);
        nbi[2 * mu + 1][i] = neighbour(g, c, hs, mu, false);
      }
      flg[i] = (tbc_flip && c.t == g.L[3] - 1 ? 1 : 0) | (tbc_flip && c.t == 0 ? 2 : 0);
    }
  }
  __syncthreads();
  // one thread per (direction, site): load the stored link, rebuild the full
  // 3x3 matrix once (reconstruct-12, time-boundary sign) in the compute
  // precision, so the rhs threads only read it
  for (int i = tid; i < 8 * TS; i += nthr) {
    const int s = i % TS;
    const int d = __builtin_amdgcn_readfirstlane(i / TS);  // one direction per wavefront
    const int hs = hs0 + s;
    if (hs < g.Vh) {
      UC U[3][3];
      const bool fl = (d == 6 && (flg[s] & 1)) || (d == 7 && (flg[s] & 2));
      gauge.ld((d & 1) ? op : out_par, d >> 1, (d & 1) ? nbi[d][s] : hs, fl, U);
      if constexpr (kDot2) {
        // M = U (forward, even d) or U^dag (backward): M_ab = conj(U_ba)
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          const int ra = k / 3, cb = k % 3;
          lnk[(d * 9 + k) * TS + s] = (d & 1) ? mdot2(U[cb][ra].x, -U[cb][ra].y)
                                              : mdot2(U[ra][cb].x, U[ra][cb].y);
        }
      } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) lnk[(d * 9 + k) * TS + s] = U[k / 3][k % 3];
      }
    }
  }
  __syncthreads();
  const int hs = hs0 + tx;
  // TS = 64 = one wavefront per rhs, so rhs is wave-uniform and the rhs
  // base pointers live in scalar registers
  static_assert(TS == 64, "one wavefront per rhs row");
  const int rhs = __builtin_amdgcn_readfirstlane(blockIdx.y * blockDim.y + ty);
  const bool live = rhs < nrhs && !(skip && skip[rhs * skip_stride]);
  double dsum[3] = {0.0, 0.0, 0.0};
  if (live && hs < g.Vh) {
    const SpIn inr = in.rhs(rhs);
    C acc[4][3];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[s][c] = cmk<C>(0, 0);
    if constexpr (kDot2) {
      // one direction at a time: issuing the next direction's loads before
      // this one's products (two directions in flight) measured slower,
      // 0.59 against 0.556 ms for 12 rhs at 32^3x64
      RawDir r0;
      RawSite own;
      load_site_raw(inr, hs, own);
#define II_LD(MU, FWD, D, R) load_dir_tile<MU, FWD, G5IN>(inr, nbi[D][tx], hs0, tx, own, R);
#define II_CP(MU, FWD, D, R)                                                             \
  hop_dir_dot2_raw<MU, FWD, G5IN>(acc, R, [&](int k) { return lnk[(D * 9 + k) * TS + tx]; }); \
  pin_acc(acc);
      II_LD(0, true, 0, r0) II_CP(0, true, 0, r0)
      II_LD(0, false, 1, r0) II_CP(0, false, 1, r0)
      II_LD(1, true, 2, r0) II_CP(1, true, 2, r0)
      II_LD(1, false, 3, r0) II_CP(1, false, 3, r0)
      II_LD(2, true, 4, r0) II_CP(2, true, 4, r0)
      II_LD(2, false, 5, r0) II_CP(2, false, 5, r0)
      II_LD(3, true, 6, r0) II_CP(3, true, 6, r0)
      II_LD(3, false, 7, r0) II_CP(3, false, 7, r0)
#undef II_LD
#undef II_CP
    } else {
#define II_MDIR(MU, FWD, D)                                                                 \
  hop_dir_with<MU, FWD, G5IN, UC>(acc, inr, nbi[D][tx], [&](UC U[3][3]) {                   \
    _Pragma("unroll") for (int k = 0; k < 9; ++k) U[k / 3][k % 3] = lnk[(D * 9 + k) * TS + tx]; \
  });                                                                                       \
  pin_acc(acc);
      II_MDIR(0, true, 0)
      II_MDIR(0, false, 1)
      II_MDIR(1, true, 2)
      II_MDIR(1, false, 3)
      II_MDIR(2, true, 4)
      II_MDIR(2, false, 5)
      II_MDIR(3, true, 6)
      II_MDIR(3, false, 7)
#undef II_MDIR
    }
    C o[12];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int so = G5OUT ? (s ^ 2) : s;
#pragma unroll
      for (int c = 0; c < 3; ++c) o[3 * so + c] = cscale(acc[s][c], a);
    }
    C xsv[12];
    if (XPAY) {
      const SpX xr = x.rhs(rhs);
      const auto xs = xr.scale(hs);
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        auto xv = xr.ld(hs, k);
        xsv[k].x = (typename SpIn::R)(xv.x * xs);
        xsv[k].y = (typename SpIn::R)(xv.y * xs);
        o[k].x += xsv[k].x;
        o[k].y += xsv[k].y;
      }
    }
    hop_epilogue<EPI, typename SpIn::R>(o, xsv, hs, dv.rhs(rhs), out.rhs(rhs),
                                        EPI == kEpiRUpd ? coef[rhs * coef_stride] : 0.0, dsum);
  }
  constexpr int NS = epi_sums(EPI);
  const size_t plane = (size_t)nrhs * ntile;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    double v = dsum[j];
    for (int off = TS / 2; off > 0; off >>= 1) v += __shfl_down(v, off, TS);
    if (tx == 0 && live) partials[j * plane + (size_t)rhs * ntile + blockIdx.x] = v;
  }
}

// Strip order (hop_block_base mode 1) over blocks of bs sites with W time
// slices per strip, or the contiguous order where the lattice does not
// tile that way (planes of whole blocks, L_t a multiple of 8 XCDs x W).
HopOrder strip_order(const Geom& g, int W, int bs) {
  const int P = g.Lxh * g.L[1];
  HopOrder o;
  if (W >= 1 && P % bs == 0 && g.L[3] % 8 == 0 && (g.L[3] / 8) % W == 0) {
    o.mode = 1;
    o.W = W;
    o.bpp = P / bs;
  }
  return o;
}

namespace {

// sites per tile of the multi-rhs kernel and its thread limit
template <typename Gauge> struct MrhsShape { static constexpr int TS = 64, MAXT = 768; };
template <> struct MrhsShape<GaugeF64> { static constexpr int TS = 64, MAXT = 768; };
// storage precision (II_PREC_*) of a gauge accessor
template <typename Gauge> struct GaugePrec;
template <> struct GaugePrec<GaugeF64> { static constexpr int value = II_PREC_DOUBLE; };
template <> struct GaugePrec<GaugeF32r12> { static constexpr int value = II_PREC_SINGLE; };
template <> struct GaugePrec<GaugeLoR12<__half2>> { static constexpr int value = II_PREC_HALF; };
template <> struct GaugePrec<GaugeLoR12<__hip_bfloat162>> {
  static constexpr int value = II_PREC_BF16;
};

template <bool G5IN, bool G5OUT, bool XPAY, int EPI, typename SpIn, typename Gauge,
          typename StOut, typename SpX>
ii_status launch_hop(ii_context* ctx, const SpIn& in, const Gauge& gauge, const StOut& out,
                     const SpX& x, const SpX& dv, typename SpIn::R a, const HopSpec& h) {
  const Geom& g = ctx->g;
  bool flip = ctx->tbc_sign < 0;
  if (h.nrhs == 1) {
    const HopOrder& o = ctx->hop_order;
    if constexpr (EPI == kEpiStore && (std::is_same<Gauge, GaugeF64>::value ||
                                       std::is_same<Gauge, GaugeF64Flat>::value)) {
      if (o.bs == 128) {
        dim3 grid(blocks_for(g.Vh, 128), 1);
        hipLaunchKernelGGL((k_hop<G5IN, G5OUT, XPAY, EPI, SpIn, Gauge, StOut, SpX, 128>), grid,
                           dim3(128), 0, ctx->stream, g, in, gauge, out, x, dv, h.out_par, a,
                           flip, h.partials, h.skip, h.skip_stride, o, h.coef, h.coef_stride);
        II_HIP(ctx, hipGetLastError());
        return II_OK;
      }
    }
    HopOrder ok = o;
    ok.bs = kBlock;
    if (ok.mode == 1 && o.bs != kBlock) ok = HopOrder();  // strip orders are built for kBlock
    dim3 grid(blocks_for(g.Vh), 1);
    hipLaunchKernelGGL((k_hop<G5IN, G5OUT, XPAY, EPI, SpIn, Gauge, StOut, SpX>), grid,
                       dim3(kBlock), 0, ctx->stream, g, in, gauge, out, x, dv, h.out_par, a, flip,
                       h.partials, h.skip, h.skip_stride, ok, h.coef, h.coef_stride);
  } else if constexpr (std::is_same<Gauge, GaugeF64Flat>::value) {
    return set_error(ctx, II_ERR_STATE, "the flat FP64 gauge form serves single-rhs hops only");
  } else {
    constexpr int TS = MrhsShape<Gauge>::TS, MAXT = MrhsShape<Gauge>::MAXT;
    constexpr int P = GaugePrec<Gauge>::value;
    if (h.rb == 0 && !h.no_tune && ctx->mrhs_nrhs[P] != h.nrhs) II_CHECK(tune_mrhs(ctx, P, h.nrhs));
    const bool tuned = ctx->mrhs_nrhs[P] == h.nrhs;
    int rb = h.rb > 0 ? h.rb : tuned ? ctx->mrhs_rb[P] : MAXT / TS;
    rb = std::max(1, std::min(rb, std::min(h.nrhs, MAXT / TS)));
    const HopOrder ord = h.order ? *h.order : tuned ? ctx->mrhs_ord[P] : ctx->mrhs_order;
    dim3 grid((g.Vh + TS - 1) / TS, (h.nrhs + rb - 1) / rb);
    hipLaunchKernelGGL((k_hop_mrhs<G5IN, G5OUT, XPAY, EPI, SpIn, Gauge, StOut, SpX, TS, MAXT>),
                       grid, dim3(TS, rb), 0, ctx->stream, g, in, gauge, out, x, dv, h.out_par,
                       a, flip, h.partials, h.skip, h.skip_stride, h.nrhs, ord, h.coef,
                       h.coef_stride);
  }
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

template <typename SpIn, typename Gauge, typename StOut, typename SpX>
ii_status dispatch_hop(ii_context* ctx, const SpIn& in, const Gauge& gauge, const StOut& out,
                       const SpX& x, const SpX& dv, const HopSpec& h) {
  using R = typename SpIn::R;
  const R a = (R)h.a;
  if (h.epi < kEpiStore || h.epi > kEpiTS)
    return set_error(ctx, II_ERR_BAD_ARG, "unknown hop epilogue");
  if (h.epi == kEpiTS && !h.xpay)
    return set_error(ctx, II_ERR_BAD_ARG, "kEpiTS needs the XPAY term");
  if (h.epi == kEpiRUpd && !h.coef)
    return set_error(ctx, II_ERR_BAD_ARG, "residual-update hop without its scalars");
  const int sel = (h.g5in ? 32 : 0) | (h.g5out ? 16 : 0) | (h.xpay ? 8 : 0) | h.epi;
#define II_CASE(A, B, CC, E)                                                                 \
  case ((A) ? 32 : 0) | ((B) ? 16 : 0) | ((CC) ? 8 : 0) | (E):                               \
    return launch_hop<A, B, CC, E>(ctx, in, gauge, out, x, dv, a, h);
  switch (sel) {
    II_CASE(false, false, false, kEpiStore)
    II_CASE(false, false, true, kEpiStore)
    II_CASE(false, false, true, kEpiDot)
    II_CASE(false, false, true, kEpiNorm)
    II_CASE(false, false, true, kEpiRUpd)
    II_CASE(false, false, true, kEpiCDot)
    II_CASE(false, false, true, kEpiTS)
    II_CASE(true, true, false, kEpiStore)
    II_CASE(true, true, true, kEpiStore)
    II_CASE(true, true, true, kEpiDot)
    II_CASE(true, false, false, kEpiStore)
    II_CASE(false, true, false, kEpiStore)
    II_CASE(false, true, true, kEpiStore)
    II_CASE(false, true, true, kEpiDot)
    II_CASE(false, true, true, kEpiRUpd)
    default:
      return set_error(ctx, II_ERR_BAD_ARG, "unsupported hop variant");
  }
#undef II_CASE
}

}  // namespace

int hop_dot_parts(const Geom& g, int nrhs, int precision) {
  if (nrhs == 1) return blocks_for(g.Vh);
  const int ts = precision == II_PREC_DOUBLE ? MrhsShape<GaugeF64>::TS
                                             : MrhsShape<GaugeF32r12>::TS;
  return (g.Vh + ts - 1) / ts;
}

// Autotuner of the single-rhs hop traversal (HopOrder): times every
// admissible order on this lattice with the FP64 hop and keeps the fastest.
// Runs once per context on the first single-rhs FP64 hop, never inside a
// stream capture.  II_HOP_ORDER=0 or II_HOP_ORDER=1:W forces an order.
ii_status tune_hop(ii_context* ctx) {
  ctx->hop_tuned = true;
  const Geom& g = ctx->g;
  const int P = g.Lxh * g.L[1];
  std::vector<HopOrder> cands(1);
  // contiguous order with 128 threads per block (FP64 hops without partial sums; 512
  // would cap the FP64 stencil's 256 registers and spill)
  if (g.Vh >= 8 * 128) {
    HopOrder o;
    o.bs = 128;
    cands.push_back(o);
  }
  const bool tiled = P % kBlock == 0 && g.L[3] % 8 == 0 && g.Vh >= 8 * kBlock;
  if (tiled)
    for (int W : {1, 2, 4, 8, 16})
      if ((g.L[3] / 8) % W == 0) cands.push_back(HopOrder{1, W, P / kBlock});
  if (const char* env = std::getenv("II_HOP_ORDER")) {
    HopOrder o;
    if (env[0] == '1' && tiled) {
      o.mode = 1;
      o.W = env[1] == ':' ? std::atoi(env + 2) : 4;
      o.bpp = P / kBlock;
      if (o.W < 1 || (g.L[3] / 8) % o.W) o = HopOrder();
    }
    ctx->hop_order = o;
    ctx->tune_report = "hop order forced by II_HOP_ORDER";
    return II_OK;
  }
  if (cands.size() == 1) {
    ctx->tune_report = "hop order: contiguous per XCD (lattice not tileable)";
    return II_OK;
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  II_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cs));
  if (cs != hipStreamCaptureStatusNone) {
    ctx->hop_tuned = false;
    return II_OK;
  }
  const size_t he = half_elems(g);
  double2 *a = nullptr, *b = nullptr;
  II_HIP(ctx, hipMalloc(&a, he * sizeof(double2)));
  if (hipMalloc(&b, he * sizeof(double2)) != hipSuccess) {
    (void)hipFree(a);
    return set_error(ctx, II_ERR_DEVICE, "autotuner: out of device memory");
  }
  ii_status st = II_OK;
  std::string rep = "hop order ms:";
  float best = 1e30f;
  HopOrder pick;
  if (hipMemsetAsync(a, 0, he * sizeof(double2), ctx->stream) != hipSuccess)
    st = set_error(ctx, II_ERR_DEVICE, "autotuner: memset failed");
  HopSpec h;
  h.a = -0.125;
  for (size_t i = 0; i < cands.size() && st == II_OK; ++i) {
    ctx->hop_order = cands[i];
    for (int r = 0; r < 2 && st == II_OK; ++r) st = hop64(ctx, h, a, b, nullptr, nullptr);
    if (st != II_OK) break;
    float ms = 0.f;
    if (hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess) st = II_ERR_DEVICE;
    for (int r = 0; r < 10 && st == II_OK; ++r) st = hop64(ctx, h, a, b, nullptr, nullptr);
    if (st == II_OK && (hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess ||
                        hipEventSynchronize(ctx->ev1) != hipSuccess ||
                        hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess))
      st = set_error(ctx, II_ERR_DEVICE, "autotuner: event timing failed");
    ms /= 10.f;
    char buf[64];
    if (cands[i].mode)
      std::snprintf(buf, sizeof buf, " strip%d=%.4f", cands[i].W, ms);
    else
      std::snprintf(buf, sizeof buf, " contig/bs%d=%.4f", cands[i].bs, ms);
    rep += buf;
    if (ms < best) {
      best = ms;
      pick = cands[i];
    }
  }
  (void)hipFree(a);
  (void)hipFree(b);
  ctx->hop_order = pick;
  ctx->tune_report = rep;
  return st;
}

// The multi-rhs hop's launch shapes: rhs per block (the tile's links are
// staged in LDS once per block, so fewer rhs per block re-read them but
// give more, smaller blocks) and the tile order (contiguous per XCD, or
// strips of W time slices, as for the single-rhs hop).  Timed with the
// plain hop (no epilogue) of this precision on scratch fields of nrhs rhs,
// one warm-up and 3 timed launches per candidate.
ii_status tune_mrhs(ii_context* ctx, int P, int nrhs) {
  if (P < 0 || P > 3 || nrhs < 2) return II_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  II_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cs));
  if (cs != hipStreamCaptureStatusNone) return II_OK;
  if (P != II_PREC_DOUBLE && !ctx->gauge_lp[P]) II_CHECK(build_lowprec_gauge(ctx, P));
  const Geom& g = ctx->g;
  constexpr int TS = 64, RBMAX = 768 / 64;
  std::vector<int> rbs;
  for (int rb : {RBMAX, 6, 4, 3, 2})
    if (rb <= std::min(nrhs, RBMAX) && std::find(rbs.begin(), rbs.end(), rb) == rbs.end())
      rbs.push_back(rb);
  if (rbs.empty() || rbs.front() != std::min(nrhs, RBMAX)) rbs.insert(rbs.begin(), std::min(nrhs, RBMAX));
  std::vector<HopOrder> ords(1);
  for (int W : {1, 2, 4}) {
    const HopOrder o = strip_order(g, W, TS);
    if (o.mode == 1) ords.push_back(o);
  }
  // scratch fields: components, then (16-bit) the per-site norms
  const size_t per = P == II_PREC_DOUBLE ? (size_t)12 * g.stride * sizeof(double2)
                     : P == II_PREC_SINGLE ? (size_t)12 * g.stride * sizeof(float2)
                                           : (size_t)12 * g.stride * 4 + (size_t)g.stride * 4;
  void *pa = nullptr, *pb = nullptr;
  II_CHECK(workspace(ctx, 92, per * nrhs, &pa));
  II_CHECK(workspace(ctx, 93, per * nrhs, &pb));
  II_HIP(ctx, hipMemsetAsync(pa, 0, per * nrhs, ctx->stream));
  LoField la, lb;
  la.v = pa;
  lb.v = pb;
  la.nrm = (float*)((char*)pa + (size_t)12 * g.stride * 4 * nrhs);
  lb.nrm = (float*)((char*)pb + (size_t)12 * g.stride * 4 * nrhs);
  std::string rep = std::string("multi-rhs hop ") +
                    (P == II_PREC_DOUBLE ? "fp64" : P == II_PREC_SINGLE ? "fp32"
                                                  : P == II_PREC_HALF ? "fp16" : "bf16") +
                    " x" + std::to_string(nrhs) + " (rhs per block / order) ms:";
  float best = 1e30f;
  int brb = rbs.front();
  HopOrder bord;
  for (int rb : rbs)
    for (const HopOrder& o : ords) {
      HopSpec h;
      h.nrhs = nrhs;
      h.out_par = 1;
      h.a = -0.125;
      h.rb = rb;
      h.order = &o;
      auto run = [&]() -> ii_status {
        switch (P) {
          case II_PREC_DOUBLE: return hop64(ctx, h, (const double2*)pa, (double2*)pb, nullptr, nullptr);
          case II_PREC_SINGLE: return hop32(ctx, h, (const float2*)pa, (float2*)pb, nullptr, nullptr);
          case II_PREC_HALF: return hop_half(ctx, h, la, lb, nullptr, nullptr);
          default: return hop_bf16(ctx, h, la, lb, nullptr, nullptr);
        }
      };
      II_CHECK(run());
      II_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
      for (int i = 0; i < 3; ++i) II_CHECK(run());
      II_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
      II_HIP(ctx, hipEventSynchronize(ctx->ev1));
      float ms = 0.f;
      II_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
      ms /= 3.f;
      char buf[48];
      if (o.mode)
        std::snprintf(buf, sizeof buf, " %d/strip%d=%.4f", rb, o.W, ms);
      else
        std::snprintf(buf, sizeof buf, " %d/contig=%.4f", rb, ms);
      rep += buf;
      if (ms < best) {
        best = ms;
        brb = rb;
        bord = o;
      }
    }
  char buf[48];
  std::snprintf(buf, sizeof buf, " -> %d/%s", brb, bord.mode ? ("strip" + std::to_string(bord.W)).c_str() : "contig");
  rep += buf;
  ctx->mrhs_rb[P] = brb;
  ctx->mrhs_ord[P] = bord;
  ctx->mrhs_nrhs[P] = nrhs;
  ctx->mrhs_report[P] = rep;
  return II_OK;
}

ii_status hop64(ii_context* ctx, const HopSpec& h, const double2* in, double2* out,
                const double2* x, const double2* dv, int rs) {
  const Geom& g = ctx->g;
  if (h.nrhs == 1 && !ctx->hop_tuned && !h.no_tune) II_CHECK(tune_hop(ctx));
  SpF64 sin{in, g.stride, rs};
  StF64 so{out, g.stride};
  SpF64 sx{x ? x : in, g.stride, rs};
  SpF64 sd{dv ? dv : in, g.stride, rs};
  if (h.nrhs == 1 && (unsigned long long)72 * g.stride * sizeof(double2) < (1ull << 32)) {
    GaugeF64Flat gf{ctx->gauge64, g.stride};
    return dispatch_hop(ctx, sin, gf, so, sx, sd, h);
  }
  GaugeF64 gf{ctx->gauge64, g.stride};
  return dispatch_hop(ctx, sin, gf, so, sx, sd, h);
}

ii_status hop32(ii_context* ctx, const HopSpec& h, const float2* in, float2* out,
                const float2* x, const float2* dv) {
  const Geom& g = ctx->g;
  SpF32 sin{in, g.stride};
  GaugeF32r12 gf{(const float2*)ctx->gauge_lp[II_PREC_SINGLE], g.stride};
  StF32 so{out, g.stride};
  SpF32 sx{x ? x : in, g.stride};
  SpF32 sd{dv ? dv : in, g.stride};
  return dispatch_hop(ctx, sin, gf, so, sx, sd, h);
}

template <typename T>
ii_status hop_lo(ii_context* ctx, const HopSpec& h, const LoField& in, const LoField& out,
                 const LoField* x, const LoField* dv) {
  const Geom& g = ctx->g;
  const int prec = std::is_same<T, __half2>::value ? II_PREC_HALF : II_PREC_BF16;
  SpLo<T> sin{(const T*)in.v, in.nrm, g.stride};
  GaugeLoR12<T> gf{(const T*)ctx->gauge_lp[prec], g.stride};
  StLo<T> so{(T*)out.v, out.nrm, g.stride};
  SpLo<T> sx = x ? SpLo<T>{(const T*)x->v, x->nrm, g.stride} : sin;
  SpLo<T> sd = dv ? SpLo<T>{(const T*)dv->v, dv->nrm, g.stride} : sin;
  return dispatch_hop(ctx, sin, gf, so, sx, sd, h);
}

ii_status hop_half(ii_context* ctx, const HopSpec& h, const LoField& in, const LoField& out,
                   const LoField* x, const LoField* dv) {
  return hop_lo<__half2>(ctx, h, in, out, x, dv);
}
ii_status hop_bf16(ii_context* ctx, const HopSpec& h, const LoField& in, const LoField& out,
                   const LoField* x, const LoField* dv) {
  return hop_lo<__hip_bfloat162>(ctx, h, in, out, x, dv);
}

}  // namespace ii
