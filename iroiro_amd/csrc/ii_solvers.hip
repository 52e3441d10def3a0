// Device-resident solvers of the Wilson even/odd solve path.
//
//  * ii_solver_cg      Solver_CG::solve on Fopr_DdagD (solver_CG.cpp:29-160)
//  * ii_multishift_cg  MultiShiftSolver_CG::solve (multiShiftSolver_CG.cpp:10-185)
//  * ii_eo_invert      Inverter_WilsonLike::invert (eoUtils.cpp:43-80)
//  * ii_propagator_solve  the production path: even/odd Schur solve of many
//    right-hand sides at once, CG on Dhat^dag Dhat in FP64, FP32 or 16-bit
//    storage with FP64 reliable updates, stopping on the FP64 true residual.
//
// All solver scalars (alpha, beta, |r|^2, stopping flags, shift recurrences)
// live in device memory and are computed by single-block kernels, so an
// iteration needs no host round trip; the host only polls a small status
// block.  Every kernel of an iteration reads the per-rhs `done` flag and
// returns at once when set, so iterations queued past convergence are free
// and leave the solution untouched: the iteration count is exact.
#include "ii_ops.h"
#include "ii_stencil.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

namespace ii {

// ============================================================ scalars
struct RUScal {
  double rr, pap, alpha, beta;
  double rmax2;    // |r|^2 at the last reliable update
  double target2;  // |r|^2 below which the true residual is checked
  double tol2;     // tol^2 |b|^2 on the true residual
  double delta2;   // reliable-update trigger on |r|^2 / rmax2
  double rho2;     // last true residual^2
  double b2;       // |b|^2 of the full source
  double snorm;    // 1/|b|^2 (reference CG)
  double prec;     // reference CG threshold on rr*snorm
  double restart;  // a recomputed |r|^2 above restart x the recursive one restarts p
  int done;        // iteration kernels skip when set
  int skip_ru;     // reliable-update kernels skip when set (= done || !update)
  int count;       // iterations done
  int iter;        // reference loop index at convergence
  int updates;     // reliable updates done
  int nan;
  int keep_p;      // 0 after a reliable update whose true residual broke with the recursion
  int restarts;    // such search-direction restarts
};
constexpr int kScalInts = sizeof(RUScal) / sizeof(int);

__device__ __forceinline__ bool finite(double v) { return isfinite(v); }

// threads per block of the per-rhs scalar kernels of the propagator solves:
// they sum ~16k tile partials per rhs at 32^3x64
constexpr int kFinThreads = 1024;

// pap -> alpha = rr/pap; with ahist, alpha is also kept at
// ahist[rhs * astride + aidx] (0 for a finished rhs) for the deferred
// solution update
__global__ void k_fin_pap(const double* __restrict__ partials, int nb, RUScal* __restrict__ sc,
                          double* __restrict__ ahist, int aidx, int astride) {
  RUScal& S = sc[blockIdx.x];
  double* ah = ahist ? ahist + (size_t)blockIdx.x * astride + aidx : nullptr;
  if (S.done) {
    if (ah && threadIdx.x == 0) *ah = 0.0;
    return;
  }
  const double* p = partials + (size_t)blockIdx.x * nb;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += p[i];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    S.pap = s;
    S.alpha = S.rr / s;
    if (!finite(S.alpha)) { S.nan = 1; S.done = 1; S.skip_ru = 1; }
    if (ah) *ah = S.done ? 0.0 : S.alpha;
  }
}

// rr_new -> beta, and the stopping / reliable-update decisions.
// mode 0: reference CG test rr*snorm < prec (solver_CG.cpp:57-60)
// mode 1: reliable-update CG: flag an update when |r|^2 < delta2*rmax2 or
//         |r|^2 < target2.
__global__ void k_fin_rr(const double* __restrict__ partials, int nb, RUScal* __restrict__ sc,
                         int mode) {
  RUScal& S = sc[blockIdx.x];
  if (S.done) return;
  const double* p = partials + (size_t)blockIdx.x * nb;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += p[i];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    double rrp = S.rr;
    S.rr = s;
    S.beta = s / rrp;
    S.count += 1;
    S.keep_p = 1;
    if (!finite(s) || !finite(S.beta)) { S.nan = 1; S.done = 1; S.skip_ru = 1; return; }
    if (mode == 0) {
      if (s * S.snorm < S.prec) { S.done = 1; S.iter = S.count - 1; }
    } else {
      bool upd = (s < S.delta2 * S.rmax2) || (s < S.target2);
      S.skip_ru = upd ? 0 : 1;
    }
  }
}

// ============================================================ generic site BLAS
// Accessor construction from an any-precision field.
struct SF {
  void* v = nullptr;
  float* nrm = nullptr;
};

template <int P> struct Acc;
template <> struct Acc<II_PREC_DOUBLE> {
  using Sp = SpF64;
  using St = StF64;
  static Sp sp(const SF& f, int stride) { return Sp{(const double2*)f.v, stride}; }
  static St st(const SF& f, int stride) { return St{(double2*)f.v, stride}; }
};
template <> struct Acc<II_PREC_SINGLE> {
  using Sp = SpF32;
  using St = StF32;
  static Sp sp(const SF& f, int stride) { return Sp{(const float2*)f.v, stride}; }
  static St st(const SF& f, int stride) { return St{(float2*)f.v, stride}; }
};
template <> struct Acc<II_PREC_HALF> {
  using Sp = SpLo<__half2>;
  using St = StLo<__half2>;
  static Sp sp(const SF& f, int stride) { return Sp{(const __half2*)f.v, f.nrm, stride}; }
  static St st(const SF& f, int stride) { return St{(__half2*)f.v, f.nrm, stride}; }
};
template <> struct Acc<II_PREC_BF16> {
  using Sp = SpLo<__hip_bfloat162>;
  using St = StLo<__hip_bfloat162>;
  static Sp sp(const SF& f, int stride) { return Sp{(const __hip_bfloat162*)f.v, f.nrm, stride}; }
  static St st(const SF& f, int stride) { return St{(__hip_bfloat162*)f.v, f.nrm, stride}; }
};

template <typename Sp>
__device__ __forceinline__ void ld_site(const Sp& a, int hs, double v[24]) {
  const double s = (double)a.scale(hs);
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    auto c = a.ld(hs, k);
    v[2 * k] = (double)c.x * s;
    v[2 * k + 1] = (double)c.y * s;
  }
}
template <typename St>
__device__ __forceinline__ void st_site(const St& a, int hs, const double v[24]) {
  typename St::C o[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    o[k].x = (typename St::R)v[2 * k];
    o[k].y = (typename St::R)v[2 * k + 1];
  }
  a.st(hs, o);
}

// y += alpha p ; r -= alpha s ; partial |r|^2      (solver_CG.cpp:131-147)
template <typename SpP, typename SpS, typename SpR, typename StR, typename SpY, typename StY>
__global__ void __launch_bounds__(kBlock)
k_cg_xr(Geom g, SpP p, SpS s, SpR rr_, StR rw, SpY yr, StY yw, RUScal* __restrict__ sc,
        double* __restrict__ partials) {
  const int rhs = blockIdx.y;
  const RUScal& S = sc[rhs];
  if (S.done) return;
  const double a = S.alpha;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (hs < g.Vh) {
    double pv[24], sv[24], rv[24], yv[24];
    ld_site(p.rhs(rhs), hs, pv);
    ld_site(s.rhs(rhs), hs, sv);
    ld_site(rr_.rhs(rhs), hs, rv);
    ld_site(yr.rhs(rhs), hs, yv);
#pragma unroll
    for (int k = 0; k < 24; ++k) {
      yv[k] = fma(a, pv[k], yv[k]);
      rv[k] = fma(-a, sv[k], rv[k]);
      acc = fma(rv[k], rv[k], acc);
    }
    st_site(yw.rhs(rhs), hs, yv);
    st_site(rw.rhs(rhs), hs, rv);
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partials[(size_t)rhs * gridDim.x + blockIdx.x] = acc;
}

// p = r + beta p        (solver_CG.cpp:152-153: p = p*cr + r)
template <typename SpR, typename SpP, typename StP>
__global__ void __launch_bounds__(kBlock)
k_cg_p(Geom g, SpR r, SpP pr, StP pw, const RUScal* __restrict__ sc) {
  const int rhs = blockIdx.y;
  const RUScal& S = sc[rhs];
  if (S.done) return;
  const double b = S.beta;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= g.Vh) return;
  double rv[24], pv[24];
  ld_site(r.rhs(rhs), hs, rv);
  ld_site(pr.rhs(rhs), hs, pv);
#pragma unroll
  for (int k = 0; k < 24; ++k) pv[k] = fma(pv[k], b, rv[k]);
  st_site(pw.rhs(rhs), hs, pv);
}

// Deferred solution update y += sum_j alpha_j p_j over up to kDefer stored
// search directions: alpha_j = ahist[rhs * astride + slot[j]] (0 for the
// iterations after an rhs finished, whose directions are not read).  One
// pass over y for kDefer iterations instead of a y read and write in each.
// With y64, the rhs not masked by skip also fold the sum into the FP64
// solution, y64 += y + sum_j alpha_j p_j, and restart y at 0 (the reliable
// update's accumulation in the same pass).
constexpr int kDefer = 8;
template <typename SpP>
struct DeferredDirs {
  SpP p[kDefer];
  int slot[kDefer];
  int n;
};
template <typename SpP, typename SpY, typename StY>
__global__ void __launch_bounds__(kBlock)
k_cg_flush(Geom g, DeferredDirs<SpP> d, const double* __restrict__ ahist, int astride, SpY yr,
           StY yw, double2* __restrict__ y64, const int* __restrict__ skip, int skip_stride) {
  const int rhs = blockIdx.y;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= g.Vh) return;
  const bool fold = y64 && !(skip && skip[rhs * skip_stride]);
  double yv[24];
  ld_site(yr.rhs(rhs), hs, yv);
#pragma unroll
  for (int j = 0; j < kDefer; ++j) {
    if (j < d.n) {
      const double a = ahist[(size_t)rhs * astride + d.slot[j]];
      if (a != 0.0) {
        double pv[24];
        ld_site(d.p[j].rhs(rhs), hs, pv);
#pragma unroll
        for (int k = 0; k < 24; ++k) yv[k] = fma(a, pv[k], yv[k]);
      }
    }
  }
  if (fold) {
    double2* y = y64 + (size_t)rhs * 12 * g.stride;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      double2 o = y[(size_t)k * g.stride + hs];
      o.x += yv[2 * k];
      o.y += yv[2 * k + 1];
      y[(size_t)k * g.stride + hs] = o;
    }
#pragma unroll
    for (int k = 0; k < 24; ++k) yv[k] = 0.0;
  }
  st_site(yw.rhs(rhs), hs, yv);
}

// dst = src (precision conversion), masked by skip flags
template <typename SpA, typename StB>
__global__ void __launch_bounds__(kBlock)
k_convert(Geom g, SpA a, StB b, const int* __restrict__ skip, int skip_stride) {
  const int rhs = blockIdx.y;
  if (skip && skip[rhs * skip_stride]) return;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= g.Vh) return;
  double v[24];
  ld_site(a.rhs(rhs), hs, v);
  st_site(b.rhs(rhs), hs, v);
}

// y64 += ys ; ys = 0  (reliable update accumulation), masked
template <typename SpY, typename StY>
__global__ void __launch_bounds__(kBlock)
k_accum(Geom g, SpY ys, StY yz, double2* __restrict__ y64, const int* __restrict__ skip,
        int skip_stride) {
  const int rhs = blockIdx.y;
  if (skip && skip[rhs * skip_stride]) return;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= g.Vh) return;
  double v[24];
  ld_site(ys.rhs(rhs), hs, v);
  double2* y = y64 + (size_t)rhs * 12 * g.stride;
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    double2 o = y[(size_t)k * g.stride + hs];
    o.x += v[2 * k];
    o.y += v[2 * k + 1];
    y[(size_t)k * g.stride + hs] = o;
  }
  double z[24];
#pragma unroll
  for (int k = 0; k < 24; ++k) z[k] = 0.0;
  st_site(yz.rhs(rhs), hs, z);
}

// masked |a|^2 partials (FP64)
__global__ void __launch_bounds__(kBlock)
k_norm2_masked(Geom g, const double2* __restrict__ a, const int* __restrict__ skip,
               int skip_stride, double* __restrict__ partials) {
  const int rhs = blockIdx.y;
  if (skip && skip[rhs * skip_stride]) return;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (hs < g.Vh) {
    size_t off = (size_t)rhs * 12 * g.stride + hs;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      double2 v = a[off + (size_t)k * g.stride];
      acc = fma(v.x, v.x, fma(v.y, v.y, acc));
    }
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partials[(size_t)rhs * gridDim.x + blockIdx.x] = acc;
}

// masked |a|^2 partials of a field in any precision (FP64 accumulation)
template <typename Sp>
__global__ void __launch_bounds__(kBlock)
k_norm2_P(Geom g, Sp a, const int* __restrict__ skip, int skip_stride,
          double* __restrict__ partials) {
  const int rhs = blockIdx.y;
  if (skip && skip[rhs * skip_stride]) return;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (hs < g.Vh) {
    double v[24];
    ld_site(a.rhs(rhs), hs, v);
#pragma unroll
    for (int k = 0; k < 24; ++k) acc = fma(v[k], v[k], acc);
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partials[(size_t)rhs * gridDim.x + blockIdx.x] = acc;
}

// after the true residual: done test
__global__ void k_fin_rho(const double* __restrict__ partials, int nb, RUScal* __restrict__ sc) {
  RUScal& S = sc[blockIdx.x];
  if (S.skip_ru) return;
  const double* p = partials + (size_t)blockIdx.x * nb;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += p[i];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    S.rho2 = s;
    if (!finite(s)) { S.nan = 1; S.done = 1; S.skip_ru = 1; return; }
    if (s <= S.tol2) { S.done = 1; S.skip_ru = 1; S.updates += 1; }
  }
}

// after the recomputed normal residual: restart the recursion scalars
__global__ void k_fin_ru(const double* __restrict__ partials, int nb, RUScal* __restrict__ sc) {
  RUScal& S = sc[blockIdx.x];
  if (S.skip_ru) return;
  const double* p = partials + (size_t)blockIdx.x * nb;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += p[i];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    // A true residual far above the recursive one means the recursion was
    // corrupted (in the sloppy arithmetic or otherwise); the old search
    // direction is then useless and keeping it makes CG diverge, so the
    // caller restarts it from the new residual (p = r).
    if (s > S.restart * S.rr) {
      S.keep_p = 0;
      S.restarts += 1;
    }
    S.rr = s;
    S.rmax2 = s;
    // |rho| / |Dhat^dag rho| calibrates when to test the true residual next
    double ratio = S.rho2 / s;
    double t = 0.81 * S.tol2 / ratio;
    S.target2 = t < S.target2 ? t : S.target2;
    S.updates += 1;
    S.skip_ru = 1;  // restores the mask; `update` consumed
    if (!finite(s)) { S.nan = 1; S.done = 1; }
  }
}

// ============================================================ precision dispatch
ii_status hopP(ii_context* ctx, int P, const HopSpec& h, const SF& in, const SF& out,
               const SF* x, const SF* dv) {
  switch (P) {
    case II_PREC_DOUBLE:
      return hop64(ctx, h, (const double2*)in.v, (double2*)out.v,
                   x ? (const double2*)x->v : nullptr, dv ? (const double2*)dv->v : nullptr);
    case II_PREC_SINGLE:
      return hop32(ctx, h, (const float2*)in.v, (float2*)out.v,
                   x ? (const float2*)x->v : nullptr, dv ? (const float2*)dv->v : nullptr);
    case II_PREC_HALF:
    case II_PREC_BF16: {
      LoField li{in.v, in.nrm}, lo{out.v, out.nrm}, lx, ld;
      if (x) lx = LoField{x->v, x->nrm};
      if (dv) ld = LoField{dv->v, dv->nrm};
      return P == II_PREC_HALF ? hop_half(ctx, h, li, lo, x ? &lx : nullptr, dv ? &ld : nullptr)
                               : hop_bf16(ctx, h, li, lo, x ? &lx : nullptr, dv ? &ld : nullptr);
    }
  }
  return set_error(ctx, II_ERR_BAD_ARG, "unknown precision");
}

// A = Dhat^dag Dhat in precision P with partials of Re<dv, A f>; skip flags.
ii_status ddagdP(ii_context* ctx, int P, double kappa, const SF& f, const SF& out, const SF& q,
                 const SF& t, const SF* dv, double* partials, int nrhs, const int* skip,
                 int skip_stride) {
  HopSpec h;
  h.nrhs = nrhs;
  h.skip = skip;
  h.skip_stride = skip_stride;
  h.out_par = 1;
  h.a = 1.0;
  II_CHECK(hopP(ctx, P, h, f, t, nullptr, nullptr));
  h.out_par = 0;
  h.a = -kappa * kappa;
  h.xpay = true;
  II_CHECK(hopP(ctx, P, h, t, q, &f, nullptr));  // q = Dhat f
  h.out_par = 1;
  h.a = 1.0;
  h.xpay = false;
  h.g5in = true;
  II_CHECK(hopP(ctx, P, h, q, t, nullptr, nullptr));
  h.out_par = 0;
  h.a = -kappa * kappa;
  h.xpay = true;
  h.g5in = false;
  h.g5out = true;
  h.epi = dv != nullptr ? kEpiDot : kEpiStore;
  h.partials = partials;
  return hopP(ctx, P, h, t, out, &q, dv);  // out = Dhat^dag q
}

// One CG step on A = Dhat^dag Dhat with the scalars and the residual update
// fused into the hops (sloppy multi-rhs solve):
//   q = Dhat p, with partial sums of |q|^2 = <p, A p>   (second hop, kEpiNorm)
//   alpha = rr / |q|^2                                   (k_fin_pap; ahist)
//   r -= alpha Dhat^dag q, with partial sums of |r|^2   (fourth hop, kEpiRUpd)
// A p itself is never stored; the k_fin_rr that follows reads the |r|^2
// partials (hop_dot_parts per rhs).
static_assert(sizeof(RUScal) % sizeof(double) == 0, "RUScal is read as a double array");
ii_status ddagd_cg_fused(ii_context* ctx, int P, double kappa, const SF& p, const SF& r,
                         const SF& q, const SF& t, RUScal* sc, double* partials, int nrhs,
                         const int* skip, int skip_stride, double* ahist, int aidx,
                         int astride) {
  HopSpec h;
  h.nrhs = nrhs;
  h.skip = skip;
  h.skip_stride = skip_stride;
  h.out_par = 1;
  h.a = 1.0;
  II_CHECK(hopP(ctx, P, h, p, t, nullptr, nullptr));
  h.out_par = 0;
  h.a = -kappa * kappa;
  h.xpay = true;
  h.epi = kEpiNorm;
  h.partials = partials;
  II_CHECK(hopP(ctx, P, h, t, q, &p, nullptr));  // q = Dhat p, |q|^2
  hipLaunchKernelGGL(k_fin_pap, dim3(nrhs), dim3(kFinThreads), 0, ctx->stream, partials,
                     hop_dot_parts(ctx->g, nrhs, P), sc, ahist, aidx, astride);
  h.out_par = 1;
  h.a = 1.0;
  h.xpay = false;
  h.g5in = true;
  h.epi = kEpiStore;
  II_CHECK(hopP(ctx, P, h, q, t, nullptr, nullptr));
  h.out_par = 0;
  h.a = -kappa * kappa;
  h.xpay = true;
  h.g5in = false;
  h.g5out = true;
  h.epi = kEpiRUpd;
  h.coef = &sc[0].alpha;
  h.coef_stride = (int)(sizeof(RUScal) / sizeof(double));
  return hopP(ctx, P, h, t, r, &q, &r);  // r -= alpha Dhat^dag q, |r|^2
}

template <int P>
ii_status cg_step_blas(ii_context* ctx, const SF& p, const SF& s, const SF& r, const SF& y,
                       int ytype, RUScal* sc, int nrhs, double* partials, int nb) {
  const Geom& g = ctx->g;
  using A = Acc<P>;
  dim3 grid(nb, nrhs);
  if (ytype == II_PREC_DOUBLE) {
    hipLaunchKernelGGL((k_cg_xr<typename A::Sp, typename A::Sp, typename A::Sp, typename A::St,
                                SpF64, StF64>),
                       grid, dim3(kBlock), 0, ctx->stream, g, A::sp(p, g.stride),
                       A::sp(s, g.stride), A::sp(r, g.stride), A::st(r, g.stride),
                       Acc<II_PREC_DOUBLE>::sp(y, g.stride), Acc<II_PREC_DOUBLE>::st(y, g.stride),
                       sc, partials);
  } else {
    hipLaunchKernelGGL((k_cg_xr<typename A::Sp, typename A::Sp, typename A::Sp, typename A::St,
                                SpF32, StF32>),
                       grid, dim3(kBlock), 0, ctx->stream, g, A::sp(p, g.stride),
                       A::sp(s, g.stride), A::sp(r, g.stride), A::st(r, g.stride),
                       Acc<II_PREC_SINGLE>::sp(y, g.stride), Acc<II_PREC_SINGLE>::st(y, g.stride),
                       sc, partials);
  }
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

// p_out = r + beta p_in (p_out may be p_in)
template <int P>
ii_status cg_p_update(ii_context* ctx, const SF& r, const SF& p, const SF& pout, RUScal* sc,
                      int nrhs) {
  const Geom& g = ctx->g;
  using A = Acc<P>;
  hipLaunchKernelGGL((k_cg_p<typename A::Sp, typename A::Sp, typename A::St>),
                     dim3(blocks_for(g.Vh), nrhs), dim3(kBlock), 0, ctx->stream, g,
                     A::sp(r, g.stride), A::sp(p, g.stride), A::st(pout, g.stride), sc);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}
// y (FP32) += the deferred directions ring[slots[j]], j < n <= kDefer; with
// y64, also fold into y64 (rhs not masked by skip)
template <int P>
ii_status cg_flush(ii_context* ctx, const SF* ring, const int* slots, int n, const double* ahist,
                   int astride, const SF& y, int nrhs, double2* y64, const int* skip,
                   int skip_stride) {
  const Geom& g = ctx->g;
  using A = Acc<P>;
  DeferredDirs<typename A::Sp> d{};
  d.n = n;
  for (int j = 0; j < n; ++j) {
    d.p[j] = A::sp(ring[slots[j]], g.stride);
    d.slot[j] = slots[j];
  }
  for (int j = n; j < kDefer; ++j) d.p[j] = d.p[0];
  hipLaunchKernelGGL((k_cg_flush<typename A::Sp, SpF32, StF32>), dim3(blocks_for(g.Vh), nrhs),
                     dim3(kBlock), 0, ctx->stream, g, d, ahist, astride,
                     Acc<II_PREC_SINGLE>::sp(y, g.stride), Acc<II_PREC_SINGLE>::st(y, g.stride),
                     y64, skip, skip_stride);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

// convert between precisions (Pa -> Pb), masked
template <int Pa, int Pb>
ii_status convert_t(ii_context* ctx, const SF& a, const SF& b, int nrhs, const int* skip,
                    int skip_stride) {
  const Geom& g = ctx->g;
  hipLaunchKernelGGL((k_convert<typename Acc<Pa>::Sp, typename Acc<Pb>::St>),
                     dim3(blocks_for(g.Vh), nrhs), dim3(kBlock), 0, ctx->stream, g,
                     Acc<Pa>::sp(a, g.stride), Acc<Pb>::st(b, g.stride), skip, skip_stride);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}

ii_status convertP(ii_context* ctx, int Pa, int Pb, const SF& a, const SF& b, int nrhs,
                   const int* skip = nullptr, int skip_stride = 0) {
#define II_CV(A, B) \
  if (Pa == A && Pb == B) return convert_t<A, B>(ctx, a, b, nrhs, skip, skip_stride);
  II_CV(0, 1) II_CV(0, 2) II_CV(0, 3) II_CV(1, 0) II_CV(2, 0) II_CV(3, 0) II_CV(0, 0)
  II_CV(1, 1) II_CV(2, 2) II_CV(3, 3)
#undef II_CV
  return set_error(ctx, II_ERR_BAD_ARG, "unsupported conversion");
}

template <int P>
ii_status norm2_t(ii_context* ctx, const SF& a, int nrhs, const int* skip, int skip_stride,
                  double* partials) {
  const Geom& g = ctx->g;
  hipLaunchKernelGGL((k_norm2_P<typename Acc<P>::Sp>), dim3(blocks_for(g.Vh), nrhs), dim3(kBlock),
                     0, ctx->stream, g, Acc<P>::sp(a, g.stride), skip, skip_stride, partials);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}
ii_status norm2P(ii_context* ctx, int P, const SF& a, int nrhs, const int* skip, int skip_stride,
                 double* partials) {
  switch (P) {
    case 0: return norm2_t<0>(ctx, a, nrhs, skip, skip_stride, partials);
    case 1: return norm2_t<1>(ctx, a, nrhs, skip, skip_stride, partials);
    case 2: return norm2_t<2>(ctx, a, nrhs, skip, skip_stride, partials);
    default: return norm2_t<3>(ctx, a, nrhs, skip, skip_stride, partials);
  }
}

ii_status cg_blas_P(ii_context* ctx, int P, const SF& p, const SF& s, const SF& r, const SF& y,
                    int ytype, RUScal* sc, int nrhs, double* partials, int nb) {
  switch (P) {
    case 0: return cg_step_blas<0>(ctx, p, s, r, y, ytype, sc, nrhs, partials, nb);
    case 1: return cg_step_blas<1>(ctx, p, s, r, y, ytype, sc, nrhs, partials, nb);
    case 2: return cg_step_blas<2>(ctx, p, s, r, y, ytype, sc, nrhs, partials, nb);
    default: return cg_step_blas<3>(ctx, p, s, r, y, ytype, sc, nrhs, partials, nb);
  }
}
ii_status cg_p_P(ii_context* ctx, int P, const SF& r, const SF& p, const SF& pout, RUScal* sc,
                 int nrhs) {
  switch (P) {
    case 0: return cg_p_update<0>(ctx, r, p, pout, sc, nrhs);
    case 1: return cg_p_update<1>(ctx, r, p, pout, sc, nrhs);
    case 2: return cg_p_update<2>(ctx, r, p, pout, sc, nrhs);
    default: return cg_p_update<3>(ctx, r, p, pout, sc, nrhs);
  }
}
ii_status cg_flush_P(ii_context* ctx, int P, const SF* ring, const int* slots, int n,
                     const double* ahist, int astride, const SF& y, int nrhs,
                     double2* y64 = nullptr, const int* skip = nullptr, int skip_stride = 0) {
  switch (P) {
    case 1: return cg_flush<1>(ctx, ring, slots, n, ahist, astride, y, nrhs, y64, skip, skip_stride);
    case 2: return cg_flush<2>(ctx, ring, slots, n, ahist, astride, y, nrhs, y64, skip, skip_stride);
    case 3: return cg_flush<3>(ctx, ring, slots, n, ahist, astride, y, nrhs, y64, skip, skip_stride);
    default: return set_error(ctx, II_ERR_BAD_ARG, "deferred update needs a sloppy precision");
  }
}

// bytes of one rhs of a half field in precision P (components + norms)
size_t field_bytes(const Geom& g, int P) {
  switch (P) {
    case II_PREC_DOUBLE: return (size_t)12 * g.stride * sizeof(double2);
    case II_PREC_SINGLE: return (size_t)12 * g.stride * sizeof(float2);
    default: return (size_t)12 * g.stride * sizeof(__half2) + (size_t)g.stride * sizeof(float);
  }
}
// a workspace field of nrhs rhs in precision P (norms after the components)
ii_status fieldP(ii_context* ctx, int slot, int P, int nrhs, SF* f) {
  const Geom& g = ctx->g;
  void* p = nullptr;
  II_CHECK(workspace(ctx, slot, field_bytes(g, P) * nrhs, &p));
  f->v = p;
  f->nrm = nullptr;
  if (P >= II_PREC_HALF)
    f->nrm = (float*)((char*)p + (size_t)12 * g.stride * sizeof(__half2) * nrhs);
  return II_OK;
}

// ============================================================ status polling
ii_status read_scalars(ii_context* ctx, int n) {
  II_HIP(ctx, hipMemcpyAsync(ctx->scal_host, ctx->scal, n * sizeof(RUScal),
                             hipMemcpyDeviceToHost, ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return II_OK;
}
RUScal* host_scal(ii_context* ctx) { return reinterpret_cast<RUScal*>(ctx->scal_host); }
RUScal* dev_scal(ii_context* ctx) { return reinterpret_cast<RUScal*>(ctx->scal); }

ii_status ensure_ru_scalars(ii_context* ctx, int n) {
  static_assert(sizeof(RUScal) <= sizeof(CGScalars) * 2, "scalar block");
  return ensure_scalars(ctx, 2 * n);
}

// When some rhs of a batch trigger a reliable update, every other rhs whose
// |r| has come within a factor 4 of its own trigger (|r|^2 < 16 delta^2
// rmax^2) joins it, so that the batch's rhs update together: one batched
// FP64 residual per event instead of one more for each rhs that would
// trigger an iteration or two later.  An early reliable update is still a
// reliable update; the rhs's next trigger is measured from it.  Needs the
// host copy of the scalars current (read_scalars) and no kernel in flight.
ii_status join_updates(ii_context* ctx, int n) {
  RUScal* H = host_scal(ctx);
  bool joined = false;
  for (int k = 0; k < n; ++k)
    if (!H[k].done && H[k].skip_ru && H[k].rr < 16.0 * H[k].delta2 * H[k].rmax2) {
      H[k].skip_ru = 0;
      joined = true;
    }
  if (joined)
    II_HIP(ctx, hipMemcpyAsync(ctx->scal, ctx->scal_host, n * sizeof(RUScal),
                               hipMemcpyHostToDevice, ctx->stream));
  return II_OK;
}

__global__ void k_scal_init(RUScal* sc, const double* __restrict__ rr, const double* __restrict__ bb,
                            double prec, double tol, double delta, int mode,
                            double restart = 16.0) {
  RUScal& S = sc[blockIdx.x];
  if (threadIdx.x != 0) return;
  RUScal z{};
  z.rr = rr[blockIdx.x];
  double bn = sqrt(bb[blockIdx.x]);
  z.b2 = bb[blockIdx.x];
  z.snorm = 1.0 / (bn * bn);
  z.prec = prec;
  z.rmax2 = z.rr;
  z.tol2 = tol * tol * z.b2;
  z.target2 = z.tol2;
  z.delta2 = delta * delta;
  z.restart = restart;
  z.iter = -1;
  z.skip_ru = 1;
  z.keep_p = 1;
  if (!(z.rr > 0.0)) { z.done = 1; }
  if (mode == 0) z.skip_ru = 1;
  S = z;
}

// ============================================================ reference CG
// Solver_CG::solve semantics on Fopr_DdagD: x0 = b (solver_CG.cpp:40).
ii_status cg_reference(ii_context* ctx, double kappa, double2* x, const double2* b, double prec,
                       int maxiter, ii_solver_output* out) {
  const Geom& g = ctx->g;
  II_CHECK(ensure_ru_scalars(ctx, 1));
  RUScal* sc = dev_scal(ctx);
  int nb = hop_blocks(g);
  II_CHECK(ensure_partials(ctx, (size_t)nb + 8));
  double* part = ctx->partials;
  double* tmp2 = nullptr;
  {
    void* p;
    II_CHECK(workspace(ctx, 30, 4 * sizeof(double), &p));
    tmp2 = (double*)p;
  }
  SF fx{x}, fr, fp, fs, fq, ft;
  II_CHECK(fieldP(ctx, 10, 0, 1, &fr));
  II_CHECK(fieldP(ctx, 11, 0, 1, &fp));
  II_CHECK(fieldP(ctx, 12, 0, 1, &fs));
  II_CHECK(fieldP(ctx, 13, 0, 1, &fq));
  II_CHECK(fieldP(ctx, 14, 0, 1, &ft));
  const int ss = kScalInts;
  const int* skip = &sc[0].done;
  II_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  II_CHECK(copy64(ctx, x, b));
  // r = b - A x
  II_CHECK(ddagdP(ctx, 0, kappa, fx, fs, fq, ft, nullptr, nullptr, 1, nullptr, 0));
  II_CHECK(copy64(ctx, (double2*)fr.v, b));
  II_CHECK(axpby64(ctx, -1.0, (double2*)fs.v, 1.0, (double2*)fr.v));
  II_CHECK(copy64(ctx, (double2*)fp.v, (double2*)fr.v));
  II_CHECK(dot64(ctx, (double2*)fr.v, (double2*)fr.v, tmp2));
  II_CHECK(dot64(ctx, b, b, tmp2 + 1));
  hipLaunchKernelGGL(k_scal_init, dim3(1), dim3(64), 0, ctx->stream, sc, tmp2, tmp2 + 1, prec,
                     0.0, 0.0, 0);
  II_HIP(ctx, hipGetLastError());
  int queued = 0;
  int chunk = 1;
  bool conv = false;
  while (queued < maxiter) {
    int n = std::min(chunk, maxiter - queued);
    for (int i = 0; i < n; ++i) {
      II_CHECK(ddagdP(ctx, 0, kappa, fp, fs, fq, ft, &fp, part, 1, skip, ss));
      hipLaunchKernelGGL(k_fin_pap, dim3(1), dim3(kBlock), 0, ctx->stream, part, nb, sc, nullptr,
                         0, 0);
      II_CHECK(cg_blas_P(ctx, 0, fp, fs, fr, fx, 0, sc, 1, part, nb));
      hipLaunchKernelGGL(k_fin_rr, dim3(1), dim3(kBlock), 0, ctx->stream, part, nb, sc, 0);
      II_CHECK(cg_p_P(ctx, 0, fr, fp, fp, sc, 1));
    }
    queued += n;
    II_CHECK(read_scalars(ctx, 1));
    const RUScal& H = host_scal(ctx)[0];
    if (H.nan) return set_error(ctx, II_ERR_NAN, "non-finite CG scalar");
    if (H.done) { conv = true; break; }
    chunk = std::min(chunk * 2, 64);
  }
  // diff = |A x - b| / |b|  (solver_CG.cpp:69-71)
  II_CHECK(ddagdP(ctx, 0, kappa, fx, fs, fq, ft, nullptr, nullptr, 1, nullptr, 0));
  II_CHECK(axpby64(ctx, -1.0, b, 1.0, (double2*)fs.v));
  II_CHECK(dot64(ctx, (double2*)fs.v, (double2*)fs.v, tmp2 + 2));
  II_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  double h[3];
  II_HIP(ctx, hipMemcpyAsync(h, tmp2, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  II_CHECK(read_scalars(ctx, 1));
  const RUScal& H = host_scal(ctx)[0];
  float ms = 0.f;
  II_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (out) {
    out->iterations = conv ? H.iter : -1;
    out->diff = std::sqrt(h[2]) * std::sqrt(H.snorm);
    out->timing_ms = ms;
    out->true_residual = -1.0;
    out->restarts = 0;
    out->inner_iterations = H.count;
  }
  if (!conv)
    return set_error(ctx, II_ERR_NOT_CONVERGED,
                     "CG solver not converged. Current residual: " +
                         std::to_string(H.rr * H.snorm));
  return II_OK;
}

// ============================================================ multishift CG
constexpr int kMaxShift = 64;
struct MSScal {
  double rr, pap, alpha, beta, alphap, betap, snorm, prec;
  double sigma[kMaxShift], csh2[kMaxShift], zeta1[kMaxShift], zeta2[kMaxShift];
  double zeta[kMaxShift], alphas[kMaxShift], betas[kMaxShift], pp[kMaxShift];
  int nshift, nshift2, done, count, nconv, nan;
};

__global__ void k_ms_init(MSScal* S, const double* __restrict__ rrbb, const double* sigma, int n,
                          double prec) {
  if (threadIdx.x != 0) return;
  S->rr = rrbb[0];
  double bn = sqrt(rrbb[1]);
  S->snorm = 1.0 / (bn * bn);
  S->prec = prec;
  S->alphap = 0.0;
  S->betap = 1.0;
  S->nshift = n;
  S->nshift2 = n;
  S->done = 0;
  S->count = 0;
  S->nconv = -1;
  S->nan = 0;
  for (int i = 0; i < n; ++i) {
    S->sigma[i] = sigma[i];
    S->csh2[i] = sigma[i] - sigma[0];
    S->zeta1[i] = 1.0;
    S->zeta2[i] = 1.0;
    S->pp[i] = 0.0;
  }
}

// The multishift kernels run on even half fields (NC = 12 complex
// components per site) or on full fields of the full-index operator
// (NC = 24: both parity blocks, [2][12][stride] = [24][stride]).
// s += sigma0 p0 ; partial <s, p0>      (multiShiftSolver_CG.cpp:50-55)
template <int NC, typename T = double2>
__global__ void __launch_bounds__(kBlock)
k_ms_sp(Geom g, T* __restrict__ s, const T* __restrict__ p0, const MSScal* S,
        double* __restrict__ partials) {
  if (S->done) return;
  const double sig = S->sigma[0];
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (hs < g.Vh) {
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      size_t i = (size_t)k * g.stride + hs;
      const T s0 = s[i], p0v = p0[i];
      double2 sv = make_double2(s0.x, s0.y), pv = make_double2(p0v.x, p0v.y);
      sv.x = fma(sig, pv.x, sv.x);
      sv.y = fma(sig, pv.y, sv.y);
      s[i].x = sv.x;
      s[i].y = sv.y;
      acc = fma(sv.x, pv.x, fma(sv.y, pv.y, acc));
    }
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ void k_ms_fin_pap(const double* __restrict__ partials, int nb, MSScal* S) {
  if (S->done) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partials[i];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    S->pap = s;
    S->beta = -S->rr / s;
    if (!isfinite(S->beta)) { S->nan = 1; S->done = 1; }
  }
}

// x0 -= beta p0 ; r += beta s ; partial |r|^2   (:57-59)
template <int NC, typename T = double2>
__global__ void __launch_bounds__(kBlock)
k_ms_xr(Geom g, T* __restrict__ x0, const T* __restrict__ p0, T* __restrict__ r,
        const T* __restrict__ s, const MSScal* S, double* __restrict__ partials) {
  if (S->done) return;
  const double b = S->beta;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (hs < g.Vh) {
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      size_t i = (size_t)k * g.stride + hs;
      const T x_ = x0[i], p_ = p0[i], r_ = r[i], s_ = s[i];
      double2 xv = make_double2(x_.x, x_.y), pv = make_double2(p_.x, p_.y);
      double2 rv = make_double2(r_.x, r_.y), sv = make_double2(s_.x, s_.y);
      xv.x = fma(-b, pv.x, xv.x);
      xv.y = fma(-b, pv.y, xv.y);
      rv.x = fma(b, sv.x, rv.x);
      rv.y = fma(b, sv.y, rv.y);
      x0[i].x = xv.x;
      x0[i].y = xv.y;
      r[i].x = rv.x;
      r[i].y = rv.y;
      acc = fma(rv.x, rv.x, fma(rv.y, rv.y, acc));
    }
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// rr, alpha and the shift recurrences (:61-83)
__global__ void k_ms_fin_rr(const double* __restrict__ partials, int nb, MSScal* S) {
  if (S->done) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partials[i];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    double rrp = S->rr;
    S->rr = s;
    double alpha = s / rrp;
    S->alpha = alpha;
    S->pp[0] = s;
    double beta = S->beta;
    double alphah = 1.0 + S->alphap * beta / S->betap;
    for (int ish = 1; ish < S->nshift2; ++ish) {
      double zeta = (alphah - S->csh2[ish] * beta) / S->zeta1[ish] + (1.0 - alphah) / S->zeta2[ish];
      zeta = 1.0 / zeta;
      double zr = zeta / S->zeta1[ish];
      S->betas[ish] = beta * zr;
      S->alphas[ish] = alpha * zr * zr;
      S->zeta[ish] = zeta;
    }
    if (!isfinite(alpha)) { S->nan = 1; S->done = 1; }
  }
}

// p0 = alpha p0 + r (:63-64) and, for ish in [1, nshift2):
// x_i -= betas p_i ; p_i = alphas p_i + zeta r ; partial |p_i|^2 (:73-79)
template <int NC, typename T = double2>
__global__ void __launch_bounds__(kBlock)
k_ms_vec(Geom g, T* __restrict__ x, T* __restrict__ p, const T* __restrict__ r,
         const MSScal* S, double* __restrict__ partials) {
  if (S->done) return;
  const int ish = blockIdx.y;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t he = (size_t)NC * g.stride;
  if (ish == 0) {
    if (hs < g.Vh) {
      const double a = S->alpha;
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        size_t i = (size_t)k * g.stride + hs;
        const T pv = p[i], rv = r[i];
        p[i].x = fma((double)pv.x, a, (double)rv.x);
        p[i].y = fma((double)pv.y, a, (double)rv.y);
      }
    }
    return;
  }
  if (ish >= S->nshift2) return;
  const double bs = S->betas[ish], as = S->alphas[ish], z = S->zeta[ish];
  double acc = 0.0;
  if (hs < g.Vh) {
    T* xi = x + ish * he;
    T* pi = p + ish * he;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      size_t i = (size_t)k * g.stride + hs;
      const T x_ = xi[i], p_ = pi[i], r_ = r[i];
      double2 xv = make_double2(x_.x, x_.y), pv = make_double2(p_.x, p_.y);
      const double2 rv = make_double2(r_.x, r_.y);
      xv.x = fma(-bs, pv.x, xv.x);
      xv.y = fma(-bs, pv.y, xv.y);
      pv.x = fma(pv.x, as, z * rv.x);
      pv.y = fma(pv.y, as, z * rv.y);
      xi[i].x = xv.x;
      xi[i].y = xv.y;
      pi[i].x = pv.x;
      pi[i].y = pv.y;
      acc = fma(pv.x, pv.x, fma(pv.y, pv.y, acc));
    }
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partials[(size_t)ish * gridDim.x + blockIdx.x] = acc;
}

// pp_i, zeta shifts, Nshift2 shrink, stopping test (:79-92, :148-152)
__global__ void k_ms_tail(const double* __restrict__ partials, int nb, MSScal* S) {
  if (S->done) return;
  __shared__ double sums[kMaxShift];
  for (int ish = 1; ish < S->nshift2; ++ish) {
    const double* p = partials + (size_t)ish * nb;
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s += p[i];
    s = block_sum(s);
    if (threadIdx.x == 0) sums[ish] = s;
  }
  if (threadIdx.x == 0) {
    for (int ish = 1; ish < S->nshift2; ++ish) {
      S->pp[ish] = sums[ish] * S->snorm;
      S->zeta2[ish] = S->zeta1[ish];
      S->zeta1[ish] = S->zeta[ish];
    }
    for (int ish = S->nshift2 - 1; ish >= 0; --ish) {
      if (S->pp[ish] > S->prec) { S->nshift2 = ish + 1; break; }
    }
    S->alphap = S->alpha;
    S->betap = S->beta;
    S->count += 1;
    if (S->rr * S->snorm < S->prec) { S->done = 1; S->nconv = S->count - 1; }
  }
}

// A = D^dag D of the full-index operator D = 1 - kappa H on full fields
// [2][12][stride] (Fopr_DdagD over Dirac_Wilson, fopr.h:146-153,
// dirac_wilson.cpp:28-36, 115-118); w: full-field temporary.
ii_status ddagd_full64(ii_context* ctx, double kappa, const double2* f, double2* out, double2* w,
                       const int* skip) {
  const size_t he = half_elems(ctx->g);
  HopSpec h;
  h.a = -kappa;
  h.xpay = true;
  h.skip = skip;
  for (int par = 0; par < 2; ++par) {
    h.out_par = par;
    II_CHECK(hop64(ctx, h, f + (1 - par) * he, w + par * he, f + par * he, nullptr));
  }
  h.g5in = h.g5out = true;
  for (int par = 0; par < 2; ++par) {
    h.out_par = par;
    II_CHECK(hop64(ctx, h, w + (1 - par) * he, out + par * he, w + par * he, nullptr));
  }
  return II_OK;
}

__global__ void k_sum2(const double* __restrict__ in, int n, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += in[i];
    out[0] = s;
  }
}

// multishift on the even Schur system (full = false: x, b even half fields)
// or on the full-index operator (full = true: full fields)
ii_status multishift_reference(ii_context* ctx, double kappa, double2* x, const double2* b,
                               const double* shifts, int nshift, double prec, int maxiter,
                               ii_solver_output* out, std::vector<double>* per_shift,
                               bool full) {
  const Geom& g = ctx->g;
  const int nh = full ? 2 : 1;  // half-field blocks per vector
  if (nshift < 1 || nshift > kMaxShift) return set_error(ctx, II_ERR_BAD_ARG, "nshift out of range");
  int nb = hop_blocks(g);
  II_CHECK(ensure_partials(ctx, (size_t)nb * nshift + 16));
  double* part = ctx->partials;
  void* wp = nullptr;
  II_CHECK(workspace(ctx, 29, sizeof(MSScal) + (kMaxShift + 8) * sizeof(double), &wp));
  MSScal* S = (MSScal*)wp;
  double* aux = (double*)((char*)wp + sizeof(MSScal));  // rr, bb, sigma[...]
  MSScal* Sh = nullptr;
  {
    void* hp = nullptr;
    II_CHECK(ensure_host_pin(ctx, sizeof(MSScal), &hp));
    Sh = (MSScal*)hp;
  }
  SF fr, fs, fq, ft, fp;
  II_CHECK(fieldP(ctx, 10, 0, nh, &fr));
  II_CHECK(fieldP(ctx, 12, 0, nh, &fs));
  II_CHECK(fieldP(ctx, 13, 0, nh, &fq));
  II_CHECK(fieldP(ctx, 14, 0, nh, &ft));
  II_CHECK(fieldP(ctx, 15, 0, nshift * nh, &fp));
  size_t he = half_elems(g) * nh;
  double2* p = (double2*)fp.v;
  // |v|^2 over all nh blocks of a vector into dst (device)
  auto norm2 = [&](const double2* v, double* dst) -> ii_status {
    II_CHECK(dot64(ctx, v, v, aux + 8 + kMaxShift - 4, nh));
    hipLaunchKernelGGL(k_sum2, dim3(1), dim3(64), 0, ctx->stream, aux + 8 + kMaxShift - 4, nh,
                       dst);
    II_HIP(ctx, hipGetLastError());
    return II_OK;
  };
  auto apply_A = [&](const double2* f, const int* skip) -> ii_status {
    if (full) return ddagd_full64(ctx, kappa, f, (double2*)fs.v, (double2*)fq.v, skip);
    return ddagdP(ctx, 0, kappa, SF{const_cast<double2*>(f)}, fs, fq, ft, nullptr, nullptr, 1,
                  skip, 0);
  };
  II_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int i = 0; i < nshift; ++i) {
    II_CHECK(copy64(ctx, p + i * he, b, nh));
    II_CHECK(zero64(ctx, x + i * he, nh));
  }
  II_CHECK(copy64(ctx, (double2*)fr.v, b, nh));
  II_CHECK(norm2(b, aux));
  II_CHECK(norm2(b, aux + 1));
  II_HIP(ctx, hipMemcpyAsync(aux + 2, shifts, nshift * sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
  hipLaunchKernelGGL(k_ms_init, dim3(1), dim3(64), 0, ctx->stream, S, aux, aux + 2, nshift, prec);
  II_HIP(ctx, hipGetLastError());
  const int* skip = &S->done;
  int queued = 0, chunk = 1;
  bool conv = false;
  while (queued < maxiter) {
    int n = std::min(chunk, maxiter - queued);
    for (int it = 0; it < n; ++it) {
      II_CHECK(apply_A(p, skip));
      if (full) {
        hipLaunchKernelGGL(k_ms_sp<24>, dim3(nb), dim3(kBlock), 0, ctx->stream, g,
                           (double2*)fs.v, p, S, part);
        hipLaunchKernelGGL(k_ms_fin_pap, dim3(1), dim3(kBlock), 0, ctx->stream, part, nb, S);
        hipLaunchKernelGGL(k_ms_xr<24>, dim3(nb), dim3(kBlock), 0, ctx->stream, g, x, p,
                           (double2*)fr.v, (const double2*)fs.v, S, part);
        hipLaunchKernelGGL(k_ms_fin_rr, dim3(1), dim3(kBlock), 0, ctx->stream, part, nb, S);
        hipLaunchKernelGGL(k_ms_vec<24>, dim3(nb, nshift), dim3(kBlock), 0, ctx->stream, g, x, p,
                           (const double2*)fr.v, S, part);
      } else {
        hipLaunchKernelGGL(k_ms_sp<12>, dim3(nb), dim3(kBlock), 0, ctx->stream, g,
                           (double2*)fs.v, p, S, part);
        hipLaunchKernelGGL(k_ms_fin_pap, dim3(1), dim3(kBlock), 0, ctx->stream, part, nb, S);
        hipLaunchKernelGGL(k_ms_xr<12>, dim3(nb), dim3(kBlock), 0, ctx->stream, g, x, p,
                           (double2*)fr.v, (const double2*)fs.v, S, part);
        hipLaunchKernelGGL(k_ms_fin_rr, dim3(1), dim3(kBlock), 0, ctx->stream, part, nb, S);
        hipLaunchKernelGGL(k_ms_vec<12>, dim3(nb, nshift), dim3(kBlock), 0, ctx->stream, g, x, p,
                           (const double2*)fr.v, S, part);
      }
      hipLaunchKernelGGL(k_ms_tail, dim3(1), dim3(kBlock), 0, ctx->stream, part, nb, S);
      II_HIP(ctx, hipGetLastError());
    }
    queued += n;
    II_HIP(ctx, hipMemcpyAsync(Sh, S, sizeof(MSScal), hipMemcpyDeviceToHost, ctx->stream));
    II_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (Sh->nan) return set_error(ctx, II_ERR_NAN, "non-finite multishift scalar");
    if (Sh->done) { conv = true; break; }
    chunk = std::min(chunk * 2, 64);
  }
  // true residuals (:165-177): diff = max_i |(A + sigma_i) x_i - b|^2 / |b|^2
  double dmax = -1.0;
  if (per_shift) per_shift->assign(nshift, 0.0);
  for (int i = 0; i < nshift; ++i) {
    II_CHECK(apply_A(x + i * he, nullptr));
    II_CHECK(axpby64(ctx, shifts[i], x + i * he, 1.0, (double2*)fs.v, nh));
    II_CHECK(axpby64(ctx, -1.0, b, 1.0, (double2*)fs.v, nh));
    II_CHECK(norm2((double2*)fs.v, aux));
    double h = 0.0;
    II_HIP(ctx, hipMemcpyAsync(&h, aux, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    II_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double d1 = h * Sh->snorm;
    if (per_shift) (*per_shift)[i] = std::sqrt(d1);
    if (d1 > dmax) dmax = d1;
  }
  II_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0.f;
  II_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (out) {
    out->iterations = conv ? Sh->nconv : -1;
    out->diff = dmax;
    out->true_residual = std::sqrt(dmax);
    out->timing_ms = ms;
    out->restarts = 0;
    out->inner_iterations = Sh->count;
  }
  if (!conv) return set_error(ctx, II_ERR_NOT_CONVERGED, "Not converged.");
  return II_OK;
}

// ============================================================ mixed multishift
// Stage 1: the multishift recursion of multishift_reference (all shifts
// fused in each launch) with FP32 vectors and FP32 operator (r12 FP32
// links), FP64 scalars, run to |r|^2/|b|^2 < prec1.  Stage 2: every shifted
// system (A + sigma_i) x_i = b is refined from the stage-1 x_i by a batched
// reliable-update CG (inner precision P, FP64 true residuals) until
// |b - (A + sigma_i) x_i| <= tol |b|.

// s += sigma_r p ; partial <p, s>   (per rhs shift of the batched CG)
template <typename SpP, typename SpS, typename StS>
__global__ void __launch_bounds__(kBlock)
k_sp_shift(Geom g, SpP p, SpS s, StS sw, const double* __restrict__ sig, RUScal* __restrict__ sc,
           double* __restrict__ partials) {
  const int rhs = blockIdx.y;
  if (sc[rhs].done) return;
  const double sg = sig[rhs];
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (hs < g.Vh) {
    double pv[24], sv[24];
    ld_site(p.rhs(rhs), hs, pv);
    ld_site(s.rhs(rhs), hs, sv);
#pragma unroll
    for (int k = 0; k < 24; ++k) {
      sv[k] = fma(sg, pv[k], sv[k]);
      acc = fma(pv[k], sv[k], acc);
    }
    st_site(sw.rhs(rhs), hs, sv);
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partials[(size_t)rhs * gridDim.x + blockIdx.x] = acc;
}

// rho = b - q - sigma_r y ; partial |rho|^2, masked (FP64; one b for all rhs)
__global__ void __launch_bounds__(kBlock)
k_resid_shift(Geom g, const double2* __restrict__ b, const double2* __restrict__ q,
              const double2* __restrict__ y, const double* __restrict__ sig,
              double2* __restrict__ rho, const int* __restrict__ skip, int skip_stride,
              double* __restrict__ partials) {
  const int rhs = blockIdx.y;
  if (skip && skip[rhs * skip_stride]) return;
  const double sg = sig[rhs];
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (hs < g.Vh) {
    const size_t off = (size_t)rhs * 12 * g.stride + hs;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const size_t i = (size_t)k * g.stride;
      const double2 bv = b[i + hs], qv = q[off + i], yv = y[off + i];
      const double2 d = make_double2(bv.x - qv.x - sg * yv.x, bv.y - qv.y - sg * yv.y);
      rho[off + i] = d;
      acc = fma(d.x, d.x, fma(d.y, d.y, acc));
    }
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partials[(size_t)rhs * gridDim.x + blockIdx.x] = acc;
}

template <int P>
ii_status sp_shift_t(ii_context* ctx, const SF& p, const SF& s, const double* sig, RUScal* sc,
                     int nrhs, double* partials) {
  const Geom& g = ctx->g;
  using A = Acc<P>;
  hipLaunchKernelGGL((k_sp_shift<typename A::Sp, typename A::Sp, typename A::St>),
                     dim3(blocks_for(g.Vh), nrhs), dim3(kBlock), 0, ctx->stream, g,
                     A::sp(p, g.stride), A::sp(s, g.stride), A::st(s, g.stride), sig, sc,
                     partials);
  II_HIP(ctx, hipGetLastError());
  return II_OK;
}
ii_status sp_shift_P(ii_context* ctx, int P, const SF& p, const SF& s, const double* sig,
                     RUScal* sc, int nrhs, double* partials) {
  switch (P) {
    case 0: return sp_shift_t<0>(ctx, p, s, sig, sc, nrhs, partials);
    case 1: return sp_shift_t<1>(ctx, p, s, sig, sc, nrhs, partials);
    case 2: return sp_shift_t<2>(ctx, p, s, sig, sc, nrhs, partials);
    default: return sp_shift_t<3>(ctx, p, s, sig, sc, nrhs, partials);
  }
}

// Stage 1 (FP32 vectors): x32 receives nshift FP32 half fields.
ii_status multishift_sloppy(ii_context* ctx, double kappa, float2* x32, const double2* b64,
                            const double* shifts, int nshift, double prec1, int maxiter,
                            int* iterations) {
  const Geom& g = ctx->g;
  if (!ctx->gauge_lp[II_PREC_SINGLE]) II_CHECK(build_lowprec_gauge(ctx, II_PREC_SINGLE));
  const int nb = hop_blocks(g);
  II_CHECK(ensure_partials(ctx, (size_t)nb * nshift + 16));
  double* part = ctx->partials;
  void* wp = nullptr;
  II_CHECK(workspace(ctx, 29, sizeof(MSScal) + (kMaxShift + 8) * sizeof(double), &wp));
  MSScal* S = (MSScal*)wp;
  double* aux = (double*)((char*)wp + sizeof(MSScal));
  MSScal* Sh = nullptr;
  {
    void* hp = nullptr;
    II_CHECK(ensure_host_pin(ctx, sizeof(MSScal), &hp));
    Sh = (MSScal*)hp;
  }
  SF fr, fs, fq, ft, fp;
  II_CHECK(fieldP(ctx, 32, II_PREC_SINGLE, 1, &fr));
  II_CHECK(fieldP(ctx, 33, II_PREC_SINGLE, 1, &fs));
  II_CHECK(fieldP(ctx, 34, II_PREC_SINGLE, 1, &fq));
  II_CHECK(fieldP(ctx, 35, II_PREC_SINGLE, 1, &ft));
  II_CHECK(fieldP(ctx, 36, II_PREC_SINGLE, nshift, &fp));
  const size_t he = half_elems(g);
  float2* p = (float2*)fp.v;
  float2* r = (float2*)fr.v;
  II_CHECK(convert_64_to_32(ctx, b64, r, 1));
  for (int i = 0; i < nshift; ++i) {
    II_HIP(ctx, hipMemcpyAsync(p + i * he, r, he * sizeof(float2), hipMemcpyDeviceToDevice,
                               ctx->stream));
    II_HIP(ctx, hipMemsetAsync(x32 + i * he, 0, he * sizeof(float2), ctx->stream));
  }
  II_CHECK(dot64(ctx, b64, b64, aux));
  II_CHECK(dot64(ctx, b64, b64, aux + 1));
  II_HIP(ctx, hipMemcpyAsync(aux + 2, shifts, nshift * sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
  hipLaunchKernelGGL(k_ms_init, dim3(1), dim3(64), 0, ctx->stream, S, aux, aux + 2, nshift,
                     prec1);
  II_HIP(ctx, hipGetLastError());
  const int* skip = &S->done;
  int queued = 0, chunk = 1;
  while (queued < maxiter) {
    const int n = std::min(chunk, maxiter - queued);
    for (int it = 0; it < n; ++it) {
      II_CHECK(ddagdP(ctx, II_PREC_SINGLE, kappa, SF{p}, fs, fq, ft, nullptr, nullptr, 1, skip,
                      0));
      hipLaunchKernelGGL((k_ms_sp<12, float2>), dim3(nb), dim3(kBlock), 0, ctx->stream, g,
                         (float2*)fs.v, p, S, part);
      hipLaunchKernelGGL(k_ms_fin_pap, dim3(1), dim3(kBlock), 0, ctx->stream, part, nb, S);
      hipLaunchKernelGGL((k_ms_xr<12, float2>), dim3(nb), dim3(kBlock), 0, ctx->stream, g, x32,
                         p, r, (const float2*)fs.v, S, part);
      hipLaunchKernelGGL(k_ms_fin_rr, dim3(1), dim3(kBlock), 0, ctx->stream, part, nb, S);
      hipLaunchKernelGGL((k_ms_vec<12, float2>), dim3(nb, nshift), dim3(kBlock), 0, ctx->stream,
                         g, x32, p, (const float2*)r, S, part);
      hipLaunchKernelGGL(k_ms_tail, dim3(1), dim3(kBlock), 0, ctx->stream, part, nb, S);
      II_HIP(ctx, hipGetLastError());
    }
    queued += n;
    II_HIP(ctx, hipMemcpyAsync(Sh, S, sizeof(MSScal), hipMemcpyDeviceToHost, ctx->stream));
    II_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (Sh->nan) return set_error(ctx, II_ERR_NAN, "non-finite multishift scalar");
    if (Sh->done) break;
    chunk = std::min(chunk * 2, 64);
  }
  *iterations = Sh->count;
  return II_OK;
}

// Stage 2: batched reliable-update CG on (A + sigma_r) x_r = b for all r,
// starting from x64 (in/out, nrhs FP64 half fields).
ii_status shifted_refine(ii_context* ctx, double kappa, const double* shifts, int nrhs,
                         const double2* b64, double2* x64, double tol, int P, int maxiter,
                         ii_solver_output* outs) {
  const Geom& g = ctx->g;
  if (P != II_PREC_DOUBLE && !ctx->gauge_lp[P]) II_CHECK(build_lowprec_gauge(ctx, P));
  const bool lo = P != II_PREC_DOUBLE;
  const int ytype = lo ? II_PREC_SINGLE : II_PREC_DOUBLE;
  const double delta = lo ? 0.1 : 0.0;
  const int nb = hop_blocks(g);
  II_CHECK(ensure_partials(ctx, (size_t)std::max(nb, hop_dot_parts(g, nrhs, P)) * nrhs + 16));
  II_CHECK(ensure_ru_scalars(ctx, nrhs));
  double* part = ctx->partials;
  RUScal* sc = dev_scal(ctx);
  const int ss = kScalInts;
  const int* skip_it = &sc[0].done;
  const int* skip_ru = &sc[0].skip_ru;
  void* auxp = nullptr;
  II_CHECK(workspace(ctx, 37, sizeof(double) * 3 * nrhs + 64, &auxp));
  double* sig = (double*)auxp;
  double* rr0 = sig + nrhs;
  double* bb = rr0 + nrhs;
  II_HIP(ctx, hipMemcpyAsync(sig, shifts, nrhs * sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
  SF q64, t64, u64, rho, ps, rs, ss_, qs, ts, ys;
  II_CHECK(fieldP(ctx, 38, 0, nrhs, &q64));  // (A) y
  II_CHECK(fieldP(ctx, 39, 0, nrhs, &t64));  // Dhat y, intermediate
  II_CHECK(fieldP(ctx, 41, 0, nrhs, &u64));  // odd temporary
  II_CHECK(fieldP(ctx, 40, 0, nrhs, &rho));
  if (lo) {
    II_CHECK(fieldP(ctx, 50, P, nrhs, &ps));
    II_CHECK(fieldP(ctx, 51, P, nrhs, &rs));
    II_CHECK(fieldP(ctx, 52, P, nrhs, &ss_));
    II_CHECK(fieldP(ctx, 53, P, nrhs, &qs));
    II_CHECK(fieldP(ctx, 54, P, nrhs, &ts));
    II_CHECK(fieldP(ctx, 55, II_PREC_SINGLE, nrhs, &ys));
  } else {
    II_CHECK(fieldP(ctx, 50, 0, nrhs, &ps));
    rs = rho;
    II_CHECK(fieldP(ctx, 52, 0, nrhs, &ss_));
    qs = t64;
    ts = u64;
    ys = SF{x64};
  }
  auto t0 = std::chrono::steady_clock::now();
  // initial true residual rho = b - (A + sigma) x and |b|^2
  II_CHECK(ddagdP(ctx, 0, kappa, SF{x64}, q64, t64, u64, nullptr, nullptr, nrhs, nullptr, 0));
  hipLaunchKernelGGL(k_resid_shift, dim3(nb, nrhs), dim3(kBlock), 0, ctx->stream, g, b64,
                     (const double2*)q64.v, (const double2*)x64, sig, (double2*)rho.v, nullptr,
                     0, part);
  II_HIP(ctx, hipGetLastError());
  II_CHECK(finalize_sums(ctx, part, nb, rr0, nrhs));
  {
    double h = 0.0;
    II_CHECK(dot64(ctx, b64, b64, bb));
    II_HIP(ctx, hipMemcpyAsync(&h, bb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    II_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<double> hb(nrhs, h);
    II_HIP(ctx, hipMemcpyAsync(bb, hb.data(), nrhs * sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream));
    II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  hipLaunchKernelGGL(k_scal_init, dim3(nrhs), dim3(64), 0, ctx->stream, sc, rr0, bb, 0.0, tol,
                     delta, 1);
  II_HIP(ctx, hipGetLastError());
  if (lo) {
    II_CHECK(convertP(ctx, 0, P, rho, rs, nrhs));
    II_CHECK(convertP(ctx, 0, P, rho, ps, nrhs));
    II_HIP(ctx, hipMemsetAsync(ys.v, 0, field_bytes(g, II_PREC_SINGLE) * nrhs, ctx->stream));
  } else {
    II_CHECK(copy64(ctx, (double2*)ps.v, (double2*)rho.v, nrhs));
  }
  int its = 0, chunk = 1;
  bool all_done = false;
  while (its < maxiter && !all_done) {
    for (int i = 0; i < chunk; ++i) {
      II_CHECK(ddagdP(ctx, P, kappa, ps, ss_, qs, ts, nullptr, nullptr, nrhs, skip_it, ss));
      II_CHECK(sp_shift_P(ctx, P, ps, ss_, sig, sc, nrhs, part));
      hipLaunchKernelGGL(k_fin_pap, dim3(nrhs), dim3(kBlock), 0, ctx->stream, part, nb, sc,
                         nullptr, 0, 0);
      II_CHECK(cg_blas_P(ctx, P, ps, ss_, rs, ys, ytype, sc, nrhs, part, nb));
      hipLaunchKernelGGL(k_fin_rr, dim3(nrhs), dim3(kBlock), 0, ctx->stream, part, nb, sc, 1);
      II_CHECK(cg_p_P(ctx, P, rs, ps, ps, sc, nrhs));
    }
    its += chunk;
    II_CHECK(read_scalars(ctx, nrhs));
    RUScal* H = host_scal(ctx);
    bool any_upd = false;
    all_done = true;
    for (int k = 0; k < nrhs; ++k) {
      if (H[k].nan) return set_error(ctx, II_ERR_NAN, "non-finite solver scalar");
      if (!H[k].done) all_done = false;
      if (!H[k].done && !H[k].skip_ru) any_upd = true;
    }
    if (all_done) break;
    if (any_upd) {
      if (lo) {
        hipLaunchKernelGGL((k_accum<SpF32, StF32>), dim3(blocks_for(g.Vh), nrhs), dim3(kBlock), 0,
                           ctx->stream, g, Acc<1>::sp(ys, g.stride), Acc<1>::st(ys, g.stride),
                           x64, skip_ru, ss);
        II_HIP(ctx, hipGetLastError());
      }
      II_CHECK(ddagdP(ctx, 0, kappa, SF{x64}, q64, t64, u64, nullptr, nullptr, nrhs, skip_ru,
                      ss));
      hipLaunchKernelGGL(k_resid_shift, dim3(nb, nrhs), dim3(kBlock), 0, ctx->stream, g, b64,
                         (const double2*)q64.v, (const double2*)x64, sig, (double2*)rho.v,
                         skip_ru, ss, part);
      hipLaunchKernelGGL(k_fin_rho, dim3(nrhs), dim3(kBlock), 0, ctx->stream, part, nb, sc);
      hipLaunchKernelGGL(k_norm2_masked, dim3(nb, nrhs), dim3(kBlock), 0, ctx->stream, g,
                         (const double2*)rho.v, skip_ru, ss, part);
      if (lo) II_CHECK(convertP(ctx, 0, P, rho, rs, nrhs, skip_ru, ss));
      hipLaunchKernelGGL(k_fin_ru, dim3(nrhs), dim3(kBlock), 0, ctx->stream, part, nb, sc);
      II_HIP(ctx, hipGetLastError());
      // restart the search direction where the recursion broke (keep_p == 0)
      II_CHECK(convertP(ctx, 0, P, rho, ps, nrhs, &sc[0].keep_p, ss));
      chunk = 1;
    } else {
      chunk = std::min(chunk * 2, 8);
    }
  }
  if (lo) {
    hipLaunchKernelGGL((k_accum<SpF32, StF32>), dim3(blocks_for(g.Vh), nrhs), dim3(kBlock), 0,
                       ctx->stream, g, Acc<1>::sp(ys, g.stride), Acc<1>::st(ys, g.stride), x64,
                       nullptr, 0);
    II_HIP(ctx, hipGetLastError());
  }
  II_CHECK(read_scalars(ctx, nrhs));
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  RUScal* H = host_scal(ctx);
  bool failed = false;
  for (int k = 0; k < nrhs; ++k) {
    outs[k].iterations = H[k].count;
    outs[k].inner_iterations = lo ? H[k].count : 0;
    outs[k].restarts = H[k].updates;
    outs[k].true_residual = std::sqrt(H[k].rho2 / H[k].b2);
    outs[k].diff = outs[k].true_residual;
    outs[k].timing_ms = ms;
    if (!H[k].done) failed = true;
  }
  if (failed)
    return set_error(ctx, II_ERR_NOT_CONVERGED, "shifted refinement did not reach the tolerance");
  return II_OK;
}

// ============================================================ propagator
// Work slots 40..63 belong to the propagator solver.
ii_status propagator(ii_context* ctx, double kappa, const ii_prop_params* prm,
                     const ii_field* src, ii_field* sol, ii_solver_output* outs) {
  const Geom& g = ctx->g;
  const int nrhs_all = src->nrhs;
  const int P = prm->precision;
  if (P < 0 || P > 3) return set_error(ctx, II_ERR_BAD_ARG, "bad precision");
  if (P != II_PREC_DOUBLE && !ctx->gauge_lp[P]) II_CHECK(build_lowprec_gauge(ctx, P));
  const int batch = prm->batched ? nrhs_all : 1;
  const double tol = prm->tol > 0 ? prm->tol : 1e-10;
  const double delta = P == II_PREC_DOUBLE ? 0.0 : (prm->delta > 0 ? prm->delta : 0.1);
  const int maxiter = prm->maxiter > 0 ? prm->maxiter : 100000;
  const size_t he = half_elems(g);
  const int nb = hop_blocks(g);
  II_CHECK(ensure_partials(ctx, (size_t)std::max(nb, hop_dot_parts(g, batch, P)) * batch + 16));
  II_CHECK(ensure_ru_scalars(ctx, batch));
  double* part = ctx->partials;
  RUScal* sc = dev_scal(ctx);
  const int ss = kScalInts;
  const int* skip_it = &sc[0].done;
  const int* skip_ru = &sc[0].skip_ru;
  void* auxp = nullptr;
  // aux: |b_e|^2 [0, batch), |b_o|^2 [batch, 2 batch), |c|^2 [2 batch, 3 batch)
  II_CHECK(workspace(ctx, 40, sizeof(double) * 3 * batch + 64, &auxp));
  double* aux = (double*)auxp;
  double* one = aux + 3 * batch;  // the device scalar 1.0 (alpha of rho = bhat - Dhat y)
  // FP64 fields
  SF be, bo, bh, y, r64, q64, t64, rho;
  II_CHECK(fieldP(ctx, 41, 0, batch, &be));
  II_CHECK(fieldP(ctx, 42, 0, batch, &bo));
  II_CHECK(fieldP(ctx, 43, 0, batch, &bh));
  II_CHECK(fieldP(ctx, 44, 0, batch, &y));
  II_CHECK(fieldP(ctx, 45, 0, batch, &r64));
  II_CHECK(fieldP(ctx, 46, 0, batch, &q64));
  II_CHECK(fieldP(ctx, 47, 0, batch, &t64));
  rho = q64;  // q64 is free once rho is formed
  // sloppy fields
  SF ps, rs, ss_, qs, ts, ys;
  const bool lo = P != II_PREC_DOUBLE;
  const int ytype = lo ? II_PREC_SINGLE : II_PREC_DOUBLE;
  if (lo) {
    II_CHECK(fieldP(ctx, 50, P, batch, &ps));
    II_CHECK(fieldP(ctx, 51, P, batch, &rs));  // (A p is not stored: fused CG step)
    II_CHECK(fieldP(ctx, 53, P, batch, &qs));
    II_CHECK(fieldP(ctx, 54, P, batch, &ts));
    II_CHECK(fieldP(ctx, 55, II_PREC_SINGLE, batch, &ys));
  } else {
    II_CHECK(fieldP(ctx, 50, 0, batch, &ps));
    rs = r64;
    II_CHECK(fieldP(ctx, 52, 0, batch, &ss_));
    qs = q64;
    ts = t64;
    ys = y;
  }
  // Deferred solution update (sloppy precisions): the search directions of
  // up to kDefer iterations stay in a ring of kDefer + 1 fields, their
  // alphas in ahist[rhs][slot], and k_cg_flush adds them to ys in one pass.
  const bool defer = lo;
  const int nring = defer ? kDefer + 1 : 1;
  std::vector<SF> ring(nring);
  ring[0] = ps;
  for (int i = 1; i < nring; ++i) II_CHECK(fieldP(ctx, 63 + i, P, batch, &ring[i]));
  double* ahist = nullptr;
  if (defer) {
    void* v = nullptr;
    II_CHECK(workspace(ctx, 62, sizeof(double) * nring * batch, &v));
    ahist = (double*)v;
  }
  HopSpec h;
  for (int r0 = 0; r0 < nrhs_all; r0 += batch) {
    const int nr = std::min(batch, nrhs_all - r0);
    auto t0 = std::chrono::steady_clock::now();
    II_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    // gather even/odd parts of the sources
    for (int k = 0; k < nr; ++k) {
      const double2* base = src->data + (size_t)(r0 + k) * 2 * he;
      II_CHECK(copy64(ctx, (double2*)be.v + k * he, base));
      II_CHECK(copy64(ctx, (double2*)bo.v + k * he, base + he));
    }
    // |b|^2 of the full sources
    II_CHECK(dot64(ctx, (double2*)be.v, (double2*)be.v, aux, nr));
    II_CHECK(dot64(ctx, (double2*)bo.v, (double2*)bo.v, aux + batch, nr));
    // bhat = b_e - mult_eo(b_o) = b_e + kappa H_eo b_o   (eoUtils.cpp:54-57)
    h = HopSpec();
    h.out_par = 0;
    h.a = kappa;
    h.xpay = true;
    h.nrhs = nr;
    II_CHECK(hop64(ctx, h, (double2*)bo.v, (double2*)bh.v, (double2*)be.v, nullptr));
    // c = Dhat^dag bhat  -> r64 ; y = 0
    II_CHECK(dhat_dag64(ctx, kappa, (double2*)bh.v, (double2*)r64.v, (double2*)t64.v, nullptr,
                        nullptr, nr));
    II_CHECK(zero64(ctx, (double2*)y.v, nr));
    // rr = |c|^2 into aux[2*batch ...]
    II_CHECK(dot64(ctx, (double2*)r64.v, (double2*)r64.v, aux + 2 * batch, nr));
    // b2 = |b_e|^2 + |b_o|^2 -> aux[0..nr)
    {
      std::vector<double> hb(2 * batch);
      II_HIP(ctx, hipMemcpyAsync(hb.data(), aux, 2 * batch * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream));
      II_HIP(ctx, hipStreamSynchronize(ctx->stream));
      for (int k = 0; k < nr; ++k) hb[k] += hb[batch + k];
      II_HIP(ctx, hipMemcpyAsync(aux, hb.data(), nr * sizeof(double), hipMemcpyHostToDevice,
                                 ctx->stream));
      const double unit = 1.0;
      II_HIP(ctx, hipMemcpyAsync(one, &unit, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      II_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    hipLaunchKernelGGL(k_scal_init, dim3(nr), dim3(64), 0, ctx->stream, sc, aux + 2 * batch, aux,
                       0.0, tol, delta, 1, prm->restart_factor > 0 ? prm->restart_factor : 16.0);
    II_HIP(ctx, hipGetLastError());
    if (lo) {
      II_CHECK(convertP(ctx, 0, P, r64, rs, nr));
      II_CHECK(convertP(ctx, 0, P, r64, ps, nr));
      II_CHECK(convertP(ctx, 0, II_PREC_SINGLE, y, ys, nr));  // zeros
    } else {
      II_CHECK(copy64(ctx, (double2*)ps.v, (double2*)r64.v, nr));
    }
    int its = 0;
    bool all_done = false;
    int chunk = 1;
    int pos = 0;             // ring slot of the current search direction
    std::vector<int> pend;   // ring slots whose alpha p is not yet in ys
    // one CG iteration on the direction in ring[cur]: 4 hops (the last with
    // <p, A p>), scalars, the r update with |r|^2 (with the y update unless
    // deferred), scalars, the p update into the next ring slot.  Every
    // argument is fixed for a ring slot and every kernel reads its rhs's done
    // flag on the device, so the sequence for each slot can be recorded once
    // into a HIP graph and replayed.
    auto iteration = [&](int cur) -> ii_status {
      const SF& p = ring[cur];
      if (defer) {
        II_CHECK(ddagd_cg_fused(ctx, P, kappa, p, rs, qs, ts, sc, part, nr, skip_it, ss, ahist,
                                cur, nring));
        hipLaunchKernelGGL(k_fin_rr, dim3(nr), dim3(kFinThreads), 0, ctx->stream, part,
                           hop_dot_parts(g, nr, P), sc, 1);
      } else {
        II_CHECK(ddagdP(ctx, P, kappa, p, ss_, qs, ts, &p, part, nr, skip_it, ss));
        hipLaunchKernelGGL(k_fin_pap, dim3(nr), dim3(kFinThreads), 0, ctx->stream, part,
                           hop_dot_parts(g, nr, P), sc, nullptr, 0, 0);
        II_CHECK(cg_blas_P(ctx, P, p, ss_, rs, ys, ytype, sc, nr, part, nb));
        hipLaunchKernelGGL(k_fin_rr, dim3(nr), dim3(kFinThreads), 0, ctx->stream, part, nb, sc, 1);
      }
      return cg_p_P(ctx, P, rs, p, ring[(cur + 1) % nring], sc, nr);
    };
    auto flush = [&]() -> ii_status {
      if (pend.empty()) return II_OK;
      II_CHECK(cg_flush_P(ctx, P, ring.data(), pend.data(), (int)pend.size(), ahist, nring, ys,
                          nr));
      pend.clear();
      return II_OK;
    };
    // the pending directions plus ys into the FP64 solution y (rhs not
    // masked by skip), ys restarted at 0
    auto flush_fold = [&](const int* skip) -> ii_status {
      II_CHECK(cg_flush_P(ctx, P, ring.data(), pend.data(), (int)pend.size(), ahist, nring, ys,
                          nr, (double2*)y.v, skip, ss));
      pend.clear();
      return II_OK;
    };
    std::vector<hipGraphExec_t> gexec(nring, nullptr);
    struct GraphGuard {
      std::vector<hipGraphExec_t>* e;
      ~GraphGuard() {
        for (hipGraphExec_t x : *e)
          if (x) (void)hipGraphExecDestroy(x);
      }
    } gguard{&gexec};
    auto launch = [&](int cur) -> ii_status {
      if (!prm->use_graph) return iteration(cur);
      if (!gexec[cur]) {
        hipGraph_t graph = nullptr;
        II_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        ii_status cst = iteration(cur);
        hipError_t ee = hipStreamEndCapture(ctx->stream, &graph);
        II_CHECK(cst);
        if (ee != hipSuccess) return set_error(ctx, II_ERR_DEVICE, "graph capture failed");
        hipError_t ie = hipGraphInstantiate(&gexec[cur], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) return set_error(ctx, II_ERR_DEVICE, "graph instantiation failed");
      }
      II_HIP(ctx, hipGraphLaunch(gexec[cur], ctx->stream));
      return II_OK;
    };
    while (its < maxiter && !all_done) {
      for (int i = 0; i < chunk; ++i) {
        II_CHECK(launch(pos));
        if (defer) {
          pend.push_back(pos);
          pos = (pos + 1) % nring;
          if ((int)pend.size() == kDefer) II_CHECK(flush());
        }
      }
      its += chunk;
      II_CHECK(read_scalars(ctx, nr));
      RUScal* H = host_scal(ctx);
      bool any_upd = false;
      all_done = true;
      for (int k = 0; k < nr; ++k) {
        if (H[k].nan)
          return set_error(ctx, II_ERR_NAN,
                           "non-finite solver scalar (rhs " + std::to_string(r0 + k) +
                               ", iteration " + std::to_string(H[k].count) +
                               ", rr " + std::to_string(H[k].rr) + ", pap " +
                               std::to_string(H[k].pap) + ", b2 " + std::to_string(H[k].b2) +
                               ")");
        if (!H[k].done) all_done = false;
        if (!H[k].done && !H[k].skip_ru) any_upd = true;
      }
      if (all_done) break;
      if (any_upd) {
        // reliable update, FP64, for the flagged rhs only
        II_CHECK(join_updates(ctx, nr));
        if (lo) II_CHECK(flush_fold(skip_ru));
        // q = Dhat y ; rho = bhat - q ; |rho|^2
        h = HopSpec();
        h.nrhs = nr;
        h.skip = skip_ru;
        h.skip_stride = ss;
        h.out_par = 1;
        h.a = 1.0;
        II_CHECK(hop64(ctx, h, (double2*)y.v, (double2*)t64.v, nullptr, nullptr));
        // rho = bhat - (y - kappa^2 H_eo t) and |rho|^2 in the hop's epilogue
        h.out_par = 0;
        h.a = -kappa * kappa;
        h.xpay = true;
        h.epi = kEpiRUpd;
        h.coef = one;
        h.coef_stride = 0;
        h.partials = part;
        II_CHECK(hop64(ctx, h, (double2*)t64.v, (double2*)rho.v, (double2*)y.v,
                       (const double2*)bh.v));
        hipLaunchKernelGGL(k_fin_rho, dim3(nr), dim3(kFinThreads), 0, ctx->stream, part,
                           hop_dot_parts(g, nr, II_PREC_DOUBLE), sc);
        // new CG residual r = Dhat^dag rho and |r|^2.  The true residual rho
        // is FP64; r only restarts the sloppy recursion, so with a sloppy
        // inner precision it is formed in that precision from rho.
        h = HopSpec();
        h.nrhs = nr;
        h.skip = skip_ru;
        h.skip_stride = ss;
        h.out_par = 1;
        h.a = 1.0;
        h.g5in = true;
        if (lo) {
          II_CHECK(convertP(ctx, 0, P, rho, qs, nr, skip_ru, ss));
          II_CHECK(hopP(ctx, P, h, qs, ts, nullptr, nullptr));
          h.out_par = 0;
          h.a = -kappa * kappa;
          h.xpay = true;
          h.g5in = false;
          h.g5out = true;
          II_CHECK(hopP(ctx, P, h, ts, rs, &qs, nullptr));
          II_CHECK(norm2P(ctx, P, rs, nr, skip_ru, ss, part));
        } else {
          // in FP64 the sloppy residual is r64 itself
          II_CHECK(hop64(ctx, h, (double2*)rho.v, (double2*)t64.v, nullptr, nullptr));
          h.out_par = 0;
          h.a = -kappa * kappa;
          h.xpay = true;
          h.g5in = false;
          h.g5out = true;
          II_CHECK(hop64(ctx, h, (double2*)t64.v, (double2*)r64.v, (double2*)rho.v, nullptr));
          hipLaunchKernelGGL(k_norm2_masked, dim3(nb, nr), dim3(kBlock), 0, ctx->stream, g,
                             (const double2*)r64.v, skip_ru, ss, part);
        }
        hipLaunchKernelGGL(k_fin_ru, dim3(nr), dim3(kFinThreads), 0, ctx->stream, part, nb, sc);
        II_HIP(ctx, hipGetLastError());
        // restart the search direction where the recursion broke (keep_p == 0)
        if (lo) II_CHECK(convertP(ctx, P, P, rs, ring[pos], nr, &sc[0].keep_p, ss));
        else II_CHECK(convertP(ctx, 0, 0, r64, ring[pos], nr, &sc[0].keep_p, ss));
        chunk = 1;
      } else {
        chunk = std::min(chunk * 2, 8);
      }
    }
    // fold the sloppy accumulator in (rhs not yet folded at convergence)
    if (lo) II_CHECK(flush_fold(nullptr));
    // x_o = b_o - mult_oe(y) = b_o + kappa H_oe y   (eoUtils.cpp:67)
    h = HopSpec();
    h.out_par = 1;
    h.a = kappa;
    h.xpay = true;
    h.nrhs = nr;
    II_CHECK(hop64(ctx, h, (double2*)y.v, (double2*)t64.v, (double2*)bo.v, nullptr));
    for (int k = 0; k < nr; ++k) {
      double2* base = sol->data + (size_t)(r0 + k) * 2 * he;
      II_CHECK(copy64(ctx, base, (double2*)y.v + k * he));
      II_CHECK(copy64(ctx, base + he, (double2*)t64.v + k * he));
    }
    II_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    II_CHECK(read_scalars(ctx, nr));
    float ms = 0.f;
    II_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    (void)t0;
    RUScal* H = host_scal(ctx);
    bool failed = false;
    for (int k = 0; k < nr; ++k) {
      ii_solver_output& o = outs[r0 + k];
      o.iterations = H[k].count;
      o.inner_iterations = lo ? H[k].count : 0;
      o.restarts = H[k].updates;
      o.recursion_restarts = H[k].restarts;
      o.diff = std::sqrt(H[k].rho2 / H[k].b2);
      o.timing_ms = ms;
      o.true_residual = -1.0;
      if (!H[k].done) failed = true;
    }
    if (failed)
      return set_error(ctx, II_ERR_NOT_CONVERGED, "propagator solve did not reach the tolerance");
  }
  return II_OK;
}

// ============================================================ BiCGStab
// Batched mixed-precision BiCGStab on Dhat (van der Vorst; no reference
// counterpart: the reference solves Dhat^dag Dhat with Solver_CG).  Per
// iteration and rhs, with complex scalars (Dhat is not Hermitian):
//   v = Dhat p, <r0, v>         (hops; kEpiCDot)     alpha = rho / <r0, v>
//   s = r - alpha v
//   t = Dhat s, <t, s>, |t|^2   (hops; kEpiTS)       omega = <t, s> / |t|^2
//   y += alpha p + omega s ; r = s - omega t ; <r0, r>, |r|^2
//   beta = (<r0, r> / rho)(alpha / omega) ; rho = <r0, r> ; p = r + beta (p - omega v)
// with FP64 reliable updates as in the CG solve: fold y into the FP64
// solution, rho_true = bhat - Dhat y in FP64, r = rho_true; if the true
// residual is far above the recursive one, restart r0 = p = r.
struct BiCoef {
  double rho[2], alpha[2], omega[2], beta[2];  // complex (re, im)
};

__device__ __forceinline__ void cdiv(double ar, double ai, double br, double bi, double& cr,
                                     double& ci) {
  const double d = br * br + bi * bi;
  cr = (ar * br + ai * bi) / d;
  ci = (ai * br - ar * bi) / d;
}
// sum over i < nb of plane j of rhs blockIdx.x
__device__ __forceinline__ double plane_sum(const double* partials, size_t pl, int j, int nb) {
  const double* p = partials + j * pl + (size_t)blockIdx.x * nb;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += p[i];
  return block_sum(s);
}

__global__ void k_bi_init(const RUScal* __restrict__ sc, BiCoef* __restrict__ bc) {
  if (threadIdx.x) return;
  BiCoef z{};
  z.rho[0] = sc[blockIdx.x].rr;  // r0 = r
  bc[blockIdx.x] = z;
}
// alpha = rho / <r0, v>; kept with omega in ahist[(rhs * astride + slot) * 4
// + 0..3] = (alpha, omega) for the deferred solution update (0 once an rhs
// has finished)
__global__ void k_bi_alpha(const double* __restrict__ partials, int nb, size_t pl,
                           RUScal* __restrict__ sc, BiCoef* __restrict__ bc,
                           double* __restrict__ ahist, int aidx, int astride) {
  RUScal& S = sc[blockIdx.x];
  double* ah = ahist + ((size_t)blockIdx.x * astride + aidx) * 4;
  if (S.done) {
    if (threadIdx.x == 0) ah[0] = ah[1] = 0.0;
    return;
  }
  const double re = plane_sum(partials, pl, 0, nb);
  const double im = plane_sum(partials, pl, 1, nb);
  if (threadIdx.x == 0) {
    BiCoef& B = bc[blockIdx.x];
    cdiv(B.rho[0], B.rho[1], re, im, B.alpha[0], B.alpha[1]);
    if (!finite(B.alpha[0]) || !finite(B.alpha[1])) { S.nan = 1; S.done = 1; S.skip_ru = 1; }
    ah[0] = S.done ? 0.0 : B.alpha[0];
    ah[1] = S.done ? 0.0 : B.alpha[1];
  }
}
// omega = <t, s> / |t|^2
__global__ void k_bi_omega(const double* __restrict__ partials, int nb, size_t pl,
                           RUScal* __restrict__ sc, BiCoef* __restrict__ bc,
                           double* __restrict__ ahist, int aidx, int astride) {
  RUScal& S = sc[blockIdx.x];
  double* ah = ahist + ((size_t)blockIdx.x * astride + aidx) * 4;
  if (S.done) {
    if (threadIdx.x == 0) ah[0] = ah[1] = ah[2] = ah[3] = 0.0;
    return;
  }
  const double re = plane_sum(partials, pl, 0, nb);
  const double im = plane_sum(partials, pl, 1, nb);
  const double tt = plane_sum(partials, pl, 2, nb);
  if (threadIdx.x == 0) {
    BiCoef& B = bc[blockIdx.x];
    B.omega[0] = re / tt;
    B.omega[1] = im / tt;
    if (!finite(B.omega[0]) || !finite(B.omega[1])) { S.nan = 1; S.done = 1; S.skip_ru = 1; }
    if (S.done) ah[0] = ah[1] = 0.0;
    ah[2] = S.done ? 0.0 : B.omega[0];
    ah[3] = S.done ? 0.0 : B.omega[1];
  }
}
// after the r update: beta, rho, |r|^2 and the reliable-update decision
__global__ void k_bi_beta(const double* __restrict__ partials, int nb, size_t pl,
                          RUScal* __restrict__ sc, BiCoef* __restrict__ bc) {
  RUScal& S = sc[blockIdx.x];
  if (S.done) return;
  const double re = plane_sum(partials, pl, 0, nb);
  const double im = plane_sum(partials, pl, 1, nb);
  const double rr = plane_sum(partials, pl, 2, nb);
  if (threadIdx.x == 0) {
    BiCoef& B = bc[blockIdx.x];
    double qr, qi, ar, ai;
    cdiv(re, im, B.rho[0], B.rho[1], qr, qi);                    // rho_new / rho
    cdiv(B.alpha[0], B.alpha[1], B.omega[0], B.omega[1], ar, ai);  // alpha / omega
    B.beta[0] = qr * ar - qi * ai;
    B.beta[1] = qr * ai + qi * ar;
    B.rho[0] = re;
    B.rho[1] = im;
    S.rr = rr;
    S.count += 1;
    // a restart (keep_p = 0 from k_bi_ru) applies to the one update event
    // that set it: the masked r -> r0, p copies of later events must not
    // fire for this rhs again
    S.keep_p = 1;
    if (!finite(rr) || !finite(B.beta[0]) || !finite(B.beta[1])) {
      S.nan = 1; S.done = 1; S.skip_ru = 1;
      return;
    }
    const bool upd = (rr < S.delta2 * S.rmax2) || (rr < S.target2);
    S.skip_ru = upd ? 0 : 1;
  }
}
// after a reliable update: |r|^2 and rho = <r0, r> of the new residual; a
// true residual far above the recursive one restarts r0 = p = r (keep_p = 0)
__global__ void k_bi_ru(const double* __restrict__ partials, int nb, size_t pl,
                        RUScal* __restrict__ sc, BiCoef* __restrict__ bc) {
  RUScal& S = sc[blockIdx.x];
  if (S.skip_ru) return;
  const double re = plane_sum(partials, pl, 0, nb);
  const double im = plane_sum(partials, pl, 1, nb);
  const double rr = plane_sum(partials, pl, 2, nb);
  if (threadIdx.x == 0) {
    BiCoef& B = bc[blockIdx.x];
    if (rr > S.restart * S.rr) {
      S.keep_p = 0;
      S.restarts += 1;
      B.rho[0] = rr;
      B.rho[1] = 0.0;
    } else {
      B.rho[0] = re;
      B.rho[1] = im;
    }
    S.rr = rr;
    S.rmax2 = rr;
    // r is the (sloppy) true residual itself: test again below 0.81 tol^2
    const double t = 0.81 * S.tol2;
    S.target2 = t < S.target2 ? t : S.target2;
    S.updates += 1;
    S.skip_ru = 1;
    if (!finite(rr) || !finite(B.rho[0])) { S.nan = 1; S.done = 1; }
  }
}

// s = r - alpha v
template <typename Sp, typename St>
__global__ void __launch_bounds__(kBlock)
k_bi_s(Geom g, Sp r, Sp v, St s, const RUScal* __restrict__ sc, const BiCoef* __restrict__ bc) {
  const int rhs = blockIdx.y;
  if (sc[rhs].done) return;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= g.Vh) return;
  const double ar = bc[rhs].alpha[0], ai = bc[rhs].alpha[1];
  double rv[24], vv[24];
  ld_site(r.rhs(rhs), hs, rv);
  ld_site(v.rhs(rhs), hs, vv);
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const double x = vv[2 * k], y = vv[2 * k + 1];
    rv[2 * k] -= ar * x - ai * y;
    rv[2 * k + 1] -= ar * y + ai * x;
  }
  st_site(s.rhs(rhs), hs, rv);
}
// r = s - omega t ; partial <r0, r> (re, im), |r|^2 (the solution update
// alpha p + omega s is deferred to k_bi_flush)
template <typename Sp, typename St>
__global__ void __launch_bounds__(kBlock)
k_bi_r(Geom g, Sp s, Sp t, Sp r0, St rw, const RUScal* __restrict__ sc,
       const BiCoef* __restrict__ bc, double* __restrict__ partials) {
  const int rhs = blockIdx.y;
  if (sc[rhs].done) return;
  const double wr = bc[rhs].omega[0], wi = bc[rhs].omega[1];
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;
  if (hs < g.Vh) {
    double sv[24], tv[24];
    ld_site(s.rhs(rhs), hs, sv);
    ld_site(t.rhs(rhs), hs, tv);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      sv[2 * k] -= wr * tv[2 * k] - wi * tv[2 * k + 1];
      sv[2 * k + 1] -= wr * tv[2 * k + 1] + wi * tv[2 * k];
    }
    st_site(rw.rhs(rhs), hs, sv);
    ld_site(r0.rhs(rhs), hs, tv);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const double x = tv[2 * k], y = tv[2 * k + 1], u = sv[2 * k], w = sv[2 * k + 1];
      d0 += x * u + y * w;
      d1 += x * w - y * u;
      d2 += u * u + w * w;
    }
  }
  const size_t pl = (size_t)gridDim.y * gridDim.x, o = (size_t)rhs * gridDim.x + blockIdx.x;
  d0 = block_sum(d0);
  if (threadIdx.x == 0) partials[o] = d0;
  d1 = block_sum(d1);
  if (threadIdx.x == 0) partials[pl + o] = d1;
  d2 = block_sum(d2);
  if (threadIdx.x == 0) partials[2 * pl + o] = d2;
}

// Deferred BiCGStab solution update: y += sum_j (alpha_j p_j + omega_j s_j)
// over up to kBiDefer stored iterations (coefficients from ahist as written
// by k_bi_alpha / k_bi_omega); with y64, the rhs not masked by skip fold the
// sum into the FP64 solution and restart y at 0 (as k_cg_flush).
constexpr int kBiDefer = 4;
template <typename SpP>
struct BiDirs {
  SpP p[kBiDefer], s[kBiDefer];
  int slot[kBiDefer];
  int n;
};
template <typename SpP, typename SpY, typename StY>
__global__ void __launch_bounds__(kBlock)
k_bi_flush(Geom g, BiDirs<SpP> d, const double* __restrict__ ahist, int astride, SpY yr, StY yw,
           double2* __restrict__ y64, const int* __restrict__ skip, int skip_stride) {
  const int rhs = blockIdx.y;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= g.Vh) return;
  const bool fold = y64 && !(skip && skip[rhs * skip_stride]);
  double yv[24];
  ld_site(yr.rhs(rhs), hs, yv);
#pragma unroll
  for (int j = 0; j < kBiDefer; ++j) {
    if (j < d.n) {
      const double* c = ahist + ((size_t)rhs * astride + d.slot[j]) * 4;
      const double ar = c[0], ai = c[1], wr = c[2], wi = c[3];
      double v[24];
      if (ar != 0.0 || ai != 0.0) {
        ld_site(d.p[j].rhs(rhs), hs, v);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          yv[2 * k] += ar * v[2 * k] - ai * v[2 * k + 1];
          yv[2 * k + 1] += ar * v[2 * k + 1] + ai * v[2 * k];
        }
      }
      if (wr != 0.0 || wi != 0.0) {
        ld_site(d.s[j].rhs(rhs), hs, v);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          yv[2 * k] += wr * v[2 * k] - wi * v[2 * k + 1];
          yv[2 * k + 1] += wr * v[2 * k + 1] + wi * v[2 * k];
        }
      }
    }
  }
  if (fold) {
    double2* y = y64 + (size_t)rhs * 12 * g.stride;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      double2 o = y[(size_t)k * g.stride + hs];
      o.x += yv[2 * k];
      o.y += yv[2 * k + 1];
      y[(size_t)k * g.stride + hs] = o;
    }
#pragma unroll
    for (int k = 0; k < 24; ++k) yv[k] = 0.0;
  }
  st_site(yw.rhs(rhs), hs, yv);
}

// p = r + beta (p - omega v)
template <typename Sp, typename St>
__global__ void __launch_bounds__(kBlock)
k_bi_p(Geom g, Sp r, Sp p, Sp v, St pw, const RUScal* __restrict__ sc,
       const BiCoef* __restrict__ bc) {
  const int rhs = blockIdx.y;
  if (sc[rhs].done) return;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= g.Vh) return;
  const double br = bc[rhs].beta[0], bi = bc[rhs].beta[1];
  const double wr = bc[rhs].omega[0], wi = bc[rhs].omega[1];
  double rv[24], pv[24], vv[24];
  ld_site(r.rhs(rhs), hs, rv);
  ld_site(p.rhs(rhs), hs, pv);
  ld_site(v.rhs(rhs), hs, vv);
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    const double qx = pv[2 * k] - (wr * vv[2 * k] - wi * vv[2 * k + 1]);
    const double qy = pv[2 * k + 1] - (wr * vv[2 * k + 1] + wi * vv[2 * k]);
    rv[2 * k] += br * qx - bi * qy;
    rv[2 * k + 1] += br * qy + bi * qx;
  }
  st_site(pw.rhs(rhs), hs, rv);
}
// <r0, r> (re, im) and |r|^2 partials, masked by skip
template <typename Sp>
__global__ void __launch_bounds__(kBlock)
k_bi_rdots(Geom g, Sp r0, Sp r, const int* __restrict__ skip, int skip_stride,
           double* __restrict__ partials) {
  const int rhs = blockIdx.y;
  if (skip && skip[rhs * skip_stride]) return;
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;
  if (hs < g.Vh) {
    double a[24], b[24];
    ld_site(r0.rhs(rhs), hs, a);
    ld_site(r.rhs(rhs), hs, b);
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const double x = a[2 * k], y = a[2 * k + 1], u = b[2 * k], w = b[2 * k + 1];
      d0 += x * u + y * w;
      d1 += x * w - y * u;
      d2 += u * u + w * w;
    }
  }
  const size_t pl = (size_t)gridDim.y * gridDim.x, o = (size_t)rhs * gridDim.x + blockIdx.x;
  d0 = block_sum(d0);
  if (threadIdx.x == 0) partials[o] = d0;
  d1 = block_sum(d1);
  if (threadIdx.x == 0) partials[pl + o] = d1;
  d2 = block_sum(d2);
  if (threadIdx.x == 0) partials[2 * pl + o] = d2;
}

template <int P>
struct BiBlas {
  using A = Acc<P>;
  static ii_status s(ii_context* ctx, const SF& r, const SF& v, const SF& so, const RUScal* sc,
                     const BiCoef* bc, int nrhs) {
    const Geom& g = ctx->g;
    hipLaunchKernelGGL((k_bi_s<typename A::Sp, typename A::St>), dim3(blocks_for(g.Vh), nrhs),
                       dim3(kBlock), 0, ctx->stream, g, A::sp(r, g.stride), A::sp(v, g.stride),
                       A::st(so, g.stride), sc, bc);
    II_HIP(ctx, hipGetLastError());
    return II_OK;
  }
  static ii_status r(ii_context* ctx, const SF& s, const SF& t, const SF& r0, const SF& rr,
                     const RUScal* sc, const BiCoef* bc, int nrhs, double* partials) {
    const Geom& g = ctx->g;
    hipLaunchKernelGGL((k_bi_r<typename A::Sp, typename A::St>), dim3(blocks_for(g.Vh), nrhs),
                       dim3(kBlock), 0, ctx->stream, g, A::sp(s, g.stride), A::sp(t, g.stride),
                       A::sp(r0, g.stride), A::st(rr, g.stride), sc, bc, partials);
    II_HIP(ctx, hipGetLastError());
    return II_OK;
  }
  static ii_status flush(ii_context* ctx, const SF* rp, const SF* rs, const int* slots, int n,
                         const double* ahist, int astride, const SF& y, int nrhs, double2* y64,
                         const int* skip, int ss) {
    const Geom& g = ctx->g;
    using Y = Acc<II_PREC_SINGLE>;
    BiDirs<typename A::Sp> d{};
    d.n = n;
    for (int j = 0; j < kBiDefer; ++j) {
      const int sl = slots[j < n ? j : 0];
      d.p[j] = A::sp(rp[sl], g.stride);
      d.s[j] = A::sp(rs[sl], g.stride);
      d.slot[j] = sl;
    }
    hipLaunchKernelGGL((k_bi_flush<typename A::Sp, typename Y::Sp, typename Y::St>),
                       dim3(blocks_for(g.Vh), nrhs), dim3(kBlock), 0, ctx->stream, g, d, ahist,
                       astride, Y::sp(y, g.stride), Y::st(y, g.stride), y64, skip, ss);
    II_HIP(ctx, hipGetLastError());
    return II_OK;
  }
  static ii_status p(ii_context* ctx, const SF& r, const SF& pin, const SF& v, const SF& pout,
                     const RUScal* sc, const BiCoef* bc, int nrhs) {
    const Geom& g = ctx->g;
    hipLaunchKernelGGL((k_bi_p<typename A::Sp, typename A::St>), dim3(blocks_for(g.Vh), nrhs),
                       dim3(kBlock), 0, ctx->stream, g, A::sp(r, g.stride), A::sp(pin, g.stride),
                       A::sp(v, g.stride), A::st(pout, g.stride), sc, bc);
    II_HIP(ctx, hipGetLastError());
    return II_OK;
  }
  static ii_status rdots(ii_context* ctx, const SF& r0, const SF& r, const int* skip, int ss,
                         int nrhs, double* partials) {
    const Geom& g = ctx->g;
    hipLaunchKernelGGL((k_bi_rdots<typename A::Sp>), dim3(blocks_for(g.Vh), nrhs), dim3(kBlock),
                       0, ctx->stream, g, A::sp(r0, g.stride), A::sp(r, g.stride), skip, ss,
                       partials);
    II_HIP(ctx, hipGetLastError());
    return II_OK;
  }
};

template <int P>
ii_status propagator_bicgstab_t(ii_context* ctx, double kappa, const ii_prop_params* prm,
                                const ii_field* src, ii_field* sol, ii_solver_output* outs) {
  const Geom& g = ctx->g;
  using BB = BiBlas<P>;
  const int nrhs_all = src->nrhs;
  if (!ctx->gauge_lp[P]) II_CHECK(build_lowprec_gauge(ctx, P));
  const int batch = prm->batched ? nrhs_all : 1;
  const double tol = prm->tol > 0 ? prm->tol : 1e-10;
  // BiCGStab drops ~3 decades in 10 iterations: a reliable update every
  // decade (0.1, the CG default) costs more than the FP16 drift it removes;
  // FP16 at 32^3x64: 0.1 / 0.03 / 0.01 / 0.001 gave 8 / 6 / 4 / 4 updates in
  // 28 / 28 / 30 / 36 iterations, 0.216 / 0.204 / 0.197 / 0.218 s.  BF16
  // (8-bit mantissa) drifts faster: 0.01 took 38-41 iterations, 0.1 27-28.
  const double delta = prm->delta > 0 ? prm->delta : (P == II_PREC_BF16 ? 0.1 : 0.01);
  const int maxiter = prm->maxiter > 0 ? prm->maxiter : 100000;
  const size_t he = half_elems(g);
  const int nb = hop_blocks(g);                   // BLAS partials per rhs
  const int nh = hop_dot_parts(g, batch, P);      // hop partials per rhs
  const int nhd = hop_dot_parts(g, batch, II_PREC_DOUBLE);
  II_CHECK(ensure_partials(ctx, (size_t)3 * std::max(std::max(nb, nh), nhd) * batch + 16));
  II_CHECK(ensure_ru_scalars(ctx, batch));
  double* part = ctx->partials;
  RUScal* sc = dev_scal(ctx);
  const int ss = kScalInts;
  const int* skip_it = &sc[0].done;
  const int* skip_ru = &sc[0].skip_ru;
  void* v = nullptr;
  II_CHECK(workspace(ctx, 40, sizeof(double) * 3 * batch + 64, &v));
  double* aux = (double*)v;
  double* one = aux + 3 * batch;
  II_CHECK(workspace(ctx, 63, sizeof(BiCoef) * batch, &v));
  BiCoef* bc = (BiCoef*)v;
  SF be, bo, bh, y, q64, t64;
  II_CHECK(fieldP(ctx, 41, 0, batch, &be));
  II_CHECK(fieldP(ctx, 42, 0, batch, &bo));
  II_CHECK(fieldP(ctx, 43, 0, batch, &bh));
  II_CHECK(fieldP(ctx, 44, 0, batch, &y));
  II_CHECK(fieldP(ctx, 46, 0, batch, &q64));
  II_CHECK(fieldP(ctx, 47, 0, batch, &t64));
  SF r, r0, vv, w, t, ys;
  II_CHECK(fieldP(ctx, 51, P, batch, &r));
  II_CHECK(fieldP(ctx, 52, P, batch, &r0));
  II_CHECK(fieldP(ctx, 53, P, batch, &vv));
  II_CHECK(fieldP(ctx, 54, P, batch, &w));
  II_CHECK(fieldP(ctx, 73, P, batch, &t));
  II_CHECK(fieldP(ctx, 55, II_PREC_SINGLE, batch, &ys));
  // p and s of the last kBiDefer iterations stay in rings of kBiDefer + 1
  // fields (their alpha, omega in ahist) until k_bi_flush adds them to ys
  constexpr int nring = kBiDefer + 1;
  SF rp[nring], rsr[nring];
  for (int i = 0; i < nring; ++i) {
    II_CHECK(fieldP(ctx, i == 0 ? 50 : 73 + i, P, batch, &rp[i]));
    II_CHECK(fieldP(ctx, i == 0 ? 72 : 77 + i, P, batch, &rsr[i]));
  }
  II_CHECK(workspace(ctx, 62, sizeof(double) * 4 * nring * batch, &v));
  double* ahist = (double*)v;
  HopSpec h;
  for (int rb0 = 0; rb0 < nrhs_all; rb0 += batch) {
    const int nr = std::min(batch, nrhs_all - rb0);
    const size_t plh = (size_t)nr * nh, plb = (size_t)nr * nb;
    II_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int k = 0; k < nr; ++k) {
      const double2* base = src->data + (size_t)(rb0 + k) * 2 * he;
      II_CHECK(copy64(ctx, (double2*)be.v + k * he, base));
      II_CHECK(copy64(ctx, (double2*)bo.v + k * he, base + he));
    }
    II_CHECK(dot64(ctx, (double2*)be.v, (double2*)be.v, aux, nr));
    II_CHECK(dot64(ctx, (double2*)bo.v, (double2*)bo.v, aux + batch, nr));
    // bhat = b_e + kappa H_eo b_o   (eoUtils.cpp:54-57)
    h = HopSpec();
    h.out_par = 0;
    h.a = kappa;
    h.xpay = true;
    h.nrhs = nr;
    II_CHECK(hop64(ctx, h, (double2*)bo.v, (double2*)bh.v, (double2*)be.v, nullptr));
    II_CHECK(zero64(ctx, (double2*)y.v, nr));
    II_CHECK(dot64(ctx, (double2*)bh.v, (double2*)bh.v, aux + 2 * batch, nr));
    {
      std::vector<double> hb(2 * batch);
      II_HIP(ctx, hipMemcpyAsync(hb.data(), aux, 2 * batch * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream));
      II_HIP(ctx, hipStreamSynchronize(ctx->stream));
      for (int k = 0; k < nr; ++k) hb[k] += hb[batch + k];
      II_HIP(ctx, hipMemcpyAsync(aux, hb.data(), nr * sizeof(double), hipMemcpyHostToDevice,
                                 ctx->stream));
      const double unit = 1.0;
      II_HIP(ctx, hipMemcpyAsync(one, &unit, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
      II_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    hipLaunchKernelGGL(k_scal_init, dim3(nr), dim3(64), 0, ctx->stream, sc, aux + 2 * batch, aux,
                       0.0, tol, delta, 1, prm->restart_factor > 0 ? prm->restart_factor : 16.0);
    hipLaunchKernelGGL(k_bi_init, dim3(nr), dim3(64), 0, ctx->stream, sc, bc);
    II_HIP(ctx, hipGetLastError());
    // r = r0 = p = bhat (x0 = 0), ys = 0
    II_CHECK(convertP(ctx, 0, P, bh, r, nr));
    II_CHECK(convertP(ctx, P, P, r, r0, nr));
    II_CHECK(convertP(ctx, P, P, r, rp[0], nr));
    II_CHECK(convertP(ctx, 0, II_PREC_SINGLE, y, ys, nr));  // zeros
    int pos = 0;            // ring slot of this iteration's p and s
    std::vector<int> pend;  // slots whose alpha p + omega s is not yet in ys
    auto flush = [&](double2* y64, const int* skip) -> ii_status {
      if (pend.empty() && !y64) return II_OK;
      II_CHECK(BB::flush(ctx, rp, rsr, pend.empty() ? &pos : pend.data(), (int)pend.size(), ahist,
                         nring, ys, nr, y64, skip, ss));
      pend.clear();
      return II_OK;
    };
    auto iteration = [&](int cur) -> ii_status {
      const SF& p = rp[cur];
      const SF& s = rsr[cur];
      HopSpec a;
      a.nrhs = nr;
      a.skip = skip_it;
      a.skip_stride = ss;
      a.out_par = 1;
      a.a = 1.0;
      II_CHECK(hopP(ctx, P, a, p, w, nullptr, nullptr));
      a.out_par = 0;
      a.a = -kappa * kappa;
      a.xpay = true;
      a.epi = kEpiCDot;
      a.partials = part;
      II_CHECK(hopP(ctx, P, a, w, vv, &p, &r0));  // v = Dhat p, <r0, v>
      hipLaunchKernelGGL(k_bi_alpha, dim3(nr), dim3(kFinThreads), 0, ctx->stream, part, nh, plh, sc,
                         bc, ahist, cur, nring);
      II_CHECK(BB::s(ctx, r, vv, s, sc, bc, nr));
      a.out_par = 1;
      a.a = 1.0;
      a.xpay = false;
      a.epi = kEpiStore;
      II_CHECK(hopP(ctx, P, a, s, w, nullptr, nullptr));
      a.out_par = 0;
      a.a = -kappa * kappa;
      a.xpay = true;
      a.epi = kEpiTS;
      II_CHECK(hopP(ctx, P, a, w, t, &s, nullptr));  // t = Dhat s, <t, s>, |t|^2
      hipLaunchKernelGGL(k_bi_omega, dim3(nr), dim3(kFinThreads), 0, ctx->stream, part, nh, plh, sc,
                         bc, ahist, cur, nring);
      II_CHECK(BB::r(ctx, s, t, r0, r, sc, bc, nr, part));
      hipLaunchKernelGGL(k_bi_beta, dim3(nr), dim3(kFinThreads), 0, ctx->stream, part, nb, plb, sc,
                         bc);
      return BB::p(ctx, r, p, vv, rp[(cur + 1) % nring], sc, bc, nr);
    };
    int its = 0, chunk = 1;
    bool all_done = false;
    while (its < maxiter && !all_done) {
      for (int i = 0; i < chunk; ++i) {
        II_CHECK(iteration(pos));
        pend.push_back(pos);
        pos = (pos + 1) % nring;
        if ((int)pend.size() == kBiDefer) II_CHECK(flush(nullptr, nullptr));
      }
      its += chunk;
      II_CHECK(read_scalars(ctx, nr));
      RUScal* H = host_scal(ctx);
      bool any_upd = false;
      all_done = true;
      for (int k = 0; k < nr; ++k) {
        if (H[k].nan)
          return set_error(ctx, II_ERR_NAN,
                           "non-finite BiCGStab scalar (rhs " + std::to_string(rb0 + k) +
                               ", iteration " + std::to_string(H[k].count) + ")");
        if (!H[k].done) all_done = false;
        if (!H[k].done && !H[k].skip_ru) any_upd = true;
      }
      if (all_done) break;
      if (!any_upd) {
        chunk = std::min(chunk * 2, 8);
        continue;
      }
      chunk = 1;
      II_CHECK(join_updates(ctx, nr));
      // reliable update for the flagged rhs: y64 += ys + pending, ys = 0
      II_CHECK(flush((double2*)y.v, skip_ru));
      // rho_true = bhat - Dhat y (FP64) and |rho_true|^2
      h = HopSpec();
      h.nrhs = nr;
      h.skip = skip_ru;
      h.skip_stride = ss;
      h.out_par = 1;
      h.a = 1.0;
      II_CHECK(hop64(ctx, h, (double2*)y.v, (double2*)t64.v, nullptr, nullptr));
      h.out_par = 0;
      h.a = -kappa * kappa;
      h.xpay = true;
      h.epi = kEpiRUpd;
      h.coef = one;
      h.coef_stride = 0;
      h.partials = part;
      II_CHECK(hop64(ctx, h, (double2*)t64.v, (double2*)q64.v, (double2*)y.v,
                     (const double2*)bh.v));
      hipLaunchKernelGGL(k_fin_rho, dim3(nr), dim3(kFinThreads), 0, ctx->stream, part, nhd, sc);
      // r = rho_true in the sloppy precision; rho = <r0, r>; restart if broken
      II_CHECK(convertP(ctx, 0, P, q64, r, nr, skip_ru, ss));
      II_CHECK(BB::rdots(ctx, r0, r, skip_ru, ss, nr, part));
      hipLaunchKernelGGL(k_bi_ru, dim3(nr), dim3(kFinThreads), 0, ctx->stream, part, nb, plb, sc, bc);
      II_HIP(ctx, hipGetLastError());
      II_CHECK(convertP(ctx, P, P, r, r0, nr, &sc[0].keep_p, ss));
      II_CHECK(convertP(ctx, P, P, r, rp[pos], nr, &sc[0].keep_p, ss));
    }
    II_CHECK(flush((double2*)y.v, nullptr));
    // x_o = b_o + kappa H_oe y   (eoUtils.cpp:67)
    h = HopSpec();
    h.out_par = 1;
    h.a = kappa;
    h.xpay = true;
    h.nrhs = nr;
    II_CHECK(hop64(ctx, h, (double2*)y.v, (double2*)t64.v, (double2*)bo.v, nullptr));
    for (int k = 0; k < nr; ++k) {
      double2* base = sol->data + (size_t)(rb0 + k) * 2 * he;
      II_CHECK(copy64(ctx, base, (double2*)y.v + k * he));
      II_CHECK(copy64(ctx, base + he, (double2*)t64.v + k * he));
    }
    II_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    II_CHECK(read_scalars(ctx, nr));
    float ms = 0.f;
    II_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    RUScal* H = host_scal(ctx);
    bool failed = false;
    for (int k = 0; k < nr; ++k) {
      ii_solver_output& o = outs[rb0 + k];
      o.iterations = H[k].count;
      o.inner_iterations = H[k].count;
      o.restarts = H[k].updates;
      o.recursion_restarts = H[k].restarts;
      o.diff = std::sqrt(H[k].rho2 / H[k].b2);
      o.timing_ms = ms;
      o.true_residual = -1.0;
      if (!H[k].done) failed = true;
    }
    if (failed)
      return set_error(ctx, II_ERR_NOT_CONVERGED, "BiCGStab solve did not reach the tolerance");
  }
  return II_OK;
}

ii_status propagator_bicgstab(ii_context* ctx, double kappa, const ii_prop_params* prm,
                              const ii_field* src, ii_field* sol, ii_solver_output* outs) {
  switch (prm->precision) {
    case II_PREC_SINGLE: return propagator_bicgstab_t<1>(ctx, kappa, prm, src, sol, outs);
    case II_PREC_HALF: return propagator_bicgstab_t<2>(ctx, kappa, prm, src, sol, outs);
    case II_PREC_BF16: return propagator_bicgstab_t<3>(ctx, kappa, prm, src, sol, outs);
    default:
      return set_error(ctx, II_ERR_BAD_ARG,
                       "BiCGStab runs in a sloppy precision (single, half, bf16) with FP64 "
                       "reliable updates");
  }
}

// true residual |b - D x|/|b| per rhs of full fields (FP64)
ii_status full_residual(ii_context* ctx, double kappa, const ii_field* src, const ii_field* sol,
                        double* res) {
  const Geom& g = ctx->g;
  const size_t he = half_elems(g);
  const int n = src->nrhs;
  double2* w;
  double* d;  // |b - D x|^2 per (parity, rhs) [0, 2n), |b|^2 per (rhs, parity) [2n, 4n), 1.0
  {
    void* p;
    II_CHECK(workspace(ctx, 56, he * n * sizeof(double2), &p));
    w = (double2*)p;
    II_CHECK(workspace(ctx, 57, (4 * (size_t)n + 1) * sizeof(double), &p));
    d = (double*)p;
  }
  const double unit = 1.0;
  II_HIP(ctx, hipMemcpyAsync(d + 4 * n, &unit, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const int nblk = hop_dot_parts(g, n, II_PREC_DOUBLE);
  II_CHECK(ensure_partials(ctx, (size_t)nblk * n + 16));
  for (int par = 0; par < 2; ++par) {
    // w = b_par - (x_par - kappa H x_other) = (b - D x)_par for all rhs at
    // once (the full fields' parity par: rhs stride 24 * stride), |w|^2 in
    // the hop's epilogue
    HopSpec h;
    h.out_par = par;
    h.a = -kappa;
    h.xpay = true;
    h.epi = kEpiRUpd;
    h.coef = d + 4 * n;
    h.partials = ctx->partials;
    h.nrhs = n;
    II_CHECK(hop64(ctx, h, sol->data + (1 - par) * he, w, sol->data + par * he,
                   src->data + par * he, (int)(2 * he)));
    II_CHECK(finalize_sums(ctx, ctx->partials, nblk, d + par * n, n));
  }
  // |b|^2 of every half field of the sources in one call
  II_CHECK(dot64(ctx, src->data, src->data, d + 2 * n, 2 * n));
  std::vector<double> hh(4 * (size_t)n);
  II_HIP(ctx, hipMemcpyAsync(hh.data(), d, 4 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < n; ++k)
    res[k] = std::sqrt((hh[k] + hh[n + k]) / (hh[2 * n + 2 * k] + hh[2 * n + 2 * k + 1]));
  return II_OK;
}

// plain streaming copy (bandwidth reference), grid-stride
__global__ void __launch_bounds__(256)
k_copy_stream(const double4* __restrict__ a, double4* __restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// Mixed-precision multishift on device fields (x: nshift even half fields,
// b: one): stage 1 runs the reference's shifted recursion with FP32 vectors
// and operator, all shifts fused per launch, to |r|/|b| < max(tol, 1e-6);
// stage 2 refines every shift at once with a batched reliable-update CG.
ii_status multishift_mixed_dev(ii_context* ctx, double kappa, double2* x, const double2* b,
                               const double* shifts, int nshift, double tol, int precision,
                               int maxiter, ii_solver_output* out) {
  if (nshift < 1 || nshift > kMaxShift || !(tol > 0) || precision < II_PREC_DOUBLE ||
      precision > II_PREC_BF16)
    return set_error(ctx, II_ERR_BAD_ARG, "multishift_mixed: bad arguments");
  auto t0 = std::chrono::steady_clock::now();
  SF x32;
  II_CHECK(fieldP(ctx, 42, II_PREC_SINGLE, nshift, &x32));
  const double t1 = std::max(tol, 1e-6);
  int its1 = 0;
  II_CHECK(multishift_sloppy(ctx, kappa, (float2*)x32.v, b, shifts, nshift, t1 * t1, maxiter,
                             &its1));
  II_CHECK(convert_32_to_64(ctx, (const float2*)x32.v, x, nshift));
  ii_status st = shifted_refine(ctx, kappa, shifts, nshift, b, x, tol, precision, maxiter, out);
  if (st != II_OK && st != II_ERR_NOT_CONVERGED) return st;
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (int i = 0; i < nshift; ++i) {
    out[i].inner_iterations = its1;
    out[i].iterations += its1;
    out[i].timing_ms = ms;
  }
  return st;
}

}  // namespace ii

using namespace ii;

extern "C" {

ii_status ii_solver_cg(ii_context* ctx, double kappa, double* x, const double* b, double prec,
                       int maxiter, ii_solver_output* out) {
  if (!x || !b || maxiter < 1) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  ii_field *fb = nullptr, *fx = nullptr;
  II_CHECK(ii_field_create(ctx, II_EVEN, 1, &fb));
  struct Del { ii_field* f; ~Del() { if (f) ii_field_destroy(f); } } d1{fb};
  II_CHECK(ii_field_create(ctx, II_EVEN, 1, &fx));
  Del d2{fx};
  II_CHECK(ii_field_upload(fb, 0, b));
  ii_status st = cg_reference(ctx, kappa, fx->data, fb->data, prec, maxiter, out);
  if (st != II_OK && st != II_ERR_NOT_CONVERGED) return st;
  std::string msg = ctx->err;
  II_CHECK(ii_field_download(fx, 0, x));
  ctx->err = msg;
  (void)g;
  return st;
}

ii_status ii_multishift_cg(ii_context* ctx, double kappa, double* x, const double* b,
                           const double* shifts, int nshift, double prec, int maxiter,
                           ii_solver_output* out) {
  if (!x || !b || !shifts || nshift < 1 || maxiter < 1) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  ii_field *fb = nullptr, *fx = nullptr;
  II_CHECK(ii_field_create(ctx, II_EVEN, 1, &fb));
  struct Del { ii_field* f; ~Del() { if (f) ii_field_destroy(f); } } d1{fb};
  II_CHECK(ii_field_create(ctx, II_EVEN, nshift, &fx));
  Del d2{fx};
  II_CHECK(ii_field_upload(fb, 0, b));
  ii_status st = multishift_reference(ctx, kappa, fx->data, fb->data, shifts, nshift, prec,
                                      maxiter, out, nullptr, false);
  if (st != II_OK && st != II_ERR_NOT_CONVERGED) return st;
  std::string msg = ctx->err;
  for (int i = 0; i < nshift; ++i) II_CHECK(ii_field_download(fx, i, x + (size_t)24 * g.Vh * i));
  ctx->err = msg;
  return st;
}

ii_status ii_multishift_mixed(ii_context* ctx, double kappa, double* x, const double* b,
                              const double* shifts, int nshift, double tol, int precision,
                              int maxiter, ii_solver_output* out) {
  if (!x || !b || !shifts || !out || nshift < 1 || nshift > kMaxShift || maxiter < 1 ||
      !(tol > 0) || precision < II_PREC_DOUBLE || precision > II_PREC_BF16)
    return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  ii_field *fb = nullptr, *fx = nullptr;
  II_CHECK(ii_field_create(ctx, II_EVEN, 1, &fb));
  struct Del { ii_field* f; ~Del() { if (f) ii_field_destroy(f); } } d1{fb};
  II_CHECK(ii_field_create(ctx, II_EVEN, nshift, &fx));
  Del d2{fx};
  II_CHECK(ii_field_upload(fb, 0, b));
  ii_status st = multishift_mixed_dev(ctx, kappa, fx->data, fb->data, shifts, nshift, tol,
                                      precision, maxiter, out);
  if (st != II_OK && st != II_ERR_NOT_CONVERGED) return st;
  std::string msg = ctx->err;
  for (int i = 0; i < nshift; ++i) II_CHECK(ii_field_download(fx, i, x + (size_t)24 * g.Vh * i));
  ctx->err = msg;
  return st;
}

ii_status ii_multishift_cg_full(ii_context* ctx, double kappa, double* x, const double* b,
                                const double* shifts, int nshift, double prec, int maxiter,
                                ii_solver_output* out) {
  if (!x || !b || !shifts || nshift < 1 || maxiter < 1) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  ii_field *fb = nullptr, *fx = nullptr;
  II_CHECK(ii_field_create(ctx, II_FULL, 1, &fb));
  struct Del { ii_field* f; ~Del() { if (f) ii_field_destroy(f); } } d1{fb};
  II_CHECK(ii_field_create(ctx, II_FULL, nshift, &fx));
  Del d2{fx};
  II_CHECK(ii_field_upload(fb, 0, b));
  ii_status st = multishift_reference(ctx, kappa, fx->data, fb->data, shifts, nshift, prec,
                                      maxiter, out, nullptr, true);
  if (st != II_OK && st != II_ERR_NOT_CONVERGED) return st;
  std::string msg = ctx->err;
  for (int i = 0; i < nshift; ++i) II_CHECK(ii_field_download(fx, i, x + (size_t)24 * g.V * i));
  ctx->err = msg;
  return st;
}

ii_status ii_qprop_multishift(ii_context* ctx, double kappa, const int site[4],
                              const double* shifts, int nshift, double prec, int maxiter,
                              ii_field* xq, ii_solver_output* out) {
  if (!site || !shifts || !xq || !out || nshift < 1 || maxiter < 1) return II_ERR_BAD_ARG;
  if (xq->ctx != ctx || xq->subset != II_FULL || xq->nrhs != 12 * nshift)
    return set_error(ctx, II_ERR_BAD_ARG, "xq must be a full field of 12*nshift rhs");
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const size_t he = half_elems(ctx->g);
  ii_field* src = nullptr;
  II_CHECK(ii_field_create(ctx, II_FULL, 12, &src));
  struct Del { ii_field* f; ~Del() { if (f) ii_field_destroy(f); } } d1{src};
  II_CHECK(ii_field_point_sources(src, site));
  ii_status worst = II_OK;
  std::string msg;
  for (int col = 0; col < 12; ++col) {
    double2* x = xq->data + (size_t)col * nshift * 2 * he;
    ii_status st = multishift_reference(ctx, kappa, x, src->data + (size_t)col * 2 * he, shifts,
                                        nshift, prec, maxiter, &out[col], nullptr, true);
    if (st == II_ERR_NOT_CONVERGED) {
      worst = st;
      msg = ctx->err;
    } else if (st != II_OK) {
      return st;
    }
    // xq[(3s+c)*nshift + m] = D x_m  (qprop_MultiShift.cpp:25-28)
    double2* w = nullptr;
    {
      void* p;
      II_CHECK(workspace(ctx, 21, 2 * he * sizeof(double2), &p));
      w = (double2*)p;
    }
    for (int m = 0; m < nshift; ++m) {
      double2* xm = x + (size_t)m * 2 * he;
      HopSpec h;
      h.a = -kappa;
      h.xpay = true;
      for (int par = 0; par < 2; ++par) {
        h.out_par = par;
        II_CHECK(hop64(ctx, h, xm + (1 - par) * he, w + par * he, xm + par * he, nullptr));
      }
      II_CHECK(copy64(ctx, xm, w, 2));
    }
  }
  if (worst != II_OK) return set_error(ctx, worst, msg);
  return II_OK;
}

ii_status ii_eo_invert(ii_context* ctx, double kappa, double* sol, const double* src, double prec,
                       int maxiter, ii_solver_output* out) {
  if (!sol || !src || maxiter < 1) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const Geom& g = ctx->g;
  const size_t he = half_elems(g);
  ii_field *fs = nullptr, *fx = nullptr, *fb = nullptr, *fy = nullptr;
  II_CHECK(ii_field_create(ctx, II_FULL, 1, &fs));
  struct Del { ii_field* f; ~Del() { if (f) ii_field_destroy(f); } } d1{fs};
  II_CHECK(ii_field_create(ctx, II_FULL, 1, &fx));
  Del d2{fx};
  II_CHECK(ii_field_create(ctx, II_EVEN, 1, &fb));
  Del d3{fb};
  II_CHECK(ii_field_create(ctx, II_EVEN, 1, &fy));
  Del d4{fy};
  II_CHECK(ii_field_upload(fs, 0, src));
  double2* se = fs->data;
  double2* so = fs->data + he;
  // be = b_e - mult_eo(b_o) = b_e + kappa H_eo b_o     (eoUtils.cpp:54-57)
  HopSpec h;
  h.out_par = 0;
  h.a = kappa;
  h.xpay = true;
  II_CHECK(hop64(ctx, h, so, fy->data, se, nullptr));
  // rhs = mult_dag(be)
  double2* t = nullptr;
  {
    void* p;
    II_CHECK(workspace(ctx, 20, he * sizeof(double2), &p));
    t = (double2*)p;
  }
  II_CHECK(dhat_dag64(ctx, kappa, fy->data, fb->data, t));
  ii_status st = cg_reference(ctx, kappa, fx->data, fb->data, prec, maxiter, out);
  if (st != II_OK && st != II_ERR_NOT_CONVERGED) return st;
  std::string msg = ctx->err;
  // x_o = b_o - mult_oe(y_e) = b_o + kappa H_oe y_e    (eoUtils.cpp:67)
  h.out_par = 1;
  h.a = kappa;
  h.xpay = true;
  II_CHECK(hop64(ctx, h, fx->data, fx->data + he, so, nullptr));
  II_CHECK(ii_field_download(fx, 0, sol));
  ctx->err = msg;
  return st;
}

ii_status ii_propagator_solve(ii_context* ctx, double kappa, const ii_prop_params* p,
                              const ii_field* src, ii_field* sol, ii_solver_output* out) {
  if (!p || !src || !sol || !out) return II_ERR_BAD_ARG;
  if (src->ctx != ctx || sol->ctx != ctx || src->subset != II_FULL || sol->subset != II_FULL ||
      src->nrhs != sol->nrhs)
    return set_error(ctx, II_ERR_BAD_ARG, "propagator fields must be full and of equal nrhs");
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  // the multi-rhs hops' launch shapes for this batch, before any iteration
  // is captured in a graph (a capture keeps the shape it records)
  if (p->batched && src->nrhs > 1) {
    const int sp = p->method == II_METHOD_FGCR ? II_PREC_SINGLE : p->precision;
    if (ctx->mrhs_nrhs[sp] != src->nrhs) II_CHECK(tune_mrhs(ctx, sp, src->nrhs));
    if (ctx->mrhs_nrhs[II_PREC_DOUBLE] != src->nrhs)
      II_CHECK(tune_mrhs(ctx, II_PREC_DOUBLE, src->nrhs));
  }
  ii_status st;
  if (p->method == II_METHOD_FGCR) {
    if (!p->batched)
      return set_error(ctx, II_ERR_BAD_ARG, "the flexible GCR method runs batched only");
    st = fgcr_batched(ctx, &ctx->fgcr, kappa, p, src, sol, out);
  } else if (p->use_learned) {
    st = learned_propagator(ctx, kappa, p, src, sol, out);
  } else if (p->method == II_METHOD_BICGSTAB) {
    st = propagator_bicgstab(ctx, kappa, p, src, sol, out);
  } else if (p->method == II_METHOD_CG) {
    st = propagator(ctx, kappa, p, src, sol, out);
  } else {
    return set_error(ctx, II_ERR_BAD_ARG, "unknown propagator method");
  }
  if (st != II_OK && st != II_ERR_NOT_CONVERGED) return st;
  std::string msg = ctx->err;
  std::vector<double> res(src->nrhs);
  II_CHECK(full_residual(ctx, kappa, src, sol, res.data()));
  for (int k = 0; k < src->nrhs; ++k) out[k].true_residual = res[k];
  ctx->err = msg;
  return st;
}

ii_status ii_autotune(ii_context* ctx, char* report, size_t len) {
  if (!ctx) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  ctx->hop_tuned = false;
  II_CHECK(tune_hop(ctx));
  std::string rep = ctx->tune_report;
  for (int P = II_PREC_DOUBLE; P <= II_PREC_BF16; ++P) {
    ctx->mrhs_nrhs[P] = 0;
    II_CHECK(tune_mrhs(ctx, P, 12));
    rep += "; " + ctx->mrhs_report[P];
  }
  if (report && len) std::snprintf(report, len, "%s", rep.c_str());
  return II_OK;
}

// Tuning file: one line per choice, for one lattice.
//   iroiro-hip tuning 1
//   lattice Lx Ly Lz Lt
//   hop mode W bpp bs                      (the single-rhs FP64 hop's order)
//   mrhs P nrhs rb mode W bpp bs           (each tuned multi-rhs precision P)
//   net nrhs rg0..rg3 by0..by3 bz0..bz3    (the network stages)
ii_status ii_tune_save(ii_context* ctx, const char* path) {
  if (!ctx || !path) return II_ERR_BAD_ARG;
  const Geom& g = ctx->g;
  std::FILE* f = std::fopen(path, "w");
  if (!f) return set_error(ctx, II_ERR_BAD_ARG, std::string("cannot create tuning file ") + path);
  std::fprintf(f, "iroiro-hip tuning 1\nlattice %d %d %d %d\n", g.L[0], g.L[1], g.L[2], g.L[3]);
  if (ctx->hop_tuned) {
    const HopOrder& o = ctx->hop_order;
    std::fprintf(f, "hop %d %d %d %d\n", o.mode, o.W, o.bpp, o.bs);
  }
  for (int P = 0; P < 4; ++P)
    if (ctx->mrhs_nrhs[P] > 0) {
      const HopOrder& o = ctx->mrhs_ord[P];
      std::fprintf(f, "mrhs %d %d %d %d %d %d %d\n", P, ctx->mrhs_nrhs[P], ctx->mrhs_rb[P], o.mode,
                   o.W, o.bpp, o.bs);
    }
  int nn = 0, rg[4], by[4], bz[4];
  bnet_shapes_get(ctx, &nn, rg, by, bz);
  if (nn > 0)
    std::fprintf(f, "net %d %d %d %d %d %d %d %d %d %d %d %d %d\n", nn, rg[0], rg[1], rg[2],
                 rg[3], by[0], by[1], by[2], by[3], bz[0], bz[1], bz[2], bz[3]);
  const bool bad = std::ferror(f) != 0;
  if (std::fclose(f) != 0 || bad)
    return set_error(ctx, II_ERR_BAD_ARG, std::string("cannot write tuning file ") + path);
  return II_OK;
}

ii_status ii_tune_load(ii_context* ctx, const char* path) {
  if (!ctx || !path) return II_ERR_BAD_ARG;
  std::FILE* f = std::fopen(path, "r");
  if (!f) return set_error(ctx, II_ERR_BAD_ARG, std::string("cannot open tuning file ") + path);
  const Geom& g = ctx->g;
  char line[256];
  int ver = 0, L[4] = {0, 0, 0, 0};
  bool ok = std::fgets(line, sizeof line, f) && std::sscanf(line, "iroiro-hip tuning %d", &ver) == 1 &&
            ver == 1 && std::fgets(line, sizeof line, f) &&
            std::sscanf(line, "lattice %d %d %d %d", &L[0], &L[1], &L[2], &L[3]) == 4;
  if (!ok) {
    std::fclose(f);
    return set_error(ctx, II_ERR_BAD_ARG, std::string("not a tuning file: ") + path);
  }
  if (L[0] != g.L[0] || L[1] != g.L[1] || L[2] != g.L[2] || L[3] != g.L[3]) {
    std::fclose(f);
    return set_error(ctx, II_ERR_BAD_ARG, "tuning file is for another lattice");
  }
  ii_status st = II_OK;
  while (st == II_OK && std::fgets(line, sizeof line, f)) {
    HopOrder o;
    int P = 0, n = 0, rb = 0;
    int v[13];
    if (std::sscanf(line, "hop %d %d %d %d", &o.mode, &o.W, &o.bpp, &o.bs) == 4) {
      ctx->hop_order = o;
      ctx->hop_tuned = true;
      ctx->tune_report = "hop order loaded from a tuning file";
    } else if (std::sscanf(line, "mrhs %d %d %d %d %d %d %d", &P, &n, &rb, &o.mode, &o.W, &o.bpp,
                           &o.bs) == 7) {
      if (P < 0 || P > 3 || n < 1 || rb < 1) {
        st = set_error(ctx, II_ERR_BAD_ARG, "bad mrhs line in tuning file");
        break;
      }
      ctx->mrhs_nrhs[P] = n;
      ctx->mrhs_rb[P] = rb;
      ctx->mrhs_ord[P] = o;
      ctx->mrhs_report[P] = "multi-rhs hop shape loaded from a tuning file";
    } else if (std::sscanf(line, "net %d %d %d %d %d %d %d %d %d %d %d %d %d", &v[0], &v[1], &v[2],
                           &v[3], &v[4], &v[5], &v[6], &v[7], &v[8], &v[9], &v[10], &v[11],
                           &v[12]) == 13) {
      st = bnet_shapes_set(ctx, v[0], v + 1, v + 5, v + 9);
    } else {
      st = set_error(ctx, II_ERR_BAD_ARG, std::string("unknown line in tuning file: ") + line);
    }
  }
  std::fclose(f);
  return st;
}

ii_status ii_bench_copy(ii_context* ctx, size_t bytes, int nrep, double* avg_ms) {
  if (!ctx || !avg_ms || nrep < 1 || bytes < 16) return II_ERR_BAD_ARG;
  II_CHECK(enter(ctx));
  void *a = nullptr, *b = nullptr;
  II_CHECK(workspace(ctx, 60, bytes, &a));
  II_CHECK(workspace(ctx, 61, bytes, &b));
  size_t n = bytes / sizeof(double4);
  const int block = 256;
  const int grid = (int)std::min<size_t>((n + block - 1) / block, 1u << 20);
  hipLaunchKernelGGL(k_copy_stream, dim3(grid), dim3(block), 0, ctx->stream,
                     (const double4*)a, (double4*)b, n);
  II_HIP(ctx, hipGetLastError());
  II_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int i = 0; i < nrep; ++i)
    hipLaunchKernelGGL(k_copy_stream, dim3(grid), dim3(block), 0, ctx->stream,
                       (const double4*)a, (double4*)b, n);
  II_HIP(ctx, hipGetLastError());
  II_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  II_HIP(ctx, hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  II_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  *avg_ms = ms / nrep;
  return II_OK;
}

ii_status ii_dirac_mult_eo_prec(ii_context* ctx, double kappa, int precision, int out_par,
                                int dag, int nrhs, double* out, const double* in) {
  if (!out || !in || nrhs < 1 || out_par < 0 || out_par > 1) return II_ERR_BAD_ARG;
  const int P = precision;
  if (P < 0 || P > 3) return set_error(ctx, II_ERR_BAD_ARG, "bad precision");
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  if (P != II_PREC_DOUBLE && !ctx->gauge_lp[P]) II_CHECK(build_lowprec_gauge(ctx, P));
  const Geom& g = ctx->g;
  const size_t n = (size_t)24 * g.Vh, he = half_elems(g);
  double *din = nullptr, *dout = nullptr;
  II_CHECK(stage_in(ctx, in, n * nrhs, &din));
  II_CHECK(stage_buf(ctx, n * nrhs, &dout));
  SF a64, b64, a, b;
  II_CHECK(fieldP(ctx, 56, II_PREC_DOUBLE, nrhs, &a64));
  II_CHECK(fieldP(ctx, 57, II_PREC_DOUBLE, nrhs, &b64));
  II_CHECK(fieldP(ctx, 58, P, nrhs, &a));
  II_CHECK(fieldP(ctx, 59, P, nrhs, &b));
  for (int k = 0; k < nrhs; ++k)
    II_CHECK(pack_half_spinor(ctx, din + n * k, (double2*)a64.v + he * k));
  II_CHECK(convertP(ctx, II_PREC_DOUBLE, P, a64, a, nrhs));
  HopSpec h;
  h.out_par = out_par;
  h.a = -kappa;
  h.g5in = dag != 0;
  h.g5out = dag != 0;
  h.nrhs = nrhs;
  II_CHECK(hopP(ctx, P, h, a, b, nullptr, nullptr));
  II_CHECK(convertP(ctx, P, II_PREC_DOUBLE, b, b64, nrhs));
  for (int k = 0; k < nrhs; ++k)
    II_CHECK(unpack_half_spinor(ctx, (const double2*)b64.v + he * k, dout + n * k));
  return stage_out(ctx, out, dout, n * nrhs);
}

ii_status ii_bench_hop(ii_context* ctx, int precision, int nrhs, int nrep, double* avg_ms) {
  if (!avg_ms || nrep < 1 || nrhs < 1) return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  const int P = precision;
  if (P != II_PREC_DOUBLE && !ctx->gauge_lp[P]) II_CHECK(build_lowprec_gauge(ctx, P));
  SF a, b;
  II_CHECK(fieldP(ctx, 58, P, nrhs, &a));
  II_CHECK(fieldP(ctx, 59, P, nrhs, &b));
  HopSpec h;
  h.out_par = 1;
  h.a = -0.1355;
  h.nrhs = nrhs;
  for (int i = 0; i < 3; ++i) II_CHECK(hopP(ctx, P, h, a, b, nullptr, nullptr));
  II_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int i = 0; i < nrep; ++i) II_CHECK(hopP(ctx, P, h, a, b, nullptr, nullptr));
  II_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  II_HIP(ctx, hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  II_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  *avg_ms = ms / nrep;
  return II_OK;
}

}  // extern "C"
