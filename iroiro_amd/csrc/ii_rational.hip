// RationalSolver_CG (lib/Solver/rationalSolver_CG.cpp:12-111) and the
// RHMC force of Action_Nf (lib/Action/action_Nf.cpp:84-110) on the device.
//
// solve / solve_inv: the multishift solver on Fopr_DdagD + pole_i, then
//   sol = a0 * src + sum_i res_i * x_i                  (:30-37, :60-67)
// formed by one kernel that reads the npoles shifted solutions once.  The
// multishift is either the reference's FP64 recursion (II_PREC_DOUBLE,
// its squared stopping test `prec`) or the mixed-precision multishift
// (FP32 fused stage, batched refinement to the true residual `prec`).
//
// Action_Nf::md_force for each pseudofermion phi: eta_i = the shifted
// solutions with the inverse poles (solve_noReconstruct), then
//   force += InvRes_i * md_force(eta_i, Dhat eta_i)
// and TracelessAntihermite at the end.  The projection is real-linear and
// idempotent, so it is applied inside every accumulation of the force
// kernel instead of in a separate pass.
#include "ii_ops.h"
#include "ii_stencil.h"

#include <chrono>
#include <string>

namespace ii {

constexpr int kMaxPoles = 64;

struct RatCoef {
  double a0;
  double res[kMaxPoles];
};

// sol = a0 * b + sum_i res_i x_i over the 12 components of each half site
__global__ void __launch_bounds__(kBlock)
k_rat_reconstruct(Geom g, const double2* __restrict__ x, const double2* __restrict__ b,
                  double2* __restrict__ sol, RatCoef c, int npoles) {
  const int hs = blockIdx.x * blockDim.x + threadIdx.x;
  if (hs >= g.Vh) return;
  const size_t he = (size_t)12 * g.stride;
#pragma unroll 4
  for (int k = 0; k < 12; ++k) {
    const size_t o = (size_t)k * g.stride + hs;
    const double2 bv = b[o];
    double2 acc = make_double2(c.a0 * bv.x, c.a0 * bv.y);
    for (int i = 0; i < npoles; ++i) {
      const double2 xv = x[(size_t)i * he + o];
      acc.x = fma(c.res[i], xv.x, acc.x);
      acc.y = fma(c.res[i], xv.y, acc.y);
    }
    sol[o] = acc;
  }
}

// The shifted solutions of one source: FP64 reference recursion or the
// mixed multishift; out receives one summary (largest iteration count and
// true residual over the shifts for the mixed path).
static ii_status shifted_solutions(ii_context* ctx, double kappa, double2* x, const double2* b,
                                   const double* poles, int npoles, double prec, int maxiter,
                                   int precision, ii_solver_output* out) {
  if (precision == II_PREC_DOUBLE)
    return multishift_reference(ctx, kappa, x, b, poles, npoles, prec, maxiter, out, nullptr,
                                false);
  ii_solver_output per[kMaxPoles];
  // the mixed path refines FP64 true residuals: its stage-2 precision is
  // the 16-bit (or FP32) storage of the reliable-update CG
  ii_status st =
      multishift_mixed_dev(ctx, kappa, x, b, poles, npoles, prec, precision, maxiter, per);
  if (out) {
    *out = per[0];
    for (int i = 1; i < npoles; ++i) {
      if (per[i].iterations > out->iterations) out->iterations = per[i].iterations;
      if (per[i].true_residual > out->true_residual) out->true_residual = per[i].true_residual;
      if (per[i].diff > out->diff) out->diff = per[i].diff;
      out->restarts += per[i].restarts;
    }
  }
  return st;
}

static ii_status check_poles(ii_context* ctx, const double* res, const double* poles,
                             int npoles) {
  if (npoles < 1 || npoles > kMaxPoles)
    return set_error(ctx, II_ERR_BAD_ARG, "number of poles must be in 1..64");
  for (int i = 0; i < npoles; ++i)
    if (!(poles[i] >= 0.0) || !std::isfinite(res[i]))
      return set_error(ctx, II_ERR_BAD_ARG, "poles must be non-negative shifts");
  return II_OK;
}

}  // namespace ii

using namespace ii;

extern "C" {

ii_status ii_rational_solve(ii_context* ctx, double kappa, double* sol, const double* src,
                            double a0, const double* res, const double* poles, int npoles,
                            double prec, int maxiter, int precision, ii_solver_output* out) {
  if (!sol || !src || !res || !poles || maxiter < 1 || precision < II_PREC_DOUBLE ||
      precision > II_PREC_BF16 || !(prec > 0))
    return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  II_CHECK(check_poles(ctx, res, poles, npoles));
  const Geom& g = ctx->g;
  auto t0 = std::chrono::steady_clock::now();
  ii_field *fb = nullptr, *fx = nullptr, *fs = nullptr;
  II_CHECK(ii_field_create(ctx, II_EVEN, 1, &fb));
  struct Del { ii_field* f; ~Del() { if (f) ii_field_destroy(f); } } d1{fb};
  II_CHECK(ii_field_create(ctx, II_EVEN, npoles, &fx));
  Del d2{fx};
  II_CHECK(ii_field_create(ctx, II_EVEN, 1, &fs));
  Del d3{fs};
  II_CHECK(ii_field_upload(fb, 0, src));
  ii_solver_output o{};
  ii_status st = shifted_solutions(ctx, kappa, fx->data, fb->data, poles, npoles, prec, maxiter,
                                   precision, &o);
  if (st != II_OK && st != II_ERR_NOT_CONVERGED) return st;
  const std::string msg = ctx->err;
  RatCoef c{};
  c.a0 = a0;
  for (int i = 0; i < npoles; ++i) c.res[i] = res[i];
  hipLaunchKernelGGL(k_rat_reconstruct, dim3(blocks_for(g.Vh)), dim3(kBlock), 0, ctx->stream, g,
                     (const double2*)fx->data, (const double2*)fb->data, fs->data, c, npoles);
  II_HIP(ctx, hipGetLastError());
  II_CHECK(ii_field_download(fs, 0, sol));
  o.timing_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (out) *out = o;
  if (st != II_OK) ctx->err = msg;
  return st;
}

ii_status ii_action_nf_md_force(ii_context* ctx, double kappa, const double* phi, int npf,
                                const double* ires, const double* ipoles, int npoles,
                                double prec, int maxiter, int precision, double* force,
                                ii_solver_output* out) {
  if (!phi || !ires || !ipoles || !force || npf < 1 || maxiter < 1 ||
      precision < II_PREC_DOUBLE || precision > II_PREC_BF16 || !(prec > 0))
    return II_ERR_BAD_ARG;
  II_CHECK(require_gauge(ctx));
  II_CHECK(enter(ctx));
  II_CHECK(check_poles(ctx, ires, ipoles, npoles));
  const Geom& g = ctx->g;
  const size_t he = half_elems(g);
  ii_field *fphi = nullptr, *fx = nullptr, *fz = nullptr;
  II_CHECK(ii_field_create(ctx, II_EVEN, npf, &fphi));
  struct Del { ii_field* f; ~Del() { if (f) ii_field_destroy(f); } } d1{fphi};
  II_CHECK(ii_field_create(ctx, II_EVEN, npoles, &fx));
  Del d2{fx};
  II_CHECK(ii_field_create(ctx, II_EVEN, 2, &fz));  // zeta and an odd temporary
  Del d3{fz};
  for (int pf = 0; pf < npf; ++pf) II_CHECK(ii_field_upload(fphi, pf, phi + (size_t)24 * g.Vh * pf));
  // the force accumulates in the device gauge layout
  void* wf = nullptr;
  II_CHECK(workspace(ctx, 9, (size_t)72 * g.stride * sizeof(double2), &wf));
  double2* F = (double2*)wf;
  void* wh = nullptr;
  II_CHECK(workspace(ctx, 8, (size_t)72 * g.V * sizeof(double), &wh));
  ii_status worst = II_OK;
  std::string msg;
  ii_solver_output o{};
  bool first = true;
  for (int pf = 0; pf < npf; ++pf) {
    ii_solver_output op{};
    ii_status st = shifted_solutions(ctx, kappa, fx->data, fphi->data + (size_t)pf * he, ipoles,
                                     npoles, prec, maxiter, precision, &op);
    if (st != II_OK && st != II_ERR_NOT_CONVERGED) return st;
    if (st != II_OK) {
      worst = st;
      msg = ctx->err;
    }
    if (op.iterations > o.iterations) o = op;
    for (int i = 0; i < npoles; ++i) {
      const double2* eta = fx->data + (size_t)i * he;
      II_CHECK(dhat64(ctx, kappa, eta, fz->data, fz->data + he));
      II_CHECK(md_force_eo64(ctx, kappa, eta, fz->data, F, true, ires[i], !first));
      first = false;
    }
  }
  II_CHECK(unpack_gauge_field(ctx, F, (double*)wh));
  II_HIP(ctx, hipMemcpyAsync(force, wh, (size_t)72 * g.V * sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
  II_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (out) *out = o;
  if (worst != II_OK) ctx->err = msg;
  return worst;
}

}  // extern "C"
