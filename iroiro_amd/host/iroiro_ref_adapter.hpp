// iroiro_ref_adapter.hpp -- drop-in classes for the IroIro++ source tree
// that run the Wilson even/odd solve path on an MI355X through the C ABI of
// libiroiro_hip.so (include/iroiro_hip.h).
//
// Include this header in the reference's build, after its own headers are
// reachable (-I<tree> -I<tree>/include -I<tree>/lib), and link
// -liroiro_hip.  The classes implement the reference's own abstract
// interfaces, so its drivers call them unchanged:
//
//   Dirac_Wilson_EvenOdd_HIP : DiracWilsonLike_EvenOdd
//       (lib/Dirac_ops/dirac_WilsonLike.hpp:66-82; replaces
//        Dirac_Wilson_EvenOdd<Dirac_Wilson>, dirac_wilson_EvenOdd.hpp:12-92)
//   Fopr_DdagD_HIP : Fopr_Herm          (lib/Fopr/fopr.h:72-77, 146-153)
//   Solver_CG_HIP : Solver              (lib/Solver/solver.hpp:24-30;
//                                        replaces Solver_CG, solver_CG.cpp:29-160)
//   MultiShiftSolver_CG_HIP : MultiShiftSolver
//                                       (lib/Solver/multiShiftSolver.hpp:15-24;
//                                        replaces multiShiftSolver_CG.cpp:10-185)
//   Dirac_Wilson_HIP : DiracWilsonLike  (the full-index Dirac_Wilson,
//                                        dirac_wilson.hpp:117-130, of
//                                        Qprop_MultiShift)
//   RationalSolver_CG_HIP : RationalSolver
//                                       (lib/Solver/rationalSolver.hpp:16-27;
//                                        replaces RationalSolver_CG,
//                                        rationalSolver_CG.cpp:12-111)
//   RationalApprox_HIP                  the coefficients of RationalApprox
//                                       (rationalapprox.hpp:56-96) from the
//                                       library's Remez (no MPFR needed)
//
// Inverter_WilsonLike (eoUtils.cpp:43-80), Qprop_EvenOdd
// (qprop_EvenOdd.cpp:11-43), Qprop_MultiShift (qprop_MultiShift.cpp:10-31),
// Action_Nf2 and RationalSolver_CG take these
// through their base-class pointers.  Error behaviour follows the
// reference: CG non-convergence calls Errors::ConvergenceErr (which exits),
// the multishift solver throws "Not converged.", device failures abort with
// the library's message.  md_force (and md_force_p/_m of the full-index
// operator) run on the device; only getDeo/getDoe, which the domain-wall
// kernels use, hand out the reference's own CPU sub-operators.
#ifndef IROIRO_REF_ADAPTER_HPP
#define IROIRO_REF_ADAPTER_HPP

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <typeinfo>
#include <vector>

#include "include/commonPrms.hpp"
#include "include/errors.hpp"
#include "include/field.h"
#include "Dirac_ops/dirac_WilsonLike.hpp"
#include "Dirac_ops/dirac_wilson.hpp"
#include "Dirac_ops/dirac_wilson_EvenOdd.hpp"
#include "Fopr/fopr.h"
#include "Solver/solver.hpp"
#include "Solver/multiShiftSolver.hpp"
#include "Solver/rationalSolver.hpp"

#include "iroiro_hip.h"

namespace iroiro_hip_detail {

inline void die(ii_context* ctx, ii_status st, const char* what) {
  std::fprintf(stderr, "iroiro_hip: %s failed (status %d): %s\n", what, (int)st,
               ctx ? ii_last_error(ctx) : "");
  std::abort();
}
inline void check(ii_context* ctx, ii_status st, const char* what) {
  if (st != II_OK) die(ctx, st, what);
}
inline double* addr(const Field& f) { return const_cast<Field&>(f).getaddr(0); }

// One device context per process and lattice (the GPU side of CommonPrms).
inline ii_context* context() {
  static ii_context* ctx = 0;
  if (!ctx) {
    int dims[4] = {CommonPrms::Nx(), CommonPrms::Ny(), CommonPrms::Nz(), CommonPrms::Nt()};
    ii_status st = ii_context_create(&ctx, dims, 0);
    if (st != II_OK) die(0, st, "ii_context_create");
  }
  return ctx;
}

// What the device Fopr and solvers need from either device operator.
class DeviceOp {
 public:
  virtual ~DeviceOp() {}
  virtual ii_context* context() const = 0;
  virtual double getKappa() const = 0;
  virtual size_t fsize() const = 0;
  virtual bool full_index() const = 0;
  // D f and D^dag f on host fields of fsize() doubles
  virtual const Field mult(const Field& f) const = 0;
  virtual const Field mult_dag(const Field& f) const = 0;
};

}  // namespace iroiro_hip_detail

class Dirac_Wilson_EvenOdd_HIP : public DiracWilsonLike_EvenOdd,
                                 public iroiro_hip_detail::DeviceOp {
 private:
  double kpp_;
  const Field* u_;
  ii_context* ctx_;
  Dirac_Wilson_EvenOdd<Dirac_Wilson> cpu_;  // getDeo / getDoe only
  Dirac_Wilson_EvenOdd_HIP(const Dirac_Wilson_EvenOdd_HIP&);

  const Field hop(ii_status (*fn)(ii_context*, double, double*, const double*),
                  const Field& f, const char* what) const {
    Field w(fsize());
    iroiro_hip_detail::check(ctx_, fn(ctx_, kpp_, w.getaddr(0), iroiro_hip_detail::addr(f)), what);
    return w;
  }

 public:
  Dirac_Wilson_EvenOdd_HIP(double mass, const Field* u)
      : kpp_(0.5 / (4.0 + mass)), u_(u), ctx_(iroiro_hip_detail::context()), cpu_(mass, u) {
    update_internal_state();
  }

  size_t fsize() const { return (size_t)CommonPrms::Nvol() / 2 * 24; }
  size_t gsize() const { return (size_t)CommonPrms::Nvol() * 18 * 4; }
  double getMass() const { return 0.5 / kpp_ - 4.0; }
  double getKappa() const { return kpp_; }
  const Field* getGaugeField_ptr() const { return u_; }
  ii_context* context() const { return ctx_; }
  bool full_index() const { return false; }

  // Dirac::update_internal_state (dirac.hpp:28): the caller changed U.
  void update_internal_state() {
    iroiro_hip_detail::check(ctx_, ii_gauge_upload(ctx_, iroiro_hip_detail::addr(*u_)),
                             "ii_gauge_upload");
  }

  const Field mult(const Field& f) const { return hop(ii_dirac_mult, f, "mult"); }
  const Field mult_dag(const Field& f) const { return hop(ii_dirac_mult_dag, f, "mult_dag"); }
  const Field mult_eo(const Field& f) const { return hop(ii_dirac_mult_eo, f, "mult_eo"); }
  const Field mult_oe(const Field& f) const { return hop(ii_dirac_mult_oe, f, "mult_oe"); }
  const Field mult_eo_dag(const Field& f) const {
    return hop(ii_dirac_mult_eo_dag, f, "mult_eo_dag");
  }
  const Field mult_oe_dag(const Field& f) const {
    return hop(ii_dirac_mult_oe_dag, f, "mult_oe_dag");
  }
  const Field mult_ee(const Field& f) const { return f; }
  const Field mult_oo(const Field& f) const { return f; }
  const Field mult_ee_inv(const Field& f) const { return f; }
  const Field mult_oo_inv(const Field& f) const { return f; }
  const Field gamma5(const Field& f) const {
    Field w(f.size());
    iroiro_hip_detail::check(ctx_, ii_dirac_gamma5(ctx_, w.getaddr(0), iroiro_hip_detail::addr(f),
                                                   (long)(f.size() / 24)),
                             "gamma5");
    return w;
  }
  const DiracWilsonLike* getDeo() const { return cpu_.getDeo(); }
  const DiracWilsonLike* getDoe() const { return cpu_.getDoe(); }
  // Dirac_Wilson_EvenOdd::md_force (dirac_wilson_EvenOdd.hpp:110-116)
  const Field md_force(const Field& eta, const Field& zeta) const {
    Field fce(gsize());
    iroiro_hip_detail::check(
        ctx_, ii_dirac_md_force(ctx_, kpp_, fce.getaddr(0), iroiro_hip_detail::addr(eta),
                                iroiro_hip_detail::addr(zeta)),
        "md_force");
    return fce;
  }
};

// Full-index Dirac_Wilson (dirac_wilson.hpp:117-130): D = 1 - kappa H on
// full lexicographic fields (mult_full, dirac_wilson.cpp:115-118).
class Dirac_Wilson_HIP : public DiracWilsonLike, public iroiro_hip_detail::DeviceOp {
 private:
  double kpp_;
  const Field* u_;
  ii_context* ctx_;
  Dirac_Wilson_HIP(const Dirac_Wilson_HIP&);
  void md_force_pm(Field& fce, const Field& eta, const Field& zeta, int minus) const {
    iroiro_hip_detail::check(
        ctx_, ii_dirac_full_md_force_pm(ctx_, fce.getaddr(0), iroiro_hip_detail::addr(eta),
                                        iroiro_hip_detail::addr(zeta), minus),
        minus ? "md_force_m" : "md_force_p");
  }

  const Field full(const Field& f, int dag, const char* what) const {
    Field w(fsize());
    iroiro_hip_detail::check(
        ctx_, ii_dirac_full_mult(ctx_, kpp_, w.getaddr(0), iroiro_hip_detail::addr(f), dag), what);
    return w;
  }

 public:
  Dirac_Wilson_HIP(double mass, const Field* u)
      : kpp_(0.5 / (4.0 + mass)), u_(u), ctx_(iroiro_hip_detail::context()) {
    update_internal_state();
  }
  size_t fsize() const { return (size_t)CommonPrms::Nvol() * 24; }
  size_t gsize() const { return (size_t)CommonPrms::Nvol() * 18 * 4; }
  double getMass() const { return 0.5 / kpp_ - 4.0; }
  double getKappa() const { return kpp_; }
  const Field* getGaugeField_ptr() const { return u_; }
  ii_context* context() const { return ctx_; }
  bool full_index() const { return true; }
  void update_internal_state() {
    iroiro_hip_detail::check(ctx_, ii_gauge_upload(ctx_, iroiro_hip_detail::addr(*u_)),
                             "ii_gauge_upload");
  }
  const Field mult(const Field& f) const { return full(f, 0, "mult"); }
  const Field mult_dag(const Field& f) const { return full(f, 1, "mult_dag"); }
  const Field gamma5(const Field& f) const {
    Field w(f.size());
    iroiro_hip_detail::check(ctx_, ii_dirac_gamma5(ctx_, w.getaddr(0), iroiro_hip_detail::addr(f),
                                                   (long)(f.size() / 24)),
                             "gamma5");
    return w;
  }
  // Dirac_Wilson::md_force_p / md_force_m / md_force (dirac_wilson.cpp:41-60)
  void md_force_p(Field& fce, const Field& eta, const Field& zeta) const {
    md_force_pm(fce, eta, zeta, 0);
  }
  void md_force_m(Field& fce, const Field& eta, const Field& zeta) const {
    md_force_pm(fce, eta, zeta, 1);
  }
  const Field md_force(const Field& eta, const Field& zeta) const {
    Field fce(gsize());
    iroiro_hip_detail::check(
        ctx_, ii_dirac_full_md_force(ctx_, kpp_, fce.getaddr(0), iroiro_hip_detail::addr(eta),
                                     iroiro_hip_detail::addr(zeta)),
        "md_force");
    return fce;
  }
};

// Fopr_DdagD on the device (fopr.h:146-153), over either device operator.
class Fopr_DdagD_HIP : public Fopr_Herm {
  const iroiro_hip_detail::DeviceOp* D_;

 public:
  explicit Fopr_DdagD_HIP(const iroiro_hip_detail::DeviceOp* D) : D_(D) {}
  double func(double x) const { return x; }
  const Field mult(const Field& f) const {
    if (D_->full_index()) return D_->mult_dag(D_->mult(f));
    Field w(fsize());
    iroiro_hip_detail::check(D_->context(),
                             ii_fopr_ddagd_mult(D_->context(), D_->getKappa(), w.getaddr(0),
                                                iroiro_hip_detail::addr(f)),
                             "Fopr_DdagD::mult");
    return w;
  }
  size_t fsize() const { return D_->fsize(); }
  const iroiro_hip_detail::DeviceOp* dirac() const { return D_; }
};

// Solver_CG with the whole iteration on the device (solver_CG.cpp:29-160):
// same parameters (Precision on |r|^2/|b|^2, MaxIter), x0 = b, same
// Iterations / diff / timing in SolverOutput.
class Solver_CG_HIP : public Solver {
  const Fopr_DdagD_HIP* opr_;
  double prec_;
  int maxiter_;

 public:
  Solver_CG_HIP(double prec, int maxiter, const Fopr_DdagD_HIP* fopr)
      : opr_(fopr), prec_(prec), maxiter_(maxiter) {}

  SolverOutput solve(Field& xq, const Field& b) const {
    const iroiro_hip_detail::DeviceOp* D = opr_->dirac();
    ii_context* ctx = D->context();
    if (D->full_index()) iroiro_hip_detail::die(ctx, II_ERR_BAD_ARG, "Solver_CG on the full index");
    if (xq.size() != b.size()) xq.resize(b.size());
    ii_solver_output o;
    ii_status st = ii_solver_cg(ctx, D->getKappa(), xq.getaddr(0), iroiro_hip_detail::addr(b),
                                prec_, maxiter_, &o);
    if (st == II_ERR_NOT_CONVERGED) {
      std::ostringstream msg;
      msg << "CG solver not converged. " << ii_last_error(ctx);
      Errors::ConvergenceErr(msg);
    }
    iroiro_hip_detail::check(ctx, st, "Solver_CG::solve");
    SolverOutput out;
    out.Msg = "CG solver";
    out.Iterations = o.iterations;
    out.diff = o.diff;
    out.timing = o.timing_ms;
    return out;
  }
  bool check_DdagD() const { return true; }
};

// MultiShiftSolver_CG on the device (multiShiftSolver_CG.cpp:95-185).
class MultiShiftSolver_CG_HIP : public MultiShiftSolver {
  const Fopr_DdagD_HIP* opr_;
  double prec_;
  int maxiter_;

 public:
  MultiShiftSolver_CG_HIP(const Fopr_DdagD_HIP* fopr, double stop_cond, int niter)
      : opr_(fopr), prec_(stop_cond), maxiter_(niter) {}
  const Fopr_DdagD_HIP* fopr() const { return opr_; }
  double stop_cond() const { return prec_; }
  int max_iter() const { return maxiter_; }

  SolverOutput solve(vector_Field& xq, const Field& b, const vector_double& sigma, double& diff,
                     int& Nconv) const {
    const iroiro_hip_detail::DeviceOp* D = opr_->dirac();
    ii_context* ctx = D->context();
    const size_t n = b.size();
    std::vector<double> x(n * sigma.size());
    ii_solver_output o;
    ii_status st = (D->full_index() ? ii_multishift_cg_full : ii_multishift_cg)(
        ctx, D->getKappa(), &x[0], iroiro_hip_detail::addr(b), &sigma[0], (int)sigma.size(),
        prec_, maxiter_, &o);
    if (st == II_ERR_NOT_CONVERGED) throw "Not converged.";
    iroiro_hip_detail::check(ctx, st, "MultiShiftSolver_CG::solve");
    for (size_t i = 0; i < sigma.size(); ++i) {
      if (xq[i].size() != n) xq[i].resize(n);
      for (size_t k = 0; k < n; ++k) xq[i].set(k, x[i * n + k]);
    }
    diff = o.diff;
    Nconv = o.iterations;
    SolverOutput out;
    out.Msg = "Multishift CG solver";
    out.Iterations = o.iterations;
    out.diff = o.true_residual;
    out.timing = o.timing_ms;
    return out;
  }
};

// The partial fractions of RationalApprox (rationalapprox.cpp:20-103),
// x^(a/b) ~ Const + sum Residuals/(x + Poles) and the inverse set, from the
// library's quad-precision Remez (ii_rational_approx) instead of the MPFR
// AlgRemez; same parameters (RationalApprox_params) and accessors.
class RationalApprox_HIP {
  int num_, den_;
  std::vector<double> res_, pole_, ires_, ipole_;
  double a0_, ia0_, err_;

 public:
  RationalApprox_HIP(int num_deg, int den_deg, int exp_num, int exp_den, double low, double high)
      : num_(exp_num), den_(exp_den), res_(num_deg), pole_(num_deg), ires_(num_deg),
        ipole_(num_deg), a0_(0), ia0_(0), err_(0) {
    ii_status st = ii_rational_approx(0, num_deg, den_deg, exp_num, exp_den, low, high, &res_[0],
                                      &pole_[0], &a0_, &ires_[0], &ipole_[0], &ia0_, &err_);
    if (st != II_OK) iroiro_hip_detail::die(0, st, "ii_rational_approx");
  }
  std::vector<double> Residuals() const { return res_; }
  std::vector<double> Poles() const { return pole_; }
  double Const() const { return a0_; }
  std::vector<double> InvResiduals() const { return ires_; }
  std::vector<double> InvPoles() const { return ipole_; }
  double InvConst() const { return ia0_; }
  double exponent() const { return (double)num_ / (double)den_; }
  double error() const { return err_; }
  // RationalApprox::rescale (rationalapprox.cpp:82-103)
  void rescale(double /*min*/, double maxIn) {
    const double mx = 1.05 * maxIn;
    double f = std::pow(mx, exponent());
    a0_ *= f;
    for (size_t i = 0; i < res_.size(); ++i) { res_[i] *= f * mx; pole_[i] *= mx; }
    f = std::pow(mx, -exponent());
    ia0_ *= f;
    for (size_t i = 0; i < ires_.size(); ++i) { ires_[i] *= f * mx; ipole_[i] *= mx; }
  }
};

// RationalSolver_CG (rationalSolver_CG.cpp:12-111) on the device: the
// multishift and the reconstruction sum both run there, so solve/solve_inv
// move one field each way instead of one per pole.  set_Approx takes the
// reference's RationalApprox (its accessors) or a RationalApprox_HIP.
class RationalSolver_CG_HIP : public RationalSolver {
  const MultiShiftSolver_CG_HIP* MS_Solver_;
  vector_double Residuals, Poles, InvResiduals, InvPoles;
  double ConstTerm, InvConstTerm;

  SolverOutput rational(Field& sol, const Field& src, double a0, const vector_double& res,
                        const vector_double& poles, const char* what) const {
    if (res.size() == 0) {
      std::fprintf(stderr, "[RationalSolver] Rational Approximation not initialized yet\n");
      std::abort();
    }
    const iroiro_hip_detail::DeviceOp* D = MS_Solver_->fopr()->dirac();
    ii_context* ctx = D->context();
    if (D->full_index()) iroiro_hip_detail::die(ctx, II_ERR_BAD_ARG, "RationalSolver_CG on the full index");
    sol.resize(src.size());
    ii_solver_output o;
    ii_status st = ii_rational_solve(ctx, D->getKappa(), sol.getaddr(0),
                                     iroiro_hip_detail::addr(src), a0, &res[0], &poles[0],
                                     (int)poles.size(), MS_Solver_->stop_cond(),
                                     MS_Solver_->max_iter(), II_PREC_DOUBLE, &o);
    if (st == II_ERR_NOT_CONVERGED) throw "Not converged.";
    iroiro_hip_detail::check(ctx, st, what);
    SolverOutput out;
    out.Msg = "Rational CG solver";
    out.Iterations = o.iterations;
    out.diff = o.true_residual;
    out.timing = o.timing_ms;
    return out;
  }

 public:
  explicit RationalSolver_CG_HIP(const MultiShiftSolver_CG_HIP* Solv)
      : MS_Solver_(Solv), ConstTerm(0), InvConstTerm(0) {}
  RationalSolver_CG_HIP(const MultiShiftSolver_CG_HIP* Solv, const RationalApprox_HIP& RA)
      : MS_Solver_(Solv), ConstTerm(0), InvConstTerm(0) {
    set_Approx(RA);
  }
  void set_Approx(RationalApprox& RA) {
    Residuals = RA.Residuals(); Poles = RA.Poles(); ConstTerm = RA.Const();
    InvResiduals = RA.InvResiduals(); InvPoles = RA.InvPoles(); InvConstTerm = RA.InvConst();
  }
  void set_Approx(const RationalApprox_HIP& RA) {
    Residuals = RA.Residuals(); Poles = RA.Poles(); ConstTerm = RA.Const();
    InvResiduals = RA.InvResiduals(); InvPoles = RA.InvPoles(); InvConstTerm = RA.InvConst();
  }
  SolverOutput solve(Field& sol, const Field& src) const {
    return rational(sol, src, ConstTerm, Residuals, Poles, "RationalSolver_CG::solve");
  }
  SolverOutput solve_inv(Field& sol, const Field& src) const {
    return rational(sol, src, InvConstTerm, InvResiduals, InvPoles, "RationalSolver_CG::solve_inv");
  }
  SolverOutput solve_noReconstruct(vector_Field& shifted_sol, const Field& src) const {
    shifted_sol.resize(InvPoles.size());
    for (size_t i = 0; i < shifted_sol.size(); ++i) shifted_sol[i].resize(src.size());
    double diff = 0;
    int nconv = 0;
    return MS_Solver_->solve(shifted_sol, src, InvPoles, diff, nconv);
  }
  SolverOutput solve_noReconstruct_inv(vector_Field& shifted_sol, const Field& src) const {
    shifted_sol.resize(Poles.size());
    for (size_t i = 0; i < shifted_sol.size(); ++i) shifted_sol[i].resize(src.size());
    double diff = 0;
    int nconv = 0;
    return MS_Solver_->solve(shifted_sol, src, Poles, diff, nconv);
  }
  bool check_DdagD() const { return true; }
};

#endif
