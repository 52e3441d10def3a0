/*
 * wilson_oracle.h -- CPU restatement of the IroIro++ Wilson even/odd solve
 * path, used ONLY as the checker of the HIP implementation.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only to check or
 * time against it.  The product path (libiroiro_hip.so) never links, loads
 * or calls it.
 *
 * Layouts are the reference's (all float64):
 *   spinor  f[24*site + 2*(c + 3*s) + {0:re,1:im}]     include/format_F.h:36-42
 *   link    u[18*(site + V*mu) + 2*(3*a + b) + {re,im}] include/format_G.h:34-40
 *   lexicographic site = x + Lx*(y + Ly*(z + Lz*t))     lib/Geometry/siteIndex.cpp:41
 *   even/odd half fields in ascending lexicographic order of the sites with
 *   (x+y+z+t)%2 == 0 / 1                                 lib/Geometry/siteIndex_EvenOdd.cpp:21-62
 *
 * Parity status of each function (how it is pinned) is listed in its
 * comment; every function is pinned by tests/test_oracle.py.
 */
#ifndef IROIRO_WILSON_ORACLE_H
#define IROIRO_WILSON_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_geom or_geom;

/* Geometry and even/odd tables (siteIndex_EvenOdd.cpp:21-62). */
or_geom* or_geom_new(const int dims[4]);
void or_geom_free(or_geom* g);
long or_geom_volume(const or_geom* g);
void or_geom_eo_tables(const or_geom* g, int* esec, int* osec);

/* Antiperiodic time boundary: negate U_t at t = Lt-1 (antiPeriodicBC.hpp:25-37). */
void or_apbc(const or_geom* g, double* u);

/* gamma5: spin 0<->2, 1<->3 (wilsonLikeUtils.cpp:42-49). */
void or_gamma5(long nsites, double* w, const double* f);

/* Even/odd hop w = -kappa * H f (Dirac_Wilson::mult with the EO constructors,
 * dirac_wilson.cpp:28-32,107-113; dirac_wilson_improved.code:5-648).
 * eotype 1 = Deo (odd input -> even output), 2 = Doe (even -> odd).
 * dag != 0 gives gamma5 H gamma5 (dirac_wilson.cpp:34-36). */
void or_hop_eo(const or_geom* g, const double* u, double* w, const double* f,
               int eotype, int dag, double kappa);

/* Schur operator on even sites: Dhat f = f - Deo Doe f, Dhat^dag f =
 * f - g5 Deo Doe g5 f (dirac_wilson_EvenOdd.hpp:79-92). */
void or_dirac_eo(const or_geom* g, const double* u, double* w, const double* f,
                 int dag, double kappa);

/* Fopr_DdagD::mult = Dhat^dag (Dhat f) (fopr.h:146-153). */
void or_ddagd(const or_geom* g, const double* u, double* w, const double* f,
              double kappa);

/* Full lexicographic Wilson operator D = 1 - kappa H (mult_full,
 * dirac_wilson.cpp:115-118); dag gives g5 D g5. */
void or_dirac_full(const or_geom* g, const double* u, double* w,
                   const double* f, int dag, double kappa);

/* MD force of the full-index Dirac_Wilson (dirac_wilson.cpp:41-92):
 * fce = -kappa (md_force_p + md_force_m), a gauge-layout field, from full
 * fields eta and zeta.  mkfrc: fce(x,mu)_{ab} += sum_s [U_mu(x)(1-gamma_mu)
 * eta(x+mu)]_{a,s} conj(zeta_{b,s}(x)); md_force_m uses (g5 zeta, g5 eta).
 * Pinned in tests/test_oracle_force.py: Re<zeta, D[U + eps X U] eta> is
 * affine in eps with slope Re sum tr(X fce), and the reference's own
 * md_force on the golden gauge fields. */
void or_md_force_full(const or_geom* g, const double* u, double* fce,
                      const double* eta, const double* zeta, double kappa);
/* md_force_p (minus = 0) or md_force_m (minus = 1) of the full-index
 * Dirac_Wilson, accumulated into fce without a kappa factor
 * (dirac_wilson.cpp:41-50). */
void or_md_force_pm_full(const or_geom* g, const double* u, double* fce,
                         const double* eta, const double* zeta, int minus);
/* MD force of Dirac_Wilson_EvenOdd (dirac_wilson_EvenOdd.hpp:94-116) from
 * even half fields eta and zeta: a gauge-layout field over all links.
 * Pinned like or_md_force_full with S = Re<zeta, Dhat eta> (quadratic in
 * the links, so a central difference is exact) and by the reference. */
void or_md_force_eo(const or_geom* g, const double* u, double* fce,
                    const double* eta, const double* zeta, double kappa);
/* FieldUtils::TracelessAntihermite (fieldUtils.cpp:264-274):
 * (m - m^dag)/2 - i Im tr(m)/3 per 3x3 link matrix. */
void or_traceless_antihermite(long nlinks, double* out, const double* in);

/* Solver_CG on Dhat^dag Dhat (solver_CG.cpp:29-160): x0 = b, stop when
 * |r|^2/|b|^2 < prec.  Returns 0 when converged, 1 otherwise.
 * iterations = loop index at exit; diff = |A x - b|/|b|. */
int or_cg(const or_geom* g, const double* u, double kappa, double* x,
          const double* b, double prec, int maxiter, int* iterations,
          double* diff);

/* MultiShiftSolver_CG on Dhat^dag Dhat + sigma_i (multiShiftSolver_CG.cpp:
 * 10-185).  x holds nshift half fields back to back.  Returns 0 when
 * converged, 1 otherwise ("Not converged.").  nconv = loop index at exit;
 * diff = max_i |(A+sigma_i) x_i - b|^2/|b|^2; out_diff = sqrt(diff). */
int or_multishift_cg(const or_geom* g, const double* u, double kappa,
                     const double* shifts, int nshift, double* x,
                     const double* b, double prec, int maxiter, int* nconv,
                     double* diff, double* out_diff);

/* Inverter_WilsonLike::invert (eoUtils.cpp:43-80): full-volume solve of
 * D x = b through the even Schur system solved by or_cg. */
int or_eo_invert(const or_geom* g, const double* u, double kappa,
                 double* sol, const double* src, double prec, int maxiter,
                 int* iterations, double* diff);

/* |b - D x| / |b| on the full lexicographic operator. */
double or_true_residual_full(const or_geom* g, const double* u, double kappa,
                             const double* x, const double* b);

/* Row Gram-Schmidt reunitarisation of one 3x3 matrix (sunMat.hpp:244-277). */
void or_reunit(double* m);

/* U_mu(x) = exp(i eps H), H traceless Hermitian with unit-variance Gaussian
 * entries, then reunit; counter-based RNG keyed by seed. */
void or_gauge_expH(const or_geom* g, double eps, uint64_t seed, double* u);

/* Complex Gaussian spinor field, unit variance per real component. */
void or_gauss_field(long ndouble, uint64_t seed, double* f);

/* Average plaquette Re tr(P)/3 over all sites and the 6 planes. */
double or_plaquette(const or_geom* g, const double* u);

/* One SU(3) Cabibbo-Marinari heat-bath sweep (Kennedy-Pendleton SU(2)
 * subgroup updates), link order mu = 0..3, parity 0,1; RNG keyed by
 * (seed, sweep, mu, site, subgroup).  No reference counterpart for SU(3):
 * HBActions_SUN.cpp:85-92 aborts for NC != 2. */
void or_heatbath_sweep(const or_geom* g, double* u, double beta,
                       uint64_t seed, int sweep);

/* Counter-based uniform deviate in (0,1], shared with the GPU heat-bath. */
double or_uniform(uint64_t seed, uint64_t stream, uint64_t counter);

/* ---- learned gauge-covariant preconditioner (no reference counterpart) --
 * One linear layer maps cin channels of full-lattice spinor fields to cout
 * channels:
 *   out[o](x) = sum_{d=0..8} sum_i W[d][o][i] (T_d in[i])(x)
 * T_0 = identity, T_{2mu+1} psi(x) = U_mu(x) psi(x+mu),
 * T_{2mu+2} psi(x) = U_mu(x-mu)^dag psi(x-mu); W[d][o][i] is a complex 4x4
 * spin matrix acting identically on colour (gauge covariant).
 * Fields: [channel][site][24] doubles (Format_F per channel, lexicographic).
 * Weights: [9][cout][cin][4][4] complex = interleaved re,im doubles.
 * round_bf16 != 0 rounds every transported value (T_d in[i])(x) to the
 * nearest bfloat16 (ties to even) before it meets W, as a device layer that
 * feeds bf16 matrix cores does.  Each of the three functions is pinned by
 * tests/test_oracle_net.py: reduction to or_dirac_full for Wilson weights,
 * the adjoint identity, finite differences of a loss, gauge covariance. */
void or_net_layer(const or_geom* g, const double* u, int cin, int cout, const double* w,
                  const double* in, double* out, int round_bf16);
/* Adjoint layer, written as the scatter transpose of or_net_layer:
 * in_grad(n_d(x)) += U_d(x)^dag W_d^dag out_grad(x); with round_bf16 the
 * colour-transported U_d(x)^dag out_grad(x) is rounded before W_d^dag. */
void or_net_layer_adj(const or_geom* g, const double* u, int cin, int cout, const double* w,
                      const double* out_grad, double* in_grad, int round_bf16);
/* Gradient of a real loss L with respect to W (dL/dRe W + i dL/dIm W), given
 * grad_out = dL/dRe out + i dL/dIm out:
 * gw[d][o][i][s][t] += sum_{x,c} grad_out[o](x)[s][c] * conj((T_d in[i])(x)[t][c]).
 * With round_bf16 both factors are rounded to bfloat16 first. */
void or_net_layer_wgrad(const or_geom* g, const double* u, int cin, int cout,
                        const double* in, const double* grad_out, double* gw, int round_bf16);

#ifdef __cplusplus
}
#endif
#endif
