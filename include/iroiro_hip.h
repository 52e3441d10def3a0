/*
 * iroiro_hip.h -- C ABI of the MI355X (gfx950) implementation of the
 * IroIro++ Wilson even/odd solve path.
 *
 * Plain C types only (pointers, sizes, doubles); no torch or HIP types.
 * Host arrays use the reference layouts, all float64:
 *   spinor   f[24*site + 2*(c + 3*s) + {re,im}]          include/format_F.h:36-42
 *   link     u[18*(site + V*mu) + 2*(3*a + b) + {re,im}]  include/format_G.h:34-40
 *   lexicographic site = x + Lx*(y + Ly*(z + Lz*t))       lib/Geometry/siteIndex.cpp:41
 *   even/odd half fields ordered as esec/osec: ascending lexicographic
 *   order of the sites with (x+y+z+t)%2 == 0 / 1     lib/Geometry/siteIndex_EvenOdd.cpp:21-62
 * The gauge field is handed over with any boundary phase already applied
 * (the reference negates U_t on the last time slice before building the
 * operator, lib/Measurements/measGeneral.cpp:205-215).
 *
 * The Wilson parameter is kappa = 0.5/(4 + mass) (lib/Dirac_ops/dirac_wilson.hpp:93).
 *
 * Every entry point returns an ii_status; ii_last_error() gives the text of
 * the last failure on a context.  A context owns one HIP stream and its
 * device buffers; calls on one context are not re-entrant, calls on
 * different contexts may run from different threads.
 */
#ifndef IROIRO_HIP_H
#define IROIRO_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  II_OK = 0,
  II_ERR_BAD_ARG = 1,        /* sizes, pointers or parameters invalid */
  II_ERR_NOT_CONVERGED = 2,  /* Errors::ConvergenceErr / "Not converged." */
  II_ERR_NAN = 3,            /* a NaN or Inf appeared in a solver scalar */
  II_ERR_DEVICE = 4,         /* HIP runtime failure */
  II_ERR_STATE = 5           /* e.g. no gauge field uploaded */
} ii_status;

typedef struct ii_context ii_context;
typedef struct ii_field ii_field;

/* Subsets of the lattice a field lives on. */
enum { II_EVEN = 0, II_ODD = 1, II_FULL = 2 };

/* Storage/compute precisions of the solvers. */
enum {
  II_PREC_DOUBLE = 0,   /* FP64 storage and arithmetic */
  II_PREC_SINGLE = 1,   /* FP32 storage and arithmetic, reconstruct-12 links */
  II_PREC_HALF = 2,     /* FP16 storage + per-site FP32 norm, r12 FP16 links, FP32 arithmetic */
  II_PREC_BF16 = 3      /* BF16 storage + per-site FP32 norm, r12 BF16 links, FP32 arithmetic */
};

/* SolverOutput (lib/Solver/solver.hpp:16-22) plus what the device solvers
 * add: true residual and timing split. */
typedef struct {
  double diff;            /* as the reference reports it (see each solver) */
  int iterations;         /* Iterations / Nconv as the reference counts them */
  double timing_ms;       /* wall time of the solve on the device stream */
  double true_residual;   /* |b - D x|/|b| on the full operator (propagator solves) */
  int restarts;           /* reliable updates / defect-correction cycles */
  int inner_iterations;   /* low-precision iterations (mixed solvers) */
  int recursion_restarts; /* reliable updates whose recomputed residual was far above the
                             recursive one, so the search direction was restarted */
} ii_solver_output;

/* ---------------------------------------------------------------- context */
/* Creates a context for a single-process lattice dims[4] = {Lx,Ly,Lz,Lt}
 * on HIP device `device`.  Lx must be even (siteIndex_EvenOdd.hpp:45). */
ii_status ii_context_create(ii_context** ctx, const int dims[4], int device);
ii_status ii_context_destroy(ii_context* ctx);
const char* ii_last_error(const ii_context* ctx);
/* Device name and HBM size, for reports. */
ii_status ii_device_info(ii_context* ctx, char* name, size_t name_len,
                         double* hbm_gib);

/* ------------------------------------------------------------ gauge field */
/* Uploads and repacks the gauge field (Format_G, full lexicographic
 * volume, 4 directions): the device copy behind Dirac::getGaugeField_ptr
 * (lib/Dirac_ops/dirac.hpp:25).  Any low-precision copies are rebuilt
 * lazily. */
ii_status ii_gauge_upload(ii_context* ctx, const double* u);
/* Dirac::update_internal_state() (lib/Dirac_ops/dirac.hpp:28): the caller
 * changed its gauge field in place; the device copy is dropped and must be
 * uploaded again before the next operator call. */
ii_status ii_gauge_invalidate(ii_context* ctx);
/* Average plaquette of the uploaded field (Re tr P / 3). */
ii_status ii_gauge_plaquette(ii_context* ctx, double* plaq);
/* Downloads the device gauge field back into the host layout. */
ii_status ii_gauge_download(ii_context* ctx, double* u);

/* Gauge configuration files (GaugeGlobal::initialize, lib/Main/gaugeGlobal.cpp:
 * 133-180; CCIO::ReadFromDisk / SaveOnDisk, lib/IO/fields_io.hpp:47-138,
 * 250-360).  File order: sites lexicographic (x fastest), 4 directions per
 * site, 3x3 matrix row by row as (re, im).
 *   II_IO_BINARY  read: big-endian doubles (BinaryReader); write: host-order
 *                 doubles, as BinaryWriter does (binary_writer.cpp:19-22)
 *   II_IO_NERSC   read: 4D_SU3_GAUGE (3x2) or 4D_SU3_GAUGE_3x3, IEEE32(BIG) or
 *                 IEEE64BIG (NERSC_reader.cpp); write: 3x3 IEEE64BIG with the
 *                 header of NERSC_writer.cpp:55-76
 *   II_IO_ILDG    LIME "ildg-binary-data" record, big-endian doubles (or
 *                 floats when an "ildg-format" record says precision 32);
 *                 write: "lime-test-text" + data record (ILDG_writer.cpp) */
typedef enum { II_IO_BINARY = 0, II_IO_NERSC = 1, II_IO_ILDG = 2 } ii_io_format;
typedef struct {
  double plaquette;          /* of the field on the device (Staples::plaquette) */
  double link_trace;         /* Staples::link_trace: <Re tr U>/3 */
  double header_plaquette;   /* NERSC header PLAQUETTE (read) / written */
  double header_link_trace;  /* NERSC header LINK_TRACE */
  unsigned int checksum;     /* NERSC: sum of the data's 32-bit words as stored */
  unsigned int header_checksum;
  int bits;                  /* 64 or 32 */
  int is3x2;                 /* NERSC 4D_SU3_GAUGE: third rows reconstructed */
} ii_io_info;
/* Reads a file into the device gauge field (which becomes valid, as after
 * ii_gauge_upload).  check != 0 applies NERSCReader::check: the plaquette
 * must match the header's to 1e-15 (64-bit) / 1e-8 (32-bit), else
 * II_ERR_STATE.  The checksum is reported, not enforced (the reference
 * prints it only, NERSC_reader.cpp:126-166). */
ii_status ii_gauge_read(ii_context* ctx, const char* path, int format, int check,
                        ii_io_info* info);
/* Writes the device gauge field to a file (SaveOnDisk with the Binary, NERSC
 * or ILDG writer). */
ii_status ii_gauge_write(ii_context* ctx, const char* path, int format, ii_io_info* info);

/* Cold start U = 1 on the device (GaugeConf_unit, lib/Main/gaugeConf.cpp:211-228). */
ii_status ii_gauge_unit(ii_context* ctx);
/* nsweeps SU(3) Cabibbo-Marinari / Kennedy-Pendleton heat-bath sweeps of
 * the Wilson plaquette action at coupling beta on the device field,
 * numbered first_sweep, first_sweep+1, ... for the counter-based RNG keyed
 * by seed.  No reference counterpart: its heat-bath is SU(2) only
 * (lib/UpdatingAlg/HeatBath/HBActions_SUN.cpp:85-92). */
ii_status ii_gauge_heatbath(ii_context* ctx, double beta, unsigned long long seed,
                            int first_sweep, int nsweeps);
/* AntiPeriodicBC<GaugeField>(TDIR)::apply_bc on the device field
 * (lib/Dirac_ops/BoundaryConditions/antiPeriodicBC.hpp:25-37). */
ii_status ii_gauge_apply_apbc(ii_context* ctx);

/* ---------------------------------------------- operators on host arrays */
/* Even/odd hops, w = -kappa H f restricted to parities
 * (Dirac_Wilson EO constructors, dirac_wilson.hpp:89-114,
 *  dirac_wilson.cpp:28-36, dirac_wilson_EvenOdd.hpp:59-77):
 *   mult_eo      = Deo.mult        odd  -> even
 *   mult_oe      = Doe.mult        even -> odd
 *   mult_eo_dag  = Doe.mult_dag    even -> odd
 *   mult_oe_dag  = Deo.mult_dag    odd  -> even                          */
ii_status ii_dirac_mult_eo(ii_context* ctx, double kappa, double* out, const double* in);
ii_status ii_dirac_mult_oe(ii_context* ctx, double kappa, double* out, const double* in);
ii_status ii_dirac_mult_eo_dag(ii_context* ctx, double kappa, double* out, const double* in);
ii_status ii_dirac_mult_oe_dag(ii_context* ctx, double kappa, double* out, const double* in);
/* Dirac_Wilson_EvenOdd::mult / mult_dag: f - Deo Doe f and its adjoint
 * (dirac_wilson_EvenOdd.hpp:79-92), on even half fields. */
ii_status ii_dirac_mult(ii_context* ctx, double kappa, double* out, const double* in);
ii_status ii_dirac_mult_dag(ii_context* ctx, double kappa, double* out, const double* in);
/* Fopr_DdagD::mult exactly as ii_fopr_ddagd_mult, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  It
 * allocates nothing: its staging and temporaries live in one static device
 * array of 32 MiB per GPU (its address taken once per device when the first
 * context on it is created), which bounds the lattice to a half volume of
 * about 34,900 sites (II_ERR_BAD_ARG above that).  It does not run the hop
 * autotuner.  All calls on one GPU share that array: calls that overlap on
 * one GPU (from several threads) race. */
ii_status ii_fopr_ddagd_mult_noalloc(ii_context* ctx, double kappa, double* out,
                                     const double* in);
/* Dirac_Wilson::gamma5 (dirac_wilson.cpp:120-126) on nsites sites. */
ii_status ii_dirac_gamma5(ii_context* ctx, double* out, const double* in, long nsites);
/* Fopr_DdagD::mult (lib/Fopr/fopr.h:146-153). */
ii_status ii_fopr_ddagd_mult(ii_context* ctx, double kappa, double* out, const double* in);
/* Full-index Dirac_Wilson::mult / mult_dag (dirac_wilson.cpp:28-36,
 * mult_full :115-118) on full lexicographic fields. */
ii_status ii_dirac_full_mult(ii_context* ctx, double kappa, double* out, const double* in,
                             int dag);
/* The hop mult_eo / mult_oe (dag = 0) or mult_oe_dag / mult_eo_dag
 * (dag = 1) in a storage precision (II_PREC_*): the nrhs FP64 half fields
 * (back to back, Format_F esec/osec order, parity 1 - out_par) are stored
 * in that precision (16-bit: per-site norm; links FP32 or 16-bit
 * reconstruct-12), hopped to parity out_par in one launch (the fused
 * multi-rhs kernel for nrhs > 1) and converted back to FP64.  The
 * reference is FP64 only; this applies the inner operator of the
 * mixed-precision solvers so its error against the FP64 operator
 * (dirac_wilson.cpp:28-36) can be measured. */
ii_status ii_dirac_mult_eo_prec(ii_context* ctx, double kappa, int precision, int out_par,
                                int dag, int nrhs, double* out, const double* in);

/* ------------------------------------------------ rational approximation */
/* RationalApprox (lib/Tools/RationalApprox/rationalapprox.cpp:20-80): the
 * minimax relative-error rational approximation of x^(exp_num/exp_den) on
 * [low, high] with num_deg = den_deg (the reference asserts equal degrees),
 * in partial fractions
 *   x^(a/b)  ~ *a0  + sum_i res[i]  / (x + pole[i])    (Residuals, Poles, Const)
 *   x^(-a/b) ~ *ia0 + sum_i ires[i] / (x + ipole[i])   (InvResiduals, ...)
 * poles ascending; error = the maximum relative error.  Host computation
 * (no device use): a levelled-error Remez exchange in 113-bit floating point
 * in place of the reference's MPFR AlgRemez (lib/Tools/Remez/alg_remez.cpp);
 * ctx may be NULL.  II_ERR_NOT_CONVERGED when the exchange does not level
 * the error (too high a degree for the arithmetic). */
ii_status ii_rational_approx(ii_context* ctx, int num_deg, int den_deg, int exp_num, int exp_den,
                             double low, double high, double* res, double* pole, double* a0,
                             double* ires, double* ipole, double* ia0, double* error);

/* RationalSolver_CG::solve / solve_inv (lib/Solver/rationalSolver_CG.cpp:
 * 12-68) on Fopr_DdagD(Dirac_Wilson_EvenOdd): the shifted solutions x_i of
 * (Dhat^dag Dhat + poles[i]) x_i = src, then sol = a0 src + sum res[i] x_i
 * (pass Poles/Residuals/Const for solve, the Inv* set for solve_inv).
 * precision II_PREC_DOUBLE: the reference's MultiShiftSolver_CG recursion,
 * prec = its squared stopping test (multiShiftSolver_CG.cpp:95-185);
 * II_PREC_SINGLE/HALF/BF16: the mixed-precision multishift, prec = the
 * target true residual |b - (A + pole_i) x_i| / |b| of every shift.
 * out: the multishift's summary (largest over the shifts). */
ii_status ii_rational_solve(ii_context* ctx, double kappa, double* sol, const double* src,
                            double a0, const double* res, const double* poles, int npoles,
                            double prec, int maxiter, int precision, ii_solver_output* out);
/* Action_Nf::md_force (lib/Action/action_Nf.cpp:84-110, no smearing) for npf
 * pseudofermions phi (even half fields back to back): per pseudofermion
 * eta_i = (Dhat^dag Dhat + ipoles[i])^-1 phi (solve_noReconstruct), then
 * force = TracelessAntihermite(sum_pf sum_i ires[i] md_force(eta_i, Dhat eta_i)),
 * with the MolecularDynApprox InvResiduals / InvPoles.  precision and prec
 * as in ii_rational_solve.  force: gauge layout, 18*4*V doubles. */
ii_status ii_action_nf_md_force(ii_context* ctx, double kappa, const double* phi, int npf,
                                const double* ires, const double* ipoles, int npoles,
                                double prec, int maxiter, int precision, double* force,
                                ii_solver_output* out);

/* ------------------------------------------------- MD force on host arrays */
/* Dirac_Wilson_EvenOdd::md_force (dirac_wilson_EvenOdd.hpp:94-116):
 * eta, zeta even half fields; fce (out) a gauge-layout field of
 * 18*4*V doubles, kappa * [Deo.md_force_p(Doe eta, zeta) +
 * Doe.md_force_m(Doe eta, zeta) + Doe.md_force_p(eta, Deo^dag zeta) +
 * Deo.md_force_m(eta, Deo^dag zeta)] with mkfrc of dirac_wilson.cpp:62-92. */
ii_status ii_dirac_md_force(ii_context* ctx, double kappa, double* fce, const double* eta,
                            const double* zeta);
/* Full-index Dirac_Wilson::md_force (dirac_wilson.cpp:52-60):
 * fce = -kappa (md_force_p + md_force_m) from full fields eta, zeta. */
ii_status ii_dirac_full_md_force(ii_context* ctx, double kappa, double* fce, const double* eta,
                                 const double* zeta);
/* Dirac_Wilson::md_force_p (minus = 0) / md_force_m (minus = 1) of the
 * full-index operator (dirac_wilson.cpp:41-50): ADDS the unscaled mkfrc sum
 * to fce (in/out, gauge layout), as the reference's Field::add does. */
ii_status ii_dirac_full_md_force_pm(ii_context* ctx, double* fce, const double* eta,
                                    const double* zeta, int minus);
/* FieldUtils::TracelessAntihermite (fieldUtils.cpp:264-274) on a gauge-layout
 * field: (m - m^dag)/2 - i Im tr(m)/3 per link.  out may equal in. */
ii_status ii_traceless_antihermite(ii_context* ctx, double* out, const double* in);
/* Action_Nf2::md_force (action_Nf2.cpp:69-86) without smearing, on the
 * uploaded gauge field: eta = (Dhat^dag Dhat)^-1 phi by Solver_CG
 * semantics (as ii_solver_cg: prec, maxiter, out), zeta = Dhat eta,
 * force = TracelessAntihermite(md_force(eta, zeta)), all on the device;
 * only phi (even half field) goes up and the force (gauge layout) comes
 * down.  Returns II_ERR_NOT_CONVERGED (force still written) when the solve
 * reaches maxiter. */
ii_status ii_action_nf2_md_force(ii_context* ctx, double kappa, const double* phi, double prec,
                                 int maxiter, double* force, ii_solver_output* out);

/* ------------------------------------------------- solvers on host arrays */
/* Solver_CG::solve on Fopr_DdagD(Dirac_Wilson_EvenOdd) (solver_CG.cpp:29-77):
 * x0 = b; stops when |r|^2/|b|^2 < prec (squared); iterations = loop index
 * at exit; diff = |A x - b|/|b|.  Returns II_ERR_NOT_CONVERGED (x still
 * written) when maxiter is reached. */
ii_status ii_solver_cg(ii_context* ctx, double kappa, double* x, const double* b,
                       double prec, int maxiter, ii_solver_output* out);
/* MultiShiftSolver_CG::solve on Fopr_DdagD + sigma_i
 * (multiShiftSolver_CG.cpp:95-185).  x: nshift half fields back to back.
 * out->diff = max_i |(A+sigma_i)x_i - b|^2/|b|^2 (the reference's `diff`),
 * out->true_residual = sqrt of it (its Out.diff); iterations = Nconv. */
ii_status ii_multishift_cg(ii_context* ctx, double kappa, double* x, const double* b,
                           const double* shifts, int nshift, double prec, int maxiter,
                           ii_solver_output* out);
/* Mixed-precision multishift on Fopr_DdagD(Dirac_Wilson_EvenOdd) + sigma_i,
 * no reference counterpart (the reference's multishift is FP64 only,
 * multiShiftSolver_CG.cpp:10-185).  Stage 1 runs the same shifted
 * recursion, all shifts fused per launch, with FP32 vectors and operator to
 * |r|/|b| < max(tol, 1e-6); stage 2 refines every shift at once with a
 * batched reliable-update CG (inner precision II_PREC_*, FP64 true
 * residuals) until |b - (A + sigma_i) x_i| <= tol |b|.  x: nshift even half
 * fields; out: nshift entries (iterations = stage 1 + stage 2 of that
 * shift, inner_iterations = stage 1, restarts = reliable updates,
 * true_residual per shift). */
ii_status ii_multishift_mixed(ii_context* ctx, double kappa, double* x, const double* b,
                              const double* shifts, int nshift, double tol, int precision,
                              int maxiter, ii_solver_output* out);
/* MultiShiftSolver_CG::solve on Fopr_DdagD of the full-index Dirac_Wilson
 * (D^dag D + sigma_i, multiShiftSolver_CG.cpp:10-185; the operator of
 * Qprop_MultiShift, qprop_MultiShift.cpp:10-31).  b: one full field
 * (24*V doubles, lexicographic); x: nshift full fields back to back.
 * Same outputs and errors as ii_multishift_cg. */
ii_status ii_multishift_cg_full(ii_context* ctx, double kappa, double* x, const double* b,
                                const double* shifts, int nshift, double prec, int maxiter,
                                ii_solver_output* out);
/* Inverter_WilsonLike::invert with a Solver_CG (eoUtils.cpp:43-80): full
 * lexicographic solve of D x = b. */
ii_status ii_eo_invert(ii_context* ctx, double kappa, double* sol, const double* src,
                       double prec, int maxiter, ii_solver_output* out);

/* ------------------------------------------------------ device fields */
/* A device-resident field (subset II_EVEN/II_ODD/II_FULL) holding nrhs
 * right-hand sides in the device layout.  Upload/download convert from/to
 * the host layout of one rhs. */
ii_status ii_field_create(ii_context* ctx, int subset, int nrhs, ii_field** f);
ii_status ii_field_destroy(ii_field* f);
ii_status ii_field_upload(ii_field* f, int rhs, const double* host);
ii_status ii_field_download(const ii_field* f, int rhs, double* host);
/* Point source of Source_local::mksrc (source_types.hpp:57-62) for every
 * (spin s, colour c) into rhs 3*s+c of a 12-rhs full field (column order of
 * Qprop_EvenOdd::calc, qprop_EvenOdd.cpp:11-23). */
ii_status ii_field_point_sources(ii_field* f, const int site[4]);

/* Complex Gaussian source (unit variance per real component, counter-based
 * RNG keyed by seed and stream) on the even sites of rhs `rhs`; the rest of
 * that rhs is zeroed.  The training sources of ii_net_train are these with
 * stream = pair index. */
ii_status ii_field_gaussian(ii_field* f, int rhs, unsigned long long seed, int stream);

/* Propagator solve parameters. */
typedef struct {
  double tol;            /* target |b - D x|/|b| per rhs on the full operator */
  int maxiter;           /* bound on total inner iterations per solve */
  int precision;         /* II_PREC_* of the inner solver */
  int batched;           /* nonzero: all rhs in one fused multi-rhs solve */
  double delta;          /* reliable-update trigger: |r| below delta times its value at the
                            last update (0 = default: 0.1 for CG, 0.01 for BiCGStab, 0.1
                            for BiCGStab in BF16).  Batched: rhs within a factor 4 of
                            their own trigger join the update of the rhs that trigger */
  int use_graph;         /* capture solver iterations in a HIP graph (II_METHOD_CG) */
  int use_learned;       /* apply the trained learned preconditioner: with II_METHOD_FGCR
                            inside the batched flexible GCR; with the other methods the
                            rhs-by-rhs FP64 flexible GCR (ii_fgcr_solve) */
  int method;            /* II_METHOD_CG: CG on Dhat^dag Dhat (Solver_CG on Fopr_DdagD);
                            II_METHOD_BICGSTAB: BiCGStab on Dhat itself (no reference
                            counterpart; sloppy precisions, batched);
                            II_METHOD_FGCR: flexible GCR on Dhat (no reference counterpart;
                            batched only): every rhs in each launch, FP32 Krylov vectors
                            and FP32 Dhat, the network (use_learned) on e4m3
                            activations with bfloat16 matrix-core arithmetic,
                            restarted every nkrylov directions from the
                            FP64 true residual; use_graph captures a restart cycle as
                            one HIP graph; batched mixed CG on the remaining defect if
                            the flexible iteration stalls */
  double restart_factor; /* a reliable update whose recomputed residual |r|^2 exceeds
                            restart_factor times the recursive one restarts the search
                            direction (0 = default 16) */
  int nkrylov;           /* flexible GCR: directions per restart cycle (0 = default 10) */
  int northo;            /* flexible GCR: each new direction is orthogonalised against
                            the last northo directions of its cycle (0 = default 1;
                            nkrylov or more: all of them, the full GCR) */
} ii_prop_params;
enum { II_METHOD_CG = 0, II_METHOD_BICGSTAB = 1, II_METHOD_FGCR = 2 };

/* Full solve D x = b for every rhs of a full field through the even/odd
 * Schur system (Inverter_WilsonLike::invert, eoUtils.cpp:43-80) with a
 * device CG on Dhat^dag Dhat or (p->method) BiCGStab on Dhat; stops on the
 * FP64 true full-D residual.
 * Sources and solutions stay on the device.  out: one entry per rhs. */
ii_status ii_propagator_solve(ii_context* ctx, double kappa, const ii_prop_params* p,
                              const ii_field* src, ii_field* sol, ii_solver_output* out);

/* ------------------------------------------------------ meson correlators */
/* MesonCorrelator::calculate / calculate_mom / calculate_PSmom
 * (lib/Measurements/FermionicM/meson_correlator.hpp:138-296) for all 18
 * MesonTypes (enum order Pion .. PS_AV4, :20-24) at once, on two device
 * propagators q1, q2 (II_FULL fields of 12 columns, column c + 3 s; q1 may
 * equal q2).
 *   mode 0: calculate along `direction` (0..3): out[18][L_direction]
 *   mode 1: calculate_mom with momentum mom[3] (time slices):
 *           out[18][2 Lt], (re, im) interleaved
 *   mode 2: calculate_PSmom with mom[3]: out[18][Lt]
 * Sums over sites run in a fixed order (deterministic). */
ii_status ii_meson_correlators(ii_context* ctx, const ii_field* q1, const ii_field* q2, int mode,
                               int direction, const double* mom, double* out);

/* ------------------------------------------- learned preconditioner
 * No counterpart in the reference (its solvers are unpreconditioned CG):
 * these calls add a trained, gauge-covariant preconditioner M for the
 * even/odd Schur operator Dhat and a flexible GCR that uses it.
 *
 * The network is linear: 4 layers of widths 1 -> 8 -> 8 -> 8 -> 1 spinor
 * channels on the full lattice.  Layer l maps its input channels by
 *   out[o](x) = sum_{d=0..8} sum_i W_l[d][o][i] (T_d in[i])(x),
 * T_0 = 1, T_{2mu+1} psi(x) = U_mu(x) psi(x+mu),
 * T_{2mu+2} psi(x) = U_mu(x-mu)^dag psi(x-mu), each W a complex 4x4 spin
 * matrix (acting alike on colour, so the map is gauge covariant).  M r is
 * the even part of N applied to r placed on the even sites.
 * Weights of layer l: float32 [9][cout][cin][4][4] complex, re/im
 * interleaved (9*cout*cin*32 floats).  Device arithmetic: bfloat16 weights
 * and bfloat16-rounded transported activations on the matrix cores, float32
 * accumulation and float32 activation storage.
 * Host full fields of nch channels: [nch][V][24] doubles, Format_F order
 * per channel (include/format_F.h:36-42), lexicographic sites. */
#define II_NET_LAYERS 4
typedef struct {
  int npairs;          /* training pairs (Gaussian source, plain-CG solution) */
  int epochs;          /* passes over the pairs */
  int batch;           /* pairs per optimizer step */
  double lr;           /* Adam step size */
  double pair_tol;     /* |bhat - Dhat y|/|bhat| of the plain-CG solutions */
  unsigned long long seed;
  int keep_weights;    /* nonzero: start from the current weights */
} ii_train_params;
typedef struct {
  double loss_first;   /* mean relative loss |M b - y|^2/|y|^2, first epoch */
  double loss_last;    /* ... last epoch */
  double pairs_ms;     /* time to produce the pairs with the plain solver */
  double train_ms;     /* time of the optimizer loop */
  int steps;           /* optimizer steps taken */
} ii_train_report;

/* Initialises the weights to the truncated Neumann series
 * sum_{k=0..4} (kappa H)^k of D^-1 (Horner form in channels 0 and 1), plus
 * noise of relative size `noise` on every weight (seeded). */
ii_status ii_net_init(ii_context* ctx, double kappa, double noise, unsigned long long seed);
ii_status ii_net_set_weights(ii_context* ctx, int layer, const float* w);
ii_status ii_net_get_weights(ii_context* ctx, int layer, float* w);
/* One layer (adjoint != 0: its adjoint, cout -> cin channels) on host full
 * fields; in has cin (cout) channels, out cout (cin). */
ii_status ii_net_layer_apply(ii_context* ctx, int layer, int adjoint, const double* in,
                             double* out);
/* Gradient of a real loss with respect to W_l (dL/dRe W + i dL/dIm W) given
 * the input of the layer and dL/dRe out + i dL/dIm out; gw as the weights. */
ii_status ii_net_layer_wgrad(ii_context* ctx, int layer, const double* in,
                             const double* grad_out, float* gw);
/* Loss sum_k |M b_k - y_k|^2/|y_k|^2 over npairs even half fields and its
 * gradient with respect to every layer's weights (concatenated). */
ii_status ii_net_loss_grad(ii_context* ctx, int npairs, const double* b_even,
                           const double* y_even, double* loss, float* grads);
/* z = M r on even half fields (esec order, 24*V/2 doubles). */
ii_status ii_net_apply(ii_context* ctx, const double* r_even, double* z_even);
/* z = M r for nrhs even half fields back to back, all rhs in one launch per
 * stage: the evaluation the batched flexible solver uses.  It regroups the
 * layers by parity (L1 and L3 are needed on odd sites only, composite
 * weights W2_0 W1_d + W2_d W1_0 etc. formed in double and rounded to
 * bfloat16), stores activations in bfloat16 and transports with bfloat16
 * links, so it equals ii_net_apply up to bfloat16 rounding. */
ii_status ii_net_apply_batched(ii_context* ctx, int nrhs, const double* r_even, double* z_even);
/* Produces npairs Gaussian sources bhat and their plain FP64 CG solutions
 * y = Dhat^-1 bhat on the device, then trains with Adam on the relative
 * loss.  Everything runs on the device. */
ii_status ii_net_train(ii_context* ctx, double kappa, const ii_train_params* p,
                       ii_train_report* rep);
ii_status ii_net_save(ii_context* ctx, const char* path);
ii_status ii_net_load(ii_context* ctx, const char* path);
/* Times every forward layer of the network (layer_ms[4], average over
 * nrep launches, HIP events on the context stream; the last layer on the
 * even output sites only, as M applies it). */
ii_status ii_net_bench(ii_context* ctx, int nrep, double* layer_ms);
/* Times the four stage kernels of ii_net_apply_batched for nrhs rhs
 * (stage_ms[4], average over nrep launches, HIP events on the context
 * stream): odd [L1 | L2], even L2, odd L3, even output. */
ii_status ii_bench_net_batched(ii_context* ctx, int nrhs, int nrep, double* stage_ms);
/* Launch-shape autotuner of ii_net_apply_batched (no counterpart in the
 * reference): times every admissible number of rhs processed side by side
 * per tile (1, 2, 3, 4, 6, 12) for each of the four stage kernels on nrhs
 * rhs and keeps the fastest (it also runs by itself before the first
 * application to a new rhs count).  force_rg (4 values >= 1, or NULL)
 * sets the choice instead of timing.  rg (4 values, or NULL) receives the
 * choice; report (len bytes, or NULL) the timings as text. */
ii_status ii_net_tune(ii_context* ctx, int nrhs, const int* force_rg, int* rg, char* report,
                      size_t len);
/* Flexible GCR on Dhat x = b (even half fields), FP64 vectors and Dhat,
 * right preconditioner M (use_net != 0) or none, restarted every `nkrylov`
 * directions with an FP64 true residual; stops when |b - Dhat x| <= tol |b|.
 * Falls back to CG on Dhat^dag Dhat if the flexible iteration stalls.
 * out->iterations counts outer directions, out->restarts restarts,
 * out->inner_iterations fallback CG iterations. */
ii_status ii_fgcr_solve(ii_context* ctx, double kappa, double* x_even, const double* b_even,
                        double tol, int maxiter, int nkrylov, int use_net,
                        ii_solver_output* out);

/* Qprop_MultiShift::calc (qprop_MultiShift.cpp:10-31) with
 * Source_local at `site`: for each column 3s+c a full-operator multishift
 * solve, then xq[(3s+c)*nshift + m] = D x_m.  xq: full field of 12*nshift
 * rhs on the device; out: 12 entries (one per column). */
ii_status ii_qprop_multishift(ii_context* ctx, double kappa, const int site[4],
                              const double* shifts, int nshift, double prec, int maxiter,
                              ii_field* xq, ii_solver_output* out);

/* ------------------------------------------------------ measurement */
/* Times one even/odd hop (Doe, V/2 output sites) in the given precision
 * with HIP events on the context stream, over nrep launches after a
 * warm-up; nrhs > 1 times the fused multi-rhs hop.  avg_ms per launch. */
ii_status ii_bench_hop(ii_context* ctx, int precision, int nrhs, int nrep, double* avg_ms);
/* Launch-shape autotuner: times the admissible site traversals of the
 * single-rhs hop on this lattice (run automatically before the first FP64
 * single-rhs hop) and, for 12 rhs in each storage precision, the fused
 * multi-rhs hop's rhs per block and tile order (run automatically before a
 * batch's first multi-rhs hop of that precision); reports the timings and
 * the choices as text. */
ii_status ii_autotune(ii_context* ctx, char* report, size_t len);

/* The autotuner's choices of this context (the single-rhs hop order, the multi-rhs hop shape
 * of each tuned precision and batch size, the network stages' shapes) to a text file, and
 * back into a context on the same lattice, so that a later run skips the timing.  No
 * counterpart in the reference (it has no autotuner).  Loading a file written for another
 * lattice returns II_ERR_BAD_ARG. */
ii_status ii_tune_save(ii_context* ctx, const char* path);
ii_status ii_tune_load(ii_context* ctx, const char* path);
/* Times Dirac_Wilson_EvenOdd::md_force on device-resident fields (no host
 * copies): total_ms = the two hops forming Doe eta and Deo^dag zeta plus the
 * force kernel; force_ms = the force kernel alone (both parities).  HIP
 * events on the context's stream, averaged over nrep launches. */
ii_status ii_bench_md_force(ii_context* ctx, int nrep, double* force_ms, double* total_ms);
/* Times a plain device copy of `bytes` bytes (read + write) with HIP
 * events: the achievable-bandwidth reference for the roofline and the
 * calibration kernel of the FETCH_SIZE / WRITE_SIZE counters. */
ii_status ii_bench_copy(ii_context* ctx, size_t bytes, int nrep, double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif
