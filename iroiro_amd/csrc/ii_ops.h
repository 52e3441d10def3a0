// Host-side building blocks shared by the API and the solvers.
#pragma once

#include "ii_internal.h"

#include <type_traits>
#include <vector>

namespace ii {

// One hop launch: out(out_par) = a * [g5] H [g5] in (+ x), optional
// partial sums of Re<dv, out> into partials[rhs * nblocks + block].
// What a hop does with its output o at a site (after the scale a and XPAY):
//   kEpiStore  store o
//   kEpiDot    store o; partial sums of Re<dv, o>     (CG <p, A p>)
//   kEpiNorm   store o; partial sums of |o|^2         (<p, Dhat^dag Dhat p> = |Dhat p|^2)
//   kEpiRUpd   r = dv - alpha o, alpha = coef[rhs * coef_stride] (a device
//              scalar); store r in place of o; partial sums of |r|^2
//              (the CG residual update fused into the last hop of A p)
//   kEpiCDot   store o; partial sums of the complex <dv, o> = sum conj(dv) o
//              (re, im)                             (BiCGStab <r0, A p>)
//   kEpiTS     XPAY only: store o; partial sums of the complex <o, x> and of
//              |o|^2 from the x it adds             (BiCGStab omega = <t, s>/|t|^2)
// Several sums are stored as planes: sum j of (rhs, block) at
// partials[j * plane + rhs * nblocks + block], plane = nrhs * nblocks.
enum HopEpilogue {
  kEpiStore = 0, kEpiDot = 1, kEpiNorm = 2, kEpiRUpd = 3, kEpiCDot = 4, kEpiTS = 5
};
constexpr int epi_sums(int epi) { return epi == kEpiStore ? 0 : epi == kEpiCDot ? 2 : epi == kEpiTS ? 3 : 1; }

struct HopSpec {
  int out_par = 0;
  double a = 1.0;
  bool g5in = false;
  bool g5out = false;
  bool xpay = false;
  int epi = kEpiStore;
  const double* coef = nullptr;  // kEpiRUpd: alpha of rhs r at coef[r * coef_stride]
  int coef_stride = 0;
  int nrhs = 1;
  double* partials = nullptr;
  // optional per-rhs skip flags: rhs r is skipped when skip[r * skip_stride] != 0
  const int* skip = nullptr;
  int skip_stride = 0;
  // do not run an autotuner (which allocates) before the first launch; the
  // current launch shape is used
  bool no_tune = false;
  // multi-rhs hops: rhs per block and tile order forced (the autotuner's
  // candidates); rb = 0: the tuned or default shape
  int rb = 0;
  const HopOrder* order = nullptr;
};

// A 16-bit half field: components [rhs][12][stride] (half2 or bf16x2 as
// raw 32-bit words) plus per-site norms [rhs][stride].
struct LoField {
  void* v = nullptr;
  float* nrm = nullptr;
};

// rs: elements between the rhs of in, x and dv (0: back to back, 12 * stride;
// 24 * stride addresses one parity of full fields); out is back to back
ii_status hop64(ii_context* ctx, const HopSpec& h, const double2* in, double2* out,
                const double2* x, const double2* dv, int rs = 0);
ii_status hop32(ii_context* ctx, const HopSpec& h, const float2* in, float2* out,
                const float2* x, const float2* dv);
ii_status hop_half(ii_context* ctx, const HopSpec& h, const LoField& in, const LoField& out,
                   const LoField* x, const LoField* dv);
ii_status hop_bf16(ii_context* ctx, const HopSpec& h, const LoField& in, const LoField& out,
                   const LoField* x, const LoField* dv);

inline int hop_blocks(const Geom& g) { return blocks_for(g.Vh); }
// autotuner of the single-rhs hop traversal (ii_dslash.hip)
ii_status tune_hop(ii_context* ctx);
// Autotuner of the fused multi-rhs hop in storage precision P (II_PREC_*)
// for nrhs rhs: times every (rhs per block, tile order) candidate and keeps
// the fastest for that precision.  Not inside a stream capture (returns
// without choosing).
ii_status tune_mrhs(ii_context* ctx, int P, int nrhs);
// number of per-rhs partial sums a DOT hop writes for nrhs right-hand sides
int hop_dot_parts(const Geom& g, int nrhs, int precision);

// ------------------------------------------------------------- FP64 operators
// Dhat f = f - Deo Doe f (dirac_wilson_EvenOdd.hpp:79-84); t: odd temp.
ii_status dhat64(ii_context* ctx, double kappa, const double2* f, double2* out, double2* t,
                 int nrhs = 1);
// Dhat^dag f = f - g5 Deo Doe g5 f (:86-92); dot: partials of Re<dv, out>.
ii_status dhat_dag64(ii_context* ctx, double kappa, const double2* f, double2* out, double2* t,
                     const double2* dv = nullptr, double* partials = nullptr, int nrhs = 1);
// A = Dhat^dag Dhat (fopr.h:151), with optional Re<dv, A f> partials.
ii_status ddagd64(ii_context* ctx, double kappa, const double2* f, double2* out, double2* q,
                  double2* t, const double2* dv = nullptr, double* partials = nullptr,
                  int nrhs = 1);

// ------------------------------------------------------------- BLAS-1 (FP64)
// All operate on nrhs half fields [rhs][12][stride]; reductions produce
// one double per rhs via partial sums and a fixed-order final sum.
ii_status dot64(ii_context* ctx, const double2* a, const double2* b, double* dst, int nrhs = 1);
ii_status finalize_sums(ii_context* ctx, const double* partials, int nblk, double* dst, int nrhs);
ii_status copy64(ii_context* ctx, double2* dst, const double2* src, int nrhs = 1);
ii_status zero64(ii_context* ctx, double2* dst, int nrhs = 1);
// y = a*x + b*y with host scalars
ii_status axpby64(ii_context* ctx, double a, const double2* x, double b, double2* y, int nrhs = 1);
ii_status gamma5_64(ii_context* ctx, const double2* in, double2* out, int nrhs = 1);

// ------------------------------------------------------------- packing
// Host-array staging (ii_api.hip): n doubles in through workspace slot 0,
// an output buffer in slot 1, and the synchronous copy back.
ii_status stage_in(ii_context* ctx, const double* host, size_t n, double** dev);
ii_status stage_buf(ii_context* ctx, size_t n, double** dev);
ii_status stage_out(ii_context* ctx, double* host, const double* dev, size_t n);
ii_status pack_half_spinor(ii_context* ctx, const double* host_dev, double2* dst);
ii_status unpack_half_spinor(ii_context* ctx, const double2* src, double* host_dev);
ii_status pack_full_spinor(ii_context* ctx, const double* host_dev, double2* even, double2* odd);
ii_status unpack_full_spinor(ii_context* ctx, const double2* even, const double2* odd,
                             double* host_dev);
ii_status pack_gauge(ii_context* ctx, const double* host_dev);
ii_status unpack_gauge(ii_context* ctx, double* host_dev);
// the same repacking for any gauge-layout field (e.g. a force):
// host layout [mu][site][3][3] <-> device [parity][mu][9][stride]
ii_status pack_gauge_field(ii_context* ctx, const double* host_dev, double2* dst);
ii_status unpack_gauge_field(ii_context* ctx, const double2* src, double* host_dev);
ii_status build_lowprec_gauge(ii_context* ctx, int prec);
ii_status plaquette64(ii_context* ctx, double* plaq);

// solvers (ii_solvers.hip)
// Solver_CG::solve semantics on Dhat^dag Dhat x = b: x0 = b, stop on
// |r|^2/|b|^2 < prec (solver_CG.cpp:29-160).
ii_status cg_reference(ii_context* ctx, double kappa, double2* x, const double2* b, double prec,
                       int maxiter, ii_solver_output* out);
// Full-field propagator solve through the even/odd Schur system.
ii_status propagator(ii_context* ctx, double kappa, const ii_prop_params* prm,
                     const ii_field* src, ii_field* sol, ii_solver_output* outs);

// MD force (ii_force.hip)
// Force field F (device gauge layout) of Dirac_Wilson_EvenOdd::md_force
// from even half fields eta, zeta (dirac_wilson_EvenOdd.hpp:94-116), times
// weight; accumulate: add to F; ta: apply the (linear, idempotent)
// traceless anti-hermitian projection on the way out.
ii_status md_force_eo64(ii_context* ctx, double kappa, const double2* eta, const double2* zeta,
                        double2* F, bool ta, double weight = 1.0, bool accumulate = false);
// multishift on the even Schur system with the reference's recursion
// (ii_solvers.hip); full = false: x and b even half fields
ii_status multishift_reference(ii_context* ctx, double kappa, double2* x, const double2* b,
                               const double* shifts, int nshift, double prec, int maxiter,
                               ii_solver_output* out, std::vector<double>* per_shift,
                               bool full);
// mixed-precision multishift: fused FP32 stage + batched refinement to the
// true residual tol (ii_solvers.hip); out: nshift entries
ii_status multishift_mixed_dev(ii_context* ctx, double kappa, double2* x, const double2* b,
                               const double* shifts, int nshift, double tol, int precision,
                               int maxiter, ii_solver_output* out);

// address of the no-allocation scratch array on the context's device, taken
// once per device at context creation (ii_noalloc.hip)
ii_status noalloc_setup(ii_context* ctx);

// learned preconditioner (ii_learned.hip)
void learned_release(ii_context* ctx);
// FP32 master weights of the 4 layers on the host, [9][cout][cin][4][4]
// complex each, and their version (bumped whenever the weights change;
// w = nullptr: the version only)
ii_status net_host_weights(ii_context* ctx, int* version, std::vector<std::vector<float2>>* w);
// The batched network's input r: bf16 (re, im) pairs, [rhs][stride][colour]
// [spin], 12 words (48 bytes) per site, one 16-byte matrix-core operand per
// colour.  Written by the producers of r next to their FP32 copy: v[k] is
// component k = 3 spin + colour of the site (round to nearest even).
__device__ __forceinline__ uint32_t net_in_word(float2 v) {
  const __bf16 re = (__bf16)v.x, im = (__bf16)v.y;
  uint16_t a, b;
  __builtin_memcpy(&a, &re, 2);
  __builtin_memcpy(&b, &im, 2);
  return (uint32_t)a | ((uint32_t)b << 16);
}
__device__ __forceinline__ void net_in_store(uint32_t* __restrict__ p, size_t site,
                                             const float2 v[12]) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
    reinterpret_cast<uint4*>(p + site * 12)[c] =
        make_uint4(net_in_word(v[c]), net_in_word(v[3 + c]), net_in_word(v[6 + c]),
                   net_in_word(v[9 + c]));
}
// batched network (ii_bnet.hip): z = M r for nrhs even half fields, r in the
// bf16 input layout above, z FP32 [rhs][12][stride]; rhs k with
// skip[k * skip_stride] != 0 is left untouched (skip = nullptr: none)
ii_status bnet_apply(ii_context* ctx, int nrhs, const uint32_t* r, float2* z,
                     const int* skip = nullptr, int skip_stride = 0);
// builds what bnet_apply needs (links, weights, buffers) so that its
// launches can be captured in a graph
ii_status bnet_prepare(ii_context* ctx, int nrhs);
// the network stages' tuned launch shapes for tuned_nrhs rhs (0: not tuned)
void bnet_shapes_get(ii_context* ctx, int* tuned_nrhs, int rg[4], int by[4], int bz[4]);
ii_status bnet_shapes_set(ii_context* ctx, int tuned_nrhs, const int rg[4], const int by[4],
                          const int bz[4]);
void bnet_release(ii_context* ctx);
// batched flexible GCR with (use_learned) the network as preconditioner
// (ii_fgcr.hip); *state holds its workspaces, released by fgcr_release
ii_status fgcr_batched(ii_context* ctx, void** state, double kappa, const ii_prop_params* prm,
                       const ii_field* src, ii_field* sol, ii_solver_output* outs);
void fgcr_release(void* state);
ii_status learned_propagator(ii_context* ctx, double kappa, const ii_prop_params* p,
                             const ii_field* src, ii_field* sol, ii_solver_output* out);

// conversions between FP64 and the low-precision sloppy fields
ii_status convert_64_to_32(ii_context* ctx, const double2* src, float2* dst, int nrhs);
ii_status convert_32_to_64(ii_context* ctx, const float2* src, double2* dst, int nrhs);

}  // namespace ii
