/* Runs every function of the CPU oracle (oracle/wilson_oracle.c) on small
 * lattices, one of them with a half volume that is not a multiple of 64,
 * under the host sanitizers (address, undefined behaviour): a build check
 * of the checker itself.  Built and run by tests/test_sanitizers.py; the
 * numbers it prints are not compared (the oracle's values are pinned by the
 * other CPU tests), only a clean run is required, plus two identities that
 * a memory error would break: the CG solution solves the system and the
 * even/odd inverter's solution has a small full-operator residual. */
#include "../../oracle/wilson_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static double* dalloc(long n) {
  double* p = (double*)calloc((size_t)n, sizeof(double));
  if (!p) exit(3);
  return p;
}

static int run(const int dims[4]) {
  or_geom* g = or_geom_new(dims);
  const long V = or_geom_volume(g), Vh = V / 2;
  const long ng = 18 * 4 * V, nf = 24 * V, nh = 24 * Vh;
  int* e = (int*)malloc(sizeof(int) * Vh);
  int* o = (int*)malloc(sizeof(int) * Vh);
  or_geom_eo_tables(g, e, o);
  double* u = dalloc(ng);
  or_gauge_expH(g, 0.5, 7, u);
  or_heatbath_sweep(g, u, 5.8, 3, 0);
  for (long l = 0; l < 4 * V; ++l) or_reunit(u + 18 * l);
  or_apbc(g, u);
  const double kappa = 0.125;
  double *f = dalloc(nh), *w = dalloc(nh), *w2 = dalloc(nh), *full = dalloc(nf), *fw = dalloc(nf);
  or_gauss_field(nh, 11, f);
  or_gauss_field(nf, 12, full);
  or_gamma5(Vh, w, f);
  or_hop_eo(g, u, w, f, 1, 0, kappa);
  or_hop_eo(g, u, w2, f, 2, 1, kappa);
  or_dirac_eo(g, u, w, f, 0, kappa);
  or_dirac_eo(g, u, w2, f, 1, kappa);
  or_ddagd(g, u, w, f, kappa);
  or_dirac_full(g, u, fw, full, 0, kappa);
  or_dirac_full(g, u, fw, full, 1, kappa);
  /* CG and its identity */
  double *x = dalloc(nh), diff = 0.0;
  int its = 0;
  if (or_cg(g, u, kappa, x, f, 1e-24, 5000, &its, &diff) != 0) return 1;
  or_ddagd(g, u, w, x, kappa);
  double rn = 0.0, bn = 0.0;
  for (long i = 0; i < nh; ++i) {
    rn += (w[i] - f[i]) * (w[i] - f[i]);
    bn += f[i] * f[i];
  }
  if (!(sqrt(rn / bn) < 1e-10)) return 2;
  /* multishift */
  const double sh[3] = {0.0, 0.05, 0.3};
  double* xs = dalloc(3 * nh);
  int nconv = 0;
  double md = 0.0, od = 0.0;
  if (or_multishift_cg(g, u, kappa, sh, 3, xs, f, 1e-22, 5000, &nconv, &md, &od) != 0) return 4;
  /* inverter */
  double* sol = dalloc(nf);
  if (or_eo_invert(g, u, kappa, sol, full, 1e-26, 5000, &its, &diff) != 0) return 5;
  if (!(or_true_residual_full(g, u, kappa, sol, full) < 1e-9)) return 6;
  /* MD force */
  double *fce = dalloc(ng), *ta = dalloc(ng), *eta = dalloc(nf), *zeta = dalloc(nf);
  or_gauss_field(nf, 13, eta);
  or_gauss_field(nf, 14, zeta);
  or_md_force_full(g, u, fce, eta, zeta, kappa);
  or_md_force_pm_full(g, u, fce, eta, zeta, 0);
  or_md_force_pm_full(g, u, fce, eta, zeta, 1);
  or_md_force_eo(g, u, fce, f, w2, kappa);
  or_traceless_antihermite(4 * V, ta, fce);
  /* learned layers: 1 -> 8 -> 1 */
  const int cin = 1, cmid = 8;
  double *w1 = dalloc(9 * cmid * cin * 32), *w4 = dalloc(9 * cin * cmid * 32);
  or_gauss_field(9 * cmid * cin * 32, 15, w1);
  or_gauss_field(9 * cin * cmid * 32, 16, w4);
  double *a0 = dalloc(nf), *a1 = dalloc((long)cmid * nf), *a2 = dalloc(nf), *g1 = dalloc((long)cmid * nf);
  double* gw = dalloc(9 * cmid * cin * 32);
  memcpy(a0, full, sizeof(double) * nf);
  for (int rb = 0; rb < 2; ++rb) {
    or_net_layer(g, u, cin, cmid, w1, a0, a1, rb);
    or_net_layer(g, u, cmid, cin, w4, a1, a2, rb);
    or_net_layer_adj(g, u, cmid, cin, w4, a2, g1, rb);
    or_net_layer_wgrad(g, u, cin, cmid, a0, g1, gw, rb);
  }
  const double p = or_plaquette(g, u);
  printf("dims %dx%dx%dx%d: cg %d iterations, multishift Nconv %d, plaquette %.6f, u %.3f\n",
         dims[0], dims[1], dims[2], dims[3], its, nconv, p, or_uniform(1, 2, 3));
  free(e); free(o); free(u); free(f); free(w); free(w2); free(full); free(fw); free(x); free(xs);
  free(sol); free(fce); free(ta); free(eta); free(zeta); free(w1); free(w4); free(a0); free(a1);
  free(a2); free(g1); free(gw);
  or_geom_free(g);
  return 0;
}

int main(void) {
  const int d1[4] = {4, 4, 4, 4}, d2[4] = {6, 2, 2, 6}, d3[4] = {4, 6, 2, 4};
  int rc = run(d1);
  if (!rc) rc = run(d2);
  if (!rc) rc = run(d3);
  if (rc) fprintf(stderr, "oracle_sanitize: check %d failed\n", rc);
  return rc;
}
