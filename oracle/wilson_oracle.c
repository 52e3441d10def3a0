/*
 * wilson_oracle.c -- CPU restatement of the IroIro++ Wilson even/odd solve
 * path.  TEST INFRASTRUCTURE ONLY (the checker of the HIP path); see
 * wilson_oracle.h for who may call it and for the layouts.
 *
 * Written from the reference's algorithm, not from its code structure: the
 * eight directional kernels of dirac_wilson_improved.code become one loop
 * over directions with the projection/reconstruction tables of the Dirac
 * basis of wilsonLikeUtils.cpp:9-49.  The summation order of the reference
 * (directions x+,x-,y+,y-,z+,z-,t+,t-, then the -kappa scale;
 * dirac_wilson.cpp:107-113) and the expression shapes of its SU(3) x
 * half-spinor products are kept so that results agree with the compiled
 * reference to the last bits on x86-64 without FMA contraction.
 */
#include "wilson_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct or_geom {
  int L[4];
  long V, Vh;
  int* esec;     /* half index -> lexicographic site, even parity */
  int* osec;     /* odd parity */
  int* hidx;     /* lexicographic site -> half index in its parity */
  int* par;      /* lexicographic site -> parity */
  int* nb_fw;    /* [mu][site] lexicographic forward neighbour (periodic) */
  int* nb_bw;    /* [mu][site] backward neighbour */
};

/* ------------------------------------------------------------------ */
/* geometry: siteIndex.cpp:41 (lexicographic), siteIndex_EvenOdd.cpp:21-62
 * (esec/osec ascending), single-process periodic wrap as in
 * communicator_single.cpp:45-66. */
or_geom* or_geom_new(const int dims[4]) {
  or_geom* g = (or_geom*)calloc(1, sizeof(or_geom));
  if (!g) return NULL;
  for (int m = 0; m < 4; ++m) g->L[m] = dims[m];
  g->V = (long)dims[0] * dims[1] * dims[2] * dims[3];
  g->Vh = g->V / 2;
  g->esec = (int*)malloc(sizeof(int) * g->Vh);
  g->osec = (int*)malloc(sizeof(int) * g->Vh);
  g->hidx = (int*)malloc(sizeof(int) * g->V);
  g->par = (int*)malloc(sizeof(int) * g->V);
  g->nb_fw = (int*)malloc(sizeof(int) * 4 * g->V);
  g->nb_bw = (int*)malloc(sizeof(int) * 4 * g->V);
  long ne = 0, no = 0;
  for (long s = 0; s < g->V; ++s) {
    int c[4];
    long r = s;
    for (int m = 0; m < 4; ++m) { c[m] = (int)(r % dims[m]); r /= dims[m]; }
    int p = (c[0] + c[1] + c[2] + c[3]) % 2;
    g->par[s] = p;
    if (p) { g->hidx[s] = (int)no; g->osec[no++] = (int)s; }
    else   { g->hidx[s] = (int)ne; g->esec[ne++] = (int)s; }
    long stride = 1;
    for (int m = 0; m < 4; ++m) {
      int up = (c[m] + 1) % dims[m];
      int dn = (c[m] + dims[m] - 1) % dims[m];
      g->nb_fw[m * g->V + s] = (int)(s + (long)(up - c[m]) * stride);
      g->nb_bw[m * g->V + s] = (int)(s + (long)(dn - c[m]) * stride);
      stride *= dims[m];
    }
  }
  return g;
}

void or_geom_free(or_geom* g) {
  if (!g) return;
  free(g->esec); free(g->osec); free(g->hidx); free(g->par);
  free(g->nb_fw); free(g->nb_bw);
  free(g);
}

long or_geom_volume(const or_geom* g) { return g->V; }

void or_geom_eo_tables(const or_geom* g, int* esec, int* osec) {
  memcpy(esec, g->esec, sizeof(int) * g->Vh);
  memcpy(osec, g->osec, sizeof(int) * g->Vh);
}

/* antiPeriodicBC.hpp:25-37: U_t(x) -> -U_t(x) on the slice t = Lt-1. */
void or_apbc(const or_geom* g, double* u) {
  long slice = (long)g->L[0] * g->L[1] * g->L[2];
  long s0 = slice * (g->L[3] - 1);
  for (long s = s0; s < s0 + slice; ++s) {
    double* m = u + 18 * (s + g->V * 3);
    for (int i = 0; i < 18; ++i) m[i] = -m[i];
  }
}

/* wilsonLikeUtils.cpp:42-49 */
void or_gamma5(long nsites, double* w, const double* f) {
  for (long s = 0; s < nsites; ++s) {
    const double* a = f + 24 * s;
    double* b = w + 24 * s;
    double tmp[24];
    memcpy(tmp, a + 12, 12 * sizeof(double));
    memcpy(tmp + 12, a, 12 * sizeof(double));
    memcpy(b, tmp, 24 * sizeof(double));
  }
}

/* ------------------------------------------------------------------ */
/* One direction of the Wilson hop at one output site.
 *
 * Spin projection (1 -/+ gamma_mu) in the Dirac basis of
 * wilsonLikeUtils.cpp:9-40 reduces the neighbour spinor v to two colour
 * vectors w1, w2 (tables from dirac_wilson_improved.code):
 *   x+: w1 = v0 + i v3, w2 = v1 + i v2     x-: w1 = v0 - i v3, w2 = v1 - i v2
 *   y+: w1 = v0 + v3,   w2 = v1 - v2       y-: w1 = v0 - v3,   w2 = v1 + v2
 *   z+: w1 = v0 + i v2, w2 = v1 - i v3     z-: w1 = v0 - i v2, w2 = v1 + i v3
 *   t+: w1 = 2 v2,      w2 = 2 v3          t-: w1 = 2 v0,      w2 = 2 v1
 * h = U w (forward, link at the output site) or U^dag w (backward, link at
 * the neighbour), then spin reconstruction adds into res:
 *   x+: r0+=h1 r1+=h2 r2+=-i h2 r3+=-i h1   x-: r2+=i h2  r3+=i h1
 *   y+: r2+=-h2 r3+=h1                      y-: r2+=h2    r3+=-h1
 *   z+: r2+=-i h1 r3+=i h2                  z-: r2+=i h1  r3+=-i h2
 *   t+: r2+=h1 r3+=h2 (r0,r1 untouched)     t-: r0+=h1 r1+=h2
 */
#define RE(s, c) (2 * (3 * (s) + (c)))
#define IM(s, c) (2 * (3 * (s) + (c)) + 1)

static void hop_dir(double* res, const double* v, const double* U, int mu,
                    int forward) {
  double w1r[3], w1i[3], w2r[3], w2i[3];
  double h1r[3], h1i[3], h2r[3], h2i[3];
  for (int c = 0; c < 3; ++c) {
    switch (mu * 2 + (forward ? 0 : 1)) {
      case 0: /* x+ */
        w1r[c] = v[RE(0,c)] - v[IM(3,c)]; w1i[c] = v[IM(0,c)] + v[RE(3,c)];
        w2r[c] = v[RE(1,c)] - v[IM(2,c)]; w2i[c] = v[IM(1,c)] + v[RE(2,c)];
        break;
      case 1: /* x- */
        w1r[c] = v[RE(0,c)] + v[IM(3,c)]; w1i[c] = v[IM(0,c)] - v[RE(3,c)];
        w2r[c] = v[RE(1,c)] + v[IM(2,c)]; w2i[c] = v[IM(1,c)] - v[RE(2,c)];
        break;
      case 2: /* y+ */
        w1r[c] = v[RE(0,c)] + v[RE(3,c)]; w1i[c] = v[IM(0,c)] + v[IM(3,c)];
        w2r[c] = v[RE(1,c)] - v[RE(2,c)]; w2i[c] = v[IM(1,c)] - v[IM(2,c)];
        break;
      case 3: /* y- */
        w1r[c] = v[RE(0,c)] - v[RE(3,c)]; w1i[c] = v[IM(0,c)] - v[IM(3,c)];
        w2r[c] = v[RE(1,c)] + v[RE(2,c)]; w2i[c] = v[IM(1,c)] + v[IM(2,c)];
        break;
      case 4: /* z+ */
        w1r[c] = v[RE(0,c)] - v[IM(2,c)]; w1i[c] = v[IM(0,c)] + v[RE(2,c)];
        w2r[c] = v[RE(1,c)] + v[IM(3,c)]; w2i[c] = v[IM(1,c)] - v[RE(3,c)];
        break;
      case 5: /* z- */
        w1r[c] = v[RE(0,c)] + v[IM(2,c)]; w1i[c] = v[IM(0,c)] - v[RE(2,c)];
        w2r[c] = v[RE(1,c)] - v[IM(3,c)]; w2i[c] = v[IM(1,c)] + v[RE(3,c)];
        break;
      case 6: /* t+ */
        w1r[c] = v[RE(2,c)] * 2.0; w1i[c] = v[IM(2,c)] * 2.0;
        w2r[c] = v[RE(3,c)] * 2.0; w2i[c] = v[IM(3,c)] * 2.0;
        break;
      default: /* t- */
        w1r[c] = v[RE(0,c)] * 2.0; w1i[c] = v[IM(0,c)] * 2.0;
        w2r[c] = v[RE(1,c)] * 2.0; w2i[c] = v[IM(1,c)] * 2.0;
        break;
    }
  }
  /* colour multiply; element (a,b) of U is U[2*(3a+b)] (+1 imaginary) */
  for (int c = 0; c < 3; ++c) {
    h1r[c] = 0.0; h1i[c] = 0.0; h2r[c] = 0.0; h2i[c] = 0.0;
    for (int c1 = 0; c1 < 3; ++c1) {
      if (forward) {
        double ur = U[2 * (3 * c + c1)], ui = U[2 * (3 * c + c1) + 1];
        h1r[c] += ur * w1r[c1] - ui * w1i[c1];
        h1i[c] += ui * w1r[c1] + ur * w1i[c1];
        h2r[c] += ur * w2r[c1] - ui * w2i[c1];
        h2i[c] += ui * w2r[c1] + ur * w2i[c1];
      } else { /* conj(U)^T */
        double ur = U[2 * (3 * c1 + c)], ui = U[2 * (3 * c1 + c) + 1];
        h1r[c] += ur * w1r[c1] + ui * w1i[c1];
        h1i[c] -= ui * w1r[c1] - ur * w1i[c1];
        h2r[c] += ur * w2r[c1] + ui * w2i[c1];
        h2i[c] -= ui * w2r[c1] - ur * w2i[c1];
      }
    }
  }
  for (int c = 0; c < 3; ++c) {
    int k = mu * 2 + (forward ? 0 : 1);
    if (k != 6) { /* spin rows 0,1 get h1,h2 except for t+ */
      res[RE(0,c)] += h1r[c]; res[IM(0,c)] += h1i[c];
      res[RE(1,c)] += h2r[c]; res[IM(1,c)] += h2i[c];
    }
    switch (k) {
      case 0: /* r2 += -i h2, r3 += -i h1 */
        res[RE(2,c)] += h2i[c]; res[IM(2,c)] -= h2r[c];
        res[RE(3,c)] += h1i[c]; res[IM(3,c)] -= h1r[c];
        break;
      case 1: /* r2 += i h2, r3 += i h1 */
        res[RE(2,c)] -= h2i[c]; res[IM(2,c)] += h2r[c];
        res[RE(3,c)] -= h1i[c]; res[IM(3,c)] += h1r[c];
        break;
      case 2: /* r2 -= h2, r3 += h1 */
        res[RE(2,c)] -= h2r[c]; res[IM(2,c)] -= h2i[c];
        res[RE(3,c)] += h1r[c]; res[IM(3,c)] += h1i[c];
        break;
      case 3: /* r2 += h2, r3 -= h1 */
        res[RE(2,c)] += h2r[c]; res[IM(2,c)] += h2i[c];
        res[RE(3,c)] -= h1r[c]; res[IM(3,c)] -= h1i[c];
        break;
      case 4: /* r2 += -i h1, r3 += i h2 */
        res[RE(2,c)] += h1i[c]; res[IM(2,c)] -= h1r[c];
        res[RE(3,c)] -= h2i[c]; res[IM(3,c)] += h2r[c];
        break;
      case 5: /* r2 += i h1, r3 += -i h2 */
        res[RE(2,c)] -= h1i[c]; res[IM(2,c)] += h1r[c];
        res[RE(3,c)] += h2i[c]; res[IM(3,c)] -= h2r[c];
        break;
      case 6: /* t+: r2 += h1, r3 += h2 */
        res[RE(2,c)] += h1r[c]; res[IM(2,c)] += h1i[c];
        res[RE(3,c)] += h2r[c]; res[IM(3,c)] += h2i[c];
        break;
      default: /* t-: only spin rows 0,1 */
        break;
    }
  }
}

/* Generic hop over a set of output sites.  out_sites[i] is the lexicographic
 * site of output i; the input value at lexicographic site s is read from
 * f + 24 * in_index[s]. */
static void hop_core(const or_geom* g, const double* u, double* w,
                     const double* f, long nout, const int* out_sites,
                     const int* in_index, double kappa) {
  for (long i = 0; i < nout; ++i) {
    int s = out_sites[i];
    double* res = w + 24 * i;
    for (int k = 0; k < 24; ++k) res[k] = 0.0;
    for (int mu = 0; mu < 4; ++mu) {
      int sp = g->nb_fw[mu * g->V + s];
      int sm = g->nb_bw[mu * g->V + s];
      hop_dir(res, f + 24 * (long)in_index[sp], u + 18 * (s + g->V * mu), mu, 1);
      hop_dir(res, f + 24 * (long)in_index[sm], u + 18 * (sm + g->V * mu), mu, 0);
    }
    for (int k = 0; k < 24; ++k) res[k] *= -kappa;
  }
}

void or_hop_eo(const or_geom* g, const double* u, double* w, const double* f,
               int eotype, int dag, double kappa) {
  const int* out = (eotype == 1) ? g->esec : g->osec;
  if (dag) {
    double* f5 = (double*)malloc(sizeof(double) * 24 * g->Vh);
    or_gamma5(g->Vh, f5, f);
    hop_core(g, u, w, f5, g->Vh, out, g->hidx, kappa);
    or_gamma5(g->Vh, w, w);
    free(f5);
  } else {
    hop_core(g, u, w, f, g->Vh, out, g->hidx, kappa);
  }
}

/* dirac_wilson_EvenOdd.hpp:79-92 */
void or_dirac_eo(const or_geom* g, const double* u, double* w, const double* f,
                 int dag, double kappa) {
  long n = 24 * g->Vh;
  double* t1 = (double*)malloc(sizeof(double) * n);
  double* t2 = (double*)malloc(sizeof(double) * n);
  if (!dag) {
    or_hop_eo(g, u, t1, f, 2, 0, kappa);   /* Doe f  */
    or_hop_eo(g, u, t2, t1, 1, 0, kappa);  /* Deo Doe f */
  } else {
    /* mult_dag(f) = f - mult_oe_dag(mult_eo_dag(f)) = f - Deo^dag(Doe^dag f) */
    or_hop_eo(g, u, t1, f, 2, 1, kappa);
    or_hop_eo(g, u, t2, t1, 1, 1, kappa);
  }
  for (long i = 0; i < n; ++i) w[i] = f[i] - t2[i];
  free(t1); free(t2);
}

void or_ddagd(const or_geom* g, const double* u, double* w, const double* f,
              double kappa) {
  double* t = (double*)malloc(sizeof(double) * 24 * g->Vh);
  or_dirac_eo(g, u, t, f, 0, kappa);
  or_dirac_eo(g, u, w, t, 1, kappa);
  free(t);
}

/* mult_full (dirac_wilson.cpp:115-118) on the lexicographic volume. */
void or_dirac_full(const or_geom* g, const double* u, double* w,
                   const double* f, int dag, double kappa) {
  long n = 24 * g->V;
  int* ident = (int*)malloc(sizeof(int) * g->V);
  for (long s = 0; s < g->V; ++s) ident[s] = (int)s;
  const double* src = f;
  double* f5 = NULL;
  if (dag) {
    f5 = (double*)malloc(sizeof(double) * n);
    or_gamma5(g->V, f5, f);
    src = f5;
  }
  double* t = (double*)malloc(sizeof(double) * n);
  hop_core(g, u, t, src, g->V, ident, ident, kappa);
  for (long i = 0; i < n; ++i) t[i] += src[i];
  if (dag) or_gamma5(g->V, w, t);
  else memcpy(w, t, sizeof(double) * n);
  free(t); free(f5); free(ident);
}

/* ------------------------------------------------------------------ */
/* MD force.  Dirac_Wilson::mkfrc (dirac_wilson.cpp:62-92): for direction
 * mu, xie = mult_p[mu](eta) is the forward hop term U_mu(x)(1-gamma_mu)
 * eta(x+mu) at the operator's output sites x (no -kappa), and
 *   fce(x,mu)_{ab} += sum_s xie(a,s,x) conj(zeta(b,s,x)),
 * spelled out in real arithmetic as the reference does it.  out_sites[i]
 * is the lexicographic site of output i (the operator's gp_ list) and
 * in_index maps a lexicographic site to its row in eta. */
static void mkfrc(const or_geom* g, const double* u, double* fce,
                  const double* eta, const double* zeta, int mu, long nout,
                  const int* out_sites, const int* in_index) {
  for (long i = 0; i < nout; ++i) {
    int s = out_sites[i];
    int sp = g->nb_fw[mu * g->V + s];
    double xie[24];
    for (int k = 0; k < 24; ++k) xie[k] = 0.0;
    hop_dir(xie, eta + 24 * (long)in_index[sp], u + 18 * (s + g->V * mu), mu, 1);
    const double* z = zeta + 24 * i;
    double* f = fce + 18 * ((long)s + (long)g->V * mu);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int sp4 = 0; sp4 < 4; ++sp4) {
          f[2 * (3 * a + b)] += z[RE(sp4, b)] * xie[RE(sp4, a)] + z[IM(sp4, b)] * xie[IM(sp4, a)];
          f[2 * (3 * a + b) + 1] += z[RE(sp4, b)] * xie[IM(sp4, a)] - z[IM(sp4, b)] * xie[RE(sp4, a)];
        }
  }
}

/* md_force_p / md_force_m of one operator (dirac_wilson.cpp:41-50). */
static void md_force_pm(const or_geom* g, const double* u, double* fce,
                        const double* eta, const double* zeta, long nsites,
                        const int* out_sites, const int* in_index, int minus) {
  if (!minus) {
    for (int mu = 0; mu < 4; ++mu)
      mkfrc(g, u, fce, eta, zeta, mu, nsites, out_sites, in_index);
    return;
  }
  double* zt5 = (double*)malloc(sizeof(double) * 24 * nsites);
  double* et5 = (double*)malloc(sizeof(double) * 24 * nsites);
  or_gamma5(nsites, zt5, zeta);
  or_gamma5(nsites, et5, eta);
  for (int mu = 0; mu < 4; ++mu)
    mkfrc(g, u, fce, zt5, et5, mu, nsites, out_sites, in_index);
  free(zt5);
  free(et5);
}

void or_md_force_full(const or_geom* g, const double* u, double* fce,
                      const double* eta, const double* zeta, double kappa) {
  long ng = 18L * 4 * g->V;
  int* ident = (int*)malloc(sizeof(int) * g->V);
  for (long s = 0; s < g->V; ++s) ident[s] = (int)s;
  for (long i = 0; i < ng; ++i) fce[i] = 0.0;
  md_force_pm(g, u, fce, eta, zeta, g->V, ident, ident, 0);
  md_force_pm(g, u, fce, eta, zeta, g->V, ident, ident, 1);
  for (long i = 0; i < ng; ++i) fce[i] *= -kappa;
  free(ident);
}

void or_md_force_pm_full(const or_geom* g, const double* u, double* fce,
                         const double* eta, const double* zeta, int minus) {
  int* ident = (int*)malloc(sizeof(int) * g->V);
  for (long s = 0; s < g->V; ++s) ident[s] = (int)s;
  md_force_pm(g, u, fce, eta, zeta, g->V, ident, ident, minus);
  free(ident);
}

/* Dirac_Wilson_EvenOdd::md_force (dirac_wilson_EvenOdd.hpp:94-116):
 *   md_force_eo(fce, Doe eta, zeta)      = Deo.md_force_p + Doe.md_force_m
 *   md_force_oe(fce, eta, Deo^dag zeta)  = Doe.md_force_p + Deo.md_force_m
 *   fce *= kappa.
 * Deo has output sites esec, Doe osec (dirac_wilson.hpp:89-114). */
void or_md_force_eo(const or_geom* g, const double* u, double* fce,
                    const double* eta, const double* zeta, double kappa) {
  long ng = 18L * 4 * g->V;
  long nh = 24L * g->Vh;
  double* eta_o = (double*)malloc(sizeof(double) * nh);
  double* zeta_o = (double*)malloc(sizeof(double) * nh);
  or_hop_eo(g, u, eta_o, eta, 2, 0, kappa);   /* mult_oe(eta) = Doe.mult   */
  or_hop_eo(g, u, zeta_o, zeta, 2, 1, kappa); /* mult_eo_dag = Doe.mult_dag */
  for (long i = 0; i < ng; ++i) fce[i] = 0.0;
  md_force_pm(g, u, fce, eta_o, zeta, g->Vh, g->esec, g->hidx, 0);  /* Deo p */
  md_force_pm(g, u, fce, eta_o, zeta, g->Vh, g->osec, g->hidx, 1);  /* Doe m */
  md_force_pm(g, u, fce, eta, zeta_o, g->Vh, g->osec, g->hidx, 0);  /* Doe p */
  md_force_pm(g, u, fce, eta, zeta_o, g->Vh, g->esec, g->hidx, 1);  /* Deo m */
  for (long i = 0; i < ng; ++i) fce[i] *= kappa;
  free(eta_o);
  free(zeta_o);
}

/* FieldUtils::TracelessAntihermite (fieldUtils.cpp:264-274) with
 * SUNmatUtils::anti_hermite_traceless (sunMatUtils.cpp:50-60) and
 * SUNmatrix::anti_hermite (sunMat.hpp:139-157): per link
 *   out = (m - m^dag)/2 - i Im tr(m)/3. */
void or_traceless_antihermite(long nlinks, double* out, const double* in) {
  for (long l = 0; l < nlinks; ++l) {
    const double* m = in + 18 * l;
    double* o = out + 18 * l;
    double imtr = (m[1] + m[9] + m[17]) / 3.0;
    double t[18];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        int ab = 2 * (3 * a + b), ba = 2 * (3 * b + a);
        t[ab] = 0.5 * (m[ab] - m[ba]);
        t[ab + 1] = 0.5 * (m[ab + 1] + m[ba + 1]) - (a == b ? imtr : 0.0);
      }
    memcpy(o, t, sizeof(t));
  }
}

/* ------------------------------------------------------------------ */
/* BLAS-1 with the reference's semantics: Field::operator* = sum over all
 * reals (field.inl:58-62, sequential valarray sum), norm = sqrt(sum a^2). */
static double dot(const double* a, const double* b, long n) {
  double s = 0.0;
  for (long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* solver_CG.cpp:29-160 */
int or_cg(const or_geom* g, const double* u, double kappa, double* x,
          const double* b, double prec, int maxiter, int* iterations,
          double* diff) {
  long n = 24 * g->Vh;
  double* r = (double*)malloc(sizeof(double) * n);
  double* p = (double*)malloc(sizeof(double) * n);
  double* s = (double*)malloc(sizeof(double) * n);
  memcpy(x, b, sizeof(double) * n);              /* x0 = b */
  or_ddagd(g, u, s, x, kappa);
  for (long i = 0; i < n; ++i) r[i] = b[i] - s[i];
  memcpy(p, r, sizeof(double) * n);
  double rr = dot(r, r, n);
  double snorm = sqrt(dot(b, b, n));
  snorm = 1.0 / (snorm * snorm);
  int iters = -1;
  for (int it = 0; it < maxiter; ++it) {
    or_ddagd(g, u, s, p, kappa);
    double pap = dot(p, s, n);
    double rrp = rr;
    double cr = rrp / pap;
    for (long i = 0; i < n; ++i) {
      x[i] += cr * p[i];
      r[i] -= cr * s[i];
    }
    rr = dot(r, r, n);
    cr = rr / rrp;
    for (long i = 0; i < n; ++i) p[i] = p[i] * cr + r[i];
    if (rr * snorm < prec) { iters = it; break; }
  }
  or_ddagd(g, u, s, x, kappa);
  for (long i = 0; i < n; ++i) s[i] -= b[i];
  *diff = sqrt(dot(s, s, n)) * sqrt(snorm);
  *iterations = iters;
  free(r); free(p); free(s);
  return iters < 0 ? 1 : 0;
}

/* multiShiftSolver_CG.cpp:10-185 */
int or_multishift_cg(const or_geom* g, const double* u, double kappa,
                     const double* sigma, int nshift, double* x,
                     const double* b, double prec, int maxiter, int* nconv,
                     double* diff, double* out_diff) {
  long n = 24 * g->Vh;
  double* r = (double*)malloc(sizeof(double) * n);
  double* s = (double*)malloc(sizeof(double) * n);
  double* p = (double*)malloc(sizeof(double) * n * nshift);
  double* zeta1 = (double*)malloc(sizeof(double) * nshift);
  double* zeta2 = (double*)malloc(sizeof(double) * nshift);
  double* csh2 = (double*)malloc(sizeof(double) * nshift);
  double* pp = (double*)calloc(nshift, sizeof(double));
  double bnorm = sqrt(dot(b, b, n));
  double snorm = 1.0 / (bnorm * bnorm);
  for (int i = 0; i < nshift; ++i) {
    zeta1[i] = 1.0; zeta2[i] = 1.0;
    csh2[i] = sigma[i] - sigma[0];
    memcpy(p + i * n, b, sizeof(double) * n);
    memset(x + i * n, 0, sizeof(double) * n);
  }
  memcpy(r, b, sizeof(double) * n);
  double rr = dot(r, r, n);
  double alphap = 0.0, betap = 1.0;
  int Nshift2 = nshift;
  int conv = -1;
  for (int it = 0; it < maxiter; ++it) {
    double* p0 = p;
    or_ddagd(g, u, s, p0, kappa);
    for (long i = 0; i < n; ++i) s[i] += sigma[0] * p0[i];
    double rrp = rr;
    double pap = dot(s, p0, n);
    double beta = -rrp / pap;
    for (long i = 0; i < n; ++i) x[i] -= beta * p0[i];
    for (long i = 0; i < n; ++i) r[i] += beta * s[i];
    rr = dot(r, r, n);
    double alpha = rr / rrp;
    for (long i = 0; i < n; ++i) p0[i] = p0[i] * alpha + r[i];
    pp[0] = rr;
    double alphah = 1.0 + alphap * beta / betap;
    for (int ish = 1; ish < Nshift2; ++ish) {
      double zeta = (alphah - csh2[ish] * beta) / zeta1[ish]
                  + (1.0 - alphah) / zeta2[ish];
      zeta = 1.0 / zeta;
      double zr = zeta / zeta1[ish];
      double betas = beta * zr;
      double alphas = alpha * zr * zr;
      double* xi = x + ish * n;
      double* pi = p + ish * n;
      for (long i = 0; i < n; ++i) xi[i] -= betas * pi[i];
      for (long i = 0; i < n; ++i) pi[i] = pi[i] * alphas + zeta * r[i];
      pp[ish] = dot(pi, pi, n) * snorm;
      zeta2[ish] = zeta1[ish];
      zeta1[ish] = zeta;
    }
    for (int ish = Nshift2 - 1; ish >= 0; --ish) {
      if (pp[ish] > prec) { Nshift2 = ish + 1; break; }
    }
    alphap = alpha;
    betap = beta;
    if (rr * snorm < prec) { conv = it; break; }
  }
  double dmax = -1.0;
  for (int i = 0; i < nshift; ++i) {
    double* xi = x + i * n;
    or_ddagd(g, u, s, xi, kappa);
    for (long k = 0; k < n; ++k) s[k] += sigma[i] * xi[k];
    for (long k = 0; k < n; ++k) s[k] -= b[k];
    double d1 = dot(s, s, n) * snorm;
    if (d1 > dmax) dmax = d1;
  }
  *nconv = conv;
  *diff = dmax;
  *out_diff = sqrt(dmax);
  free(r); free(s); free(p); free(zeta1); free(zeta2); free(csh2); free(pp);
  return conv < 0 ? 1 : 0;
}

/* eoUtils.cpp:43-80 */
int or_eo_invert(const or_geom* g, const double* u, double kappa,
                 double* sol, const double* src, double prec, int maxiter,
                 int* iterations, double* diff) {
  long nh = 24 * g->Vh;
  double* be = (double*)malloc(sizeof(double) * nh);
  double* bo = (double*)malloc(sizeof(double) * nh);
  double* t = (double*)malloc(sizeof(double) * nh);
  double* ye = (double*)malloc(sizeof(double) * nh);
  for (long h = 0; h < g->Vh; ++h) {
    memcpy(be + 24 * h, src + 24 * (long)g->esec[h], 24 * sizeof(double));
    memcpy(bo + 24 * h, src + 24 * (long)g->osec[h], 24 * sizeof(double));
  }
  or_hop_eo(g, u, t, bo, 1, 0, kappa);           /* mult_eo(bo) */
  for (long i = 0; i < nh; ++i) be[i] -= t[i];
  or_dirac_eo(g, u, t, be, 1, kappa);            /* mult_dag(be) */
  int rc = or_cg(g, u, kappa, ye, t, prec, maxiter, iterations, diff);
  or_hop_eo(g, u, t, ye, 2, 0, kappa);           /* mult_oe(ye) */
  for (long i = 0; i < nh; ++i) bo[i] -= t[i];
  for (long h = 0; h < g->Vh; ++h) {
    memcpy(sol + 24 * (long)g->esec[h], ye + 24 * h, 24 * sizeof(double));
    memcpy(sol + 24 * (long)g->osec[h], bo + 24 * h, 24 * sizeof(double));
  }
  free(be); free(bo); free(t); free(ye);
  return rc;
}

double or_true_residual_full(const or_geom* g, const double* u, double kappa,
                             const double* x, const double* b) {
  long n = 24 * g->V;
  double* t = (double*)malloc(sizeof(double) * n);
  or_dirac_full(g, u, t, x, 0, kappa);
  double rn = 0.0, bn = 0.0;
  for (long i = 0; i < n; ++i) {
    double d = b[i] - t[i];
    rn += d * d;
    bn += b[i] * b[i];
  }
  free(t);
  return sqrt(rn / bn);
}

/* ------------------------------------------------------------------ */
/* Counter-based random numbers (SplitMix64 finaliser).  Identical integer
 * arithmetic is used by the GPU heat-bath. */
static uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double or_uniform(uint64_t seed, uint64_t stream, uint64_t counter) {
  uint64_t k = mix64(seed ^ mix64(stream));
  uint64_t z = mix64(k + counter * 0xD1B54A32D192ED03ULL);
  return (double)((z >> 11) + 1) * (1.0 / 9007199254740992.0);
}

static double gauss(uint64_t seed, uint64_t stream, uint64_t c) {
  /* Box-Muller on two counter-based deviates: unit variance */
  double u1 = or_uniform(seed, stream, 2 * c);
  double u2 = or_uniform(seed, stream, 2 * c + 1);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
}

void or_gauss_field(long ndouble, uint64_t seed, double* f) {
  for (long i = 0; i < ndouble; ++i) f[i] = gauss(seed, (uint64_t)i, 0);
}

/* sunMat.hpp:244-277: row-wise Gram-Schmidt. */
void or_reunit(double* m) {
  for (int a = 0; a < 3; ++a) {
    double nrm = 0.0;
    for (int cc = 0; cc < 6; ++cc) nrm += m[6 * a + cc] * m[6 * a + cc];
    nrm = 1.0 / sqrt(nrm);
    for (int cc = 0; cc < 6; cc += 2) {
      m[6 * a + cc] *= nrm;
      m[6 * a + cc + 1] *= nrm;
    }
    for (int b = a + 1; b < 3; ++b) {
      double prr = 0.0, pri = 0.0;
      for (int c = 0; c < 3; ++c) {
        int ac = 3 * a + c, bc = 3 * b + c;
        prr += m[2 * ac] * m[2 * bc] + m[2 * ac + 1] * m[2 * bc + 1];
        pri += m[2 * ac] * m[2 * bc + 1] - m[2 * ac + 1] * m[2 * bc];
      }
      for (int c = 0; c < 3; ++c) {
        int ac = 3 * a + c, bc = 3 * b + c;
        m[2 * bc] -= prr * m[2 * ac] - pri * m[2 * ac + 1];
        m[2 * bc + 1] -= prr * m[2 * ac + 1] + pri * m[2 * ac];
      }
    }
  }
}

/* 3x3 complex helpers (row-major, interleaved re/im) */
static void mm(double* c, const double* a, const double* b) {
  double t[18];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double sr = 0.0, si = 0.0;
      for (int k = 0; k < 3; ++k) {
        double ar = a[2 * (3 * i + k)], ai = a[2 * (3 * i + k) + 1];
        double br = b[2 * (3 * k + j)], bi = b[2 * (3 * k + j) + 1];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
      }
      t[2 * (3 * i + j)] = sr; t[2 * (3 * i + j) + 1] = si;
    }
  memcpy(c, t, sizeof(t));
}

static void dagger(double* c, const double* a) {
  double t[18];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      t[2 * (3 * i + j)] = a[2 * (3 * j + i)];
      t[2 * (3 * i + j) + 1] = -a[2 * (3 * j + i) + 1];
    }
  memcpy(c, t, sizeof(t));
}

void or_gauge_expH(const or_geom* g, double eps, uint64_t seed, double* u) {
  for (long l = 0; l < 4 * g->V; ++l) {
    /* H traceless Hermitian: diagonal real, off-diagonal complex, with
     * unit-variance Gaussian entries; trace removed. */
    double H[18];
    uint64_t st = (uint64_t)l;
    double d0 = gauss(seed, st, 0), d1 = gauss(seed, st, 1), d2 = gauss(seed, st, 2);
    double tr = (d0 + d1 + d2) / 3.0;
    memset(H, 0, sizeof(H));
    H[0] = d0 - tr; H[8] = d1 - tr; H[16] = d2 - tr;
    int k = 3;
    for (int i = 0; i < 3; ++i)
      for (int j = i + 1; j < 3; ++j) {
        double re = gauss(seed, st, k++), im = gauss(seed, st, k++);
        H[2 * (3 * i + j)] = re; H[2 * (3 * i + j) + 1] = im;
        H[2 * (3 * j + i)] = re; H[2 * (3 * j + i) + 1] = -im;
      }
    /* X = i eps H; exp by Taylor series to convergence */
    double X[18], term[18], res[18];
    for (int i = 0; i < 9; ++i) {
      X[2 * i] = -eps * H[2 * i + 1];
      X[2 * i + 1] = eps * H[2 * i];
    }
    memset(res, 0, sizeof(res));
    memset(term, 0, sizeof(term));
    for (int i = 0; i < 3; ++i) { res[8 * i] = 1.0; term[8 * i] = 1.0; }
    for (int n = 1; n < 40; ++n) {
      mm(term, term, X);
      for (int i = 0; i < 18; ++i) term[i] /= (double)n;
      for (int i = 0; i < 18; ++i) res[i] += term[i];
    }
    or_reunit(res);
    memcpy(u + 18 * l, res, sizeof(res));
  }
}

double or_plaquette(const or_geom* g, const double* u) {
  double sum = 0.0;
  for (long s = 0; s < g->V; ++s)
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu + 1; nu < 4; ++nu) {
        double a[18], b[18], c[18];
        long sp_mu = g->nb_fw[mu * g->V + s];
        long sp_nu = g->nb_fw[nu * g->V + s];
        mm(a, u + 18 * (s + g->V * mu), u + 18 * (sp_mu + g->V * nu));
        dagger(b, u + 18 * (sp_nu + g->V * mu));
        mm(a, a, b);
        dagger(c, u + 18 * (s + g->V * nu));
        mm(a, a, c);
        sum += (a[0] + a[8] + a[16]) / 3.0;
      }
  return sum / (6.0 * (double)g->V);
}

/* ------------------------------------------------------------------ */
/* SU(3) heat-bath: Cabibbo-Marinari over the subgroups (0,1),(0,2),(1,2)
 * with the Kennedy-Pendleton SU(2) generator.  Action S = beta sum_P
 * (1 - Re tr U_P / 3); for a link U with staple sum A the local weight is
 * exp((beta/3) Re tr(U A)). */
static void staple_sum(const or_geom* g, const double* u, long s, int mu,
                       double* A) {
  memset(A, 0, 18 * sizeof(double));
  long V = g->V;
  long sp = g->nb_fw[mu * V + s];
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == mu) continue;
    double t[18], d[18];
    /* upper: U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag */
    long snp = g->nb_fw[nu * V + s];
    dagger(d, u + 18 * (snp + V * mu));
    mm(t, u + 18 * (sp + V * nu), d);
    dagger(d, u + 18 * (s + V * nu));
    mm(t, t, d);
    for (int i = 0; i < 18; ++i) A[i] += t[i];
    /* lower: U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu) */
    long snm = g->nb_bw[nu * V + s];
    long spm = g->nb_bw[nu * V + sp];
    dagger(t, u + 18 * (spm + V * nu));
    dagger(d, u + 18 * (snm + V * mu));
    mm(t, t, d);
    mm(t, t, u + 18 * (snm + V * nu));
    for (int i = 0; i < 18; ++i) A[i] += t[i];
  }
}

/* Kennedy-Pendleton: quaternion X with density ~ exp(alpha X0) on SU(2).
 * Returns 0 and leaves X = (1,0,0,0) if no candidate was accepted within
 * the attempt budget (probability < 1e-300 for the couplings used). */
static void kp_su2(double alpha, uint64_t seed, uint64_t stream,
                   uint64_t base, double* X) {
  const double twopi = 6.283185307179586476925;
  X[0] = 1.0; X[1] = X[2] = X[3] = 0.0;
  for (int att = 0; att < 64; ++att) {
    uint64_t c = base + (uint64_t)att * 8;
    double r1 = or_uniform(seed, stream, c);
    double r2 = or_uniform(seed, stream, c + 1);
    double r3 = or_uniform(seed, stream, c + 2);
    double r4 = or_uniform(seed, stream, c + 3);
    double cs = cos(twopi * r2);
    double lam2 = -(log(r1) + cs * cs * log(r3)) / (2.0 * alpha);
    if (lam2 > 1.0 || r4 * r4 > 1.0 - lam2) continue;
    double x0 = 1.0 - 2.0 * lam2;
    double rad = sqrt(fmax(0.0, 1.0 - x0 * x0));
    double ct = 2.0 * or_uniform(seed, stream, c + 4) - 1.0;
    double st = sqrt(fmax(0.0, 1.0 - ct * ct));
    double ph = twopi * or_uniform(seed, stream, c + 5);
    X[0] = x0;
    X[1] = rad * st * cos(ph);
    X[2] = rad * st * sin(ph);
    X[3] = rad * ct;
    return;
  }
}

/* SU(3) projection for drift control: normalise row 0, orthogonalise and
 * normalise row 1, row 2 = conj(row0 x row1). */
static void su3_project(double* m) {
  double n0 = 0.0;
  for (int k = 0; k < 6; ++k) n0 += m[k] * m[k];
  n0 = 1.0 / sqrt(n0);
  for (int k = 0; k < 6; ++k) m[k] *= n0;
  double pr = 0.0, pi = 0.0;
  for (int c = 0; c < 3; ++c) {
    pr += m[2 * c] * m[6 + 2 * c] + m[2 * c + 1] * m[6 + 2 * c + 1];
    pi += m[2 * c] * m[6 + 2 * c + 1] - m[2 * c + 1] * m[6 + 2 * c];
  }
  for (int c = 0; c < 3; ++c) {
    m[6 + 2 * c] -= pr * m[2 * c] - pi * m[2 * c + 1];
    m[6 + 2 * c + 1] -= pr * m[2 * c + 1] + pi * m[2 * c];
  }
  double n1 = 0.0;
  for (int k = 6; k < 12; ++k) n1 += m[k] * m[k];
  n1 = 1.0 / sqrt(n1);
  for (int k = 6; k < 12; ++k) m[k] *= n1;
  for (int j = 0; j < 3; ++j) {
    int k = (j + 1) % 3, l = (j + 2) % 3;
    double ar = m[2 * k], ai = m[2 * k + 1], br = m[6 + 2 * l], bi = m[6 + 2 * l + 1];
    double cr = ar * br - ai * bi, ci = ar * bi + ai * br;
    ar = m[2 * l]; ai = m[2 * l + 1]; br = m[6 + 2 * k]; bi = m[6 + 2 * k + 1];
    cr -= ar * br - ai * bi; ci -= ar * bi + ai * br;
    m[12 + 2 * j] = cr; m[12 + 2 * j + 1] = -ci;
  }
}

void or_heatbath_link(double* U, const double* A, double beta, uint64_t seed,
                      uint64_t stream) {
  static const int sub[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  double W[18];
  mm(W, U, A);
  for (int sg = 0; sg < 3; ++sg) {
    int i = sub[sg][0], j = sub[sg][1];
    double w11r = W[2 * (3 * i + i)], w11i = W[2 * (3 * i + i) + 1];
    double w12r = W[2 * (3 * i + j)], w12i = W[2 * (3 * i + j) + 1];
    double w21r = W[2 * (3 * j + i)], w21i = W[2 * (3 * j + i) + 1];
    double w22r = W[2 * (3 * j + j)], w22i = W[2 * (3 * j + j) + 1];
    double a0 = 0.5 * (w11r + w22r);
    double a1 = 0.5 * (w12i + w21i);
    double a2 = 0.5 * (w12r - w21r);
    double a3 = 0.5 * (w11i - w22i);
    double k = sqrt(a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3);
    if (k < 1e-300) continue;
    double v0 = a0 / k, v1 = a1 / k, v2 = a2 / k, v3 = a3 / k;
    double X[4];
    kp_su2(2.0 * beta * k / 3.0, seed, stream, (uint64_t)sg * 1024, X);
    /* R = X v^dag as 2x2 complex matrices, with the quaternion (q0..q3)
     * represented as [[q0 + i q3, q2 + i q1], [-q2 + i q1, q0 - i q3]]. */
    double Xm[8] = {X[0], X[3], X[2], X[1], -X[2], X[1], X[0], -X[3]};
    double Vm[8] = {v0, v3, v2, v1, -v2, v1, v0, -v3};
    double Rm[8];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        double sr = 0.0, si = 0.0;
        for (int c = 0; c < 2; ++c) {
          /* Xm[a][c] * conj(Vm[b][c]) */
          double xr = Xm[2 * (2 * a + c)], xi = Xm[2 * (2 * a + c) + 1];
          double vr = Vm[2 * (2 * b + c)], vi = Vm[2 * (2 * b + c) + 1];
          sr += xr * vr + xi * vi;
          si += xi * vr - xr * vi;
        }
        Rm[2 * (2 * a + b)] = sr; Rm[2 * (2 * a + b) + 1] = si;
      }
    double R11r = Rm[0], R11i = Rm[1], R12r = Rm[2], R12i = Rm[3];
    double R21r = Rm[4], R21i = Rm[5], R22r = Rm[6], R22i = Rm[7];
    double* mats[2] = {U, W};
    for (int q = 0; q < 2; ++q) {
      double* M = mats[q];
      for (int c = 0; c < 3; ++c) {
        double xr = M[2 * (3 * i + c)], xi = M[2 * (3 * i + c) + 1];
        double yr = M[2 * (3 * j + c)], yi = M[2 * (3 * j + c) + 1];
        M[2 * (3 * i + c)] = R11r * xr - R11i * xi + R12r * yr - R12i * yi;
        M[2 * (3 * i + c) + 1] = R11r * xi + R11i * xr + R12r * yi + R12i * yr;
        M[2 * (3 * j + c)] = R21r * xr - R21i * xi + R22r * yr - R22i * yi;
        M[2 * (3 * j + c) + 1] = R21r * xi + R21i * xr + R22r * yi + R22i * yr;
      }
    }
  }
  su3_project(U);
}

void or_heatbath_sweep(const or_geom* g, double* u, double beta,
                       uint64_t seed, int sweep) {
  for (int mu = 0; mu < 4; ++mu)
    for (int p = 0; p < 2; ++p)
      for (long s = 0; s < g->V; ++s) {
        if (g->par[s] != p) continue;
        double A[18];
        staple_sum(g, u, s, mu, A);
        uint64_t stream = ((uint64_t)sweep * 4 + (uint64_t)mu) * (uint64_t)g->V
                        + (uint64_t)s;
        or_heatbath_link(u + 18 * (s + g->V * mu), A, beta, seed, stream);
      }
}

/* ------------------------------------------------------------------ */
/* Learned gauge-covariant layer (no reference counterpart).  Path d of a
 * site x: d = 0 identity; d = 2mu+1 reads x+mu through U_mu(x); d = 2mu+2
 * reads x-mu through U_mu(x-mu)^dag. */

/* bfloat16 round to nearest, ties to even (finite inputs). */
static double bf16_round(double v) {
  float f = (float)v;
  uint32_t b;
  memcpy(&b, &f, 4);
  b += 0x7FFFu + ((b >> 16) & 1u);
  b &= 0xFFFF0000u;
  memcpy(&f, &b, 4);
  return (double)f;
}

/* Colour transport of one 24-double spinor: r[t][a] = sum_b M[a][b] v[t][b]
 * (dag = 0) or conj(M[b][a]) v[t][b] (dag = 1). */
static void colour_mul(double* r, const double* M, const double* v, int dag) {
  for (int t = 0; t < 4; ++t)
    for (int a = 0; a < 3; ++a) {
      double re = 0.0, im = 0.0;
      for (int b = 0; b < 3; ++b) {
        double mr, mi;
        if (dag) { mr = M[2 * (3 * b + a)]; mi = -M[2 * (3 * b + a) + 1]; }
        else     { mr = M[2 * (3 * a + b)]; mi =  M[2 * (3 * a + b) + 1]; }
        double vr = v[RE(t, b)], vi = v[IM(t, b)];
        re += mr * vr - mi * vi;
        im += mr * vi + mi * vr;
      }
      r[RE(t, a)] = re;
      r[IM(t, a)] = im;
    }
}

/* Neighbour site n_d(x) and the matrix of path d (dag flag) at site s. */
static long net_path(const or_geom* g, const double* u, int d, long s,
                     const double** M, int* dag) {
  if (d == 0) { *M = NULL; *dag = 0; return s; }
  int mu = (d - 1) / 2;
  if ((d - 1) % 2 == 0) {
    *M = u + 18 * (s + g->V * mu);
    *dag = 0;
    return g->nb_fw[mu * g->V + s];
  }
  long n = g->nb_bw[mu * g->V + s];
  *M = u + 18 * (n + g->V * mu);
  *dag = 1;
  return n;
}

static void net_transport(const or_geom* g, const double* u, int d, long s,
                          const double* field, double* r, int round_bf16) {
  const double* M;
  int dag;
  long n = net_path(g, u, d, s, &M, &dag);
  if (!M) memcpy(r, field + 24 * n, 24 * sizeof(double));
  else colour_mul(r, M, field + 24 * n, dag);
  if (round_bf16)
    for (int k = 0; k < 24; ++k) r[k] = bf16_round(r[k]);
}

/* acc[s][c] += W[s][t] v[t][c]  (dag = 0)  or  conj(W[t][s]) v[t][c]  (dag = 1) */
static void spin_mul_add(double* acc, const double* W, const double* v, int dag) {
  for (int s = 0; s < 4; ++s)
    for (int t = 0; t < 4; ++t) {
      double wr, wi;
      if (dag) { wr = W[2 * (4 * t + s)]; wi = -W[2 * (4 * t + s) + 1]; }
      else     { wr = W[2 * (4 * s + t)]; wi =  W[2 * (4 * s + t) + 1]; }
      for (int c = 0; c < 3; ++c) {
        double vr = v[RE(t, c)], vi = v[IM(t, c)];
        acc[RE(s, c)] += wr * vr - wi * vi;
        acc[IM(s, c)] += wr * vi + wi * vr;
      }
    }
}

void or_net_layer(const or_geom* g, const double* u, int cin, int cout, const double* w,
                  const double* in, double* out, int round_bf16) {
  const long V = g->V;
  double tv[24];
  for (int o = 0; o < cout; ++o) memset(out + (long)o * 24 * V, 0, 24 * V * sizeof(double));
  for (long s = 0; s < V; ++s)
    for (int d = 0; d < 9; ++d)
      for (int i = 0; i < cin; ++i) {
        net_transport(g, u, d, s, in + (long)i * 24 * V, tv, round_bf16);
        for (int o = 0; o < cout; ++o)
          spin_mul_add(out + (long)o * 24 * V + 24 * s,
                       w + 32 * ((long)(d * cout + o) * cin + i), tv, 0);
      }
}

void or_net_layer_adj(const or_geom* g, const double* u, int cin, int cout, const double* w,
                      const double* out_grad, double* in_grad, int round_bf16) {
  const long V = g->V;
  double tv[24];
  for (int i = 0; i < cin; ++i) memset(in_grad + (long)i * 24 * V, 0, 24 * V * sizeof(double));
  for (long s = 0; s < V; ++s)
    for (int d = 0; d < 9; ++d) {
      const double* M;
      int dag;
      long n = net_path(g, u, d, s, &M, &dag);
      for (int o = 0; o < cout; ++o) {
        const double* gx = out_grad + (long)o * 24 * V + 24 * s;
        /* transpose of "read n through M": carry g(x) back to n through M^dag */
        if (!M) memcpy(tv, gx, sizeof tv);
        else colour_mul(tv, M, gx, !dag);
        if (round_bf16)
          for (int k = 0; k < 24; ++k) tv[k] = bf16_round(tv[k]);
        for (int i = 0; i < cin; ++i)
          spin_mul_add(in_grad + (long)i * 24 * V + 24 * n,
                       w + 32 * ((long)(d * cout + o) * cin + i), tv, 1);
      }
    }
}

void or_net_layer_wgrad(const or_geom* g, const double* u, int cin, int cout,
                        const double* in, const double* grad_out, double* gw, int round_bf16) {
  const long V = g->V;
  double tv[24], gv[24];
  for (long s = 0; s < V; ++s)
    for (int d = 0; d < 9; ++d)
      for (int i = 0; i < cin; ++i) {
        net_transport(g, u, d, s, in + (long)i * 24 * V, tv, round_bf16);
        for (int o = 0; o < cout; ++o) {
          memcpy(gv, grad_out + (long)o * 24 * V + 24 * s, sizeof gv);
          if (round_bf16)
            for (int k = 0; k < 24; ++k) gv[k] = bf16_round(gv[k]);
          double* G = gw + 32 * ((long)(d * cout + o) * cin + i);
          for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
              double re = 0.0, im = 0.0;
              for (int c = 0; c < 3; ++c) {
                double xr = gv[RE(a, c)], xi = gv[IM(a, c)];
                double yr = tv[RE(b, c)], yi = tv[IM(b, c)];
                re += xr * yr + xi * yi;
                im += xi * yr - xr * yi;
              }
              G[2 * (4 * a + b)] += re;
              G[2 * (4 * a + b) + 1] += im;
            }
        }
      }
}
