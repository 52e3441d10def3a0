// Minimax rational approximation of x^(a/b) on [low, high] in relative
// error, for RationalApprox (lib/Tools/RationalApprox/rationalapprox.cpp:
// 20-80): the coefficients of
//   x^(a/b)  ~ a0  + sum_i res[i]  / (x + pole[i])
//   x^(-a/b) ~ ia0 + sum_i ires[i] / (x + ipole[i])
// that RationalSolver_CG feeds to the multishift solver.
//
// The reference computes them with Clark's AlgRemez in MPFR
// (lib/Tools/Remez/alg_remez.cpp), which moves the zeros of the error
// between its extrema.  This is a different, levelled-error exchange in
// __float128 (113-bit mantissa) arithmetic, host code only:
//   * work in u = x / high on [low/high, 1] (rescaled at the end);
//   * reference set: N = n + d + 2 points, log-Chebyshev initial spacing
//     (the extrema of a relative-error minimax of a power cluster
//     geometrically towards the lower end);
//   * each step solves P(u_i) - f_i (1 + (-1)^i h) Q(u_i) = 0 for the
//     numerator P, the monic denominator Q and the levelled error h by
//     Newton's method (F is bilinear in Q and h), started from the previous
//     step's solution;
//   * wide intervals are reached by continuation from a narrower one;
//   * the new reference set takes, between consecutive sign changes of the
//     error e(u) = P/(f Q) - 1, the point of largest |e| (scanned on a log
//     grid, refined by golden-section search), the ends included;
//   * stop when the extrema are level: (max|e| - min|e|) / min|e| < 1e-15
//     over the set, the reference's spread criterion (alg_remez.cpp:41).
// The minimax approximation is unique (Chebyshev alternation), so the
// result equals the reference's to the convergence tolerance; the checker
// is a restatement of AlgRemez in 50-digit arithmetic.
//
// Zeros and poles of P and Q are real and negative for these functions;
// they are bracketed by sign changes on a log grid over (-1e8, -1e-18) and
// refined by bisection + Newton on the full polynomial (no deflation).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/iroiro_hip.h"

namespace {

typedef __float128 qf;

inline qf qabs(qf x) { return x < 0 ? -x : x; }

// x^p for integer p
qf ipow(qf x, int p) {
  bool inv = p < 0;
  unsigned e = inv ? (unsigned)(-p) : (unsigned)p;
  qf r = 1, b = x;
  while (e) {
    if (e & 1u) r *= b;
    b *= b;
    e >>= 1;
  }
  return inv ? 1 / r : r;
}

// u^(num/den) for u > 0, den > 0: Newton on y^den = u^num from the double
// result (quadratic convergence: three steps reach the quad precision)
qf qpow_rational(qf u, int num, int den) {
  if (num == 0 || u == 1) return 1;
  const qf t = ipow(u, num);
  qf y = (qf)std::pow((double)u, (double)num / (double)den);
  for (int it = 0; it < 4; ++it) {
    const qf yd1 = ipow(y, den - 1);
    y -= (yd1 * y - t) / ((qf)den * yd1);
  }
  return y;
}

struct Remez {
  int n, d, num, den;
  qf lo;                  // low / high
  std::vector<qf> p;      // numerator coefficients p[0..n]
  std::vector<qf> q;      // denominator q[0..d], q[d] = 1
  std::vector<qf> xs;     // reference set (N points)
  qf h = 0;
  int iterations = 0;
  std::string err;

  qf f(qf u) const { return qpow_rational(u, num, den); }
  static qf horner(const std::vector<qf>& c, qf u) {
    qf r = c.back();
    for (int i = (int)c.size() - 2; i >= 0; --i) r = r * u + c[i];
    return r;
  }
  qf e(qf u) const { return horner(p, u) / (f(u) * horner(q, u)) - 1; }

  // Gaussian elimination with partial pivoting on a dense m x m system.
  static bool solve(std::vector<qf>& A, std::vector<qf>& b, int m) {
    for (int k = 0; k < m; ++k) {
      int piv = k;
      qf best = qabs(A[(size_t)k * m + k]);
      for (int i = k + 1; i < m; ++i)
        if (qabs(A[(size_t)i * m + k]) > best) {
          best = qabs(A[(size_t)i * m + k]);
          piv = i;
        }
      if (best == 0) return false;
      if (piv != k) {
        for (int j = 0; j < m; ++j) std::swap(A[(size_t)k * m + j], A[(size_t)piv * m + j]);
        std::swap(b[k], b[piv]);
      }
      for (int i = k + 1; i < m; ++i) {
        const qf l = A[(size_t)i * m + k] / A[(size_t)k * m + k];
        if (l == 0) continue;
        for (int j = k; j < m; ++j) A[(size_t)i * m + j] -= l * A[(size_t)k * m + j];
        b[i] -= l * b[k];
      }
    }
    for (int i = m - 1; i >= 0; --i) {
      qf s = b[i];
      for (int j = i + 1; j < m; ++j) s -= A[(size_t)i * m + j] * b[j];
      b[i] = s / A[(size_t)i * m + i];
    }
    return true;
  }

  // Linear start: P - f Q = 0 at the N - 1 log-Chebyshev nodes of
  // [lo, 1] (Q monic).  The error then vanishes at N - 1 points, so its
  // N extrema alternate, and the levelled Newton below starts on the
  // branch whose Q has no zero in the interval.
  bool interpolate() {
    const int M = n + d + 1;
    std::vector<qf> A((size_t)M * M), b(M);
    const double llo = std::log((double)lo);
    for (int i = 0; i < M; ++i) {
      const double c = 0.5 * (1.0 + std::cos(M_PI * (2.0 * i + 1.0) / (2.0 * M)));
      const qf u = (qf)std::exp(llo * c), fu = f(u);
      qf z = 1;
      for (int j = 0; j <= n; ++j) { A[(size_t)i * M + j] = z; z *= u; }
      z = 1;
      for (int j = 0; j < d; ++j) { A[(size_t)i * M + n + 1 + j] = -fu * z; z *= u; }
      b[i] = fu * z;
    }
    if (!solve(A, b, M)) { err = "singular interpolation system"; return false; }
    for (int j = 0; j <= n; ++j) p[j] = b[j];
    for (int j = 0; j < d; ++j) q[j] = b[n + 1 + j];
    q[d] = 1;
    return true;
  }

  // P, Q and h levelled on the current reference set:
  //   F_i = P(u_i) - f_i (1 + s_i h) Q(u_i) = 0,  s_i = (-1)^i, Q monic,
  // by Newton on (p, q, h) from the current solution; F is bilinear in
  // (q, h), so Newton converges quadratically near the solution.
  bool level() {
    const int N = n + d + 2;
    std::vector<qf> A((size_t)N * N), b(N);
    for (int it = 0; it < 30; ++it) {
      qf fmax = 0;
      for (int i = 0; i < N; ++i) {
        const qf u = xs[i], fu = f(u);
        const qf si = (i & 1) ? -1 : 1;
        const qf g = fu * (1 + si * h);
        const qf Qu = horner(q, u);
        b[i] = -(horner(p, u) - g * Qu);
        if (qabs(b[i]) > fmax) fmax = qabs(b[i]);
        qf z = 1;
        for (int j = 0; j <= n; ++j) { A[(size_t)i * N + j] = z; z *= u; }
        z = 1;
        for (int j = 0; j < d; ++j) { A[(size_t)i * N + n + 1 + j] = -g * z; z *= u; }
        A[(size_t)i * N + N - 1] = -si * fu * Qu;
      }
      if (!solve(A, b, N)) { err = "singular Newton system"; return false; }
      for (int j = 0; j <= n; ++j) p[j] += b[j];
      for (int j = 0; j < d; ++j) q[j] += b[n + 1 + j];
      h += b[N - 1];
      if (qabs(b[N - 1]) <= qabs(h) * (qf)1e-28) break;
    }
    return true;
  }

  // location of max |e| on [a, b] (|e| unimodal there), golden section in log u
  qf argmax(qf a, qf b, qf* val) const {
    const qf g = (qf)0.6180339887498948482045868343656381L;
    qf la = (qf)std::log((double)a), lb = (qf)std::log((double)b);
    auto ev = [&](qf l) { return qabs(e((qf)std::exp((double)l))); };
    qf c = lb - g * (lb - la), dd = la + g * (lb - la);
    qf fc = ev(c), fd = ev(dd);
    for (int it = 0; it < 60; ++it) {
      if (fc > fd) {
        lb = dd; dd = c; fd = fc;
        c = lb - g * (lb - la); fc = ev(c);
      } else {
        la = c; c = dd; fc = fd;
        dd = la + g * (lb - la); fd = ev(dd);
      }
    }
    const qf lm = (la + lb) / 2;
    const qf um = (qf)std::exp((double)lm);
    // the ends of the interval can hold the maximum (endpoints of [lo, 1])
    qf best = um, bv = qabs(e(um));
    const qf ea = qabs(e(a)), eb = qabs(e(b));
    if (ea > bv) { best = a; bv = ea; }
    if (eb > bv) { best = b; bv = eb; }
    *val = bv;
    return best;
  }

  // new reference set from the extrema between sign changes; returns the
  // spread (max - min)/min of |e| over it, or a negative value on failure
  qf exchange() {
    const int N = n + d + 2;
    const int G = 64 * N;
    std::vector<qf> ug(G), eg(G);
    const double llo = std::log((double)lo);
    for (int i = 0; i < G; ++i) {
      ug[i] = i == 0 ? lo : (i == G - 1 ? (qf)1 : (qf)std::exp(llo * (1.0 - (double)i / (G - 1))));
      eg[i] = e(ug[i]);
    }
    // sign-constant runs of the error
    std::vector<int> start;
    start.push_back(0);
    for (int i = 1; i < G; ++i)
      if ((eg[i] < 0) != (eg[i - 1] < 0)) start.push_back(i);
    const int runs = (int)start.size();
    std::vector<qf> loc(runs), val(runs);
    for (int r = 0; r < runs; ++r) {
      const int i0 = start[r], i1 = (r + 1 < runs ? start[r + 1] : G) - 1;
      int im = i0;
      for (int i = i0; i <= i1; ++i)
        if (qabs(eg[i]) > qabs(eg[im])) im = i;
      const qf a = im > 0 ? ug[im - 1] : ug[0];
      const qf b = im < G - 1 ? ug[im + 1] : ug[G - 1];
      loc[r] = argmax(a, b, &val[r]);
    }
    if (runs < N) {
      err = "error curve has fewer alternations than the reference set";
      return -1;
    }
    // keep N alternating extrema: while too many, drop the smaller end run
    // or merge the weakest adjacent pair
    std::vector<qf> L = loc, Vv = val;
    while ((int)L.size() > N) {
      const int m = (int)L.size();
      if (Vv[0] < Vv[m - 1]) {
        L.erase(L.begin());
        Vv.erase(Vv.begin());
      } else {
        L.pop_back();
        Vv.pop_back();
      }
    }
    qf mx = 0, mn = Vv[0];
    for (int i = 0; i < N; ++i) {
      xs[i] = L[i];
      if (Vv[i] > mx) mx = Vv[i];
      if (Vv[i] < mn) mn = Vv[i];
    }
    maxerr = mx;
    return mn > 0 ? (mx - mn) / mn : (qf)1e30;
  }
  qf maxerr = 0;

  // Exchange steps on [lo_k, 1] from the current reference set and
  // solution until the extrema are level.
  bool converge(qf lo_k, bool cold) {
    lo = lo_k;
    if (cold) {
      if (!interpolate()) return false;
      if (exchange() < 0) return false;
      // levelled-error estimate with the sign of the first extremum
      qf acc = 0;
      for (int i = 0; i < n + d + 2; ++i) acc += ((i & 1) ? -1 : 1) * e(xs[i]);
      h = acc / (n + d + 2);
    }
    qf spread = 1e30;
    for (int it = 0; it < 100; ++it, ++iterations) {
      if (!level()) return false;
      spread = exchange();
      if (spread < 0) return false;
      if (spread < (qf)1e-15) break;
    }
    if (spread > (qf)1e-12) {
      err = "Remez exchange did not level the error";
      return false;
    }
    return level();
  }

  // Continuation in the width of the interval: start on [lo0, 1], where the
  // minimax error of degree n is about 1e-20 (so 113-bit arithmetic resolves
  // it; estimate exp(-2 pi^2 n / log(4/lo0)) for fractional powers), and
  // widen by factors of lo^0.2 in log width, warm-starting each step from
  // the previous solution and its reference set mapped by u -> u^(log
  // lo_next / log lo_k).
  bool run() {
    const int N = n + d + 2;
    const qf target = lo;
    double l0 = std::log(4.0) - 2.0 * M_PI * M_PI * n / 46.0;  // log lo0
    if (l0 > std::log(0.25)) l0 = std::log(0.25);
    if (l0 < std::log((double)target)) l0 = std::log((double)target);
    p.assign(n + 1, 0);
    q.assign(d + 1, 0);
    q[d] = 1;
    xs.resize(N);
    for (int i = 0; i < N; ++i) {
      const double c = 0.5 * (1.0 + std::cos(M_PI * (double)i / (N - 1)));
      xs[i] = i == 0 ? (qf)std::exp(l0) : (i == N - 1 ? (qf)1 : (qf)std::exp(l0 * c));
    }
    const qf lo0 = l0 <= std::log((double)target) ? target : (qf)std::exp(l0);
    xs[0] = lo0;
    iterations = 0;
    h = 0;
    if (!converge(lo0, true)) return false;
    const double lt = std::log((double)target);
    double lk = std::log((double)lo0);
    while (lk > lt) {
      double ln = lk + 0.2 * lt;
      if (ln < lt) ln = lt;
      const qf lo_next = ln <= lt ? target : (qf)std::exp(ln);
      const double r = ln / lk;
      for (int i = 0; i < N; ++i) xs[i] = (qf)std::exp(std::log((double)xs[i]) * r);
      xs[0] = lo_next;
      xs[N - 1] = 1;
      if (!converge(lo_next, false)) return false;
      lk = ln;
    }
    return true;
  }
};

// Real roots of c (degree m) in (-1e8, -1e-18): sign changes on a log grid,
// then bisection and Newton polishing on the polynomial itself.
bool negative_roots(const std::vector<qf>& c, std::vector<qf>& roots, std::string& err) {
  const int m = (int)c.size() - 1;
  auto P = [&](qf x) { return Remez::horner(c, x); };
  auto dP = [&](qf x) {
    qf r = (qf)m * c[m];
    for (int i = m - 1; i >= 1; --i) r = r * x + (qf)i * c[i];
    return r;
  };
  roots.clear();
  const int per_decade = 200;
  const double l0 = std::log(1e8), l1 = std::log(1e-18);
  const int G = (int)((l0 - l1) / std::log(10.0) * per_decade);
  qf xa = -(qf)std::exp(l0), fa = P(xa);
  for (int i = 1; i <= G; ++i) {
    const qf xb = -(qf)std::exp(l0 + (l1 - l0) * (double)i / G);
    const qf fb = P(xb);
    if (fa == 0) {
      roots.push_back(xa);
    } else if ((fa < 0) != (fb < 0) && fb != 0) {
      qf a = xa, b = xb, fl = fa;
      for (int it = 0; it < 200; ++it) {
        const qf mid = (a + b) / 2;
        const qf fm = P(mid);
        if ((fm < 0) == (fl < 0)) { a = mid; fl = fm; } else { b = mid; }
        if (qabs(b - a) <= qabs(mid) * (qf)1e-30) break;
      }
      qf r = (a + b) / 2;
      for (int it = 0; it < 4; ++it) {
        const qf dp = dP(r);
        if (dp == 0) break;
        const qf nr = r - P(r) / dp;
        if (nr <= a || nr >= b) break;  // stay inside the bracket (a < b < 0)
        r = nr;
      }
      roots.push_back(r);
    }
    xa = xb;
    fa = fb;
  }
  if ((int)roots.size() != m) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "found %d of %d real negative roots", (int)roots.size(), m);
    err = buf;
    return false;
  }
  return true;
}

// c * prod(u - z)/prod(u - w) = c + sum R_k/(u - w_k); returns the R_k.
std::vector<qf> residues(qf c, const std::vector<qf>& z, const std::vector<qf>& w) {
  std::vector<qf> R(w.size());
  for (size_t k = 0; k < w.size(); ++k) {
    qf num = c, den = 1;
    for (size_t j = 0; j < z.size(); ++j) num *= w[k] - z[j];
    for (size_t m = 0; m < w.size(); ++m)
      if (m != k) den *= w[k] - w[m];
    R[k] = num / den;
  }
  return R;
}

}  // namespace

namespace ii {
ii_status set_error(ii_context* ctx, ii_status st, const std::string& msg);
}

extern "C" ii_status ii_rational_approx(ii_context* ctx, int num_deg, int den_deg, int exp_num,
                                        int exp_den, double low, double high, double* res,
                                        double* pole, double* a0, double* ires, double* ipole,
                                        double* ia0, double* error) {
  if (num_deg < 1 || num_deg != den_deg || num_deg > 64 || exp_den <= 0 || !(low > 0) ||
      !(high > low) || !res || !pole || !a0 || !ires || !ipole || !ia0)
    return II_ERR_BAD_ARG;
  Remez R;
  R.n = num_deg;
  R.d = den_deg;
  R.num = exp_num;
  R.den = exp_den;
  R.lo = (qf)low / (qf)high;
  if (!R.run())
    return ii::set_error(ctx, II_ERR_NOT_CONVERGED,
                         "ii_rational_approx: " + R.err +
                             " (degree too high for the interval in 113-bit arithmetic)");
  std::vector<qf> z, w;
  std::string err;
  if (!negative_roots(R.p, z, err) || !negative_roots(R.q, w, err))
    return ii::set_error(ctx, II_ERR_NOT_CONVERGED, "ii_rational_approx: " + err);
  const qf c = R.p[R.n];  // leading coefficient of P (Q is monic)
  std::vector<qf> Rk = residues(c, z, w);
  std::vector<qf> IRk = residues(1 / c, w, z);
  // back to x = high * u: x^a ~ high^a [c + sum R/(x/high + s)]
  const qf hq = (qf)high;
  const qf ha = qpow_rational(hq, exp_num, exp_den);
  const qf hia = 1 / ha;
  // poles as positive shifts, ascending (w, z are ascending negative: the
  // grid runs from -1e8 towards 0, so reverse)
  const int n = R.n;
  for (int k = 0; k < n; ++k) {
    const int s = n - 1 - k;
    pole[k] = (double)(-w[s] * hq);
    res[k] = (double)(Rk[s] * ha * hq);
    ipole[k] = (double)(-z[s] * hq);
    ires[k] = (double)(IRk[s] * hia * hq);
  }
  *a0 = (double)(c * ha);
  *ia0 = (double)(hia / c);
  if (error) *error = (double)R.maxerr;
  return II_OK;
}
