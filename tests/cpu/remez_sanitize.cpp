// The library's quad-precision Remez (iroiro_amd/csrc/ii_remez.cpp, host
// code) under the host sanitizers (address, undefined behaviour), on the
// reference's own RHMC approximations (degree 4 on [0.1, 4] of x^(1/2),
// x^(-1/4); degree 10 on [0.001, 1] of x^(1/2)), one beyond the exchange's
// envelope (II_ERR_NOT_CONVERGED, with its error text) and bad arguments.
// Built and run by tests/test_sanitizers.py.  Checks the partial fractions
// against x^(a/b) at a few points to the reported error.
#include "../../include/iroiro_hip.h"

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

namespace ii {
// the library's error sink (ii_api.hip); here: print the message
ii_status set_error(ii_context*, ii_status st, const std::string& msg) {
  std::printf("remez_sanitize: status %d: %s\n", (int)st, msg.c_str());
  return st;
}
}  // namespace ii

static int check(int deg, int a, int b, double lo, double hi) {
  std::vector<double> res(deg), pole(deg), ires(deg), ipole(deg);
  double a0 = 0, ia0 = 0, err = 0;
  const ii_status st = ii_rational_approx(nullptr, deg, deg, a, b, lo, hi, res.data(), pole.data(),
                                          &a0, ires.data(), ipole.data(), &ia0, &err);
  if (st != II_OK) return 1;
  for (int k = 0; k <= 20; ++k) {
    const double x = lo * std::pow(hi / lo, k / 20.0);
    double r = a0, ir = ia0;
    for (int i = 0; i < deg; ++i) {
      r += res[i] / (x + pole[i]);
      ir += ires[i] / (x + ipole[i]);
    }
    const double f = std::pow(x, (double)a / b);
    if (!(std::fabs(r / f - 1.0) <= 1.01 * err + 1e-14)) return 2;
    if (!(std::fabs(ir * f - 1.0) <= 1.01 * err + 1e-14)) return 3;
  }
  return 0;
}

int main() {
  int rc = check(4, 1, 2, 0.1, 4.0);
  if (!rc) rc = check(4, -1, 4, 0.1, 4.0);
  if (!rc) rc = check(10, 1, 2, 0.001, 1.0);
  if (!rc) rc = check(6, -1, 2, 1e-4, 5.0);
  if (rc) {
    std::fprintf(stderr, "remez_sanitize: check failed (%d)\n", rc);
    return rc;
  }
  // beyond the envelope: a clean NOT_CONVERGED, no memory error
  std::vector<double> v(40);
  double a0, ia0, err;
  const ii_status st = ii_rational_approx(nullptr, 20, 20, 1, 2, 1e-9, 1.0, v.data(), v.data() + 20,
                                          &a0, v.data(), v.data() + 20, &ia0, &err);
  if (st != II_ERR_NOT_CONVERGED && st != II_OK) return 5;
  if (ii_rational_approx(nullptr, 4, 5, 1, 2, 0.1, 1.0, v.data(), v.data(), &a0, v.data(),
                         v.data(), &ia0, &err) != II_ERR_BAD_ARG)
    return 6;
  if (ii_rational_approx(nullptr, 4, 4, 1, 2, 1.0, 0.1, v.data(), v.data(), &a0, v.data(),
                         v.data(), &ia0, &err) != II_ERR_BAD_ARG)
    return 7;
  std::printf("remez_sanitize: ok\n");
  return 0;
}
