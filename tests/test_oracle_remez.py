"""Pins of the oracle's Remez (oracle/remez.py, a restatement of the
reference's AlgRemez / RationalApprox in 50-digit arithmetic) and of the
library's own quad-precision Remez (ii_rational_approx, host code: no GPU)
against it.

The reference's Remez needs GMP/MPFR, which this image does not have, so it
cannot be compiled; the pins are:
 * the reference test's own checks (tests/test_RationalApprox.cpp:76-100,
   170-190): the approximation at x = 0.5 against pow, and applying the
   x^(1/2) approximation twice (here on scalars spanning the interval);
 * the Chebyshev alternation theorem with de la Vallee Poussin's bound: an
   approximant whose relative error reaches +-E at 2n+2 alternating points
   and nowhere exceeds E is THE minimax approximation (unique), so its
   poles and residues are fixed;
 * the partial-fraction identities: c + sum r/(x+p) equals the ratio of the
   Remez polynomials P/Q, and the inverse expansion equals Q/P;
 * a planted mistake (one pole moved by 1e-6 relative) breaks the
   alternation.
"""
import functools
import math

import numpy as np
import pytest

from oracle import remez

# (num_deg, exp_num, exp_den, low, high): the reference test's approximation
# (test_RationalApprox.xml) and the three approximations of its Wilson RHMC
# test (test_HMC.wilson.rational.xml: Nf = 2, 2 pseudofermions, degree 4,
# [0.1, 4]: Metropolis and MD x^(2/4), pseudofermions x^(2/8))
CASES = [
    (10, 1, 2, 0.001, 1.0),
    (4, 2, 4, 0.1, 4.0),
    (4, 2, 8, 0.1, 4.0),
    (6, -1, 2, 0.01, 1.0),
]


@functools.lru_cache(maxsize=None)
def oracle_approx(n, a, b, lo, hi):
    return remez.RationalApprox(n, n, a, b, 50, lo, hi)


def pfe_eval(c, res, poles, x):
    """c + sum r/(x + p) in extended precision: at the lower end the sum
    cancels the constant to a few parts in 1e3, which double precision
    would turn into 1e-13 noise on a 1e-9 relative error."""
    x = np.asarray(x, dtype=np.longdouble)
    r = np.asarray(res, dtype=np.longdouble)
    p = np.asarray(poles, dtype=np.longdouble)
    return np.longdouble(c) + np.sum(r[None, :] / (x[:, None] + p[None, :]), axis=1)


def alternation(c, res, poles, alpha, lo, hi, npts=200001):
    """Extrema of the relative error per sign run on a dense log grid:
    returns (list of (sign, |e|) per run, max |e|)."""
    x = np.exp(np.linspace(math.log(lo), math.log(hi), npts)).astype(np.longdouble)
    e = pfe_eval(c, res, poles, x) / x ** np.longdouble(alpha) - 1
    runs = []
    start = 0
    sgn = np.sign(e)
    for i in range(1, npts + 1):
        if i == npts or sgn[i] != sgn[start]:
            runs.append((int(sgn[start]), float(np.max(np.abs(e[start:i])))))
            start = i
    return runs, float(np.max(np.abs(e)))


@pytest.mark.parametrize("case", CASES)
def test_oracle_is_the_minimax_approximation(case):
    n, a, b, lo, hi = case
    ra = oracle_approx(*case)
    alpha = a / b
    for c, res, poles, al in [(ra.Const(), ra.Residuals(), ra.Poles(), alpha),
                              (ra.InvConst(), ra.InvResiduals(), ra.InvPoles(), -alpha)]:
        runs, emax = alternation(c, res, poles, al, lo, hi)
        assert len(runs) == 2 * n + 2
        signs = [s for s, _ in runs]
        assert all(signs[i] == -signs[i + 1] for i in range(len(signs) - 1))
        emin = min(v for _, v in runs)
        # de la Vallee Poussin: emin <= E* <= emax, and they agree (to the
        # rounding of the coefficients to double and the grid spacing)
        assert (emax - emin) / emax < 1e-4
        assert abs(emax / float(ra.error) - 1) < 1e-4
        assert all(p > 0 for p in poles)


def test_oracle_reference_test_checks():
    """tests/test_RationalApprox.cpp: r(0.5) vs pow(0.5, 1/2), and the x^(1/2)
    approximation applied twice gives x (the operator check on scalars)."""
    ra = oracle_approx(10, 1, 2, 0.001, 1.0)
    r = float(pfe_eval(ra.Const(), ra.Residuals(), ra.Poles(), [0.5])[0])
    assert abs(r - math.sqrt(0.5)) <= float(ra.error) * math.sqrt(0.5) * (1 + 1e-9)
    assert abs(r - math.sqrt(0.5)) > 1e-10   # an approximation, not the function
    x = np.exp(np.linspace(math.log(0.001), 0.0, 101))
    once = pfe_eval(ra.Const(), ra.Residuals(), ra.Poles(), x)
    assert float(np.max(np.abs(once * once / x - 1))) < 2.1 * float(ra.error)


@pytest.mark.parametrize("case", CASES[:2])
def test_oracle_pfe_identities(case):
    ra = oracle_approx(*case)
    rz = ra.remez
    lo, hi = case[3], case[4]
    xs = [lo, math.sqrt(lo * hi), 0.37 * hi, hi]
    for x in xs:
        pq = float(rz.approx(remez.mp.mpf(x)))
        assert abs(float(pfe_eval(ra.Const(), ra.Residuals(), ra.Poles(), [x])[0]) / pq - 1) \
            < 1e-13
        assert abs(float(pfe_eval(ra.InvConst(), ra.InvResiduals(), ra.InvPoles(), [x])[0]) * pq
                   - 1) < 1e-13


def test_planted_pole_shift_breaks_alternation():
    n, a, b, lo, hi = CASES[0]
    ra = oracle_approx(*CASES[0])
    poles = np.array(ra.Poles())
    poles[3] *= 1 + 1e-6
    runs, emax = alternation(ra.Const(), ra.Residuals(), poles, a / b, lo, hi)
    emin = min(v for _, v in runs)
    assert len(runs) < 2 * n + 2 or (emax - emin) / emax > 1e-2


def test_rescale_follows_the_reference_formula():
    """rationalapprox.cpp:82-103 on the oracle and the library mirror."""
    import iroiro_amd as ii
    ra = remez.RationalApprox(4, 4, 2, 4, 40, 0.1, 4.0)   # a fresh object: rescale mutates
    lib = ii.RationalApprox(4, 4, 2, 4, 40, 0.1, 4.0)
    c0, r0, p0 = ra.Const(), ra.Residuals(), ra.Poles()
    ra.rescale(0.2, 3.0)
    lib.rescale(0.2, 3.0)
    mx = 1.05 * 3.0
    f = mx ** 0.5
    assert ra.Const() == c0 * f
    assert ra.Poles() == [p * mx for p in p0]
    assert ra.Residuals() == [r * (f * mx) for r in r0]
    np.testing.assert_allclose(lib.Poles(), ra.Poles(), rtol=1e-10)
    np.testing.assert_allclose(lib.Residuals(), ra.Residuals(), rtol=1e-9)


@pytest.mark.parametrize("case", CASES)
def test_library_remez_matches_oracle(case):
    """The library's levelled-error exchange in 113-bit arithmetic finds the
    same (unique) minimax approximation as the 50-digit AlgRemez."""
    import iroiro_amd as ii
    n, a, b, lo, hi = case
    ra = oracle_approx(*case)
    lib = ii.RationalApprox(n, n, a, b, 50, lo, hi)
    assert abs(lib.error / float(ra.error) - 1) < 1e-8
    np.testing.assert_allclose(lib.Poles(), ra.Poles(), rtol=1e-9)
    np.testing.assert_allclose(lib.Residuals(), ra.Residuals(), rtol=1e-8)
    np.testing.assert_allclose(lib.Const(), ra.Const(), rtol=1e-12)
    np.testing.assert_allclose(lib.InvPoles(), ra.InvPoles(), rtol=1e-9)
    np.testing.assert_allclose(lib.InvResiduals(), ra.InvResiduals(), rtol=1e-8)
    np.testing.assert_allclose(lib.InvConst(), ra.InvConst(), rtol=1e-12)


def test_library_remez_reports_failure_beyond_its_range():
    import iroiro_amd as ii
    with pytest.raises(ii.IIError):
        ii.RationalApprox(24, 24, 1, 4, 50, 1e-9, 1.0)
    with pytest.raises(ValueError):
        ii.RationalApprox(6, 5, 1, 2, 50, 0.01, 1.0)
