"""Error of the mixed-precision solvers' inner operator against the FP64
oracle hop, through ii_dirac_mult_eo_prec: the hop (mult_eo, mult_oe and
their adjoints) with the fields stored in FP32, FP16 or BF16 (per-site
norm) and the links in FP32 or 16-bit reconstruct-12, single-rhs and fused
multi-rhs launches.

The reference has no low-precision operator, so the bar is the storage
format's own rounding.  A stored component carries a relative error of at
most u = 2^-24 (FP32), 2^-11 (FP16) or 2^-8 (BF16) of its site's largest
component; the hop sums 8 such terms and its output is rounded again when
stored, so the relative 2-norm error of the result is a few u.  The bounds
below are 6 u (about twice the largest error measured); a wrong index,
sign or dropped direction gives an error of order 1.
"""
import numpy as np
import pytest

import iroiro_amd as ii
from oracle import oracle

pytestmark = pytest.mark.gpu

U_ROUND = {"single": 2.0 ** -24, "half": 2.0 ** -11, "bf16": 2.0 ** -8}
TOL = {p: 6 * u for p, u in U_ROUND.items()}
KAPPA = 0.1355


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def _sources(lat, nrhs, seed):
    """Gaussian half fields whose site magnitudes spread over six decades,
    so every site's stored norm differs."""
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((nrhs, lat.Vh, 24))
    f *= np.exp(rng.uniform(-7.0, 7.0, (nrhs, lat.Vh, 1)))
    return f.reshape(nrhs, -1)


@pytest.fixture(scope="module", params=[(8, 8, 8, 16), (6, 4, 4, 6)],
                ids=["8c16", "6x4x4x6"])
def setup(request):
    dims = request.param
    lat = oracle.Lattice(dims)
    u = lat.apbc(lat.gauge_expH(0.6, 17))
    ctx = ii.Context(dims)
    D = ii.Dirac_Wilson_EvenOdd(0.5 / KAPPA - 4.0, u, ctx)
    yield lat, u, D
    ctx.close()


# (out_par, dag) -> the oracle's eotype: 1 = odd -> even, 2 = even -> odd
CASES = [(0, False, 1), (1, False, 2), (1, True, 2), (0, True, 1)]


@pytest.mark.parametrize("precision", ["single", "half", "bf16"])
@pytest.mark.parametrize("nrhs", [1, 12, 13])
def test_hop_error_within_storage_rounding(setup, precision, nrhs):
    lat, u, D = setup
    f = _sources(lat, nrhs, seed=nrhs)
    worst = 0.0
    for out_par, dag, eot in CASES:
        got = D.mult_eo_prec(f.ravel(), precision, out_par, dag).reshape(nrhs, -1)
        for k in range(nrhs):
            want = lat.hop_eo(u, f[k], eot, int(dag), KAPPA)
            e = rel(got[k], want)
            worst = max(worst, e)
            assert e < TOL[precision], (precision, nrhs, out_par, dag, k, e)
    print("hop %s nrhs=%d: worst relative error %.3g (bound %.3g)"
          % (precision, nrhs, worst, TOL[precision]))


def test_double_matches_fp64_operator(setup):
    """precision "double" is the FP64 operator itself: the single-rhs launch
    is the kernel mult_eo runs, the multi-rhs one agrees to rounding."""
    lat, u, D = setup
    f = _sources(lat, 3, seed=4)
    one = D.mult_eo_prec(f[0], "double", 0, False)
    assert np.array_equal(one, D.mult_eo(f[0]))
    many = D.mult_eo_prec(f.ravel(), "double", 1, True).reshape(3, -1)
    for k in range(3):
        assert rel(many[k], D.mult_eo_dag(f[k])) < 1e-13


def test_zero_and_tiny_sites_stay_finite(setup):
    """Sites of zero norm (a point source's far region) store as zero."""
    lat, u, D = setup
    f = np.zeros((2, lat.Vh * 24))
    f[0, :24] = 1.0
    f[1, 24 * 5:24 * 6] = 1e-35
    for p in ("half", "bf16", "single"):
        got = D.mult_eo_prec(f.ravel(), p, 1, False).reshape(2, -1)
        assert np.all(np.isfinite(got))
        want = lat.hop_eo(u, f[0], 2, 0, KAPPA)
        assert rel(got[0], want) < TOL[p]
