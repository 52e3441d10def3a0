"""Pins of the oracle's meson correlator (oracle/meson.py) against the
reference's own MesonCorrelator (golden fixture tests/golden/
meson_2x4x2x6.npz: calculate in all four directions, calculate_mom and
calculate_PSmom for all 18 MesonTypes on two Gaussian propagators) and
against identities the contraction must satisfy."""
import numpy as np
import pytest

from oracle import meson


@pytest.fixture(scope="module")
def mz(golden):
    z = golden("meson_2x4x2x6")
    return tuple(int(d) for d in z["dims"]), z


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("t", range(18))
def test_all_types_match_reference(mz, t):
    dims, z = mz
    calc = np.concatenate([meson.calculate(z["q1"], z["q2"], dims, t, d) for d in range(4)])
    assert rel(calc, z["calc"][t]) < 1e-13
    assert rel(meson.calculate_mom(z["q1"], z["q2"], dims, t, z["mom"]), z["mom_out"][t]) < 1e-13
    assert rel(meson.calculate_PSmom(z["q1"], z["q2"], dims, t, z["mom"]), z["psmom_out"][t]) \
        < 1e-13


def test_identities(mz):
    dims, z = mz
    q = z["q1"]
    V = int(np.prod(dims))
    # Pion with q1 = q2: unit gammas, C(t) = sum over the slice of |q|^2
    pion = meson.calculate(q, q, dims, 0)
    a = np.asarray(q).reshape(12, V, 24)
    t = meson.coords(dims)[:, 3]
    want = np.bincount(t, weights=np.sum(a * a, axis=(0, 2)), minlength=dims[3])
    assert np.allclose(pion, want, rtol=1e-13)
    assert np.all(pion > 0)
    # zero momentum: calculate_mom's real part and calculate_PSmom are calculate
    for ty in (0, 5, 12, 17):
        c = meson.calculate(z["q1"], z["q2"], dims, ty)
        m = meson.calculate_mom(z["q1"], z["q2"], dims, ty, (0.0, 0.0, 0.0))
        assert np.allclose(m[0::2], c, rtol=1e-13, atol=1e-12)
        assert np.allclose(meson.calculate_PSmom(z["q1"], z["q2"], dims, ty, (0, 0, 0)), c,
                           rtol=1e-13, atol=1e-12)
        # every direction's slices add up to the same total
        tot = [np.sum(meson.calculate(z["q1"], z["q2"], dims, ty, d)) for d in range(4)]
        assert np.allclose(tot, tot[3], rtol=1e-12)
    # the Pion of q1 = q2 is real at zero momentum
    m = meson.calculate_mom(q, q, dims, 0, (0.0, 0.0, 0.0))
    assert np.max(np.abs(m[1::2])) < 1e-12 * np.max(np.abs(m[0::2]))
