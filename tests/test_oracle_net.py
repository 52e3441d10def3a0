"""Pins the oracle's learned gauge-covariant layer (oracle/wilson_oracle.c
or_net_layer*), which has no counterpart in the reference, by mathematics:
  * with Wilson spin weights a single-channel layer IS the Wilson operator
    (or_dirac_full, itself pinned against the compiled reference);
  * the adjoint satisfies <g, L x> = <L^dag g, x>;
  * weight and data gradients agree with central finite differences of a loss;
  * gauge covariance: L[U^g](g x) = g L[U](x);
  * bfloat16 rounding: known values and ties to even.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.wilson_math import GAMMAS, random_su3, shift_index

DIMS = (4, 4, 4, 4)


@pytest.fixture(scope="module")
def lat():
    return orc.Lattice(DIMS)


@pytest.fixture(scope="module")
def gauge(lat):
    return lat.apbc(lat.gauge_expH(0.4, 77))


def wilson_weights(kappa):
    w = np.zeros((9, 1, 1, 4, 4), complex)
    w[0, 0, 0] = np.eye(4)
    for mu in range(4):
        w[2 * mu + 1, 0, 0] = -kappa * (np.eye(4) - GAMMAS[mu])
        w[2 * mu + 2, 0, 0] = -kappa * (np.eye(4) + GAMMAS[mu])
    return w


def rand_w(rng, cout, cin, scale=0.3):
    return scale * (rng.normal(size=(9, cout, cin, 4, 4)) + 1j * rng.normal(size=(9, cout, cin, 4, 4)))


def test_single_channel_wilson_weights_give_dirac_full(lat, gauge):
    kappa = 0.1355
    x = orc.gauss_field(lat.full_size, 5)
    ref = lat.dirac_full(gauge, x, 0, kappa)
    got = lat.net_layer(gauge, wilson_weights(kappa), x)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-14


def test_adjoint_identity(lat, gauge):
    rng = np.random.default_rng(1)
    w = rand_w(rng, 2, 3)
    x = rng.normal(size=3 * lat.full_size)
    g = rng.normal(size=2 * lat.full_size)
    lhs = np.dot(g, lat.net_layer(gauge, w, x))
    rhs = np.dot(lat.net_layer_adj(gauge, w, g), x)
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


def _loss(lat, gauge, w, x, y):
    e = lat.net_layer(gauge, w, x) - y
    return float(np.dot(e, e))


def test_weight_gradient_finite_differences(lat, gauge):
    rng = np.random.default_rng(2)
    cin, cout = 2, 2
    w = rand_w(rng, cout, cin)
    x = rng.normal(size=cin * lat.full_size)
    y = rng.normal(size=cout * lat.full_size)
    g_out = 2.0 * (lat.net_layer(gauge, w, x) - y)
    gw = lat.net_layer_wgrad(gauge, cin, cout, x, g_out)
    h = 1e-5
    for idx in [(0, 0, 0, 0, 0), (3, 1, 0, 2, 1), (8, 0, 1, 3, 3), (6, 1, 1, 1, 2)]:
        for part, unit in (("re", 1.0), ("im", 1j)):
            wp = w.copy()
            wm = w.copy()
            wp[idx] += h * unit
            wm[idx] -= h * unit
            fd = (_loss(lat, gauge, wp, x, y) - _loss(lat, gauge, wm, x, y)) / (2 * h)
            an = gw[idx].real if part == "re" else gw[idx].imag
            assert abs(fd - an) < 1e-6 * max(1.0, abs(an)), (idx, part, fd, an)


def test_network_gradient_finite_differences(lat, gauge):
    """Loss of a 3-layer composition on even-site pairs: analytic gradient
    (wgrad + adjoint chain) against central differences."""
    rng = np.random.default_rng(3)
    ws = [rand_w(rng, 2, 1), rand_w(rng, 2, 2), rand_w(rng, 1, 2)]
    pairs = [(rng.normal(size=lat.half_size), rng.normal(size=lat.half_size)) for _ in range(2)]
    net = orc.Net(lat, gauge, ws)
    _, grads = net.loss_and_grad(pairs)
    h = 1e-6
    for li, idx in [(0, (0, 1, 0, 2, 3)), (1, (4, 0, 1, 1, 0)), (2, (7, 0, 1, 3, 2)),
                    (1, (0, 1, 1, 0, 0))]:
        for unit in (1.0, 1j):
            wp = [w.copy() for w in ws]
            wm = [w.copy() for w in ws]
            wp[li][idx] += h * unit
            wm[li][idx] -= h * unit
            fp = orc.Net(lat, gauge, wp).loss_and_grad(pairs)[0]
            fm = orc.Net(lat, gauge, wm).loss_and_grad(pairs)[0]
            fd = (fp - fm) / (2 * h)
            an = grads[li][idx].real if unit == 1.0 else grads[li][idx].imag
            assert abs(fd - an) < 1e-5 * max(1.0, abs(an)), (li, idx, unit, fd, an)


def test_gauge_covariance(lat, gauge):
    rng = np.random.default_rng(4)
    V = lat.V
    w = rand_w(rng, 2, 2)
    x = rng.normal(size=2 * lat.full_size)
    gt = random_su3(V, rng)
    U = gauge.reshape(4, V, 3, 3, 2)
    Uc = U[..., 0] + 1j * U[..., 1]
    Ug = np.empty_like(Uc)
    for mu in range(4):
        fw = shift_index(DIMS, mu, +1)
        Ug[mu] = gt @ Uc[mu] @ np.conj(np.transpose(gt[fw], (0, 2, 1)))
    ug = np.stack([Ug.real, Ug.imag], axis=-1).ravel()

    def rot(f):
        c = f.reshape(-1, V, 4, 3, 2)
        c = c[..., 0] + 1j * c[..., 1]
        r = np.einsum("xab,nxsb->nxsa", gt, c)
        return np.stack([r.real, r.imag], axis=-1).ravel()

    lhs = lat.net_layer(ug, w, rot(x))
    rhs = rot(lat.net_layer(gauge, w, x))
    assert np.linalg.norm(lhs - rhs) / np.linalg.norm(rhs) < 1e-13


def test_bf16_rounding_known_values_and_c_agrees(lat, gauge):
    vals = np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8 + 2.0 ** -20),
                     0.1, 3.0e-5])
    exp = np.array([1.0, 1.0 + 2.0 ** -6, -(1.0 + 2.0 ** -7), 0.10009765625, 252 / 128 * 2.0 ** -16])
    assert np.array_equal(orc.bf16_round(vals), exp)
    # the C rounding inside the layer: identity path only
    w = np.zeros((9, 1, 1, 4, 4), complex)
    w[0, 0, 0] = np.eye(4)
    x = np.random.default_rng(5).normal(size=lat.full_size)
    assert np.array_equal(lat.net_layer(gauge, w, x, round_bf16=True), orc.bf16_round(x))
