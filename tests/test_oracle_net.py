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


def _e4m3_table():
    """The non-negative finite OCP e4m3 values below 448 from the bit layout
    (sign, 4 exponent bits with bias 7, 3 mantissa bits; exponent 0 is
    subnormal), in code order."""
    out = []
    for code in range(0x7F):
        e, m = code >> 3, code & 7
        out.append(m * 2.0 ** -9 if e == 0 else (1 + m / 8) * 2.0 ** (e - 7))
    return np.array(out)


def test_e4m3_round_grid_ties_and_known_values():
    t = _e4m3_table()
    assert np.all(np.diff(t) > 0) and t[-1] == 448.0
    assert np.array_equal(orc.e4m3_round(t), t)
    assert np.array_equal(orc.e4m3_round(-t), -t)
    mid = (t[1:] + t[:-1]) / 2
    codes = np.searchsorted(t, orc.e4m3_round(mid))
    assert np.all(codes % 2 == 0)  # ties go to the even mantissa
    nudge = 1e-9 * t[1:]
    assert np.array_equal(orc.e4m3_round(mid + nudge), t[1:])
    assert np.array_equal(orc.e4m3_round(mid - nudge), t[:-1])
    known = {1.0625: 1.0, 1.1875: 1.25, 0.3: 0.3125, 2.5 * 2.0 ** -9: 2.0 ** -8,
             3.5 * 2.0 ** -9: 2.0 ** -7, 200.0: 192.0, -13.0: -13.0}
    for x, want in known.items():
        assert orc.e4m3_round(x) == want, x


def test_fp8_record_round_scale_error_and_idempotence(lat):
    rng = np.random.default_rng(11)
    x = rng.normal(size=16 * lat.full_size) * np.exp(rng.normal(size=16 * lat.full_size))
    e, o = lat.eo_tables()
    y = orc.fp8_record_round(x, 16, o)
    xr, yr = x.reshape(16, lat.V, 24), y.reshape(16, lat.V, 24)
    assert np.array_equal(yr[:, e, :], xr[:, e, :])  # other sites untouched
    # rounding again changes only groups whose largest value rounded up to
    # 256 x scale (their scale doubles)
    y2 = orc.fp8_record_round(y, 16, o).reshape(16, lat.V, 24)
    for h in range(2):
        grp, grp2 = yr[8 * h:8 * h + 8][:, o, :], y2[8 * h:8 * h + 8][:, o, :]
        mx = np.abs(xr[8 * h:8 * h + 8][:, o, :]).max(axis=(0, 2))
        ds = np.ldexp(1.0, np.frexp(mx)[1] - 8)
        same_scale = np.abs(grp).max(axis=(0, 2)) < 256 * ds
        assert same_scale.mean() > 0.5
        assert np.array_equal(grp2[:, same_scale, :], grp[:, same_scale, :])
    for h in range(2):
        blk, got = xr[8 * h:8 * h + 8][:, o, :], yr[8 * h:8 * h + 8][:, o, :]
        mx = np.abs(blk).max(axis=(0, 2))
        ds = np.ldexp(1.0, np.frexp(mx)[1] - 8)
        assert np.all((mx / ds >= 128) & (mx / ds < 256))
        # normal range: half a unit in the 4th significant bit; below it the
        # subnormal spacing 2^-9 ds
        scaled = np.abs(blk) / ds[None, :, None]
        err = np.abs(got - blk) / ds[None, :, None]
        normal = scaled >= 2.0 ** -6
        assert np.all(err[normal] <= 2.0 ** -4 * scaled[normal] + 1e-12)
        assert np.all(err[~normal] <= 2.0 ** -10 + 1e-15)
        # the scale is per 8-channel group: a group 2^20 times smaller keeps
        # its relative precision
        z = x.copy().reshape(16, lat.V, 24)
        z[8:] *= 2.0 ** -20
        zq = orc.fp8_record_round(z.ravel(), 16, o).reshape(16, lat.V, 24)
        lo, loq = z[8:][:, o, :], zq[8:][:, o, :]
        assert np.linalg.norm(loq - lo) / np.linalg.norm(lo) < 5e-2


def test_apply_fp8_records_is_the_network_when_rounding_is_exact(lat, gauge, monkeypatch):
    """With the record rounding replaced by the identity, the regrouped data
    flow of apply_fp8_records (L2(odd) from the unrounded L1, L2(even) from
    the rounded L1(odd), L3 and z from the rounded L2) is the network; with
    the rounding on it differs from it by the e4m3 rounding, a few percent."""
    rng = np.random.default_rng(2)
    ws = [rand_w(rng, 8, 1), rand_w(rng, 8, 8, 0.1), rand_w(rng, 8, 8, 0.1), rand_w(rng, 1, 8, 0.1)]
    net = orc.Net(lat, gauge, ws)
    b = rng.normal(size=lat.half_size)
    exact = net.apply(b)
    rounded = net.apply_fp8_records(b)
    d = np.linalg.norm(rounded - exact) / np.linalg.norm(exact)
    assert 3e-3 < d < 6e-2, d
    monkeypatch.setattr(orc, "fp8_record_round", lambda x, nch, sites: np.array(x, dtype=np.float64))
    same = net.apply_fp8_records(b)
    assert np.linalg.norm(same - exact) <= 1e-12 * np.linalg.norm(exact)


def test_device_form_is_the_network_without_rounding(lat, gauge):
    """Net.apply_device_form, the oracle of the batched device evaluation:
    with its roundings left out, the regrouped evaluation (composite weights
    of the first stage, the identity products, the Y mixes) is the network to
    1e-12; with them (bf16 links, weights, inputs and transports, e4m3
    records) it stays within a few percent, like apply_fp8_records."""
    rng = np.random.default_rng(4)
    ws = [rand_w(rng, 8, 1), rand_w(rng, 8, 8, 0.1), rand_w(rng, 8, 8, 0.1), rand_w(rng, 1, 8, 0.1)]
    net = orc.Net(lat, gauge, ws)
    b = rng.normal(size=lat.half_size)
    exact = net.apply(b)
    plain = net.apply_device_form(b, rounding=False)
    assert np.linalg.norm(plain - exact) <= 1e-12 * np.linalg.norm(exact)
    dev = net.apply_device_form(b)
    d = np.linalg.norm(dev - exact) / np.linalg.norm(exact)
    assert 3e-3 < d < 6e-2, d
    # a dropped path of the second layer moves it far beyond that
    wrong = [w.copy() for w in ws]
    wrong[1][3] = 0.0
    bad = orc.Net(lat, gauge, wrong).apply_device_form(b)
    assert np.linalg.norm(bad - dev) / np.linalg.norm(dev) > 0.1
