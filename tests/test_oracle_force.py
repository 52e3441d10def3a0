"""Pins of the oracle's MD force (or_md_force_full / or_md_force_eo /
or_md_force_pm_full / or_traceless_antihermite) before it checks the device
kernels.

 * against the compiled reference: tests/golden/*_mdforce.npz hold
   Dirac_Wilson_EvenOdd::md_force, Dirac_Wilson::md_force/_p/_m and
   FieldUtils::TracelessAntihermite outputs of IroIro++ itself;
 * against what the mathematics fixes: S(U) = Re<zeta, D[U] eta> is affine in
   the links for the full operator and quadratic for the Schur operator, so
   under U -> U + eps X U (X any complex 3x3 per link) the central difference
   (S(eps) - S(-eps)) / (2 eps) equals Re sum_{x,mu} tr(X_mu(x) F_mu(x))
   exactly (up to rounding) when F is the force.  A dropped md_force_m term,
   a swapped pair, a missing gamma5, a transposed outer product or a wrong
   kappa factor all break this identity.
"""
import numpy as np
import pytest

from oracle import oracle


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def as_mats(u, V):
    a = np.asarray(u).reshape(4, V, 3, 3, 2)
    return a[..., 0] + 1j * a[..., 1]


def from_mats(m):
    out = np.empty(m.shape + (2,))
    out[..., 0] = m.real
    out[..., 1] = m.imag
    return out.ravel()


def random_x(V, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((4, V, 3, 3)) + 1j * rng.standard_normal((4, V, 3, 3))


def tr_pairing(x, f_real, V):
    """Re sum_{x,mu} tr(X F) = Re sum_{ab} X_ab F_ba."""
    f = as_mats(f_real, V)
    return float(np.real(np.einsum("mxab,mxba->", x, f)))


@pytest.fixture(scope="module")
def aniso(golden):
    base = golden("wilson_aniso_4x6x2x4")
    z = golden("wilson_aniso_4x6x2x4_mdforce")
    lat = oracle.Lattice(tuple(base["dims"]))
    k = oracle.kappa_from_mass(float(base["mass"]))
    return lat, k, base["u_apbc"], z


def test_md_force_bit_exact_vs_reference(aniso):
    lat, k, u, z = aniso
    f_eo = lat.md_force_eo(u, z["eta_e"], z["zeta_e"], k)
    f_full = lat.md_force_full(u, z["eta_f"], z["zeta_f"], k)
    fp = lat.md_force_pm_full(u, np.zeros(lat.gauge_size), z["eta_f"], z["zeta_f"], 0)
    fm = lat.md_force_pm_full(u, np.zeros(lat.gauge_size), z["eta_f"], z["zeta_f"], 1)
    ta = lat.traceless_antihermite(f_eo)
    for got, key in [(f_eo, "mdf_eo"), (f_full, "mdf_full"), (fp, "mdf_p"), (fm, "mdf_m"),
                     (ta, "ta_eo")]:
        assert np.array_equal(got, z[key]), key


def test_full_force_is_the_link_derivative(aniso):
    lat, k, u, z = aniso
    V = lat.V
    eta, zeta = z["eta_f"], z["zeta_f"]
    F = lat.md_force_full(u, eta, zeta, k)
    U = as_mats(u, V)
    X = random_x(V, 11)
    eps = 0.25   # S is affine in eps: any step is exact

    def S(e):
        ue = from_mats(U + e * (X @ U))
        return float(np.dot(zeta, lat.dirac_full(ue, eta, 0, k)))
    slope = (S(eps) - S(-eps)) / (2 * eps)
    want = tr_pairing(X, F, V)
    assert abs(slope - want) < 1e-10 * abs(want)
    # affine: the second difference vanishes
    assert abs(S(eps) - 2 * S(0.0) + S(-eps)) < 1e-10 * abs(S(0.0))


def test_eo_force_is_the_link_derivative(aniso):
    lat, k, u, z = aniso
    V = lat.V
    eta, zeta = z["eta_e"], z["zeta_e"]
    F = lat.md_force_eo(u, eta, zeta, k)
    U = as_mats(u, V)
    X = random_x(V, 12)
    eps = 0.125  # S is quadratic in eps: the central difference is exact

    def S(e):
        ue = from_mats(U + e * (X @ U))
        return float(np.dot(zeta, lat.dirac_eo(ue, eta, 0, k)))
    slope = (S(eps) - S(-eps)) / (2 * eps)
    want = tr_pairing(X, F, V)
    assert abs(slope - want) < 1e-10 * abs(want)


def test_force_pieces_and_planted_mistakes(aniso):
    """md_force = -kappa (p + m); dropping either piece or flipping the
    kappa sign moves the derivative identity far off."""
    lat, k, u, z = aniso
    V = lat.V
    eta, zeta = z["eta_f"], z["zeta_f"]
    fp = lat.md_force_pm_full(u, np.zeros(lat.gauge_size), eta, zeta, 0)
    fm = lat.md_force_pm_full(u, np.zeros(lat.gauge_size), eta, zeta, 1)
    F = lat.md_force_full(u, eta, zeta, k)
    assert rel(F, -k * (fp + fm)) < 1e-15
    U = as_mats(u, V)
    X = random_x(V, 13)

    def S(e):
        return float(np.dot(zeta, lat.dirac_full(from_mats(U + e * (X @ U)), eta, 0, k)))
    slope = (S(0.5) - S(-0.5))
    want = tr_pairing(X, F, V)
    assert abs(slope - want) < 1e-10 * abs(want)
    for wrong in (-k * fp, -k * fm, k * (fp + fm)):
        assert abs(slope - tr_pairing(X, wrong, V)) > 1e-2 * abs(want)


def test_traceless_antihermite_properties(aniso):
    lat, k, u, z = aniso
    V = lat.V
    m = z["mdf_eo"]
    t = as_mats(lat.traceless_antihermite(m), V)
    M = as_mats(m, V)
    assert np.abs(t + np.conj(np.swapaxes(t, -1, -2))).max() < 1e-15
    assert np.abs(np.trace(t, axis1=-2, axis2=-1)).max() < 1e-14
    # closed form: (M - M^dag)/2 - (tr(M - M^dag)/6) 1
    ah = 0.5 * (M - np.conj(np.swapaxes(M, -1, -2)))
    want = ah - (np.trace(ah, axis1=-2, axis2=-1) / 3.0)[..., None, None] * np.eye(3)
    assert np.abs(t - want).max() < 1e-14
    # idempotent
    t2 = as_mats(lat.traceless_antihermite(from_mats(t)), V)
    assert np.abs(t2 - t).max() < 1e-15
