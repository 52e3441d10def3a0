"""Parity of the device MD force (Dirac_Wilson(_EvenOdd)::md_force,
md_force_p/_m, FieldUtils::TracelessAntihermite, Action_Nf2::md_force)
with the reference's own outputs (golden fixtures), with the oracle, and at
larger sizes with the link-derivative identity

    (S(eps) - S(-eps)) / (2 eps) = Re sum_{x,mu} tr(X_mu(x) F_mu(x)),
    S(eps) = Re<zeta, Dhat[U + eps X U] eta>,

exact for the even/odd operator because S is quadratic in eps.  FP64
throughout: the bar is 1e-13 relative (2-norm) against the reference and
the oracle; the derivative identity holds to 1e-9 relative (the two sides
are sums of ~1e6 terms computed in different orders)."""
import numpy as np
import pytest

import iroiro_amd as ii
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-13


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


@pytest.fixture(scope="module")
def aniso(golden):
    base = golden("wilson_aniso_4x6x2x4")
    z = golden("wilson_aniso_4x6x2x4_mdforce")
    ctx = ii.Context(tuple(base["dims"]))
    yield base, z, ctx
    ctx.close()


def test_eo_md_force_vs_reference(aniso):
    base, z, ctx = aniso
    D = ii.Dirac_Wilson_EvenOdd(float(base["mass"]), base["u_apbc"], ctx)
    f = D.md_force(z["eta_e"], z["zeta_e"])
    assert rel(f, z["mdf_eo"]) < TOL, rel(f, z["mdf_eo"])
    ta = ii.TracelessAntihermite(ctx, z["mdf_eo"])
    assert rel(ta, z["ta_eo"]) < TOL, rel(ta, z["ta_eo"])


def test_full_md_force_and_pieces_vs_reference(aniso):
    base, z, ctx = aniso
    Df = ii.Dirac_Wilson(float(base["mass"]), base["u_apbc"], ctx)
    f = Df.md_force(z["eta_f"], z["zeta_f"])
    assert rel(f, z["mdf_full"]) < TOL, rel(f, z["mdf_full"])
    # md_force_p / md_force_m add into an existing field (Field::add)
    g = np.random.default_rng(3).standard_normal(Df.gsize())
    fp = g.copy()
    Df.md_force_p(fp, z["eta_f"], z["zeta_f"])
    assert rel(fp - g, z["mdf_p"]) < 1e-12, rel(fp - g, z["mdf_p"])
    fm = np.zeros(Df.gsize())
    Df.md_force_m(fm, z["eta_f"], z["zeta_f"])
    assert rel(fm, z["mdf_m"]) < TOL, rel(fm, z["mdf_m"])


def test_eo_md_force_vs_oracle_c1(golden):
    z = golden("wilson_c1_4x4x4x8")
    dims = tuple(z["dims"])
    lat = oracle.Lattice(dims)
    k = oracle.kappa_from_mass(float(z["mass"]))
    eta = oracle.gauss_field(lat.half_size, 901)
    zeta = oracle.gauss_field(lat.half_size, 902)
    ctx = ii.Context(dims)
    D = ii.Dirac_Wilson_EvenOdd(float(z["mass"]), z["u_apbc"], ctx)
    f = D.md_force(eta, zeta)
    want = lat.md_force_eo(z["u_apbc"], eta, zeta, k)
    assert rel(f, want) < TOL, rel(f, want)
    # a second call on the same context reuses its buffers
    f2 = D.md_force(zeta, eta)
    assert rel(f2, lat.md_force_eo(z["u_apbc"], zeta, eta, k)) < TOL
    ctx.close()


def test_action_nf2_md_force_vs_oracle(golden):
    """Action_Nf2::md_force: eta = (Dhat^dag Dhat)^-1 phi, zeta = Dhat eta,
    TracelessAntihermite(md_force(eta, zeta)), against the oracle chain."""
    z = golden("wilson_c1_4x4x4x8")
    dims = tuple(z["dims"])
    lat = oracle.Lattice(dims)
    k = oracle.kappa_from_mass(float(z["mass"]))
    u = z["u_apbc"]
    ctx = ii.Context(dims)
    D = ii.Dirac_Wilson_EvenOdd(float(z["mass"]), u, ctx)
    act = ii.Action_Nf2(D, ii.Solver_CG(1e-26, 1000, ii.Fopr_DdagD(D)))
    phi = oracle.gauss_field(lat.half_size, 903)
    act.set_phi(phi)
    F = act.md_force()
    eta, info = lat.cg(u, k, phi, 1e-26, 1000)
    assert info["converged"]
    assert act.last_solve.Iterations == info["iterations"]
    want = lat.traceless_antihermite(lat.md_force_eo(u, eta, lat.dirac_eo(u, eta, 0, k), k))
    assert rel(F, want) < 1e-11, rel(F, want)
    # traceless and anti-hermitian
    M = as_mats(F, lat.V)
    assert np.abs(M + np.conj(np.swapaxes(M, -1, -2))).max() < 1e-14 * np.abs(M).max()
    assert np.abs(np.trace(M, axis1=-2, axis2=-1)).max() < 1e-13 * np.abs(M).max()
    ctx.close()


def test_md_force_derivative_identity_16c32():
    """At 16^3 x 32 on a beta = 5.8 heat-bath field, the device force is the
    derivative of the device operator's bilinear form."""
    dims = (16, 16, 16, 32)
    ctx = ii.Context(dims)
    ctx.unit_gauge()
    ctx.heatbath(5.8, 4242, 4)
    ctx.apply_apbc()
    u = ctx.download_gauge()
    mass = 1.0 / (2 * 0.1355) - 4.0
    D = ii.Dirac_Wilson_EvenOdd(mass, u, ctx)
    rng = np.random.default_rng(77)
    eta = rng.standard_normal(D.fsize())
    zeta = rng.standard_normal(D.fsize())
    F = D.md_force(eta, zeta)
    V = ctx.V
    U = as_mats(u, V)
    X = rng.standard_normal((4, V, 3, 3)) + 1j * rng.standard_normal((4, V, 3, 3))
    want = float(np.real(np.einsum("mxab,mxba->", X, as_mats(F, V))))
    eps = 1.0 / 64
    s = []
    for e in (eps, -eps):
        ctx.upload_gauge(from_mats(U + e * (X @ U)))
        s.append(float(np.dot(zeta, D.mult(eta))))
    ctx.upload_gauge(u)
    slope = (s[0] - s[1]) / (2 * eps)
    assert abs(slope - want) < 1e-9 * abs(want), (slope, want)
    # the same identity fails for the force with a flipped sign or a
    # dropped half (sanity of the test's sensitivity)
    assert abs(slope + want) > 0.5 * abs(want)
    ctx.close()
