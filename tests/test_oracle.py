"""Pins of the CPU oracle (oracle/liboracle.so) before it is trusted as the
checker of the device path.

 * against the compiled reference: golden fixtures written by
   tests/golden/make_golden.py from IroIro++'s own Dirac_Wilson(_EvenOdd),
   Fopr_DdagD, Solver_CG, MultiShiftSolver_CG, Inverter_WilsonLike and
   AntiPeriodicBC (bit-exact for the operators);
 * against mathematics that fixes the answer independently: Clifford algebra
   of the basis, free-field plane waves, gamma5-hermiticity, gauge
   covariance, a dense textbook Wilson matrix and dense linear solves,
   literature plaquettes of the SU(3) heat-bath.
"""
import numpy as np
import pytest

from oracle import oracle
from tests import wilson_math as wm


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


# ------------------------------------------------------------------ golden
@pytest.mark.parametrize("name", ["wilson_c1_4x4x4x8", "wilson_aniso_4x6x2x4"])
def test_operators_bit_exact_vs_reference(golden, name):
    z = golden(name)
    lat = oracle.Lattice(tuple(z["dims"]))
    k = oracle.kappa_from_mass(float(z["mass"]))
    u = z["u_apbc"]
    src = z["src"]
    fe, fo = lat.split(src)
    got = {
        "deo": lat.hop_eo(u, fo, 1, 0, k), "doe": lat.hop_eo(u, fe, 2, 0, k),
        "deo_dag": lat.hop_eo(u, fe, 2, 1, k), "doe_dag": lat.hop_eo(u, fo, 1, 1, k),
        "dhat": lat.dirac_eo(u, fe, 0, k), "dhat_dag": lat.dirac_eo(u, fe, 1, k),
        "g5": oracle.gamma5(fe), "ddagd": lat.ddagd(u, fe, k),
        "dfull": lat.dirac_full(u, src, 0, k), "dfull_dag": lat.dirac_full(u, src, 1, k),
    }
    for key, v in got.items():
        assert np.array_equal(v, z[key]), key


@pytest.mark.parametrize("name", ["wilson_c1_4x4x4x8", "wilson_aniso_4x6x2x4"])
def test_gauge_generator_and_apbc_vs_reference(golden, name):
    z = golden(name)
    lat = oracle.Lattice(tuple(z["dims"]))
    u_raw = lat.gauge_expH(float(z["eps"]), int(z["seed"]))
    assert np.array_equal(lat.apbc(u_raw), z["u_apbc"])
    # only the last time slice of U_t changed sign
    V = lat.V
    d = (lat.apbc(u_raw) != u_raw).reshape(4, V, 18).any(axis=2)
    t = wm.coords(lat.dims)[:, 3]
    assert not d[:3].any()
    assert np.array_equal(d[3], t == lat.dims[3] - 1)


def test_cg_vs_reference(golden):
    z = golden("wilson_c1_4x4x4x8")
    lat = oracle.Lattice(tuple(z["dims"]))
    k = oracle.kappa_from_mass(float(z["mass"]))
    fe, _ = lat.split(z["src"])
    x, info = lat.cg(z["u_apbc"], k, fe, float(z["cg_prec"]), 1000)
    assert info["converged"]
    assert info["iterations"] == int(z["cg_iterations"])
    # the reference sums with 8 OpenMP threads; only rounding differs
    assert rel(x, z["cg_x"]) < 1e-12
    assert abs(info["diff"] - float(z["cg_diff"])) < 1e-6 * float(z["cg_diff"])


def test_multishift_vs_reference(golden):
    z = golden("wilson_c1_4x4x4x8")
    lat = oracle.Lattice(tuple(z["dims"]))
    k = oracle.kappa_from_mass(float(z["mass"]))
    fe, _ = lat.split(z["src"])
    xs, info = lat.multishift_cg(z["u_apbc"], k, z["ms_shifts"], fe, float(z["cg_prec"]), 1000)
    assert info["converged"]
    assert info["nconv"] == int(z["ms_nconv"])
    for i in range(len(z["ms_shifts"])):
        assert rel(xs[i], z["ms_x"][i]) < 1e-12
    assert abs(info["diff"] - float(z["ms_diff"])) < 1e-6 * float(z["ms_diff"])
    assert abs(info["out_diff"] - float(z["ms_out_diff"])) < 1e-6 * float(z["ms_out_diff"])


def test_invert_vs_reference(golden):
    z = golden("wilson_c1_4x4x4x8")
    lat = oracle.Lattice(tuple(z["dims"]))
    k = oracle.kappa_from_mass(float(z["mass"]))
    sol, info = lat.eo_invert(z["u_apbc"], k, z["src"], float(z["cg_prec"]), 1000)
    assert info["converged"]
    assert rel(sol, z["inv_x"]) < 1e-12


# ------------------------------------------------------------------ basis
def test_dirac_basis_clifford():
    I4 = np.eye(4)
    for m, gm in enumerate(wm.GAMMAS):
        assert np.allclose(gm, gm.conj().T)
        for n, gn in enumerate(wm.GAMMAS):
            assert np.allclose(gm @ gn + gn @ gm, 2 * I4 * (m == n))
    # gamma5 anticommutes with all and squares to one; it equals
    # -g_x g_y g_z g_t in this basis
    g5 = wm.G5
    for gm in wm.GAMMAS:
        assert np.allclose(g5 @ gm + gm @ g5, 0)
    assert np.allclose(g5 @ g5, I4)
    assert np.allclose(g5, -wm.GX @ wm.GY @ wm.GZ @ wm.GT)


# ------------------------------------------------------------------ free field
@pytest.mark.parametrize("apbc", [False, True])
def test_free_field_plane_waves(apbc):
    dims = (4, 6, 4, 8)
    lat = oracle.Lattice(dims)
    u = lat.unit_gauge()
    if apbc:
        u = lat.apbc(u)
    kappa = 0.1355
    rng = np.random.default_rng(3)
    c = wm.coords(dims)
    for n in [(0, 0, 0, 0), (1, 2, 3, 1), (3, 5, 1, 6)]:
        p = np.array([2 * np.pi * n[m] / dims[m] for m in range(4)])
        if apbc:
            p[3] = (2 * n[3] + 1) * np.pi / dims[3]
        spin = rng.normal(size=(4, 3)) + 1j * rng.normal(size=(4, 3))
        phase = np.exp(1j * (c @ p))
        psi = phase[:, None, None] * spin[None]
        M = (1 - 2 * kappa * np.cos(p).sum()) * np.eye(4) + 2j * kappa * sum(
            wm.GAMMAS[m] * np.sin(p[m]) for m in range(4))
        expect = phase[:, None, None] * (M @ spin)[None]
        got = wm.to_cplx_spinor(lat.dirac_full(u, wm.from_cplx_spinor(psi), 0, kappa), lat.V)
        assert np.allclose(got, expect, atol=1e-13)


# ------------------------------------------------------------------ structure
def _random_config(dims, seed):
    lat = oracle.Lattice(dims)
    u = lat.apbc(lat.gauge_expH(0.5, seed))
    return lat, u


def test_gamma5_hermiticity():
    lat, u = _random_config((4, 4, 2, 4), 11)
    k = 0.137
    rng = np.random.default_rng(5)
    psi = rng.normal(size=lat.full_size)
    chi = rng.normal(size=lat.full_size)
    Dpsi = wm.spinor_vec(lat.dirac_full(u, psi, 0, k), lat.V)
    Ddchi = wm.spinor_vec(lat.dirac_full(u, chi, 1, k), lat.V)
    lhs = np.vdot(wm.spinor_vec(chi, lat.V), Dpsi)
    rhs = np.vdot(Ddchi, wm.spinor_vec(psi, lat.V))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # the same on the Schur operator
    fe = rng.normal(size=lat.half_size)
    ge = rng.normal(size=lat.half_size)
    a = np.vdot(wm.spinor_vec(ge, lat.Vh), wm.spinor_vec(lat.dirac_eo(u, fe, 0, k), lat.Vh))
    b = np.vdot(wm.spinor_vec(lat.dirac_eo(u, ge, 1, k), lat.Vh), wm.spinor_vec(fe, lat.Vh))
    assert abs(a - b) < 1e-12 * abs(a)


def test_gauge_covariance():
    dims = (4, 4, 2, 4)
    lat, u = _random_config(dims, 21)
    k = 0.12
    V = lat.V
    rng = np.random.default_rng(9)
    g = wm.random_su3(V, rng)
    U = wm.to_cplx_gauge(u, V)
    Ug = np.empty_like(U)
    for mu in range(4):
        fw = wm.shift_index(dims, mu, +1)
        Ug[mu] = g @ U[mu] @ g[fw].conj().transpose(0, 2, 1)
    ug = wm.from_cplx_gauge(Ug)
    psi = rng.normal(size=lat.full_size)
    P = wm.to_cplx_spinor(psi, V)
    gpsi = wm.from_cplx_spinor(np.einsum("xab,xsb->xsa", g, P))
    lhs = wm.to_cplx_spinor(lat.dirac_full(ug, gpsi, 0, k), V)
    rhs = np.einsum("xab,xsb->xsa", g, wm.to_cplx_spinor(lat.dirac_full(u, psi, 0, k), V))
    assert np.allclose(lhs, rhs, atol=1e-12)


def test_dense_textbook_matrix_and_eo_blocks():
    dims = (4, 2, 2, 4)
    lat, u = _random_config(dims, 31)
    k = 0.125
    V = lat.V
    D = wm.dense_wilson(dims, u, k)
    rng = np.random.default_rng(1)
    psi = rng.normal(size=lat.full_size)
    got = wm.spinor_vec(lat.dirac_full(u, psi, 0, k), V)
    assert np.allclose(got, D @ wm.spinor_vec(psi, V), atol=1e-13)
    # D^dag of the oracle is the conjugate transpose of the dense matrix
    gotd = wm.spinor_vec(lat.dirac_full(u, psi, 1, k), V)
    assert np.allclose(gotd, D.conj().T @ wm.spinor_vec(psi, V), atol=1e-13)
    # even/odd blocks: D = [[1, Deo], [Doe, 1]] restricted to parities
    e, o = lat.eo_tables()
    idx = lambda sites: (12 * sites[:, None] + np.arange(12)[None, :]).ravel()
    De_o = D[np.ix_(idx(e), idx(o))]
    Do_e = D[np.ix_(idx(o), idx(e))]
    fe = rng.normal(size=lat.half_size)
    fo = rng.normal(size=lat.half_size)
    assert np.allclose(wm.spinor_vec(lat.hop_eo(u, fo, 1, 0, k), lat.Vh),
                       De_o @ wm.spinor_vec(fo, lat.Vh), atol=1e-13)
    assert np.allclose(wm.spinor_vec(lat.hop_eo(u, fe, 2, 0, k), lat.Vh),
                       Do_e @ wm.spinor_vec(fe, lat.Vh), atol=1e-13)
    Dhat = np.eye(12 * lat.Vh) - De_o @ Do_e
    assert np.allclose(wm.spinor_vec(lat.dirac_eo(u, fe, 0, k), lat.Vh),
                       Dhat @ wm.spinor_vec(fe, lat.Vh), atol=1e-13)
    assert np.allclose(wm.spinor_vec(lat.dirac_eo(u, fe, 1, k), lat.Vh),
                       Dhat.conj().T @ wm.spinor_vec(fe, lat.Vh), atol=1e-13)


def test_solvers_vs_dense_lu():
    dims = (4, 2, 2, 4)
    lat, u = _random_config(dims, 41)
    k = 0.13
    V = lat.V
    D = wm.dense_wilson(dims, u, k)
    rng = np.random.default_rng(2)
    b = rng.normal(size=lat.full_size)
    x_lu = np.linalg.solve(D, wm.spinor_vec(b, V))
    sol, info = lat.eo_invert(u, k, b, 1e-26, 2000)
    assert info["converged"]
    assert np.allclose(wm.spinor_vec(sol, V), x_lu, atol=1e-10)
    assert lat.true_residual_full(u, k, sol, b) < 1e-11
    # multishift on A = Dhat^dag Dhat
    e, o = lat.eo_tables()
    idx = lambda sites: (12 * sites[:, None] + np.arange(12)[None, :]).ravel()
    Dhat = np.eye(12 * lat.Vh) - D[np.ix_(idx(e), idx(o))] @ D[np.ix_(idx(o), idx(e))]
    A = Dhat.conj().T @ Dhat
    be = rng.normal(size=lat.half_size)
    shifts = np.array([0.0, 0.01, 0.1, 1.0])
    xs, info = lat.multishift_cg(u, k, shifts, be, 1e-24, 2000)
    assert info["converged"]
    for i, s in enumerate(shifts):
        ref = np.linalg.solve(A + s * np.eye(A.shape[0]), wm.spinor_vec(be, lat.Vh))
        assert np.allclose(wm.spinor_vec(xs[i], lat.Vh), ref, atol=1e-9)


# ------------------------------------------------------------------ gauge fields
def _su3_check(u, V):
    U = wm.to_cplx_gauge(u, V).reshape(-1, 3, 3)
    eye = np.eye(3)
    assert np.abs(U @ U.conj().transpose(0, 2, 1) - eye).max() < 1e-12
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-12


def test_expH_is_su3_and_reunit_is_gram_schmidt():
    lat = oracle.Lattice((4, 4, 4, 4))
    u = lat.gauge_expH(0.3, 5)
    _su3_check(u, lat.V)
    rng = np.random.default_rng(0)
    m = rng.normal(size=18)
    r = oracle.reunit(m)
    M = r[0::2] + 1j * r[1::2]
    M = M.reshape(3, 3)
    assert np.allclose(M @ M.conj().T, np.eye(3), atol=1e-14)
    # row 0 is the normalised input row 0 (Gram-Schmidt order)
    m0 = (m[0::2] + 1j * m[1::2]).reshape(3, 3)[0]
    assert np.allclose(M[0], m0 / np.linalg.norm(m0))


def test_uniform_rng_moments():
    u = np.array([oracle.uniform(7, 3, c) for c in range(20000)])
    assert u.min() > 0 and u.max() <= 1
    assert abs(u.mean() - 0.5) < 0.01
    assert abs(u.var() - 1 / 12) < 0.003


def test_heatbath_strong_coupling():
    """<P> = beta/18 + beta^2/216 + O(beta^3) for SU(3) at small beta."""
    lat = oracle.Lattice((4, 4, 4, 4))
    u = lat.unit_gauge()
    beta = 0.5
    vals = []
    for sw in range(60):
        lat.heatbath_sweep(u, beta, 123, sw)
        if sw >= 20:
            vals.append(lat.plaquette(u))
    _su3_check(u, lat.V)
    expect = beta / 18 + beta ** 2 / 216
    assert abs(np.mean(vals) - expect) < 0.004


@pytest.mark.slow
@pytest.mark.parametrize("beta,plaq", [(5.8, 0.5677), (6.0, 0.5937)])
def test_heatbath_literature_plaquette(beta, plaq):
    """Quenched Wilson-action plaquette (large-volume literature values)."""
    lat = oracle.Lattice((6, 6, 6, 6))
    u = lat.unit_gauge()
    vals = []
    for sw in range(120):
        lat.heatbath_sweep(u, beta, 99, sw)
        if sw >= 60:
            vals.append(lat.plaquette(u))
    _su3_check(u, lat.V)
    assert abs(np.mean(vals) - plaq) < 0.006
