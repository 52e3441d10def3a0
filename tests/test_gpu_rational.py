"""RationalSolver_CG and Action_Nf::md_force on the device, against the
oracle (its multishift, Schur operator and MD force, each pinned elsewhere)
and against what the mathematics fixes:
 * applying the x^(1/2) approximation twice gives Dhat^dag Dhat src, and
   x^(-1/2) after x^(1/2) gives src back, to the approximation's error
   (the reference's own check, tests/test_RationalApprox.cpp:160-190);
 * the RHMC force is the derivative of its action: with
   S(U) = sum_i r_i phi . (A[U] + s_i)^-1 phi and U -> U + eps X U for
   traceless anti-hermitian X, dS/deps = -2 Re sum tr(X F) where F is the
   projected force (the projection drops nothing that pairs with such X).
FP64 paths: 1e-10 relative against the oracle (two multishift solves
stopped at |r|^2/|b|^2 < 1e-26); mixed precision: to its true-residual
target."""
import ctypes

import numpy as np
import pytest

import iroiro_amd as ii
from oracle import oracle

pytestmark = pytest.mark.gpu


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
def c1(golden):
    z = golden("wilson_c1_4x4x4x8")
    dims = tuple(z["dims"])
    ctx = ii.Context(dims)
    D = ii.Dirac_Wilson_EvenOdd(float(z["mass"]), z["u_apbc"], ctx)
    lat = oracle.Lattice(dims)
    k = oracle.kappa_from_mass(float(z["mass"]))
    # the spectrum of Dhat^dag Dhat at kappa = 0.12 lies in
    # [(1 - 64 k^2)^2, (1 + 64 k^2)^2] = [0.0061, 3.70]
    ra = ii.RationalApprox(10, 10, 1, 2, 50, 1e-3, 4.0)
    yield z, ctx, D, lat, k, ra
    ctx.close()


def oracle_rational(lat, u, k, a0, res, poles, b, prec=1e-26):
    xs, info = lat.multishift_cg(u, k, poles, b, prec, 5000)
    assert info["converged"]
    return a0 * b + np.tensordot(np.asarray(res), xs, axes=1)


def test_rational_solve_vs_oracle(c1):
    z, ctx, D, lat, k, ra = c1
    b = oracle.gauss_field(lat.half_size, 1201)
    slv = ii.RationalSolver_CG(ii.MultiShiftSolver_CG(ii.Fopr_DdagD(D), 1e-26, 5000), ra)
    sol = np.empty_like(b)
    out = slv.solve(sol, b)
    want = oracle_rational(lat, z["u_apbc"], k, ra.Const(), ra.Residuals(), ra.Poles(), b)
    assert rel(sol, want) < 1e-10, rel(sol, want)
    assert out.Iterations > 0
    sol_inv = np.empty_like(b)
    slv.solve_inv(sol_inv, b)
    want = oracle_rational(lat, z["u_apbc"], k, ra.InvConst(), ra.InvResiduals(), ra.InvPoles(),
                           b)
    assert rel(sol_inv, want) < 1e-10, rel(sol_inv, want)
    # no-reconstruct variants: the multishift solutions themselves
    shifted = []
    slv.solve_noReconstruct(shifted, b)
    xs, _ = lat.multishift_cg(z["u_apbc"], k, ra.InvPoles(), b, 1e-26, 5000)
    assert len(shifted) == len(ra.InvPoles())
    for a, w in zip(shifted, xs):
        assert rel(a, w) < 1e-10


def test_square_root_applied_twice(c1):
    """The reference test's check: (x^(1/2))^2 src = Dhat^dag Dhat src."""
    z, ctx, D, lat, k, ra = c1
    b = oracle.gauss_field(lat.half_size, 1202)
    slv = ii.RationalSolver_CG(ii.MultiShiftSolver_CG(ii.Fopr_DdagD(D), 1e-26, 5000), ra)
    s1, s2, back = np.empty_like(b), np.empty_like(b), np.empty_like(b)
    slv.solve(s1, b)
    slv.solve(s2, s1)
    ddagd = lat.ddagd(z["u_apbc"], b, k)
    assert rel(s2, ddagd) < 10 * ra.error, (rel(s2, ddagd), ra.error)
    slv.solve_inv(back, s1)
    assert rel(back, b) < 10 * ra.error
    # and the approximation is not exact: a planted wrong constant term shows
    slv.ConstTerm *= 1 + 1e-6
    slv.solve(s1, b)
    slv.solve(s2, s1)
    assert rel(s2, ddagd) > 10 * ra.error


@pytest.mark.parametrize("precision", ["single", "half"])
def test_rational_solve_mixed_precision(c1, precision):
    z, ctx, D, lat, k, ra = c1
    b = oracle.gauss_field(lat.half_size, 1203)
    ref = np.empty_like(b)
    ii.RationalSolver_CG(ii.MultiShiftSolver_CG(ii.Fopr_DdagD(D), 1e-26, 5000), ra).solve(ref, b)
    sol = np.empty_like(b)
    out = ii.RationalSolver_CG(ii.MultiShiftSolver_CG(ii.Fopr_DdagD(D), 1e-10, 5000), ra,
                               precision=precision).solve(sol, b)
    assert out.true_residual <= 1e-10
    # each shifted solution to 1e-10 true residual; the sum of positive-
    # and negative-residue terms keeps the error at that level
    assert rel(sol, ref) < 1e-8, rel(sol, ref)


def rhmc_action(D, ctx, k):
    ms = ii.MultiShiftSolver_CG(ii.Fopr_DdagD(D), 1e-26, 5000)
    slv = ii.RationalSolver_CG(ms)
    # Nf = 1 with one pseudofermion: Metropolis and MD x^(1/2), pseudofermions
    # x^(1/4); degree 8 on [1e-3, 4]
    act = ii.Action_Nf(D, slv, n_flav=1, n_pseudof=1, degree=[8, 8, 8],
                       precision=[50, 50, 50], b_low=[1e-3] * 3, b_high=[4.0] * 3)
    return act, slv


def test_action_nf_md_force_vs_oracle(c1):
    z, ctx, D, lat, k, ra = c1
    u = z["u_apbc"]
    act, slv = rhmc_action(D, ctx, k)
    phi = oracle.gauss_field(lat.half_size, 1204)
    act.phi_[0][...] = phi
    F = act.md_force()
    md = act.MolecularDynApprox_
    xs, info = lat.multishift_cg(u, k, md.InvPoles(), phi, 1e-26, 5000)
    want = np.zeros(lat.gauge_size)
    for r, x in zip(md.InvResiduals(), xs):
        want += r * lat.md_force_eo(u, x, lat.dirac_eo(u, x, 0, k), k)
    want = lat.traceless_antihermite(want)
    assert rel(F, want) < 1e-10, rel(F, want)


def test_action_nf_md_force_is_the_action_derivative(c1):
    z, ctx, D, lat, k, ra = c1
    u = z["u_apbc"]
    V = ctx.V
    act, slv = rhmc_action(D, ctx, k)
    phi = oracle.gauss_field(lat.half_size, 1205)
    act.phi_[0][...] = phi
    F = act.md_force()
    md = act.MolecularDynApprox_
    rng = np.random.default_rng(5)
    H = rng.standard_normal((4, V, 3, 3)) + 1j * rng.standard_normal((4, V, 3, 3))
    X = 0.5 * (H - np.conj(np.swapaxes(H, -1, -2)))
    X -= (np.trace(X, axis1=-2, axis2=-1) / 3.0)[..., None, None] * np.eye(3)
    U = as_mats(u, V)
    slv.set_Approx(md)

    def S(eps):
        ctx.upload_gauge(from_mats(U + eps * (X @ U)))
        t = np.empty_like(phi)
        slv.solve_inv(t, phi)
        return float(np.dot(phi, t))
    eps = 1e-4
    slope = (S(eps) - S(-eps)) / (2 * eps)
    ctx.upload_gauge(u)
    want = -2.0 * float(np.real(np.einsum("mxab,mxba->", X, as_mats(F, V))))
    assert abs(slope - want) < 1e-6 * abs(want), (slope, want)


def test_action_nf_single_zero_pole_is_action_nf2(c1):
    """With one pole at 0 and residue 1, the RHMC force is Action_Nf2's."""
    z, ctx, D, lat, k, ra = c1
    phi = oracle.gauss_field(lat.half_size, 1206)
    f1 = np.empty(D.gsize())
    out = ii._lib.SolverOutputC()
    one, zero = np.array([1.0]), np.array([0.0])
    st = ii._lib.load().ii_action_nf_md_force(ctx.handle, D.kpp, ii._lib.dptr(phi), 1,
                                              ii._lib.dptr(one), ii._lib.dptr(zero), 1, 1e-28,
                                              5000, ii._lib.II_PREC_DOUBLE, ii._lib.dptr(f1),
                                              ctypes.byref(out))
    assert st == 0
    act2 = ii.Action_Nf2(D, ii.Solver_CG(1e-28, 5000, ii.Fopr_DdagD(D)))
    act2.set_phi(phi)
    f2 = act2.md_force()
    assert rel(f1, f2) < 1e-10, rel(f1, f2)
