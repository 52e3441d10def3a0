"""Parity of the HIP path with the reference (golden fixtures) and with the
oracle, through the C ABI.  FP64 operators: 1e-13 relative (the stated
bar); solvers: same iteration counts as the reference and solutions to
rounding; mixed precision: the FP64 true residual target is met."""
import numpy as np
import pytest

import iroiro_amd as ii
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL_OP = 1e-13   # FP64 operator parity bar (relative, 2-norm)


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def c1(golden):
    z = golden("wilson_c1_4x4x4x8")
    ctx = ii.Context(tuple(z["dims"]))
    D = ii.Dirac_Wilson_EvenOdd(float(z["mass"]), z["u_apbc"], ctx)
    yield z, ctx, D
    ctx.close()


@pytest.mark.parametrize("name", ["wilson_c1_4x4x4x8", "wilson_aniso_4x6x2x4"])
def test_operators_vs_reference(golden, name):
    z = golden(name)
    dims = tuple(z["dims"])
    ctx = ii.Context(dims)
    D = ii.Dirac_Wilson_EvenOdd(float(z["mass"]), z["u_apbc"], ctx)
    Df = ii.Dirac_Wilson(float(z["mass"]), z["u_apbc"], ctx)
    lat = oracle.Lattice(dims)
    fe, fo = lat.split(z["src"])
    got = {
        "deo": D.mult_eo(fo), "doe": D.mult_oe(fe), "deo_dag": D.mult_eo_dag(fe),
        "doe_dag": D.mult_oe_dag(fo), "dhat": D.mult(fe), "dhat_dag": D.mult_dag(fe),
        "g5": D.gamma5(fe), "ddagd": ii.Fopr_DdagD(D).mult(fe),
        "dfull": Df.mult(z["src"]), "dfull_dag": Df.mult_dag(z["src"]),
    }
    for k, v in got.items():
        assert rel(v, z[k]) < TOL_OP, (k, rel(v, z[k]))
    ctx.close()


def test_cg_vs_reference(c1):
    z, ctx, D = c1
    lat = oracle.Lattice(tuple(z["dims"]))
    fe, _ = lat.split(z["src"])
    x = np.empty_like(fe)
    out = ii.Solver_CG(float(z["cg_prec"]), 1000, ii.Fopr_DdagD(D)).solve(x, fe)
    assert out.Iterations == int(z["cg_iterations"])
    assert rel(x, z["cg_x"]) < 1e-12
    assert abs(out.diff - float(z["cg_diff"])) < 1e-4 * float(z["cg_diff"])


def test_cg_is_deterministic(c1):
    z, ctx, D = c1
    lat = oracle.Lattice(tuple(z["dims"]))
    fe, _ = lat.split(z["src"])
    s = ii.Solver_CG(1e-20, 1000, ii.Fopr_DdagD(D))
    x1, x2 = np.empty_like(fe), np.empty_like(fe)
    s.solve(x1, fe)
    s.solve(x2, fe)
    assert np.array_equal(x1, x2)


def test_multishift_vs_reference(c1):
    z, ctx, D = c1
    lat = oracle.Lattice(tuple(z["dims"]))
    fe, _ = lat.split(z["src"])
    shifts = z["ms_shifts"]
    xs = [np.empty_like(fe) for _ in shifts]
    so, diff, nconv = ii.MultiShiftSolver_CG(ii.Fopr_DdagD(D), float(z["cg_prec"]), 1000).solve(
        xs, fe, shifts)
    assert nconv == int(z["ms_nconv"])
    for i in range(len(shifts)):
        assert rel(xs[i], z["ms_x"][i]) < 1e-12
    assert abs(diff - float(z["ms_diff"])) < 1e-4 * float(z["ms_diff"])
    assert abs(so.diff - float(z["ms_out_diff"])) < 1e-4 * float(z["ms_out_diff"])


def test_invert_vs_reference(c1):
    z, ctx, D = c1
    sol = np.empty(24 * ctx.V)
    ii.Inverter_WilsonLike(D, ii.Solver_CG(float(z["cg_prec"]), 1000, ii.Fopr_DdagD(D))).invert(
        sol, z["src"])
    assert rel(sol, z["inv_x"]) < 1e-12


def test_not_converged_raises(c1):
    z, ctx, D = c1
    lat = oracle.Lattice(tuple(z["dims"]))
    fe, _ = lat.split(z["src"])
    x = np.zeros_like(fe)
    with pytest.raises(ii.ConvergenceErr):
        ii.Solver_CG(1e-20, 3, ii.Fopr_DdagD(D)).solve(x, fe)
    assert np.linalg.norm(x) > 0  # the iterate is still written back
    xs = [np.empty_like(fe) for _ in range(2)]
    with pytest.raises(ii.NotConverged):
        ii.MultiShiftSolver_CG(ii.Fopr_DdagD(D), 1e-20, 3).solve(xs, fe, [0.0, 0.1])


def test_invalidated_gauge_is_an_error(golden):
    z = golden("wilson_c1_4x4x4x8")
    ctx = ii.Context(tuple(z["dims"]))
    D = ii.Dirac_Wilson_EvenOdd(float(z["mass"]), z["u_apbc"], ctx)
    ctx.invalidate_gauge()
    with pytest.raises(ii.IIError):
        D.mult(np.zeros(D.fsize()))
    D.update_internal_state()
    assert np.all(D.mult(np.zeros(D.fsize())) == 0)
    ctx.close()


@pytest.mark.parametrize("dims", [(8, 8, 8, 16), (6, 4, 10, 8), (16, 8, 8, 4)])
def test_operators_vs_oracle_larger(dims):
    lat = oracle.Lattice(dims)
    u = lat.apbc(lat.gauge_expH(0.7, 1234))
    kappa = 0.1355
    mass = 0.5 / kappa - 4.0
    ctx = ii.Context(dims)
    D = ii.Dirac_Wilson_EvenOdd(mass, u, ctx)
    k = D.getKappa()
    rng = np.random.default_rng(8)
    fe = rng.normal(size=lat.half_size)
    fo = rng.normal(size=lat.half_size)
    assert rel(D.mult_eo(fo), lat.hop_eo(u, fo, 1, 0, k)) < TOL_OP
    assert rel(D.mult_oe(fe), lat.hop_eo(u, fe, 2, 0, k)) < TOL_OP
    assert rel(D.mult_eo_dag(fe), lat.hop_eo(u, fe, 2, 1, k)) < TOL_OP
    assert rel(D.mult(fe), lat.dirac_eo(u, fe, 0, k)) < TOL_OP
    assert rel(D.mult_dag(fe), lat.dirac_eo(u, fe, 1, k)) < TOL_OP
    assert rel(ii.Fopr_DdagD(D).mult(fe), lat.ddagd(u, fe, k)) < TOL_OP
    full = rng.normal(size=lat.full_size)
    Df = ii.Dirac_Wilson(mass, u, ctx)
    assert rel(Df.mult(full), lat.dirac_full(u, full, 0, k)) < TOL_OP
    # gauge round trip and plaquette
    assert np.array_equal(ctx.download_gauge(), u)
    assert abs(ctx.plaquette() - lat.plaquette(u)) < 1e-13
    ctx.close()


def _prop_setup(dims, seed=5, eps=0.6):
    lat = oracle.Lattice(dims)
    u = lat.apbc(lat.gauge_expH(eps, seed))
    return lat, u


@pytest.mark.parametrize("precision,batched", [("double", False), ("double", True),
                                               ("single", True), ("half", True), ("bf16", True),
                                               ("half", False)])
def test_propagator_true_residual(precision, batched):
    dims = (8, 8, 8, 8)
    lat, u = _prop_setup(dims)
    kappa = 0.1355
    ctx = ii.Context(dims)
    ctx.upload_gauge(u)
    src = ii.DeviceField(ctx, "full", 12)
    sol = ii.DeviceField(ctx, "full", 12)
    src.point_sources((0, 0, 0, 0))
    outs = ii.PropagatorSolver(ctx, kappa, tol=1e-10, precision=precision,
                               batched=batched).solve(src, sol)
    for rhs in (0, 5, 11):
        b = src.download(rhs)
        x = sol.download(rhs)
        s, c = divmod(rhs, 3)
        expect_b = np.zeros(lat.full_size)
        expect_b[2 * (c + 3 * s)] = 1.0
        assert np.array_equal(b, expect_b)
        r_oracle = lat.true_residual_full(u, kappa, x, b)
        assert r_oracle <= 1e-10, (rhs, r_oracle)
        assert abs(outs[rhs].true_residual - r_oracle) < 1e-12
        # error against the FP64 oracle solution
        xo, info = lat.eo_invert(u, kappa, b, 1e-28, 5000)
        assert rel(x, xo) < 1e-8
    ctx.close()


@pytest.mark.parametrize("precision,batched", [("half", True), ("bf16", True), ("single", True),
                                               ("half", False)])
def test_propagator_bicgstab(precision, batched):
    """BiCGStab on Dhat (sloppy precision, FP64 reliable updates): the 12
    point-source columns reach the true residual 1e-10 on the full operator
    (checked with the oracle) and agree with the oracle's FP64 solution."""
    dims = (8, 8, 8, 8)
    lat, u = _prop_setup(dims)
    kappa = 0.1355
    ctx = ii.Context(dims)
    ctx.upload_gauge(u)
    src = ii.DeviceField(ctx, "full", 12)
    sol = ii.DeviceField(ctx, "full", 12)
    src.point_sources((0, 0, 0, 0))
    outs = ii.PropagatorSolver(ctx, kappa, tol=1e-10, precision=precision, batched=batched,
                               method="bicgstab").solve(src, sol)
    for rhs in (0, 4, 11):
        b = src.download(rhs)
        x = sol.download(rhs)
        r_oracle = lat.true_residual_full(u, kappa, x, b)
        assert r_oracle <= 1e-10, (rhs, r_oracle)
        assert abs(outs[rhs].true_residual - r_oracle) < 1e-12
        xo, info = lat.eo_invert(u, kappa, b, 1e-28, 5000)
        assert rel(x, xo) < 1e-8
    ctx.close()


def test_bicgstab_fewer_iterations_than_cg_and_maxiter():
    """On the same point sources BiCGStab (two Dhat per iteration) needs fewer
    Dhat applications than CG on Dhat^dag Dhat (two Dhat and two Dhat^dag per
    iteration); a maxiter below that count is reported as not converged."""
    dims = (8, 8, 8, 8)
    lat, u = _prop_setup(dims)
    ctx = ii.Context(dims)
    ctx.upload_gauge(u)
    src = ii.DeviceField(ctx, "full", 12)
    sol = ii.DeviceField(ctx, "full", 12)
    src.point_sources((0, 0, 0, 0))
    cg = ii.PropagatorSolver(ctx, 0.1355, precision="half", batched=True).solve(src, sol)
    bi = ii.PropagatorSolver(ctx, 0.1355, precision="half", batched=True,
                             method="bicgstab").solve(src, sol)
    n_cg = max(o.Iterations for o in cg)
    n_bi = max(o.Iterations for o in bi)
    assert 2 * n_bi < 2 * n_cg, (n_bi, n_cg)
    with pytest.raises(ii.IIError):
        ii.PropagatorSolver(ctx, 0.1355, precision="half", batched=True, method="bicgstab",
                            maxiter=max(2, n_bi // 3)).solve(src, sol)
    ctx.close()


def test_bicgstab_rejects_double():
    ctx = ii.Context((4, 4, 4, 4))
    ctx.unit_gauge()
    src = ii.DeviceField(ctx, "full", 12)
    sol = ii.DeviceField(ctx, "full", 12)
    src.point_sources((0, 0, 0, 0))
    with pytest.raises(ii.IIError):
        ii.PropagatorSolver(ctx, 0.12, precision="double", method="bicgstab").solve(src, sol)
    ctx.close()


@pytest.mark.parametrize("nrhs,method", [(5, "cg"), (13, "cg"), (5, "bicgstab"), (13, "bicgstab")])
def test_propagator_ragged_batches(nrhs, method):
    """Batched half-precision solves of 5 and 13 Gaussian sources (13: two
    rhs row-blocks of the multi-rhs kernel); every column against the
    oracle's FP64 operator and solution."""
    dims = (8, 8, 4, 8)
    lat, u = _prop_setup(dims, seed=8)
    kappa = 0.1355
    ctx = ii.Context(dims)
    ctx.upload_gauge(u)
    src = ii.DeviceField(ctx, "full", nrhs)
    sol = ii.DeviceField(ctx, "full", nrhs)
    rng = np.random.default_rng(nrhs)
    bs = rng.standard_normal((nrhs, lat.full_size))
    for k in range(nrhs):
        src.upload(k, bs[k])
    outs = ii.PropagatorSolver(ctx, kappa, tol=1e-10, precision="half",
                               batched=True, method=method).solve(src, sol)
    for k in range(nrhs):
        x = sol.download(k)
        r_oracle = lat.true_residual_full(u, kappa, x, bs[k])
        assert r_oracle <= 1e-10, (k, r_oracle)
        assert abs(outs[k].true_residual - r_oracle) < 1e-12
        xo, info = lat.eo_invert(u, kappa, bs[k], 1e-28, 5000)
        assert rel(x, xo) < 1e-8
    ctx.close()


@pytest.mark.parametrize("method", ["cg", "bicgstab"])
def test_forced_recursion_restarts_converge(method):
    """restart_factor = 1e-300 makes every reliable update take the branch that
    restarts the search direction (a recomputed residual above restart_factor x
    the recursive one).  The solve still reaches the true residual, every
    update but the converging one restarted, and the solutions match the
    oracle's FP64 solution: after a restart the recursion scalars (rho = |r|^2
    for BiCGStab, r0 = p = r) stay consistent with the copied vectors."""
    dims = (8, 8, 8, 8)
    lat, u = _prop_setup(dims, seed=9)
    kappa = 0.1355
    ctx = ii.Context(dims)
    ctx.upload_gauge(u)
    src = ii.DeviceField(ctx, "full", 12)
    sol = ii.DeviceField(ctx, "full", 12)
    src.point_sources((0, 0, 0, 0))
    outs = ii.PropagatorSolver(ctx, kappa, tol=1e-10, precision="half", batched=True,
                               method=method, restart_factor=1e-300).solve(src, sol)
    for k, o in enumerate(outs):
        assert o.recursion_restarts >= 1, (k, o.recursion_restarts)
        assert o.recursion_restarts == o.restarts - 1, (k, o.recursion_restarts, o.restarts)
        assert o.true_residual <= 1e-10
    for rhs in (0, 5, 11):
        b = src.download(rhs)
        x = sol.download(rhs)
        assert lat.true_residual_full(u, kappa, x, b) <= 1e-10
        xo, _ = lat.eo_invert(u, kappa, b, 1e-28, 5000)
        assert rel(x, xo) < 1e-8
    ctx.close()


def test_r12_rejects_non_su3():
    dims = (4, 4, 4, 4)
    lat = oracle.Lattice(dims)
    u = lat.gauge_expH(0.3, 3) * 1.01  # not unitary
    ctx = ii.Context(dims)
    ctx.upload_gauge(u)
    src = ii.DeviceField(ctx, "full", 12)
    sol = ii.DeviceField(ctx, "full", 12)
    src.point_sources((0, 0, 0, 0))
    with pytest.raises(ii.IIError):
        ii.PropagatorSolver(ctx, 0.12, precision="half", batched=True).solve(src, sol)
    ctx.close()


def test_heatbath_matches_oracle():
    """Same counter-based random numbers on both sides: two sweeps from a
    cold start agree link by link up to rounding (a rare accept/reject flip
    from last-bit differences of log/cos is tolerated on a few links)."""
    dims = (4, 4, 4, 8)
    lat = oracle.Lattice(dims)
    u = lat.unit_gauge()
    for sw in range(2):
        lat.heatbath_sweep(u, 6.0, 77, sw)
    ctx = ii.Context(dims)
    ctx.unit_gauge()
    ctx.heatbath(6.0, 77, 2)
    ug = ctx.download_gauge()
    d = np.abs(ug - u).reshape(-1, 18).max(axis=1)
    assert np.mean(d < 1e-10) > 0.99, np.sort(d)[-5:]
    assert abs(ctx.plaquette() - lat.plaquette(u)) < 1e-3
    ctx.close()


def test_heatbath_plaquette_and_su3():
    dims = (8, 8, 8, 8)
    ctx = ii.Context(dims)
    ctx.unit_gauge()
    vals = []
    for sw in range(0, 160, 10):
        ctx.heatbath(6.0, 5, 10, first_sweep=sw)
        if sw >= 60:
            vals.append(ctx.plaquette())
    assert abs(np.mean(vals) - 0.5937) < 0.004
    u = ctx.download_gauge().reshape(4 * ctx.V, 3, 3, 2)
    U = u[..., 0] + 1j * u[..., 1]
    assert np.abs(U @ U.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-12
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-12
    # apbc flips exactly the last t-slice of U_t
    ctx.apply_apbc()
    u2 = ctx.download_gauge().reshape(4, ctx.V, 18)
    t = np.arange(ctx.V) // (8 * 8 * 8)
    flipped = np.all(u2 == -u.reshape(4, ctx.V, 18), axis=2)
    assert np.array_equal(flipped[3], t == 7) and not flipped[:3].any()
    ctx.close()


def test_batched_fp64_equals_rhs_by_rhs():
    """The fused multi-rhs kernels do per-site arithmetic identical to the
    single-rhs ones: batched and rhs-by-rhs FP64 solves take the same
    iterations and agree to rounding of the reductions."""
    dims = (8, 8, 8, 8)
    lat, u = _prop_setup(dims, seed=9)
    ctx = ii.Context(dims)
    ctx.upload_gauge(u)
    src = ii.DeviceField(ctx, "full", 12)
    s1 = ii.DeviceField(ctx, "full", 12)
    s2 = ii.DeviceField(ctx, "full", 12)
    src.point_sources((1, 2, 3, 4))
    o1 = ii.PropagatorSolver(ctx, 0.13, precision="double", batched=False).solve(src, s1)
    o2 = ii.PropagatorSolver(ctx, 0.13, precision="double", batched=True).solve(src, s2)
    for k in range(12):
        assert abs(o1[k].Iterations - o2[k].Iterations) <= 1
        assert rel(s2.download(k), s1.download(k)) < 1e-11
    ctx.close()


def test_copy_bandwidth_runs():
    ctx = ii.Context((8, 8, 8, 8))
    ms = ctx.bench_copy(64 << 20, 5)
    assert ms > 0
    ctx.close()


@pytest.mark.parametrize("precision,method", [("half", "cg"), ("bf16", "cg"), ("single", "cg"),
                                              ("half", "bicgstab"), ("bf16", "bicgstab")])
def test_point_source_heatbath_16c32(precision, method):
    """Point sources on a beta=6.0 heat-bath field at 16^3 x 32: far from the
    source the early Krylov vectors underflow to FP32 denormals (values
    ~kappa^40), the regime where a 16-bit site norm must not overflow."""
    dims = (16, 16, 16, 32)
    ctx = ii.Context(dims)
    ctx.unit_gauge()
    ctx.heatbath(6.0, 11, 30)
    ctx.apply_apbc()
    src = ii.DeviceField(ctx, "full", 12)
    sol = ii.DeviceField(ctx, "full", 12)
    src.point_sources((0, 0, 0, 0))
    outs = ii.PropagatorSolver(ctx, 0.1355, tol=1e-10, precision=precision,
                               batched=True, method=method).solve(src, sol)
    assert max(o.true_residual for o in outs) <= 1e-10
    # check one column against the oracle on the same (downloaded) field
    u = ctx.download_gauge()
    lat = oracle.Lattice(dims)
    b = src.download(7)
    r = lat.true_residual_full(u, 0.1355, sol.download(7), b)
    assert r <= 1e-10
    ctx.close()


def test_multishift_full_operator_vs_oracle():
    """MultiShiftSolver_CG on Fopr_DdagD of the full-index Dirac_Wilson (the
    operator of Qprop_MultiShift): every shifted system is solved, checked
    with the oracle's full operator."""
    dims = (4, 4, 4, 8)
    lat = oracle.Lattice(dims)
    u = lat.apbc(lat.gauge_expH(0.3, 31))
    mass = 1.0 / 6.0
    kappa = oracle.kappa_from_mass(mass)
    ctx = ii.Context(dims)
    try:
        Df = ii.Dirac_Wilson(mass, u, ctx)
        b = oracle.gauss_field(lat.full_size, 5)
        shifts = [0.0, 0.02, 0.1, 0.5]
        xs = [np.empty(lat.full_size) for _ in shifts]
        so, diff, nconv = ii.MultiShiftSolver_CG(ii.Fopr_DdagD(Df), 1e-22, 2000).solve(
            xs, b, shifts)
        assert nconv > 0
        worst = 0.0
        for x, s in zip(xs, shifts):
            ax = lat.dirac_full(u, lat.dirac_full(u, x, 0, kappa), 1, kappa) + s * x
            worst = max(worst, rel(ax, b) ** 2)
            assert rel(ax, b) < 1e-10
        assert abs(diff - worst) <= 1e-3 * worst + 1e-30
    finally:
        ctx.close()


def test_qprop_multishift_columns():
    """Qprop_MultiShift::calc: xq[(3s+c)n + m] = D x_m with
    (D^dag D + sigma_m) x_m = src(s,c); checked as (D D^dag + sigma_m) xq = D src."""
    dims = (4, 4, 4, 8)
    lat = oracle.Lattice(dims)
    u = lat.apbc(lat.gauge_expH(0.3, 32))
    mass = 1.0 / 6.0
    kappa = oracle.kappa_from_mass(mass)
    ctx = ii.Context(dims)
    try:
        Df = ii.Dirac_Wilson(mass, u, ctx)
        shifts = [0.01, 0.2]
        src = ii.Source_local((1, 2, 3, 4), dims)
        q = ii.Qprop_MultiShift(Df, ii.MultiShiftSolver_CG(ii.Fopr_DdagD(Df), 1e-22, 2000))
        xq = q.calc(src, shifts)
        assert len(xq) == 12 * len(shifts)
        for s in range(4):
            for c in range(3):
                dsrc = lat.dirac_full(u, src.mksrc(s, c), 0, kappa)
                for m, sig in enumerate(shifts):
                    y = xq[(3 * s + c) * len(shifts) + m]
                    lhs = lat.dirac_full(u, lat.dirac_full(u, y, 1, kappa), 0, kappa) + sig * y
                    assert rel(lhs, dsrc) < 1e-9, (s, c, m)
    finally:
        ctx.close()


@pytest.mark.parametrize("precision", ["single", "half"])
def test_mixed_multishift_12_poles(precision):
    """12 shifts log-spaced in [1e-4, 1] on an 8^3x16 heat-bath field: the
    fused FP32 multishift plus batched refinement reaches |b - (A+s)x|/|b|
    <= 1e-10 for every shift (checked with the oracle's operator) and agrees
    with the FP64 multishift."""
    dims = (8, 8, 8, 16)
    kappa = 0.1350
    mass = 0.5 / kappa - 4.0
    ctx = ii.Context(dims)
    try:
        ctx.unit_gauge()
        ctx.heatbath(5.8, 13, 30)
        ctx.apply_apbc()
        u = ctx.download_gauge()
        D = ii.Dirac_Wilson_EvenOdd(mass, u, ctx)
        lat = oracle.Lattice(dims)
        b = oracle.gauss_field(lat.half_size, 77)
        shifts = [10.0 ** (-4 + 4 * k / 11.0) for k in range(12)]
        xs = [np.empty_like(b) for _ in shifts]
        outs = ii.MultiShiftSolver_Mixed(ii.Fopr_DdagD(D), tol=1e-10,
                                         precision=precision).solve(xs, b, shifts)
        x64 = [np.empty_like(b) for _ in shifts]
        ii.MultiShiftSolver_CG(ii.Fopr_DdagD(D), 1e-24, 5000).solve(x64, b, shifts)
        for i, s in enumerate(shifts):
            ax = lat.ddagd(u, xs[i], kappa) + s * xs[i]
            assert rel(ax, b) < 1.2e-10, (i, rel(ax, b))
            assert outs[i].true_residual <= 1e-10
            assert rel(xs[i], x64[i]) < 1e-7, (i, rel(xs[i], x64[i]))
    finally:
        ctx.close()


@pytest.mark.parametrize("precision,batched", [("half", True), ("double", False)])
def test_graph_replayed_iterations_equal_eager(precision, batched):
    """The CG iteration recorded once into a HIP graph and replayed (between
    eager reliable updates and host synchronisations) gives bit-identical
    solutions and iteration counts to launching it eagerly."""
    dims = (8, 8, 8, 16)
    lat, u = _prop_setup(dims, seed=21)
    ctx = ii.Context(dims)
    try:
        ctx.upload_gauge(u)
        src = ii.DeviceField(ctx, "full", 12)
        s1 = ii.DeviceField(ctx, "full", 12)
        s2 = ii.DeviceField(ctx, "full", 12)
        src.point_sources((0, 1, 2, 3))
        o1 = ii.PropagatorSolver(ctx, 0.13, precision=precision, batched=batched).solve(src, s1)
        o2 = ii.PropagatorSolver(ctx, 0.13, precision=precision, batched=batched,
                                 use_graph=True).solve(src, s2)
        for k in range(12):
            assert o1[k].Iterations == o2[k].Iterations
            assert o1[k].restarts == o2[k].restarts
            assert np.array_equal(s1.download(k), s2.download(k))
            assert o2[k].true_residual <= 1e-10
    finally:
        ctx.close()
