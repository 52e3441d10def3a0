"""The benchmark configurations at their full sizes, every result checked by
the oracle on the CPU (lat.true_residual_full: |b - D x|/|b| with the oracle's
full Wilson operator; lat.ddagd for the shifted systems):

  C2  16^3 x 32, heat-bath beta = 5.8 (200 sweeps), kappa = 0.1350, one
      Gaussian source: FP64 against FP32 / FP16 / BF16 inner storage, each
      solution's error against the FP64 solution reported
  C3  24^3 x 48, beta = 5.8, kappa = 0.1350: the 12 point-source columns in
      one fused multi-rhs CG (FP32 inner, FP64 reliable updates), and the
      multishift on Dhat^dag Dhat with 12 shifts log-spaced in [1e-4, 1]
      against the FP64 multishift
  C4  16^3 x 32, beta = 5.8, kappa = 0.1355: the learned preconditioner
      trained on 64 plain-CG pairs, 16 fresh Gaussian sources solved by the
      batched flexible GCR with it, against the same solver without it
  C5  32^3 x 64, beta = 6.0, kappa = 0.1355: every column of the batched
      mixed-precision propagator solves (the flexible GCR with the learned
      preconditioner, CG and BiCGStab), error against the plain FP64 CG
      solution (one rhs after the other, as Qprop_EvenOdd does)

The caller chain they stand for is Qprop_EvenOdd::calc
(lib/Measurements/FermionicM/qprop_EvenOdd.cpp:11-23) over
Inverter_WilsonLike::invert (lib/Dirac_ops/eoUtils.cpp:43-80), and
MultiShiftSolver_CG::solve (lib/Solver/multiShiftSolver_CG.cpp:95-185)."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import iroiro_amd as ii
from oracle import oracle

pytestmark = pytest.mark.gpu

SEED = 20260601


def heatbath_ctx(dims, beta, sweeps=200):
    ctx = ii.Context(dims)
    ctx.unit_gauge()
    ctx.heatbath(beta, SEED, sweeps)
    plaq = ctx.plaquette()
    ctx.apply_apbc()
    return ctx, plaq


def oracle_residuals(lat, u, kappa, pairs):
    """Full-operator true residuals of (x, b) pairs, the oracle's C code run
    on host threads (ctypes releases the interpreter lock)."""
    with ThreadPoolExecutor(max_workers=min(12, len(pairs))) as ex:
        return list(ex.map(lambda xb: lat.true_residual_full(u, kappa, xb[0], xb[1]), pairs))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ------------------------------------------------------------------ C2
@pytest.fixture(scope="module")
def c2():
    dims, beta, kappa = (16, 16, 16, 32), 5.8, 0.1350
    print("C2 setup %s beta=%.1f kappa=%.4f" % (dims, beta, kappa), flush=True)
    ctx, plaq = heatbath_ctx(dims, beta)
    lat = oracle.Lattice(dims)
    u = ctx.download_gauge()
    b = oracle.gauss_field(lat.full_size, 31)
    src = ii.DeviceField(ctx, "full", 1)
    src.upload(0, b)
    sol = ii.DeviceField(ctx, "full", 1)
    out = ii.PropagatorSolver(ctx, kappa, tol=1e-10, precision="double").solve(src, sol)
    x64 = sol.download(0)
    yield dict(ctx=ctx, lat=lat, u=u, b=b, src=src, sol=sol, x64=x64, kappa=kappa,
               plaq=plaq, fp64=out[0])
    ctx.close()


def test_c2_fp64_reference(c2):
    r = c2["lat"].true_residual_full(c2["u"], c2["kappa"], c2["x64"], c2["b"])
    print("C2 plaquette %.5f, FP64 CG %d iterations, oracle residual %.2e"
          % (c2["plaq"], c2["fp64"].Iterations, r))
    assert abs(c2["plaq"] - 0.5677) < 0.004   # quenched Wilson beta = 5.8
    assert r <= 1e-10


@pytest.mark.parametrize("precision", ["single", "half", "bf16"])
def test_c2_mixed_vs_fp64(c2, precision):
    out = ii.PropagatorSolver(c2["ctx"], c2["kappa"], tol=1e-10, precision=precision,
                              batched=True).solve(c2["src"], c2["sol"])
    x = c2["sol"].download(0)
    r = c2["lat"].true_residual_full(c2["u"], c2["kappa"], x, c2["b"])
    err = rel(x, c2["x64"])
    print("C2 %s: %d iterations, %d reliable updates, oracle residual %.2e, "
          "|x - x_fp64|/|x_fp64| = %.2e" % (precision, out[0].Iterations, out[0].restarts, r, err))
    assert r <= 1e-10
    # both solutions carry a residual of 1e-10 through D^-1 (|D^-1| ~ 10^2 here)
    assert err < 1e-7


# ------------------------------------------------------------------ C3
@pytest.fixture(scope="module")
def c3():
    dims, beta, kappa = (24, 24, 24, 48), 5.8, 0.1350
    print("C3 setup %s beta=%.1f kappa=%.4f" % (dims, beta, kappa), flush=True)
    ctx, plaq = heatbath_ctx(dims, beta)
    lat = oracle.Lattice(dims)
    u = ctx.download_gauge()
    yield dict(ctx=ctx, lat=lat, u=u, kappa=kappa, plaq=plaq)
    ctx.close()


def test_c3_fused_12rhs_cg(c3):
    ctx, lat, u, kappa = c3["ctx"], c3["lat"], c3["u"], c3["kappa"]
    src = ii.DeviceField(ctx, "full", 12)
    sol = ii.DeviceField(ctx, "full", 12)
    src.point_sources((0, 0, 0, 0))
    outs = ii.PropagatorSolver(ctx, kappa, tol=1e-10, precision="single",
                               batched=True).solve(src, sol)
    pairs = [(sol.download(k), src.download(k)) for k in range(12)]
    res = oracle_residuals(lat, u, kappa, pairs)
    print("C3 fused 12-rhs CG (FP32 inner): iterations %s, oracle residuals max %.2e"
          % ([o.Iterations for o in outs], max(res)))
    assert max(res) <= 1e-10
    for o, r in zip(outs, res):
        assert abs(o.true_residual - r) < 1e-12
    src.close()
    sol.close()


@pytest.mark.parametrize("precision,batched", [("double", False), ("single", True)])
def test_c3_solve_is_deterministic(c3, precision, batched):
    """The same sources solved twice give bitwise identical solutions and iteration
    counts: every reduction in the solvers sums in a fixed order (no atomics)."""
    ctx, kappa = c3["ctx"], c3["kappa"]
    src = ii.DeviceField(ctx, "full", 12)
    sol = ii.DeviceField(ctx, "full", 12)
    src.point_sources((0, 0, 0, 0))
    solver = ii.PropagatorSolver(ctx, kappa, tol=1e-10, precision=precision,
                                 batched=batched)
    runs = []
    for _ in range(2):
        outs = solver.solve(src, sol)
        runs.append(([o.Iterations for o in outs],
                     np.stack([sol.download(k) for k in range(12)])))
    print("C3 %s batched=%s: iterations %s" % (precision, batched, runs[0][0]))
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1], runs[1][1])
    src.close()
    sol.close()


def test_c3_multishift_12_poles(c3):
    ctx, lat, u, kappa = c3["ctx"], c3["lat"], c3["u"], c3["kappa"]
    shifts = [10.0 ** (-4 + 4 * k / 11.0) for k in range(12)]
    D = ii.Dirac_Wilson_EvenOdd(1.0 / (2 * kappa) - 4.0, u, ctx)
    A = ii.Fopr_DdagD(D)
    b = oracle.gauss_field(lat.half_size, 17)
    xs_mixed = [np.empty_like(b) for _ in shifts]
    outs = ii.MultiShiftSolver_Mixed(A, tol=1e-10, precision="single").solve(xs_mixed, b, shifts)
    assert max(o.true_residual for o in outs) <= 1e-10
    xs64 = [np.empty_like(b) for _ in shifts]
    ii.MultiShiftSolver_CG(A, 1e-24, 20000).solve(xs64, b, shifts)
    bn = np.linalg.norm(b)
    errs, resid = [], []
    for s, xm, x64 in zip(shifts, xs_mixed, xs64):
        r = b - (lat.ddagd(u, xm, kappa) + s * xm)
        resid.append(float(np.linalg.norm(r) / bn))
        errs.append(rel(xm, x64))
    print("C3 multishift, 12 poles in [1e-4, 1]: oracle residuals max %.2e, "
          "|x - x_fp64|/|x_fp64| max %.2e" % (max(resid), max(errs)))
    # the device stops at 1e-10 on its own residual; the oracle recomputes
    # (A + s) x in another summation order
    assert max(resid) <= 1.2e-10
    assert max(errs) < 1e-7


# ------------------------------------------------------------------ C5
@pytest.fixture(scope="module")
def c5():
    dims, beta, kappa = (32, 32, 32, 64), 6.0, 0.1355
    print("C5 setup %s beta=%.1f kappa=%.4f" % (dims, beta, kappa), flush=True)
    ctx, plaq = heatbath_ctx(dims, beta)
    lat = oracle.Lattice(dims)
    u = ctx.download_gauge()
    src = ii.DeviceField(ctx, "full", 12)
    src.point_sources((0, 0, 0, 0))
    x64 = ii.DeviceField(ctx, "full", 12)
    outs = ii.PropagatorSolver(ctx, kappa, tol=1e-10, precision="double").solve(src, x64)
    print("C5 plain FP64 CG: iterations %s" % [o.Iterations for o in outs], flush=True)
    sol = ii.DeviceField(ctx, "full", 12)
    yield dict(ctx=ctx, lat=lat, u=u, kappa=kappa, plaq=plaq, src=src, x64=x64, sol=sol)
    ctx.close()


@pytest.mark.parametrize("method", ["fgcr-learned", "cg", "bicgstab"])
def test_c5_every_column_vs_oracle(c5, method):
    ctx, lat, u, kappa = c5["ctx"], c5["lat"], c5["u"], c5["kappa"]
    if method == "fgcr-learned":
        # the network trained on this gauge field (64 plain-CG pairs)
        net = ii.LearnedPreconditioner(ctx)
        rep = net.train(kappa, npairs=64, epochs=8, batch=4, lr=2e-3, pair_tol=1e-8, seed=SEED)
        print("C5 network trained: loss %.4f -> %.4f" % (rep["loss_first"], rep["loss_last"]),
              flush=True)
        solver = ii.PropagatorSolver(ctx, kappa, tol=1e-10, batched=True, method="fgcr",
                                     use_learned=True, use_graph=True)
    else:
        solver = ii.PropagatorSolver(ctx, kappa, tol=1e-10, precision="half", batched=True,
                                     method=method)
    outs = solver.solve(c5["src"], c5["sol"])
    errs, pairs = [], []
    for k in range(12):
        x = c5["sol"].download(k)
        errs.append(rel(x, c5["x64"].download(k)))
        pairs.append((x, c5["src"].download(k)))
    res = oracle_residuals(lat, u, kappa, pairs)
    print("C5 batched %s: iterations %s, oracle residuals max %.2e, "
          "|x - x_fp64|/|x_fp64| per rhs %s" % (method, [o.Iterations for o in outs], max(res),
                                                ["%.1e" % e for e in errs]))
    assert max(res) <= 1e-10
    for o, r in zip(outs, res):
        assert abs(o.true_residual - r) < 1e-12
    assert max(errs) < 1e-7


# ------------------------------------------------------------------ C4
def test_c4_learned_16_fresh_sources():
    """C4: 16^3 x 32, heat-bath beta = 5.8, kappa = 0.1355.  The network is
    trained on 64 (Gaussian source, plain-CG solution) pairs of this gauge
    field, then 16 fresh Gaussian sources are solved in one batch by the
    flexible GCR with the network as preconditioner.  Every solution meets
    the full-operator true residual 1e-10 by the oracle and agrees with the
    plain FP64 CG solution; the network takes fewer directions than the
    same flexible GCR without it (times of both, and of the plain FP64 CG,
    are printed)."""
    import time
    dims, beta, kappa = (16, 16, 16, 32), 5.8, 0.1355
    print("C4 setup %s beta=%.1f kappa=%.4f" % (dims, beta, kappa), flush=True)
    ctx, plaq = heatbath_ctx(dims, beta)
    try:
        lat = oracle.Lattice(dims)
        u = ctx.download_gauge()
        net = ii.LearnedPreconditioner(ctx)
        rep = net.train(kappa, npairs=64, epochs=8, batch=4, lr=2e-3, pair_tol=1e-8, seed=SEED)
        print("C4 trained on 64 pairs: loss %.4f -> %.4f" % (rep["loss_first"], rep["loss_last"]),
              flush=True)
        assert rep["loss_last"] < rep["loss_first"]
        n = 16
        bs = [oracle.gauss_field(lat.full_size, 1000 + k) for k in range(n)]
        src = ii.DeviceField(ctx, "full", n)
        sol = ii.DeviceField(ctx, "full", n)
        for k in range(n):
            src.upload(k, bs[k])
        runs = {}
        for name, kw in [("learned", dict(method="fgcr", use_learned=True, use_graph=True)),
                         ("none", dict(method="fgcr", use_learned=False, use_graph=True)),
                         ("fp64_cg", dict(method="cg", precision="double"))]:
            s = ii.PropagatorSolver(ctx, kappa, tol=1e-10, batched=(name != "fp64_cg"), **kw)
            s.solve(src, sol)                      # warm-up
            t0 = time.time()
            outs = s.solve(src, sol)
            dt = time.time() - t0
            runs[name] = (dt, outs, [sol.download(k) for k in range(n)])
            print("C4 %s: %.3f s for %d sources, directions/iterations %s"
                  % (name, dt, n, [o.Iterations for o in outs]), flush=True)
        xs = runs["learned"][2]
        res = oracle_residuals(lat, u, kappa, list(zip(xs, bs)))
        errs = [rel(x, y) for x, y in zip(xs, runs["fp64_cg"][2])]
        print("C4 learned: oracle residuals max %.2e, |x - x_fp64|/|x_fp64| max %.1e"
              % (max(res), max(errs)))
        assert max(res) <= 1e-10
        for o, r in zip(runs["learned"][1], res):
            assert abs(o.true_residual - r) < 1e-12
        assert max(errs) < 1e-7
        d_net = max(o.Iterations for o in runs["learned"][1])
        d_none = max(o.Iterations for o in runs["none"][1])
        assert d_net < d_none, (d_net, d_none)
    finally:
        ctx.close()
