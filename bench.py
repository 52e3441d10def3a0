"""Benchmark of the Wilson even/odd solve path on one MI355X.

Workload (default): lattice 32^3 x 64, quenched SU(3) heat-bath gauge field
at beta = 6.0 (200 sweeps from a cold start, fixed seed), kappa = 0.1355,
antiperiodic time boundary, the 12 point-source right-hand sides of a
propagator (delta at the origin, one per spin x colour).  One step = one
12-rhs propagator solve to the FP64 true residual |b - D x|/|b| <= 1e-10 per
rhs with mixed precision + batching + the learned preconditioner: the batched
flexible GCR on the even/odd Schur operator with the gauge-covariant network
(trained on this gauge field; e4m3 activations, bf16 links and matrix-core
arithmetic for the channel mixing and the 8-channel colour transports)
as preconditioner, FP32 Krylov vectors, FP64 true-residual restarts, each
restart cycle replayed as a captured HIP graph; sources already resident in
HBM.  The network's training (64 plain-CG pairs) is timed apart.

Reported beside it, on the same problem: the plain FP64 GPU CG (one rhs after
the other, as the reference's Qprop_EvenOdd does; warmed up like the timed
method) with each rhs's error of the timed solution against it, the same
flexible GCR without the network, the batched mixed BiCGStab and CG, the
network's stage kernels against the HBM roofline (the solve's dominant kernel
is `roofline`), the FP64 even/odd hop (Dslash) bandwidth (metric part 2,
`dslash_fp64_roofline`) with its measured HBM traffic, and the reference CPU
path (oracle/_ref, the compiled IroIro++) timed on this host on the same
gauge field.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--method M]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E peak
FP64_HOP_BYTES_PER_SITE = 1536  # 192 in + 8*144 links + 192 out
HOP_FLOPS_PER_SITE = 1320
# MD force, both parities: 4 links + 4 force matrices (144 B each) + 4 spinors (192 B)
MD_FORCE_BYTES_PER_SITE = 1920
# meson contraction of one 12-column FP64 propagator: 12 x 24 doubles read,
# 18 channel values written and read back by the slice sums
MESON_BYTES_PER_SITE = 12 * 24 * 8 + 2 * 18 * 8
# algorithmic bytes of the network's four stage kernels per output site:
# (per rhs, shared by all rhs).  Activation records are e4m3 (fp8) bytes,
# 24 B per channel, plus one FP32 scale per site and 8-channel half (4 B);
# Y_d is e4m3 too, 32 B per site and path (24 B and 8 pad bytes) plus one
# scale per site; the input r is read as its bf16 copy (48 B per site), z
# is written as an FP32 half field (96 B per site); each launch reads the 8
# link records of its sites once (8 x 36 B) for every rhs.
#   first: r at the even neighbours 48 + [L1|L2](odd) written 384 + 8
#   mid: L1(odd) 192 + 4 + r 48 + L2(even) written 192 + 4
#   third: L2(even) 196 + L2(odd) 196 + Y(odd) written 8 x 32 + 4
#   last: Y(odd) (one path of each of 8 neighbours, with its scale) 8 x 36 +
#         L2(even) 196 + z written 96
NET_STAGE_BYTES = [(48 + 392, 288), (196 + 48 + 196, 288), (196 + 196 + 260, 288),
                   (288 + 196 + 96, 288)]
# MFMA flops per output site and rhs as issued (v_mfma_f32_16x16x32_bf16,
# 16384 flops, 16 sites per instruction, x 3 colours): first 8 x 2, mid
# 8 x 4 x 2 + 4, third 8 x 4 x 2 + 4 x 2 + 4 x 4 (the Y mixes), last 2
# (identity; the first and last stages transport on v_dot2_f32_bf16); plus,
# in the middle stages, the colour transports on v_mfma_f32_4x4x4_16b_bf16
# (2048 flops, one site per 4x4 block, 16 per instruction): 8 paths x 32
# instructions per 16 sites = 32 Ki per site (the 8 x 8 real form of a 3 x 3
# complex matrix, 56% of it nonzero)
NET_STAGE_MFMA_FLOPS = [48 * 1024, (204 + 32) * 1024, (264 + 32) * 1024, 6 * 1024]

CONFIGS = {
    "c5": dict(dims=(32, 32, 32, 64), beta=6.0, kappa=0.1355, sweeps=200,
               workload="wilson_eo_propagator_32c64_b6.0_k0.1355_12rhs"),
    "c3": dict(dims=(24, 24, 24, 48), beta=5.8, kappa=0.1350, sweeps=200,
               workload="wilson_eo_propagator_24c48_b5.8_k0.1350_12rhs"),
    "c2": dict(dims=(16, 16, 16, 32), beta=5.8, kappa=0.1350, sweeps=200,
               workload="wilson_eo_propagator_16c32_b5.8_k0.1350_12rhs"),
    "tiny": dict(dims=(8, 8, 8, 16), beta=6.0, kappa=0.1355, sweeps=20,
                 workload="wilson_eo_propagator_8c16_b6.0_k0.1355_12rhs"),
}


def measured_traffic(dims, key="k_hop_fp64_32c64", fname="wilson_hop_32c64.md"):
    """HBM bytes per launch of a kernel (key: k_hop_fp64_32c64, the FP64 hop;
    k_hop_mrhs_half_32c64, the 12-rhs FP16 hop; k_bnet_dominant_32c64 in
    learned_net_batched_32c64.md, the network's dominant stage) from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised under profiles/
    (scaled by the copy-kernel calibration stated there); None for other
    lattices or when no summary holds the key."""
    if tuple(dims) != (32, 32, 32, 64):
        return None, None
    path = os.path.join(ROOT, "profiles", fname)
    try:
        text = open(path).read()
    except OSError:
        return None, None
    m = re.search(r"pmc %s: fetch_kib=(\d+) write_kib=(\d+) "
                  r"fetch_scale=([0-9.]+) write_scale=([0-9.]+)" % re.escape(key), text)
    if not m:
        return None, None
    fk, wk, fs, ws = int(m.group(1)), int(m.group(2)), float(m.group(3)), float(m.group(4))
    return (fk * fs + wk * ws) * 1024.0, "profiles/" + fname


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(ctx, cfg, n_iter_per_rhs, nrhs=12):
    """Times the reference CPU path (compiled IroIro++, oracle/_ref) on this
    host, on this run's own heat-bath gauge field (downloaded from the device
    with its boundary phase): one EO hop, then column 0 of the point-source
    propagator through Solver_CG on Fopr_DdagD as Inverter_WilsonLike calls
    it, stopped after about 12 iterations (a squared stopping precision
    scaled from the FP64 GPU iteration count: the point source's first
    iterations converge faster than the average rate, so the precision is
    that of 20 average iterations).  The 12-rhs solve time is
    extrapolated from the measured seconds per CG iteration and the per-rhs
    iteration count of the FP64 GPU solve."""
    import tempfile
    ref = os.path.join(ROOT, "oracle", "_ref", "iroiro_ref")
    if not os.path.exists(ref):
        return None
    dims = [str(d) for d in cfg["dims"]]
    mass = repr(0.5 / cfg["kappa"] - 4.0)
    threads = int(os.environ.get("OMP_NUM_THREADS", "8"))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    target = 20
    prec = 10.0 ** (-20.0 * target / max(n_iter_per_rhs, target))
    fd, path = tempfile.mkstemp(suffix=".gauge")
    os.close(fd)
    try:
        ctx.download_gauge().tofile(path)
        cmd = [ref, "benchprop"] + dims + [mass, path, "2", repr(prec)]
        t0 = time.time()
        out = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600).stdout
        wall = time.time() - t0
    except (subprocess.TimeoutExpired, OSError) as e:
        log("cpu baseline failed:", e)
        return None
    finally:
        try:
            os.unlink(path)
        except OSError:
            pass
    m = re.search(r"IIREF_BENCH (\{.*\})", out)
    if not m:
        return None
    r = json.loads(m.group(1))
    iters = r["cg_iterations"] + 1          # loop index at exit -> steps taken
    # Solver_CG::solve = 1 initial A-apply + iters steps + 1 final A-apply
    per_iter_ms = (r["cg_ms"] - 2 * r["ddagd_ms"]) / iters
    per_iter_ms = max(per_iter_ms, r["ddagd_ms"])
    value_s = per_iter_ms * n_iter_per_rhs * nrhs / 1000.0
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": value_s, "unit": "s per 12-rhs propagator solve (extrapolated)",
        "cores": threads, "kind": "reference",
        "sample": ("compiled IroIro++ (oracle/_ref) on %s, %d OpenMP threads (its Dslash is "
                   "single-threaded): EO hop %.1f ms; Solver_CG on Fopr_DdagD for column 0 of "
                   "the point-source propagator on this run's heat-bath gauge field, %d "
                   "iterations in %.1f s (%.2f s/iteration, stopping precision %.1e); x %d "
                   "iterations/rhs (FP64 GPU count) x %d rhs; nproc=%d; sample wall %.1f s"
                   % (model or "host CPU", threads, r["hop_ms"], iters, r["cg_ms"] / 1000.0,
                      per_iter_ms / 1000.0, prec, n_iter_per_rhs, nrhs, os.cpu_count() or 0,
                      wall)),
        "hop_ms": r["hop_ms"], "s_per_cg_iteration": per_iter_ms / 1000.0,
        "cg_iterations_timed": iters,
    }


def rhmc_timings(ii, ctx, cfg):
    """RationalSolver_CG::solve_inv with a degree-10 approximation of
    x^(-1/2) on [1e-4, 5] (the spectrum of Dhat^dag Dhat at this kappa lies
    inside), FP64 reference multishift against the mixed-precision one, and
    Action_Nf::md_force (Nf = 1, one pseudofermion) on the device."""
    import numpy as np
    kappa = cfg["kappa"]
    D = ii.Dirac_Wilson_EvenOdd(1.0 / (2 * kappa) - 4.0, ctx.download_gauge(), ctx)
    t0 = time.time()
    ra = ii.RationalApprox(10, 10, 1, 2, 50, 1e-4, 5.0)
    remez_s = time.time() - t0
    b = np.random.default_rng(7).standard_normal(D.fsize())
    rows = {"remez_s": remez_s, "remez_error": ra.error, "poles": len(ra.InvPoles())}
    for prec, stop in [("double", 1e-20), ("half", 1e-10)]:
        slv = ii.RationalSolver_CG(ii.MultiShiftSolver_CG(ii.Fopr_DdagD(D), stop, 20000), ra,
                                   precision=prec)
        x = np.empty_like(b)
        slv.solve_inv(x, b)            # warm-up
        t0 = time.time()
        out = slv.solve_inv(x, b)
        rows["solve_inv_" + prec] = {"s": time.time() - t0, "iterations": out.Iterations,
                                     "true_residual": out.true_residual}
        log("rational solve_inv (%s): %.3f s, %d iterations" % (prec, rows["solve_inv_" + prec]["s"],
                                                               out.Iterations))
    act = ii.Action_Nf(D, ii.RationalSolver_CG(ii.MultiShiftSolver_CG(ii.Fopr_DdagD(D), 1e-10,
                                                                       20000), precision="half"),
                       n_flav=1, n_pseudof=1, degree=[10, 10, 10], precision=[50] * 3,
                       b_low=[1e-4] * 3, b_high=[5.0] * 3)
    act.phi_[0][...] = b
    act.md_force()
    t0 = time.time()
    act.md_force()
    rows["action_nf_md_force_s"] = time.time() - t0
    log("Action_Nf md_force (10 poles, half multishift): %.3f s" % rows["action_nf_md_force_s"])
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c5", choices=sorted(CONFIGS))
    ap.add_argument("--precision", default="half", choices=["single", "half", "bf16"],
                    help="storage of the batched BiCGStab / CG methods")
    ap.add_argument("--method", default="fgcr-learned",
                    choices=["fgcr-learned", "fgcr", "bicgstab", "cg"],
                    help="timed solver: flexible GCR with the learned preconditioner (default), "
                         "the same without it, BiCGStab on Dhat or CG on Dhat^dag Dhat")
    ap.add_argument("--nkrylov", type=int, default=10,
                    help="flexible GCR directions per restart cycle")
    ap.add_argument("--northo", type=int, default=0,
                    help="flexible GCR: directions each new one is orthogonalised against "
                         "(0: the library default, 1; --nkrylov or more: the full GCR)")
    ap.add_argument("--no-others", action="store_true",
                    help="skip timing the other methods")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--seed", type=int, default=20260601)
    ap.add_argument("--hop-only", action="store_true",
                    help="only time the hop kernels and the copy reference")
    ap.add_argument("--net-only", action="store_true",
                    help="only time the layers of the learned network")
    ap.add_argument("--force-only", action="store_true",
                    help="only time the MD force (Dirac_Wilson_EvenOdd::md_force)")
    ap.add_argument("--rhmc-only", action="store_true",
                    help="only time RationalSolver_CG and the RHMC force (Action_Nf)")
    ap.add_argument("--kappa", type=float, default=0.0,
                    help="hopping parameter instead of the configuration's (0: keep it)")
    ap.add_argument("--train-epochs", type=int, default=8)
    ap.add_argument("--train-pairs", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (one MI355X per box); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import iroiro_amd as ii

    cfg = dict(CONFIGS[args.config])
    if args.kappa > 0:
        cfg["kappa"] = args.kappa
        cfg["workload"] = re.sub(r"_k[0-9.]+_", "_k%.4f_" % args.kappa, cfg["workload"])
    dims = cfg["dims"]
    kappa = cfg["kappa"]
    ctx = ii.Context(dims)
    name, hbm = ctx.device_info()
    log("device", name, "%.0f GiB" % hbm)
    t0 = time.time()
    ctx.unit_gauge()
    ctx.heatbath(cfg["beta"], args.seed, cfg["sweeps"])
    plaq = ctx.plaquette()
    ctx.apply_apbc()
    log("heat-bath %d sweeps beta=%.2f: plaquette %.5f (%.1f s)"
        % (cfg["sweeps"], cfg["beta"], plaq, time.time() - t0))

    if args.hop_only:
        # kernel-level measurement only (used under rocprofv3): every hop
        # variant of the solve path, and the copy reference
        rows = {}
        for prec, nrhs, bpsr in [("double", 1, 1536.0), ("double", 12, 384 + 1152 / 12.0),
                                 ("single", 12, 192 + 384 / 12.0), ("half", 12, 104 + 192 / 12.0),
                                 ("bf16", 12, 104 + 192 / 12.0)]:
            ms = ctx.bench_hop(prec, nrhs, 20)
            rows["%s_x%d" % (prec, nrhs)] = {
                "ms": ms, "alg_gbs": bpsr * ctx.Vh * nrhs / (ms * 1e-3) / 1e9,
                "site_rhs_per_s": ctx.Vh * nrhs / (ms * 1e-3)}
            log("hop %-6s x%-2d %.4f ms  %.0f GB/s algorithmic" % (prec, nrhs, ms,
                                                                   rows["%s_x%d" % (prec, nrhs)]["alg_gbs"]))
        nbytes = 2 << 30
        rows["copy_gbs"] = 2 * nbytes / (ctx.bench_copy(nbytes, 10) * 1e-3) / 1e9
        log("copy %.0f GB/s" % rows["copy_gbs"])
        print(json.dumps({"hop_only": rows, "lattice": list(dims)}), flush=True)
        ctx.close()
        return

    if args.net_only:
        # kernel-level measurement of the learned network's layers (under
        # rocprofv3): FP32 activations, bf16 matrix-core channel mixing
        net = ii.LearnedPreconditioner(ctx)
        net.init_neumann(kappa, noise=0.05, seed=args.seed)
        ms = net.bench(10)
        rows = {}
        for l, t in enumerate(ms):
            cin, cout = net.WIDTHS[l], net.WIDTHS[l + 1]
            sites = ctx.V if l < 3 else ctx.Vh
            # bytes: input channels + output channels (FP32 spinors) + 4 r12 FP32 links per site
            nbytes = (96 * cin * (ctx.V / sites) + 96 * cout + 192) * sites
            flops = 9 * (8 * cout) * (8 * cin) * 3 * 2 * sites
            rows["layer%d_%dto%d" % (l, cin, cout)] = {
                "ms": t, "alg_gbs": nbytes / (t * 1e-3) / 1e9,
                "mfma_tflops": flops / (t * 1e-3) / 1e12}
            log("net layer %d (%d->%d): %.3f ms, %.0f GB/s, %.1f TFLOP/s" %
                (l, cin, cout, t, rows["layer%d_%dto%d" % (l, cin, cout)]["alg_gbs"],
                 rows["layer%d_%dto%d" % (l, cin, cout)]["mfma_tflops"]))
        # the batched evaluation the flexible solver uses: 12 rhs per launch
        st = net.bench_batched(12, 10)
        for i, t in enumerate(st):
            nb = NET_STAGE_BYTES[i]
            alg = ctx.Vh * (12 * nb[0] + nb[1])
            flops = ctx.Vh * 12 * NET_STAGE_MFMA_FLOPS[i]
            rows["batched_stage%d" % i] = {"ms": t, "alg_gbs": alg / (t * 1e-3) / 1e9,
                                           "mfma_tflops": flops / (t * 1e-3) / 1e12}
            log("batched net stage %d (12 rhs): %.3f ms, %.0f GB/s, %.0f TFLOP/s"
                % (i, t, rows["batched_stage%d" % i]["alg_gbs"],
                   rows["batched_stage%d" % i]["mfma_tflops"]))
        print(json.dumps({"net_only": rows, "lattice": list(dims)}), flush=True)
        ctx.close()
        return

    if args.force_only:
        # kernel-level measurement of the MD force (under rocprofv3)
        f_ms, t_ms = ctx.bench_md_force(20)
        gbs = MD_FORCE_BYTES_PER_SITE * ctx.V / (f_ms * 1e-3) / 1e9
        log("md_force kernel %.3f ms = %.0f GB/s (%.1f%% of HBM peak); with its two hops %.3f ms"
            % (f_ms, gbs, 100 * gbs / HBM_PEAK_GBS, t_ms))
        print(json.dumps({"force_only": {"kernel_ms": f_ms, "alg_gbs": gbs, "total_ms": t_ms},
                          "lattice": list(dims)}), flush=True)
        ctx.close()
        return

    if args.rhmc_only:
        print(json.dumps({"rhmc_only": rhmc_timings(ii, ctx, cfg), "lattice": list(dims)}),
              flush=True)
        ctx.close()
        return

    src = ii.DeviceField(ctx, "full", 12)
    sol = ii.DeviceField(ctx, "full", 12)
    x64 = ii.DeviceField(ctx, "full", 12)
    src.point_sources((0, 0, 0, 0))
    tune = ctx.autotune()
    log("autotune:", tune)

    # ---- metric (2): Dslash roofline (FP64 hop, one launch = V/2 output sites)
    Vh = ctx.Vh
    hop_ms = ctx.bench_hop("double", 1, 50)
    hop_gbs = FP64_HOP_BYTES_PER_SITE * Vh / (hop_ms * 1e-3) / 1e9
    copy_bytes = 2 << 30
    copy_gbs = 2 * copy_bytes / (ctx.bench_copy(copy_bytes, 10) * 1e-3) / 1e9
    log("FP64 hop %.3f ms = %.0f GB/s (%.1f%% of HBM peak); plain copy %.0f GB/s"
        % (hop_ms, hop_gbs, 100 * hop_gbs / HBM_PEAK_GBS, copy_gbs))
    traffic, traffic_src = measured_traffic(dims)
    force_ms, force_total_ms = ctx.bench_md_force(20)
    force_gbs = MD_FORCE_BYTES_PER_SITE * ctx.V / (force_ms * 1e-3) / 1e9
    log("MD force kernel %.3f ms = %.0f GB/s (%.1f%% of HBM peak)"
        % (force_ms, force_gbs, 100 * force_gbs / HBM_PEAK_GBS))

    # ---- plain FP64 GPU CG, one rhs after the other (Qprop_EvenOdd's
    # shape), warmed up like the optimized path; its solution is kept
    fp64 = None
    if not args.no_fp64:
        s64 = ii.PropagatorSolver(ctx, kappa, tol=1e-10, precision="double", batched=False)
        s64.solve(src, x64)
        t0 = time.time()
        outs64 = s64.solve(src, x64)
        fp64_s = time.time() - t0
        its = [o.Iterations for o in outs64]
        fp64 = {"s": fp64_s, "iterations_per_rhs": its,
                "max_true_residual": max(o.true_residual for o in outs64)}
        log("FP64 CG 12 rhs: %.3f s, iterations %s, max true residual %.2e"
            % (fp64_s, its, fp64["max_true_residual"]))

    # ---- the learned preconditioner, trained on this gauge field from
    # plain-CG pairs (its cost is reported apart from the solve time)
    net = None
    train = None
    if args.method == "fgcr-learned":
        net = ii.LearnedPreconditioner(ctx)
        t0 = time.time()
        rep = net.train(kappa, npairs=args.train_pairs, epochs=args.train_epochs, batch=4,
                        lr=2e-3, pair_tol=1e-8, seed=args.seed)
        train = dict(rep, s=time.time() - t0, pairs=args.train_pairs, epochs=args.train_epochs)
        log("learned: trained on %d pairs, %d epochs in %.1f s (pairs %.1f s): loss %.4f -> %.4f"
            % (args.train_pairs, args.train_epochs, train["s"], rep["pairs_ms"] / 1e3,
               rep["loss_first"], rep["loss_last"]))
        # the network's launch shapes for 12 rhs (otherwise chosen at the
        # first application, inside the warm-up solve)
        rg, rep_rg = net.tune(12)
        tune = {"hop": tune, "network": rep_rg}
        log("autotune:", rep_rg)

    def make_solver(method, use_graph=True):
        if method in ("fgcr-learned", "fgcr"):
            return ii.PropagatorSolver(ctx, kappa, tol=1e-10, batched=True, method="fgcr",
                                       use_learned=(method == "fgcr-learned"),
                                       use_graph=use_graph, nkrylov=args.nkrylov,
                                       northo=args.northo)
        return ii.PropagatorSolver(ctx, kappa, tol=1e-10, precision=args.precision, batched=True,
                                   method=method)

    # ---- the timed method
    solver = make_solver(args.method)
    for _ in range(args.warmup):
        solver.solve(src, sol)
    times, outs = [], None
    t_all = time.time()
    for _ in range(args.steps):
        t0 = time.time()
        outs = solver.solve(src, sol)  # returns after the device has finished
        times.append(time.time() - t0)
    t_all = time.time() - t_all
    ms_per_step = 1000.0 * t_all / args.steps
    max_res = max(o.true_residual for o in outs)
    log("timed %s 12-rhs solve: %.3f s/step (%s), directions/iterations %s, cycles/updates %s, "
        "max true residual %.2e" % (args.method, ms_per_step / 1000,
                                   ["%.3f" % t for t in times], [o.Iterations for o in outs],
                                   [o.restarts for o in outs], max_res))
    err_fp64 = None
    if fp64 is not None:
        import numpy as np
        err_fp64 = []
        for k in range(12):
            a, b = sol.download(k), x64.download(k)
            err_fp64.append(float(np.linalg.norm(a - b) / np.linalg.norm(b)))
        log("|x - x_fp64|/|x_fp64| per rhs: %s" % ["%.1e" % e for e in err_fp64])

    # ---- the other methods on the same problem (one warm-up, one timed solve)
    others = {}
    if not args.no_others:
        for m in ["fgcr-learned", "fgcr", "bicgstab", "cg"]:
            if m == args.method or (m == "fgcr-learned" and net is None):
                continue
            s2 = make_solver(m)
            s2.solve(src, sol)
            t0 = time.time()
            o2 = s2.solve(src, sol)
            dt2 = time.time() - t0
            others[m] = {"s": dt2, "iterations_per_rhs": [o.Iterations for o in o2],
                         "restarts": [o.restarts for o in o2],
                         "max_true_residual": max(o.true_residual for o in o2)}
            log("%s: %.3f s, iterations %s" % (m, dt2, others[m]["iterations_per_rhs"]))
        outs = solver.solve(src, sol)  # the timed method's solution for the contraction

    # ---- the network's stage kernels (HIP events, the rhs count of the solve)
    net_stages = None
    dominant = None
    if net is not None:
        st = net.bench_batched(12, 5)
        names = ["k_bnet_first (odd, [L1|L2] from 8 transports of r)",
                 "k_bnet_mid (even, L2: 8 transported channels)",
                 "k_bnet_third (odd, L3: 8 transported channels, then the 8 Y mixes)",
                 "k_bnet_last (even, z: one transport of Y per path)"]
        net_stages = []
        for i, t in enumerate(st):
            nb = NET_STAGE_BYTES[i]
            alg = Vh * (12 * nb[0] + nb[1])
            flops = Vh * 12 * NET_STAGE_MFMA_FLOPS[i]
            net_stages.append({"kernel": names[i], "ms": t, "alg_bytes": alg,
                               "alg_gbs": alg / (t * 1e-3) / 1e9,
                               "mfma_tflops": flops / (t * 1e-3) / 1e12})
            log("net stage %d: %.3f ms, %.0f GB/s algorithmic, %.0f TFLOP/s MFMA"
                % (i, t, net_stages[-1]["alg_gbs"], net_stages[-1]["mfma_tflops"]))
        dominant = max(net_stages, key=lambda r: r["ms"])

    # ---- meson correlators of the solved propagator (stays on the device)
    import numpy as np
    ii.meson_correlators(ctx, sol, sol, 0, 3)
    t0 = time.time()
    nrep_m = 5
    for _ in range(nrep_m):
        corr = ii.meson_correlators(ctx, sol, sol, 0, 3)
    meson_ms = (time.time() - t0) * 1000.0 / nrep_m
    meson = {"ms": meson_ms, "pion": [float(v) for v in corr[0][:8]],
             "alg_gbs": MESON_BYTES_PER_SITE * ctx.V / (meson_ms * 1e-3) / 1e9}
    log("meson correlators (18 channels, time slices): %.3f ms, pion C(0..3) = %s"
        % (meson_ms, np.array2string(corr[0][:4], precision=4)))

    cpu = None
    if not args.no_cpu and fp64 is not None:
        n_it = int(round(sum(fp64["iterations_per_rhs"]) / 12.0))
        cpu = cpu_baseline(ctx, cfg, n_it)
        if cpu:
            log("cpu baseline:", cpu["sample"])

    fp64_hop_roofline = {
        "bound": "hbm", "achieved": hop_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": hop_gbs / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
        "algorithmic_bytes": FP64_HOP_BYTES_PER_SITE * Vh,
        "kernel": "k_hop FP64 even/odd Wilson hop, V/2 = %d sites/launch, 1536 B/site "
                  "algorithmic" % Vh,
        "avg_ms": hop_ms, "gflops": HOP_FLOPS_PER_SITE * Vh / (hop_ms * 1e-3) / 1e9,
        "copy_kernel_gbs": copy_gbs}
    if dominant is not None:
        ntraffic, ntraffic_src = measured_traffic(dims, "k_bnet_dominant_32c64",
                                                  "learned_net_batched_32c64.md")
        roofline = {"bound": "hbm", "achieved": dominant["alg_gbs"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": dominant["alg_gbs"] / HBM_PEAK_GBS,
                    "traffic": ntraffic, "traffic_source": ntraffic_src,
                    "algorithmic_bytes": dominant["alg_bytes"],
                    "kernel": dominant["kernel"] + ", the timed solve's dominant kernel, "
                              "12 rhs x V/2 = %d sites per launch" % Vh,
                    "avg_ms": dominant["ms"], "mfma_tflops": dominant["mfma_tflops"],
                    "mfma_frac_of_bf16_peak": dominant["mfma_tflops"] / 2500.0}
    else:
        roofline = fp64_hop_roofline

    method_desc = {
        "fgcr-learned": "batched 12-rhs flexible GCR on Dhat with the trained gauge-covariant "
                        "network as preconditioner (e4m3 activations, bf16 links, MFMA channel "
                        "mixing and colour transports), FP32 Krylov vectors and FP32 Dhat, "
                        "each direction "
                        "orthogonalised against the previous %d, restart cycles of %d "
                        "directions captured as HIP graphs, FP64 true-residual restarts"
                        % (args.northo if args.northo > 0 else 1, args.nkrylov),
        "fgcr": "batched 12-rhs flexible GCR on Dhat without preconditioner, FP32 Krylov "
                "vectors, FP64 true-residual restarts",
        "bicgstab": "batched 12-rhs BiCGStab on Dhat, %s storage, fp64 reliable updates"
                    % args.precision,
        "cg": "batched 12-rhs CG on Dhat^dag Dhat, %s storage, fp64 reliable updates"
              % args.precision}[args.method]
    value_s = ms_per_step / 1000.0
    result = {
        "metric": "solve time @1e-10 true resid: mixed+batched+learned vs FP64 CG; Dslash GB/s",
        "value": value_s,
        "unit": "s per 12-rhs propagator solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("bf16 network / fp32 Krylov / fp64 restarts" if args.method.startswith("fgcr")
                  else "%s storage / fp32 compute / fp64 reliable updates" % args.precision),
        "data": "synthetic: SU(3) heat-bath gauge field generated on the device, point sources",
        "config": {"workload": cfg["workload"], "lattice": list(dims), "beta": cfg["beta"],
                   "kappa": kappa, "nrhs": 12, "tol": 1e-10, "method": method_desc,
                   "plaquette": plaq, "device": name},
        "iterations_per_rhs": [o.Iterations for o in outs],
        "restarts_per_rhs": [o.restarts for o in outs],
        "max_true_residual": max_res,
        "rel_error_vs_fp64_per_rhs": err_fp64,
        "fp64_cg": fp64,
        "speedup_vs_fp64_cg": (fp64["s"] / value_s) if fp64 else None,
        "other_methods": others,
        "speedup_vs_same_without_network": (others["fgcr"]["s"] / value_s
                                            if args.method == "fgcr-learned" and "fgcr" in others
                                            else None),
        "learned_training": train,
        "roofline": roofline,
        "net_stages": net_stages,
        "dslash_fp64_roofline": fp64_hop_roofline,
        "md_force": {"kernel": "k_md_force, Dirac_Wilson_EvenOdd::md_force over V = %d sites, "
                               "1920 B/site algorithmic" % ctx.V,
                     "avg_ms": force_ms, "achieved": force_gbs, "unit": "GB/s",
                     "frac": force_gbs / HBM_PEAK_GBS, "with_hops_ms": force_total_ms},
        "meson": meson,
        "cpu_baseline": cpu,
        "autotune": tune,
    }
    print(json.dumps(result), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
