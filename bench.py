"""Benchmark of the Wilson even/odd solve path on one MI355X.

Workload (default): lattice 32^3 x 64, quenched SU(3) heat-bath gauge field
at beta = 6.0 (200 sweeps from a cold start, fixed seed), kappa = 0.1355,
antiperiodic time boundary, the 12 point-source right-hand sides of a
propagator (delta at the origin, one per spin x colour).  One step = one
12-rhs propagator solve to the FP64 true residual |b - D x|/|b| <= 1e-10 per
rhs with the optimized method (batched multi-rhs BiCGStab on the even/odd
Schur operator, 16-bit storage with per-site norms and reconstruct-12 links,
FP32 arithmetic, FP64 reliable updates; --method cg selects the batched mixed
CG on Dhat^dag Dhat instead), sources already resident in HBM.

Reported beside it: the plain FP64 GPU CG (one rhs after the other, as the
reference's Qprop_EvenOdd does) on the same problem, the FP64 even/odd hop
(Dslash) bandwidth against the HBM roofline with its measured HBM traffic, the
learned gauge-covariant preconditioner (trained on this gauge field from 64
plain-CG pairs, then the same 12 columns through its flexible GCR; training
time reported apart), and the reference CPU path (oracle/_ref, the compiled
IroIro++) timed on this host.  `value` is the time of the fastest full-accuracy
method, the batched mixed-precision BiCGStab; the batched mixed CG is in
"other_mixed_method", the learned method's time in "learned".

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
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


def measured_traffic(dims, key="k_hop_fp64_32c64"):
    """HBM bytes per launch of a hop kernel (key: k_hop_fp64_32c64, the FP64
    hop, or k_hop_mrhs_half_32c64, the 12-rhs FP16 hop) from the rocprofv3
    FETCH_SIZE / WRITE_SIZE passes summarised in profiles/wilson_hop_32c64.md
    (scaled by the copy-kernel calibration stated there); None for other
    lattices."""
    if tuple(dims) != (32, 32, 32, 64):
        return None, None
    path = os.path.join(ROOT, "profiles", "wilson_hop_32c64.md")
    try:
        text = open(path).read()
    except OSError:
        return None, None
    m = re.search(r"pmc %s: fetch_kib=(\d+) write_kib=(\d+) "
                  r"fetch_scale=([0-9.]+) write_scale=([0-9.]+)" % re.escape(key), text)
    if not m:
        return None, None
    fk, wk, fs, ws = int(m.group(1)), int(m.group(2)), float(m.group(3)), float(m.group(4))
    return (fk * fs + wk * ws) * 1024.0, "profiles/wilson_hop_32c64.md"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(cfg, n_iter_per_rhs, nrhs=12):
    """Times the reference CPU path (compiled IroIro++, oracle/_ref) on this
    host: one EO hop and a bounded Solver_CG sample on the same lattice; the
    12-rhs solve time is extrapolated from the measured seconds per CG
    iteration and the per-rhs iteration count of the FP64 GPU solve."""
    ref = os.path.join(ROOT, "oracle", "_ref", "iroiro_ref")
    if not os.path.exists(ref):
        return None
    dims = [str(d) for d in cfg["dims"]]
    mass = repr(0.5 / cfg["kappa"] - 4.0)
    threads = int(os.environ.get("OMP_NUM_THREADS", "8"))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    # a loose precision bounds the CG sample to a handful of iterations
    cmd = [ref, "bench"] + dims + [mass, "2", "1e-3"]
    t0 = time.time()
    try:
        out = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600).stdout
    except (subprocess.TimeoutExpired, OSError) as e:
        log("cpu baseline failed:", e)
        return None
    wall = time.time() - t0
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
                   "single-threaded): EO hop %.1f ms, Solver_CG %d iterations in %.1f s "
                   "(%.1f s/iteration) on a random SU(3) field of the same lattice; "
                   "x %d iterations/rhs (FP64 GPU count) x %d rhs; nproc=%d; sample wall %.1f s"
                   % (model or "host CPU", threads, r["hop_ms"], iters, r["cg_ms"] / 1000.0,
                      per_iter_ms / 1000.0, n_iter_per_rhs, nrhs, os.cpu_count() or 0, wall)),
        "hop_ms": r["hop_ms"], "s_per_cg_iteration": per_iter_ms / 1000.0,
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
    ap.add_argument("--precision", default="half", choices=["single", "half", "bf16"])
    ap.add_argument("--method", default="bicgstab", choices=["bicgstab", "cg"],
                    help="optimized solver: BiCGStab on Dhat or CG on Dhat^dag Dhat")
    ap.add_argument("--no-cg-mixed", action="store_true",
                    help="skip the secondary mixed-precision CG measurement")
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
    ap.add_argument("--no-learned", action="store_true",
                    help="skip training and timing the learned preconditioner")
    ap.add_argument("--train-epochs", type=int, default=8)
    ap.add_argument("--train-pairs", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (one MI355X per box); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import iroiro_amd as ii

    cfg = CONFIGS[args.config]
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
    src.point_sources((0, 0, 0, 0))
    tune = ctx.autotune()
    log("autotune:", tune)

    # ---- Dslash roofline (FP64 hop, one launch = V/2 output sites)
    Vh = ctx.Vh
    hop_ms = ctx.bench_hop("double", 1, 50)
    hop_gbs = FP64_HOP_BYTES_PER_SITE * Vh / (hop_ms * 1e-3) / 1e9
    hop_lp_ms = ctx.bench_hop(args.precision, 12, 20)
    lp_bytes = (104 + 192 / 12.0) * Vh * 12  # 16-bit spinor+norm in/out, r12 links shared
    copy_bytes = 2 << 30
    copy_gbs = 2 * copy_bytes / (ctx.bench_copy(copy_bytes, 10) * 1e-3) / 1e9
    log("FP64 hop %.3f ms = %.0f GB/s (%.1f%% of HBM peak); %s 12-rhs hop %.3f ms; "
        "plain copy %.0f GB/s" % (hop_ms, hop_gbs, 100 * hop_gbs / HBM_PEAK_GBS, args.precision,
                                  hop_lp_ms, copy_gbs))

    traffic, traffic_src = measured_traffic(dims)
    lp_traffic, lp_traffic_src = (measured_traffic(dims, "k_hop_mrhs_half_32c64")
                                  if args.precision == "half" else (None, None))
    force_ms, force_total_ms = ctx.bench_md_force(20)
    force_gbs = MD_FORCE_BYTES_PER_SITE * ctx.V / (force_ms * 1e-3) / 1e9
    log("MD force kernel %.3f ms = %.0f GB/s (%.1f%% of HBM peak)"
        % (force_ms, force_gbs, 100 * force_gbs / HBM_PEAK_GBS))

    # ---- plain FP64 GPU CG, rhs after rhs (reference Qprop_EvenOdd shape)
    fp64 = None
    if not args.no_fp64:
        s64 = ii.PropagatorSolver(ctx, kappa, tol=1e-10, precision="double", batched=False)
        t0 = time.time()
        outs = s64.solve(src, sol)
        fp64_s = time.time() - t0
        its = [o.Iterations for o in outs]
        fp64 = {"s": fp64_s, "iterations_per_rhs": its,
                "max_true_residual": max(o.true_residual for o in outs)}
        log("FP64 CG 12 rhs: %.3f s, iterations %s, max true residual %.2e"
            % (fp64_s, its, fp64["max_true_residual"]))

    # ---- optimized solve: batched, mixed precision
    solver = ii.PropagatorSolver(ctx, kappa, tol=1e-10, precision=args.precision, batched=True,
                                 method=args.method)
    for _ in range(args.warmup):
        solver.solve(src, sol)
    import ctypes  # noqa: F401
    times, outs = [], None
    t_all = time.time()
    for _ in range(args.steps):
        t0 = time.time()
        outs = solver.solve(src, sol)
        times.append(time.time() - t0)
    t_all = time.time() - t_all
    ms_per_step = 1000.0 * t_all / args.steps
    max_res = max(o.true_residual for o in outs)
    log("optimized 12-rhs solve: %.3f s/step (%s), iterations %s, reliable updates %s, "
        "max true residual %.2e" % (ms_per_step / 1000, ["%.3f" % t for t in times],
                                   [o.Iterations for o in outs], [o.restarts for o in outs],
                                   max_res))

    # ---- the other optimized method (batched mixed CG on Dhat^dag Dhat or
    # BiCGStab), for comparison
    other = None
    if not args.no_cg_mixed:
        om = "cg" if args.method == "bicgstab" else "bicgstab"
        s2 = ii.PropagatorSolver(ctx, kappa, tol=1e-10, precision=args.precision, batched=True,
                                 method=om)
        s2.solve(src, sol)
        t0 = time.time()
        o2 = s2.solve(src, sol)
        dt2 = time.time() - t0
        other = {"method": om, "s": dt2, "iterations_per_rhs": [o.Iterations for o in o2],
                 "reliable_updates": [o.restarts for o in o2],
                 "max_true_residual": max(o.true_residual for o in o2)}
        log("batched mixed %s: %.3f s, iterations %s" % (om, dt2, other["iterations_per_rhs"]))
        outs = solver.solve(src, sol)  # the headline method's solution for the contraction

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

    # ---- learned preconditioner: trained on this gauge field, flexible GCR
    learned = None
    if not args.no_learned:
        net = ii.LearnedPreconditioner(ctx)
        t0 = time.time()
        rep = net.train(kappa, npairs=args.train_pairs, epochs=args.train_epochs, batch=4,
                        lr=2e-3, pair_tol=1e-8, seed=args.seed)
        train_s = time.time() - t0
        log("learned: trained on %d pairs, %d epochs in %.1f s (pairs %.1f s): loss %.4f -> %.4f"
            % (args.train_pairs, args.train_epochs, train_s, rep["pairs_ms"] / 1e3,
               rep["loss_first"], rep["loss_last"]))
        ls = ii.PropagatorSolver(ctx, kappa, tol=1e-10, precision="double", batched=False,
                                 use_learned=True)
        t0 = time.time()
        louts = ls.solve(src, sol)
        l_s = time.time() - t0
        learned = {"s": l_s, "directions_per_rhs": [o.Iterations for o in louts],
                   "cg_fallback_iterations": [o.inner_iterations for o in louts],
                   "max_true_residual": max(o.true_residual for o in louts),
                   "train": rep, "train_s": train_s}
        log("learned 12-rhs solve: %.3f s, directions %s, max true residual %.2e"
            % (l_s, learned["directions_per_rhs"], learned["max_true_residual"]))

    cpu = None
    if not args.no_cpu and fp64 is not None:
        n_it = int(round(sum(fp64["iterations_per_rhs"]) / 12.0))
        cpu = cpu_baseline(cfg, n_it)
        if cpu:
            log("cpu baseline:", cpu["sample"])

    result = {
        "metric": "solve time @1e-10 true resid: mixed+batched+learned vs FP64 CG; Dslash GB/s",
        "value": ms_per_step / 1000.0,
        "unit": "s per 12-rhs propagator solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "%s storage / fp32 compute / fp64 reliable updates" % args.precision,
        "data": "synthetic: SU(3) heat-bath gauge field generated on the device, point sources",
        "config": {"workload": cfg["workload"], "lattice": list(dims), "beta": cfg["beta"],
                   "kappa": kappa, "nrhs": 12, "tol": 1e-10, "method":
                   ("batched 12-rhs BiCGStab on Dhat" if args.method == "bicgstab" else
                    "batched 12-rhs CG on Dhat^dag Dhat")
                   + ", %s storage, fp64 reliable updates" % args.precision,
                   "plaquette": plaq, "device": name},
        "iterations_per_rhs": [o.Iterations for o in outs],
        "max_true_residual": max_res,
        "other_mixed_method": other,
        "fp64_cg": fp64,
        "speedup_vs_fp64_cg": (fp64["s"] / (ms_per_step / 1000.0)) if fp64 else None,
        "roofline": {"bound": "hbm", "achieved": hop_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": hop_gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes": FP64_HOP_BYTES_PER_SITE * Vh,
                     "kernel": "k_hop FP64 even/odd Wilson hop, V/2 = %d sites/launch, "
                               "1536 B/site algorithmic" % Vh,
                     "avg_ms": hop_ms,
                     "gflops": HOP_FLOPS_PER_SITE * Vh / (hop_ms * 1e-3) / 1e9,
                     "copy_kernel_gbs": copy_gbs},
        "roofline_batched_lowprec": {
            "kernel": "k_hop_mrhs %s 12-rhs (the solve's dominant kernel), V/2 x 12 site-rhs "
                      "per launch, 120 B/site-rhs algorithmic" % args.precision,
            "bound": "vector-memory path (rocprof counters: texture addresser ~85% busy)",
            "avg_ms": hop_lp_ms,
            "achieved": lp_bytes / (hop_lp_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
            "frac": lp_bytes / (hop_lp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "unit": "GB/s", "bytes_per_site_per_rhs": 104 + 16,
            "traffic": lp_traffic, "traffic_source": lp_traffic_src},
        "md_force": {"kernel": "k_md_force, Dirac_Wilson_EvenOdd::md_force over V = %d sites, "
                               "1920 B/site algorithmic" % ctx.V,
                     "avg_ms": force_ms, "achieved": force_gbs, "unit": "GB/s",
                     "frac": force_gbs / HBM_PEAK_GBS, "with_hops_ms": force_total_ms},
        "meson": meson,
        "cpu_baseline": cpu,
        "learned": learned,
        "autotune": tune,
    }
    print(json.dumps(result), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
