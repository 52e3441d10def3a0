"""Writes the golden fixtures of the Wilson even/odd solve path.

The expected outputs come from the reference IroIro++ itself, compiled from
its own sources by oracle/ref_build.sh into oracle/_ref/iroiro_ref and run in
this container (the reference tree is not available elsewhere).  Inputs are
generated here:
  * gauge field U = exp(i eps H) reunitarised (eps = 0.3), H traceless
    Hermitian with unit-variance Gaussian entries, fixed seed; the reference
    then applies its antiperiodic time boundary (AntiPeriodicBC, TDIR);
  * a complex Gaussian spinor source with unit variance per real component.
Reference calls made (see oracle/ref_driver.cpp): Deo/Doe mult and mult_dag,
Dirac_Wilson_EvenOdd mult/mult_dag/gamma5, Fopr_DdagD, the full-index
Dirac_Wilson mult/mult_dag, Solver_CG, MultiShiftSolver_CG and
Inverter_WilsonLike::invert.

The MD-force fixtures (<case>_mdforce.npz) reuse the gauge field of <case>
and add Gaussian eta/zeta fields (even half and full); the reference calls
are Dirac_Wilson_EvenOdd::md_force, FieldUtils::TracelessAntihermite of it,
and the full-index Dirac_Wilson md_force, md_force_p and md_force_m.

The gauge-IO fixture (<case>_io.npz) holds the files the reference's own
BinaryWriter and NERSCWriter write for the gauge field of <case>, and the
field its BinaryReader decodes from a big-endian file made here (the
writers and reader are driven row by row as ReadFromDisk / SaveOnDisk do).

The meson fixture (meson_2x4x2x6.npz) holds two Gaussian 12-column
propagators and the reference's MesonCorrelator::calculate (all four
directions), calculate_mom and calculate_PSmom for all 18 MesonTypes.

Usage: python tests/golden/make_golden.py [mdforce|io|meson]
       (writes tests/golden/*.npz; a name writes only those files)
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "iroiro_ref")

CASES = {
    # config 1 of the solve path: 4^3 x 8, kappa = 0.12 (mass 1/6)
    "wilson_c1_4x4x4x8": dict(dims=(4, 4, 4, 8), mass=1.0 / 6.0, eps=0.3, seed=20240611,
                              src_seed=0x123, solvers=True),
    # anisotropic lattice to pin the index arithmetic of every direction
    "wilson_aniso_4x6x2x4": dict(dims=(4, 6, 2, 4), mass=0.05, eps=0.3, seed=77,
                                 src_seed=0x234, solvers=False),
}
CG_PREC = 1e-20
MS_SHIFTS = [0.05, 0.10, 0.20, 0.30]   # tests/test_MultiShiftSolver.cpp shifts


def run(args):
    out = subprocess.run([REF] + [str(a) for a in args], capture_output=True, text=True,
                         check=True)
    return out.stdout


def parse(stdout, tag):
    for line in stdout.splitlines():
        if line.startswith(tag):
            return json.loads(line[len(tag):].strip())
    raise RuntimeError("no %s line in reference output" % tag)


def make(name, dims, mass, eps, seed, src_seed, solvers):
    lat = oracle.Lattice(dims)
    u_raw = lat.gauge_expH(eps, seed)
    src = oracle.gauss_field(lat.full_size, src_seed)
    d = [str(x) for x in dims]
    m = repr(float(mass))
    res = {"dims": np.array(dims), "mass": mass, "eps": eps, "seed": seed,
           "src_seed": src_seed, "src": src}
    with tempfile.TemporaryDirectory() as tmp:
        u_raw.tofile(os.path.join(tmp, "u_raw.bin"))
        src.tofile(os.path.join(tmp, "src.bin"))
        run(["apbc"] + d + [os.path.join(tmp, "u_raw.bin"), os.path.join(tmp, "u.bin")])
        u = np.fromfile(os.path.join(tmp, "u.bin"))
        res["u_apbc"] = u
        out = run(["ops"] + d + [m, os.path.join(tmp, "u.bin"), os.path.join(tmp, "src.bin"), tmp])
        res["kappa_line"] = out.strip().splitlines()[-1]
        for k in ["deo", "doe", "deo_dag", "doe_dag", "dhat", "dhat_dag", "g5", "ddagd",
                  "dfull", "dfull_dag"]:
            res[k] = np.fromfile(os.path.join(tmp, k + ".bin"))
        if solvers:
            se, so = lat.split(src)
            se.tofile(os.path.join(tmp, "se.bin"))
            out = run(["cg"] + d + [m, os.path.join(tmp, "u.bin"), os.path.join(tmp, "se.bin"),
                                    repr(CG_PREC), 1000, tmp])
            info = parse(out, "IIREF_CG")
            res["cg_x"] = np.fromfile(os.path.join(tmp, "cg_x.bin"))
            res["cg_iterations"] = info["iterations"]
            res["cg_diff"] = info["diff"]
            res["cg_prec"] = CG_PREC
            out = run(["mshift"] + d + [m, os.path.join(tmp, "u.bin"), os.path.join(tmp, "se.bin"),
                                        repr(CG_PREC), 1000, tmp] + MS_SHIFTS)
            info = parse(out, "IIREF_MS")
            res["ms_shifts"] = np.array(MS_SHIFTS)
            res["ms_x"] = np.stack([np.fromfile(os.path.join(tmp, "ms_x%d.bin" % i))
                                    for i in range(len(MS_SHIFTS))])
            res["ms_nconv"] = info["nconv"]
            res["ms_diff"] = info["diff"]
            res["ms_out_diff"] = info["out_diff"]
            out = run(["invert"] + d + [m, os.path.join(tmp, "u.bin"), os.path.join(tmp, "src.bin"),
                                        repr(CG_PREC), 1000, tmp])
            res["inv_x"] = np.fromfile(os.path.join(tmp, "inv_x.bin"))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **res)
    print("wrote", name, {k: (v.shape if hasattr(v, "shape") else v) for k, v in res.items()
                          if k not in ("kappa_line",)})


MDFORCE_CASES = {"wilson_aniso_4x6x2x4": 0x5151}


def make_mdforce(base, seed):
    z = dict(np.load(os.path.join(HERE, base + ".npz")))
    dims = tuple(int(v) for v in z["dims"])
    lat = oracle.Lattice(dims)
    res = {"base": base, "seed": seed}
    res["eta_e"] = oracle.gauss_field(lat.half_size, seed)
    res["zeta_e"] = oracle.gauss_field(lat.half_size, seed + 1)
    res["eta_f"] = oracle.gauss_field(lat.full_size, seed + 2)
    res["zeta_f"] = oracle.gauss_field(lat.full_size, seed + 3)
    d = [str(x) for x in dims]
    with tempfile.TemporaryDirectory() as tmp:
        z["u_apbc"].tofile(os.path.join(tmp, "u.bin"))
        names = []
        for k in ("eta_e", "zeta_e", "eta_f", "zeta_f"):
            res[k].tofile(os.path.join(tmp, k + ".bin"))
            names.append(os.path.join(tmp, k + ".bin"))
        run(["mdforce"] + d + [repr(float(z["mass"])), os.path.join(tmp, "u.bin")] + names
            + [tmp])
        for k in ("mdf_eo", "ta_eo", "mdf_full", "mdf_p", "mdf_m"):
            res[k] = np.fromfile(os.path.join(tmp, k + ".bin"))
    name = base + "_mdforce"
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **res)
    print("wrote", name)


IO_CASES = ["wilson_aniso_4x6x2x4"]


def make_io(base):
    z = dict(np.load(os.path.join(HERE, base + ".npz")))
    dims = tuple(int(v) for v in z["dims"])
    V = int(np.prod(dims))
    d = [str(x) for x in dims]
    u = z["u_apbc"]
    res = {"base": base}
    with tempfile.TemporaryDirectory() as tmp:
        ub = os.path.join(tmp, "u.bin")
        u.tofile(ub)
        for wid in ("Binary", "NERSC"):
            out = os.path.join(tmp, "out." + wid)
            run(["iowrite"] + d + [wid, ub, out])
            res["written_" + wid] = np.fromfile(out, dtype=np.uint8)
        # a big-endian file in the reader's (in, ex, site) order, made here
        be = u.reshape(4, V, 18).transpose(1, 0, 2).astype(">f8")
        fin = os.path.join(tmp, "in.be")
        be.tofile(fin)
        res["binary_input_be"] = np.fromfile(fin, dtype=np.uint8)
        dec = os.path.join(tmp, "dec.bin")
        run(["ioread"] + d + ["Binary", fin, dec])
        res["binary_read"] = np.fromfile(dec)
    np.savez_compressed(os.path.join(HERE, base + "_io.npz"), **res)
    print("wrote", base + "_io")


MESON_DIMS = (2, 4, 2, 6)
MESON_MOM = (0.3, 0.7, 1.1)


def make_meson():
    dims = MESON_DIMS
    V = int(np.prod(dims))
    q1 = oracle.gauss_field(12 * 24 * V, 0x7101)
    q2 = oracle.gauss_field(12 * 24 * V, 0x7102)
    res = {"dims": np.array(dims), "mom": np.array(MESON_MOM), "q1": q1, "q2": q2}
    d = [str(x) for x in dims]
    with tempfile.TemporaryDirectory() as tmp:
        q1.tofile(os.path.join(tmp, "q1.bin"))
        q2.tofile(os.path.join(tmp, "q2.bin"))
        run(["meson"] + d + [os.path.join(tmp, "q1.bin"), os.path.join(tmp, "q2.bin")]
            + [repr(m) for m in MESON_MOM] + [tmp])
        res["calc"] = np.stack([np.concatenate([np.fromfile(os.path.join(
            tmp, "meson_%d_dir%d.bin" % (t, dr))) for dr in range(4)]) for t in range(18)])
        res["mom_out"] = np.stack([np.fromfile(os.path.join(tmp, "meson_%d_mom.bin" % t))
                                   for t in range(18)])
        res["psmom_out"] = np.stack([np.fromfile(os.path.join(tmp, "meson_%d_psmom.bin" % t))
                                     for t in range(18)])
    np.savez_compressed(os.path.join(HERE, "meson_2x4x2x6.npz"), **res)
    print("wrote meson_2x4x2x6")


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/iroiro_ref is not built (run oracle/ref_build.sh)")
    only = sys.argv[1] if len(sys.argv) > 1 else None
    if only is None:
        for name, kw in CASES.items():
            make(name, **kw)
    if only in (None, "mdforce"):
        for base, seed in MDFORCE_CASES.items():
            make_mdforce(base, seed)
    if only in (None, "io"):
        for base in IO_CASES:
            make_io(base)
    if only in (None, "meson"):
        make_meson()


if __name__ == "__main__":
    main()
