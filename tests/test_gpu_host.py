"""Drop-in test of the C++ host side: the reference's own
Inverter_WilsonLike, Qprop_EvenOdd and Qprop_MultiShift (with Source_local), Solver and
MultiShiftSolver base-class calls, Action_Nf2::md_force's sequence of
Dirac/Solver calls and FieldUtils::TracelessAntihermite, and the full-index
md_force/_p/_m, run once with the reference CPU classes
and once with the device classes of iroiro_amd/host/iroiro_ref_adapter.hpp,
in one process (oracle/_ref/iroiro_ref_hip, built by oracle/ref_build.sh)."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "iroiro_ref_hip")

pytestmark = pytest.mark.gpu


def test_reference_drivers_on_device_classes(golden, tmp_path):
    if not os.path.exists(BIN):
        pytest.skip("iroiro_ref_hip not built (its build needs the reference sources)")
    z = golden("wilson_c1_4x4x4x8")
    z["u_apbc"].tofile(str(tmp_path / "u.bin"))
    z["src"].tofile(str(tmp_path / "src.bin"))
    dims = [str(int(d)) for d in z["dims"]]
    out = subprocess.run([BIN] + dims + [repr(float(z["mass"])), str(tmp_path / "u.bin"),
                                          str(tmp_path / "src.bin")],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    r = {}
    for line in out.stdout.splitlines():
        if line.startswith("IIHIP "):
            r.update(json.loads(line[len("IIHIP "):]))
    assert r["ops_rel"] < 1e-13
    assert r["cg_iter_gpu"] == r["cg_iter_cpu"]
    assert r["cg_rel"] < 1e-12
    assert r["ms_nconv_gpu"] == r["ms_nconv_cpu"]
    assert r["ms_rel"] < 1e-12
    assert r["inv_rel"] < 1e-12
    assert r["ncol"] == 12 and r["prop_rel"] < 1e-12
    # Qprop_MultiShift (qprop_MultiShift.cpp:10-31) on the full-index operator.
    # It pushes sol.size() = Nc*Nd = 12 fields per column whatever the number
    # of shifts (:16-29); the solver fills only the first Nshift of them
    # (multiShiftSolver_CG.cpp:179), so 12 x 12 entries, 2 per column nonzero.
    assert r["full_ops_rel"] < 1e-13
    assert r["qprop_ms_n"] == 144 and r["qprop_ms_rel"] < 1e-10
    # MD force: Action_Nf2::md_force's sequence (solve, mult, md_force,
    # TracelessAntihermite) and the full-index md_force / md_force_p / _m
    assert r["force_op_rel"] < 1e-13
    assert r["force_rel"] < 1e-11
    assert r["full_force_rel"] < 1e-13
    # RationalSolver_CG_HIP through the RationalSolver base class, against the
    # reference's multishift and reconstruction
    assert r["rational_rel"] < 1e-10
