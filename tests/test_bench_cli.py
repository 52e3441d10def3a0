"""The benchmark's command-line contract and its host-side helpers, on the
CPU (no GPU call is made): a multi-GPU request exits with a message before
the library is loaded, and the traffic figures bench.py reports come from
the rocprofv3 counter summaries committed under profiles/.
"""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def test_traffic_from_committed_counters():
    b = _bench()
    text = open(os.path.join(ROOT, "profiles", "wilson_hop_32c64.md")).read()
    for key in ("k_hop_fp64_32c64", "k_hop_mrhs_half_32c64"):
        line = next(l for l in text.splitlines() if ("pmc %s:" % key) in l)
        f = dict(kv.split("=") for kv in line.split(":", 1)[1].split() if "=" in kv)
        want = (int(f["fetch_kib"]) * float(f["fetch_scale"])
                + int(f["write_kib"]) * float(f["write_scale"])) * 1024.0
        got, src = b.measured_traffic((32, 32, 32, 64), key)
        assert got == want and src.endswith("wilson_hop_32c64.md")
    # the FP64 hop moves at least its algorithmic bytes (1536 B per output site)
    got, _ = b.measured_traffic((32, 32, 32, 64), "k_hop_fp64_32c64")
    assert got >= 1536 * 32 ** 3 * 64 // 2
    # no counters for other lattices or unknown kernels
    assert b.measured_traffic((24, 24, 24, 48)) == (None, None)
    assert b.measured_traffic((32, 32, 32, 64), "k_unknown") == (None, None)
