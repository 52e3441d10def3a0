"""Host sanitizer builds (AddressSanitizer + UndefinedBehaviorSanitizer), run
on the CPU:

  * the CPU oracle (oracle/wilson_oracle.c), every function on small lattices
    (tests/cpu/oracle_sanitize.c);
  * the library's quad-precision Remez, host C++ behind ii_rational_approx
    (iroiro_amd/csrc/ii_remez.cpp; tests/cpu/remez_sanitize.cpp);
  * the library's site-index and tile arithmetic, the inline functions the
    kernels use (decode_site, neighbour, hop_block_base), compiled for the
    host from iroiro_amd/csrc/ii_stencil.h and checked exhaustively for
    every traversal and tiling on lattices with ragged last tiles
    (tests/cpu/index_check.cpp).

Each program must exit 0 with no sanitizer report.  GPU code is not built
with sanitizers (the host flags go after -Xarch_host and the index check is
a host-only compilation)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
HIPCC = "/opt/rocm/bin/hipcc"


def _run(build, exe):
    b = subprocess.run(build, capture_output=True, text=True, cwd=ROOT, timeout=600)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    report = r.stderr
    assert r.returncode == 0, report[-3000:]
    assert "AddressSanitizer" not in report and "runtime error" not in report, report[-3000:]
    return r.stdout


def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_san")
    out = _run(["gcc", "-std=c99", "-O1", "-g"] + SAN +
               ["oracle/wilson_oracle.c", "tests/cpu/oracle_sanitize.c", "-lm", "-o", exe], exe)
    assert out.count("dims") == 3


def test_remez_under_sanitizers(tmp_path):
    exe = str(tmp_path / "remez_san")
    out = _run(["g++", "-std=c++17", "-O1", "-g"] + SAN +
               ["iroiro_amd/csrc/ii_remez.cpp", "tests/cpu/remez_sanitize.cpp", "-lquadmath",
                "-o", exe], exe)
    assert "remez_sanitize: ok" in out


def test_index_arithmetic_under_sanitizers(tmp_path):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "index_check")
    host_san = []
    for f in SAN:
        host_san += ["-Xarch_host", f]
    out = _run([HIPCC, "-x", "hip", "--cuda-host-only", "-O1", "-g", "-std=c++17"] + host_san +
               ["tests/cpu/index_check.cpp", "-o", exe], exe)
    assert "index_check: ok" in out
