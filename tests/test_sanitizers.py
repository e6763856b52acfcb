"""The host planner under sanitizers, on the CPU (ket_amd/cpp/san.mk builds
the planner sources straight into the test programs with g++):

* ASan + UBSan: the exact-run planner and its CPU interpreter
  (tests/cpp/test_runs.cpp) over circuits with every gate kind, both
  precisions, one pass per gate against runs and ket's Fuser with runs off
  and on, and the learned planner's feature recording;
* ASan + UBSan and TSan: the batched sweep's planning on up to 16 threads
  (tests/cpp/test_batch_plan.cpp), which must also equal planning one
  instance at a time byte for byte, and the upload pool's parallel write.

Any sanitizer report fails the program (UBSan does not recover).
"""
import os
import subprocess
import tempfile

import pytest

from ket_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(ROOT, "ket_amd", "_san")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "ket_amd", "cpp"), "-f", "san.mk",
                        "-j4"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return OUT


def run(exe, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               TSAN_OPTIONS="halt_on_error=1")
    return subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=600,
                          env=env)


@pytest.mark.parametrize("name,make", [
    ("qft14", lambda: W.qft_u(14, seed=1)),
    ("mixed14", lambda: W.random_mixed(14, 10, seed=3)),
    ("diag13", lambda: W.diag_mix(13, 8, seed=4)),
    ("bwm14", lambda: W.brickwork_mixed(14, 6, seed=6)),
])
def test_run_planner_under_asan_ubsan(built, name, make):
    c = make()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c.bin")
        c.records().tofile(path)
        for precision in (0, 1):
            for mode in ("none", "reference", "features"):
                r = run(os.path.join(built, "test_runs_asan"), path, c.n_qubits, precision, mode)
                assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr[-4000:]
                assert "runtime error" not in r.stderr


@pytest.mark.parametrize("exe", ["test_batch_plan_asan", "test_batch_plan_tsan"])
def test_batch_planning_under_sanitizers(built, exe):
    r = run(os.path.join(built, exe))
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-4000:]
    assert r.stdout.strip().endswith("0 failures")
    assert "WARNING: ThreadSanitizer" not in r.stderr and "runtime error" not in r.stderr
