"""The host planner under sanitizers, on the CPU (ket_amd/cpp/san.mk builds
the planner sources straight into the test programs with g++):

* ASan + UBSan: the exact-run planner and its CPU interpreter
  (tests/cpp/test_runs.cpp) over circuits with every gate kind, both
  precisions, one pass per gate against runs, against the run packer, and
  ket's Fuser with runs off and on, and the learned planner's candidate
  features;
* ASan + UBSan and TSan: the learned planner (its candidate region plans on
  up to 16 threads) with a fixed function of the features standing in for
  the GPU model (KET_SAN_STUB_MODEL), planned twice and compared;
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


def run(exe, *args, **extra_env):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               TSAN_OPTIONS="halt_on_error=1", **extra_env)
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
            for mode in ("none", "reference", "runs", "features"):
                r = run(os.path.join(built, "test_runs_asan"), path, c.n_qubits, precision, mode)
                assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr[-4000:]
                assert "runtime error" not in r.stderr


@pytest.mark.parametrize("exe", ["test_runs_asan", "test_runs_tsan"])
def test_learned_planner_under_sanitizers(built, exe):
    from ket_amd import costmodel as CM
    circs = CM.config5_circuits(2, seed=5, lo=20, hi=22, depth=24) + [W.qft_u(18, seed=2)]
    blocks = runs = 0
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c.bin")
        for c in circs:
            c.records().tofile(path)
            for precision in (0, 1):
                r = run(os.path.join(built, exe), path, c.n_qubits, precision, "learned",
                        KET_SAN_STUB_MODEL="1")
                assert r.returncode == 0 and r.stdout.startswith("ok learned"), \
                    r.stdout + r.stderr[-4000:]
                assert "WARNING: ThreadSanitizer" not in r.stderr
                assert "runtime error" not in r.stderr
                _, _, passes, nr, nb = r.stdout.split()
                runs += int(nr)
                blocks += int(nb)
    assert runs > 0 and blocks > 0  # the stand-in model picks both kinds of region


@pytest.mark.parametrize("exe", ["test_batch_plan_asan", "test_batch_plan_tsan"])
def test_batch_planning_under_sanitizers(built, exe):
    r = run(os.path.join(built, exe))
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-4000:]
    assert r.stdout.strip().endswith("0 failures")
    assert "WARNING: ThreadSanitizer" not in r.stderr and "runtime error" not in r.stderr
