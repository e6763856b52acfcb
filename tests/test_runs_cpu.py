"""Exact runs on the CPU (tests/cpp/test_runs.cpp, no GPU): circuits planned
one pass per gate and with consecutive passes grouped into runs on LDS tiles
are both executed on the CPU, the runs through their records exactly as
k_run_tile walks them (segment layouts thread by thread, ops in their
register form, through the LDS swizzle), and must give equal bits; for KET_FUSE_NONE against
KET_FUSE_EXACT and for ket's Fuser with runs off against on; KET_FUSE_RUNS
(reordered passes) against one pass per gate within rounding."""
import os
import subprocess
import tempfile

import pytest

from ket_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ket_amd", "test_runs")

CIRCUITS = {
    "qft14": lambda: W.qft_u(14, seed=1),
    "brick13": lambda: W.brickwork(13, 6, seed=2, measure=False),
    "mixed14": lambda: W.random_mixed(14, 10, seed=3),
    "diag13": lambda: W.diag_mix(13, 8, seed=4),
    "mixed12": lambda: W.random_mixed(12, 12, seed=5),
    "bwm14": lambda: W.brickwork_mixed(14, 8, seed=6),
    "qft12": lambda: W.qft_u(12, seed=7),
    # one tile holds the whole state (n = 11 complex128, 12 complex64)
    "mixed11": lambda: W.random_mixed(11, 14, seed=8),
    "clifford13": lambda: W.random_clifford(13, 60, seed=9),
    "hea15": lambda: W.hea(15, 8, seed=10),
}


@pytest.mark.skipif(not os.path.exists(EXE), reason="ket_amd/test_runs not built")
@pytest.mark.parametrize("mode", ["none", "reference"])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", sorted(CIRCUITS))
def test_runs_equal_per_gate_passes(name, precision, mode):
    c = CIRCUITS[name]()
    if c.n_qubits < (11 if precision == 0 else 12):
        pytest.skip("below the tile size: no runs")
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c.bin")
        c.records().tofile(path)
        r = subprocess.run([EXE, path, str(c.n_qubits), str(precision), mode],
                           capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ok, passes, runs, in_runs, conflicts = r.stdout.split()
    assert ok == "ok" and int(runs) >= 1 and int(in_runs) >= 2


@pytest.mark.skipif(not os.path.exists(EXE), reason="ket_amd/test_runs not built")
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", sorted(CIRCUITS))
def test_packed_runs_equal_per_gate_passes_within_rounding(name, precision):
    """KET_FUSE_RUNS reorders commuting passes into fewer runs: executed on
    the CPU (runs through their records) it must equal one pass per gate in
    program order within rounding (an illegal reordering changes amplitudes
    by O(1)), with no more runs than KET_FUSE_EXACT makes."""
    c = CIRCUITS[name]()
    if c.n_qubits < (11 if precision == 0 else 12):
        pytest.skip("below the tile size: no runs")
    counts = {}
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c.bin")
        c.records().tofile(path)
        for mode in ("none", "runs"):
            r = subprocess.run([EXE, path, str(c.n_qubits), str(precision), mode],
                               capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stdout + r.stderr
            ok, passes, runs, in_runs, conflicts = r.stdout.split()
            assert ok == "ok"
            counts[mode] = int(runs)
    assert 1 <= counts["runs"] <= counts["none"]


@pytest.mark.skipif(not os.path.exists(EXE), reason="ket_amd/test_runs not built")
@pytest.mark.parametrize("low", [3, 4, 7, 9])
@pytest.mark.parametrize("precision", [0, 1])
def test_runs_equal_per_gate_passes_other_tile_shapes(precision, low):
    """Runs planned with other numbers of low tile bits (ket_tuning.run_low_bits)."""
    for name in ("qft14", "mixed14", "bwm14"):
        c = CIRCUITS[name]()
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "c.bin")
            c.records().tofile(path)
            r = subprocess.run([EXE, path, str(c.n_qubits), str(precision), "none", str(low)],
                               capture_output=True, text=True, timeout=120)
        assert r.returncode == 0 and r.stdout.startswith("ok "), r.stdout + r.stderr


@pytest.mark.skipif(not os.path.exists(EXE), reason="ket_amd/test_runs not built")
@pytest.mark.parametrize("precision", [0, 1])
def test_learned_planner_recording_features_match_full_packing(precision):
    """The learned planner predicts region plans whose block passes carry no
    matrices; for every form it weighs, their cost-model features must equal
    those of the full plan it then executes."""
    from ket_amd import costmodel as CM
    circs = CM.config5_circuits(4, seed=99, lo=20, hi=24) + [W.qft_u(20, 1), W.diag_mix(18, 10, 2),
                                                             W.hea(16, 8, 3)]
    with tempfile.TemporaryDirectory() as d:
        for c in circs:
            path = os.path.join(d, "c.bin")
            c.records().tofile(path)
            r = subprocess.run([EXE, path, str(c.n_qubits), str(precision), "features"],
                               capture_output=True, text=True, timeout=120)
            assert r.returncode == 0 and r.stdout.strip() == "ok features", r.stdout + r.stderr


def every_gate_circuit(n: int, depth: int, seed: int):
    """Random layers of all 24 gate types, angles uniform [0, 2pi)."""
    import numpy as np
    from ket_amd.gates import ARITY, NPARAMS, Circuit
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    names = sorted(ARITY)
    for _ in range(depth):
        for name in names:
            qs = [int(q) for q in rng.choice(n, ARITY[name], replace=False)]
            getattr(c, name)(*qs, *rng.uniform(0.0, 2 * np.pi, NPARAMS.get(name, 0)))
    return c


@pytest.mark.skipif(not os.path.exists(EXE), reason="ket_amd/test_runs not built")
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", ["every13", "every7", "qft14", "hea15", "mixed14"])
def test_gpu_build_records_equal_host_matrices(name, precision):
    """Block matrices left to the GPU (ket_tuning.device_build): the build
    records, run through k_build's arithmetic on the CPU, must give block
    records equal byte for byte to those built on the host, for ket's Fuser
    and the block packer at widths 2..6 (every gate type, folded phase terms,
    complex64 conversion)."""
    c = {"every13": lambda: every_gate_circuit(13, 6, 11),
         "every7": lambda: every_gate_circuit(7, 10, 12)}.get(name, CIRCUITS.get(name))()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c.bin")
        c.records().tofile(path)
        r = subprocess.run([EXE, path, str(c.n_qubits), str(precision), "build"],
                           capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    ok, what, records = r.stdout.split()
    assert ok == "ok" and what == "build" and int(records) > 10
