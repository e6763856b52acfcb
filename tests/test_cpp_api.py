"""The ket:: C++ API: the same test program and the same driver source built
against the reference library (CPU, checks the tests themselves) and against
this repository's GPU backend (must agree with the reference)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from ket_amd import workloads as W
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPU_TEST = os.path.join(ROOT, "ket_amd", "test_api")
GPU_DRIVER = os.path.join(ROOT, "ket_amd", "ket_driver")
REF_TEST = os.path.join(ROOT, "oracle", "_ref", "test_api")


def run(cmd, **kw):
    return subprocess.run([str(c) for c in cmd], capture_output=True, text=True, timeout=300, **kw)


@pytest.mark.skipif(not os.path.exists(REF_TEST), reason="oracle/_ref not built")
def test_api_suite_passes_on_reference():
    r = run([REF_TEST])
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_api_suite_passes_on_gpu_backend():
    r = run([GPU_TEST])
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_amd_extensions():
    r = run([os.path.join(ROOT, "ket_amd", "test_amd_ext")])
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def _write(c, d):
    p = os.path.join(d, "c.txt")
    with open(p, "w") as f:
        f.write(c.to_text())
    return p


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2])
def test_driver_run_and_sample_match_reference(seed):
    c = W.random_mixed(14, 12, seed)
    c.measure_all()
    with tempfile.TemporaryDirectory() as d:
        cp = _write(c, d)
        out = os.path.join(d, "s.bin")
        r = run([GPU_DRIVER, "run", cp, out])
        assert r.returncode == 0, r.stderr
        got = np.fromfile(out, dtype=np.complex128)
        want = O.ref_run(c) if O.ref_available() else O.run(c)
        assert np.array_equal(got, want)
        r = run([GPU_DRIVER, "sample", cp, seed, 3000, out])
        assert r.returncode == 0, r.stderr
        shots = np.fromfile(out, dtype=np.uint64)
        want_s = O.ref_sample(c, seed, 3000) if O.ref_available() else O.sample(want, seed, 3000)
        assert np.array_equal(shots, want_s)
        r = run([GPU_DRIVER, "sample-calls", cp, seed, 50, out])
        assert r.returncode == 0, r.stderr
        assert np.array_equal(np.fromfile(out, dtype=np.uint64), want_s[:50])
        paulis = ["Z" * 14, "XY" * 7, "I" * 13 + "Y"]
        r = run([GPU_DRIVER, "expval", cp] + paulis)
        assert r.returncode == 0, r.stderr
        lines = r.stdout.strip().split("\n")
        assert lines[0] == "state vector"
        for p, v in zip(paulis, lines[1:]):
            assert abs(float(v) - O.expval(want, p)) <= 1e-12


@pytest.mark.gpu
def test_qasm_front_end():
    qdir = os.path.join(ROOT, "tests", "qasm")
    files = [os.path.join(qdir, f)
             for f in ["ghz5.qasm", "grover3.qasm", "qft_basis.qasm", "user_gates.qasm"]]
    r = run([os.path.join(ROOT, "ket_amd", "test_qasm")] + files)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    res = {}
    for line in r.stdout.strip().split("\n"):
        parts = line.split()
        res[os.path.basename(parts[0])] = dict(p.split("=") for p in parts[1:])
    assert res["ghz5.qasm"]["qubits"] == "5" and abs(float(res["ghz5.qasm"]["pmax"]) - 0.5) < 1e-12
    assert res["grover3.qasm"]["argmax"] == "5"
    assert abs(float(res["grover3.qasm"]["pmax"]) - 0.9453125) < 1e-12
    assert abs(float(res["qft_basis.qasm"]["pmax"]) - 1 / 16) < 1e-12
    assert float(res["user_gates.qasm"]["roundtrip"]) <= 1e-12
    bad = [os.path.join(qdir, f) for f in sorted(os.listdir(qdir))
           if f.startswith("err_") or f == "bad.qasm.txt"]
    r = run([os.path.join(ROOT, "ket_amd", "test_qasm")] + bad)
    assert r.returncode == len(bad) == 5
    assert r.stdout.count(" error: qasm: ") == len(bad)
    # Written angles follow ket's text rule: within 1e-12 of 0 or 1e-9 of a
    # pi fraction they snap, others keep 15 digits. The re-read state moves
    # by about the snapped amounts (crx(1e-13) -> crx(0) here).
    r = run([os.path.join(ROOT, "ket_amd", "test_qasm"), os.path.join(qdir, "angles.qasm")])
    rt = float(r.stdout.split("roundtrip=")[1].split()[0])
    assert 0 < rt <= 1e-9, r.stdout
