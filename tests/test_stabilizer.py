"""The tableau backend for Clifford circuits (ket_amd/cpp/stabilizer.cpp,
reference src/backends/stabilizer.cpp) through the shared C++ driver.

It runs on the CPU: Method::Stabilizer never touches the GPU. Shots and
expectation values must equal the reference library's (built from its own
sources in oracle/_ref) value for value for the same std::mt19937 seed, and
expectation values must equal the dense oracle's on every Pauli string of
small circuits, which pins the tableau without the reference.
"""
import itertools
import os
import subprocess
import tempfile

import numpy as np
import pytest

from ket_amd import workloads as W
from ket_amd.gates import Circuit
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OURS = os.path.join(ROOT, "ket_amd", "ket_driver")
REF = os.path.join(ROOT, "oracle", "_ref", "ket_driver")

needs_ours = pytest.mark.skipif(not os.path.exists(OURS), reason="ket_amd/ket_driver not built")
needs_ref = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")


def drive(exe, *args):
    return subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True,
                          timeout=120)


def write(c, d, name="c.txt"):
    p = os.path.join(d, name)
    with open(p, "w") as f:
        f.write(c.to_text())
    return p


def shots(exe, path, seed, count, d):
    out = os.path.join(d, os.path.basename(exe) + ".bin")
    r = drive(exe, "stab-sample", path, seed, count, out)
    assert r.returncode == 0, r.stderr
    return np.fromfile(out, dtype=np.uint64)


def random_paulis(n, count, seed):
    rng = np.random.default_rng(seed)
    letters = np.array(list("IXYZixyz"))
    return ["".join(rng.choice(letters, n)) for _ in range(count)]


CASES = [(1, 20, 4), (2, 30, 5), (3, 60, 6), (5, 80, 9), (8, 150, 12), (17, 300, 20),
         (63, 400, 30), (64, 500, 40), (65, 500, 40), (130, 900, 60)]


@needs_ours
@needs_ref
@pytest.mark.parametrize("n,depth,measures", CASES)
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_shots_and_expvals_equal_reference(n, depth, measures, seed):
    c = W.random_clifford(n, depth, seed, measures)
    with tempfile.TemporaryDirectory() as d:
        p = write(c, d)
        got = shots(OURS, p, seed + 10, 300, d)
        want = shots(REF, p, seed + 10, 300, d)
        assert np.array_equal(got, want)
        paulis = random_paulis(n, 12, seed) + ["Z" * n, "X" * n]
        a = drive(OURS, "stab-expval", p, *paulis)
        b = drive(REF, "stab-expval", p, *paulis)
        assert a.returncode == 0 and b.returncode == 0, a.stderr + b.stderr
        assert a.stdout == b.stdout
        assert a.stdout.split("\n")[0] == "stabilizer"


@needs_ours
@needs_ref
def test_errors_equal_reference():
    with tempfile.TemporaryDirectory() as d:
        nc = Circuit(2)
        nc.h(0)
        nc.t(1)
        p = write(nc, d, "nc.txt")
        for args in (("stab-sample", p, 1, 4, os.path.join(d, "o.bin")), ("stab-expval", p, "ZZ")):
            a, b = drive(OURS, *args), drive(REF, *args)
            assert a.returncode == b.returncode == 1
            assert a.stderr == b.stderr and "Clifford" in a.stderr
        c = W.random_clifford(3, 10, 4)
        p = write(c, d)
        for bad in ("ZZ", "ZQZ", "XYZW"):
            a, b = drive(OURS, "stab-expval", p, bad), drive(REF, "stab-expval", p, bad)
            assert a.returncode == b.returncode == 1
            assert a.stderr == b.stderr


@needs_ours
@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_expval_equals_dense_oracle_on_every_pauli(n):
    """All 4^n Pauli strings on random Clifford circuits: the tableau's
    +-1 / 0 against <psi|P|psi> of the dense CPU oracle state."""
    paulis = ["".join(t) for t in itertools.product("IXYZ", repeat=n)]
    with tempfile.TemporaryDirectory() as d:
        for seed in range(6):
            c = W.random_clifford(n, 25, 100 * n + seed)
            state = O.run(c, fuse=False)
            p = write(c, d)
            r = drive(OURS, "stab-expval", p, *paulis)
            assert r.returncode == 0, r.stderr
            got = [float(v) for v in r.stdout.split("\n")[1:len(paulis) + 1]]
            for pauli, v in zip(paulis, got):
                assert abs(v - O.expval(state, pauli)) <= 1e-9, (seed, pauli)


@needs_ours
def test_shots_are_possible_outcomes_and_ghz_correlated():
    """Every shot has nonzero probability in the dense oracle state; GHZ
    shots agree on all qubits and show both outcomes."""
    with tempfile.TemporaryDirectory() as d:
        for seed in range(4):
            c = W.random_clifford(6, 40, 50 + seed)
            c.measure_all()
            state = O.run(c, fuse=False)
            s = shots(OURS, write(c, d), seed, 200, d)
            assert np.all(np.abs(state[s.astype(np.int64)]) ** 2 > 1e-12)
        g = Circuit(5)
        g.h(0)
        for q in range(4):
            g.cx(q, q + 1)
        g.measure_all()
        s = shots(OURS, write(g, d), 7, 400, d)
        assert set(np.unique(s)) == {0, 31}


@needs_ours
def test_wide_circuit_runs_on_cpu():
    """1000 qubits: far beyond any state vector, quadratic in the tableau."""
    c = W.random_clifford(1000, 3000, 5, measures=50)
    with tempfile.TemporaryDirectory() as d:
        s = shots(OURS, write(c, d), 3, 5, d)
        assert s.size == 5
