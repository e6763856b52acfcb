"""Pins the CPU oracle (oracle/ket_oracle.c) before anything is checked
against it.

* golden fixtures written by the compiled reference (tests/golden/golden.npz):
  exact equality of states, sampled outcomes, uniforms and expvals;
* the compiled reference itself (oracle/_ref), when built: random circuits,
  fused and unfused, several thread counts, bit for bit;
* mathematics that fixes the answer independently of any ket code: textbook
  gate matrices applied by dense Kronecker products (n <= 8), closed-form
  amplitudes (the cases of the reference's tests/test_simulator.cpp), the
  C++ standard's mt19937 known answer, Born statistics, Pauli eigenstates.
"""
import math
import os

import numpy as np
import pytest

from ket_amd import workloads as W
from ket_amd.gates import Circuit
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "golden.npz")


def circuit_from_text(text: str) -> Circuit:
    lines = [l for l in text.strip().split("\n") if l and not l.startswith("#")]
    c = Circuit(int(lines[0]))
    for l in lines[1:]:
        t = l.split()
        name = t[0]
        if name == "measure":
            c.measure(int(t[1]), int(t[2]))
        elif name == "barrier":
            c.barrier()
        else:
            from ket_amd.gates import ARITY
            k = ARITY[name]
            getattr(c, name)(*[int(x) for x in t[1:1 + k]], *[float(x) for x in t[1 + k:]])
    return c


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def gold_circuits(g):
    names = sorted({k.split("__")[0] for k in g.files if k.endswith("__state")})
    return names


def test_golden_states_exact(gold):
    names = gold_circuits(gold)
    assert len(names) >= 7
    for name in names:
        c = circuit_from_text(gold[f"{name}__text"].tobytes().decode())
        got = O.run(c, fuse=True)
        assert np.array_equal(got.view(np.uint64), gold[f"{name}__state"].view(np.uint64)), name


def test_golden_config1_digest_and_shots(gold):
    c = circuit_from_text(gold["brick20__text"].tobytes().decode())
    assert c.n_unitary() == 590
    s = O.run(c)
    assert np.array_equal(s[gold["brick20__idx"]], gold["brick20__amps"])
    p = s.real ** 2 + s.imag ** 2
    assert np.array_equal(p.reshape(-1, 1 << 14).sum(axis=1), gold["brick20__block_sums"])
    shots = O.sample(s, 1, 10_000)
    assert np.array_equal(shots.astype(np.uint32), gold["brick20__shots_seed1"])


def test_golden_uniforms(gold):
    u = O.uniforms(1, 20_000)
    assert np.array_equal(u[:256], gold["uniforms_seed1_first256"])
    assert math.fsum(u) == gold["uniforms_seed1_sum"][0]


def test_golden_expval(gold):
    c = circuit_from_text(gold["mixed10__text"].tobytes().decode())
    s = O.run(c)
    paulis = gold["mixed10__paulis"].tobytes().decode().split("\n")
    for p, want in zip(paulis, gold["mixed10__expval"]):
        assert abs(O.expval(s, p) - want) <= 1e-15, p


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_against_reference_build(seed):
    c = W.random_mixed(13, 10, seed)
    ref = O.ref_run(c, threads=4)
    for threads in (1, 4):
        O.set_threads(threads)
        for fuse in (True,):
            got = O.run(c, fuse=fuse)
            assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    O.set_threads(1)
    got_u = O.run(c, fuse=False)
    assert np.abs(got_u - ref).max() <= 1e-12


# ---- textbook linear algebra (independent of ket's code) ----

SQ2 = 1 / math.sqrt(2)


def mat1(name, p):
    c, s = (math.cos(p[0] / 2), math.sin(p[0] / 2)) if p else (0, 0)
    return {
        "h": np.array([[1, 1], [1, -1]]) * SQ2,
        "x": np.array([[0, 1], [1, 0]]),
        "y": np.array([[0, -1j], [1j, 0]]),
        "z": np.diag([1, -1]),
        "s": np.diag([1, 1j]),
        "sdg": np.diag([1, -1j]),
        "t": np.diag([1, np.exp(1j * math.pi / 4)]),
        "tdg": np.diag([1, np.exp(-1j * math.pi / 4)]),
        "rx": np.array([[c, -1j * s], [-1j * s, c]]),
        "ry": np.array([[c, -s], [s, c]]),
        "rz": np.diag([np.exp(-1j * p[0] / 2), np.exp(1j * p[0] / 2)]) if p else None,
        "u": np.array([[c, -np.exp(1j * p[2]) * s],
                       [np.exp(1j * p[1]) * s, np.exp(1j * (p[1] + p[2])) * c]]) if len(p) == 3 else None,
    }[name].astype(np.complex128)


def full_matrix(n, name, qs, p):
    """Dense 2^n matrix of one gate by its action on basis states."""
    d = 1 << n
    M = np.zeros((d, d), dtype=np.complex128)
    for col in range(d):
        bits = [(col >> q) & 1 for q in range(n)]
        out = {}
        def put(b, amp):
            idx = sum(bb << q for q, bb in enumerate(b))
            out[idx] = out.get(idx, 0) + amp
        if name in ("h", "x", "y", "z", "s", "sdg", "t", "tdg", "rx", "ry", "rz", "u"):
            m = mat1(name, p)
            t = qs[0]
            for nb in (0, 1):
                b = list(bits); b[t] = nb
                put(b, m[nb, bits[t]])
        elif name in ("ch", "cx", "cy", "crx", "cry", "crz", "cu"):
            base = {"ch": "h", "cx": "x", "cy": "y", "crx": "rx", "cry": "ry", "crz": "rz", "cu": "u"}[name]
            ctl, t = qs
            if bits[ctl] == 0:
                put(bits, 1)
            else:
                m = mat1(base, p)
                for nb in (0, 1):
                    b = list(bits); b[t] = nb
                    put(b, m[nb, bits[t]])
        elif name in ("cz", "cp"):
            ph = -1 if name == "cz" else np.exp(1j * p[0])
            put(bits, ph if bits[qs[0]] and bits[qs[1]] else 1)
        elif name == "swap":
            b = list(bits); b[qs[0]], b[qs[1]] = bits[qs[1]], bits[qs[0]]
            put(b, 1)
        elif name == "ccx":
            b = list(bits)
            if bits[qs[0]] and bits[qs[1]]:
                b[qs[2]] ^= 1
            put(b, 1)
        elif name == "cswap":
            b = list(bits)
            if bits[qs[0]]:
                b[qs[1]], b[qs[2]] = bits[qs[2]], bits[qs[1]]
            put(b, 1)
        for idx, amp in out.items():
            M[idx, col] += amp
    return M


def brute(c: Circuit) -> np.ndarray:
    s = np.zeros(1 << c.n_qubits, dtype=np.complex128)
    s[0] = 1
    for name, qs, ps in c.ops:
        if name in ("measure", "barrier"):
            continue
        s = full_matrix(c.n_qubits, name, qs, list(ps)) @ s
    return s


@pytest.mark.parametrize("seed", range(6))
def test_brute_force_kronecker(seed):
    c = W.random_mixed(7, 8, 300 + seed)
    want = brute(c)
    for fuse in (False, True):
        got = O.run(c, fuse=fuse)
        assert np.abs(got - want).max() <= 1e-12


def test_every_gate_against_textbook():
    rng = np.random.default_rng(0)
    from ket_amd.gates import ARITY, NPARAMS
    for name, k in ARITY.items():
        for _ in range(4):
            qs = [int(x) for x in rng.permutation(5)[:k]]
            ps = list(rng.uniform(0, 2 * np.pi, NPARAMS.get(name, 0)))
            c = Circuit(5)
            c.h(0); c.ry(1, 0.4); c.u(2, 0.3, 0.2, 0.1); c.rx(3, 1.2); c.h(4)
            getattr(c, name)(*qs, *ps)
            assert np.abs(O.run(c, fuse=False) - brute(c)).max() <= 1e-13, name


def amp(s, bits: str):
    return s[int(bits, 2)]


def test_closed_forms():
    c = Circuit(1); c.x(0); s = O.run(c); assert s[1] == 1 and s[0] == 0
    c = Circuit(1); c.h(0); s = O.run(c); assert abs(s[0] - SQ2) < 1e-12 and abs(s[1] - SQ2) < 1e-12
    c = Circuit(1); c.h(0); c.h(0); s = O.run(c); assert abs(s[0] - 1) < 1e-12 and abs(s[1]) < 1e-12
    c = Circuit(2); c.h(0); c.cx(0, 1); s = O.run(c)
    assert abs(amp(s, "00") - SQ2) < 1e-12 and abs(amp(s, "11") - SQ2) < 1e-12
    for g, ph in [("s", 1j), ("sdg", -1j), ("t", np.exp(1j * math.pi / 4)), ("tdg", np.exp(-1j * math.pi / 4)), ("z", -1)]:
        c = Circuit(1); c.x(0); getattr(c, g)(0); s = O.run(c); assert abs(s[1] - ph) < 1e-12, g
    c1 = Circuit(1); c1.h(0); c1.t(0); c1.t(0)
    c2 = Circuit(1); c2.h(0); c2.s(0)
    assert np.abs(O.run(c1) - O.run(c2)).max() < 1e-12
    c = Circuit(1); c.y(0); s = O.run(c); assert abs(s[1] - 1j) < 1e-12
    c = Circuit(2); c.x(0); c.x(1); c.cz(0, 1); s = O.run(c); assert abs(amp(s, "11") + 1) < 1e-12
    c = Circuit(2); c.x(0); c.x(1); c.cp(0, 1, math.pi / 2); s = O.run(c); assert abs(amp(s, "11") - 1j) < 1e-12
    th = 0.7
    c = Circuit(1); c.ry(0, th); s = O.run(c); assert abs(s[0] - math.cos(th / 2)) < 1e-12 and abs(s[1] - math.sin(th / 2)) < 1e-12
    c = Circuit(1); c.rx(0, th); s = O.run(c); assert abs(s[1] + 1j * math.sin(th / 2)) < 1e-12
    c = Circuit(1); c.rz(0, th); s = O.run(c); assert abs(s[0] - np.exp(-1j * th / 2)) < 1e-12
    for a, b in [(0, 0), (0, 1), (1, 0), (1, 1)]:
        c = Circuit(3)
        if a: c.x(0)
        if b: c.x(1)
        c.ccx(0, 1, 2); s = O.run(c)
        idx = a | (b << 1) | ((a & b) << 2)
        assert s[idx] == 1
    c = Circuit(3); c.x(0); c.x(1); c.cswap(0, 1, 2); s = O.run(c); assert s[0b101] == 1
    c = Circuit(3); c.x(1); c.cswap(0, 1, 2); s = O.run(c); assert s[0b010] == 1
    c = Circuit(2); c.x(0); c.swap(0, 1); s = O.run(c); assert s[0b10] == 1
    c1 = Circuit(3); c1.h(0); c1.ry(2, 0.3); c1.swap(0, 2)
    c2 = Circuit(3); c2.h(0); c2.ry(2, 0.3); c2.cx(0, 2); c2.cx(2, 0); c2.cx(0, 2)
    assert np.abs(O.run(c1) - O.run(c2)).max() < 1e-12
    c1 = Circuit(1); c1.u(0, math.pi / 2, 0, math.pi)
    c2 = Circuit(1); c2.h(0)
    assert np.abs(O.run(c1) - O.run(c2)).max() < 1e-12
    c1 = Circuit(1); c1.u(0, math.pi, 0, math.pi)
    c2 = Circuit(1); c2.x(0)
    assert np.abs(O.run(c1) - O.run(c2)).max() < 1e-12
    c = Circuit(2); c.x(0); c.ch(0, 1); s = O.run(c)
    assert abs(amp(s, "01") - SQ2) < 1e-12 and abs(amp(s, "11") - SQ2) < 1e-12
    c = Circuit(2); c.x(0); c.cy(0, 1); s = O.run(c); assert abs(amp(s, "11") - 1j) < 1e-12


def test_fusion_matches_unfused_20q():
    rng = np.random.default_rng(99)
    c = Circuit(20)
    for _ in range(60):
        k = int(rng.integers(5))
        q = int(rng.integers(20))
        if k == 0: c.h(q)
        elif k == 1: c.ry(q, rng.uniform(0, 6.28))
        elif k == 2: c.rz(q, rng.uniform(0, 6.28))
        else:
            r = int(rng.integers(19)); r = r if r < q else r + 1
            (c.cx if k == 3 else c.cz)(q, r)
    assert np.abs(O.run(c, fuse=True) - O.run(c, fuse=False)).max() <= 1e-9


def test_threads_bit_identical():
    c = W.random_mixed(16, 30, 2024)
    O.set_threads(1)
    a = O.run(c)
    for t in (2, 4, 8):
        O.set_threads(t)
        assert np.array_equal(O.run(c).view(np.uint64), a.view(np.uint64))
    O.set_threads(1)


def test_mt19937_known_answer_and_canonical():
    r = O.MT19937(5489)
    v = [r.next() for _ in range(10000)]
    assert v[-1] == 4123659995  # [rand.predef]
    r1, r2 = O.MT19937(1), O.MT19937(1)
    a, b = r1.next(), r1.next()
    want = (a + b * 2.0 ** 32) / 2.0 ** 64
    assert r2.canonical() == want


def test_born_statistics_bell():
    c = Circuit(2); c.h(0); c.cx(0, 1)
    s = O.run(c)
    out = O.sample(s, 42, 10_000)
    assert set(np.unique(out)) <= {0, 3}
    assert 4800 < int((out == 0).sum()) < 5200
    c = Circuit(1); s = O.run(c)
    assert (O.sample(s, 42, 100) == 0).all()
    c.x(0); s = O.run(c)
    assert (O.sample(s, 42, 100) == 1).all()


def test_measure_sequential_scan_and_fallback():
    s = np.array([0.6, 0.8j, 0, 0], dtype=np.complex128)
    assert O.measure_u(s, 0.0) == 0
    assert O.measure_u(s, 0.3599) == 0
    assert O.measure_u(s, 0.36000000000000004) == 1
    short = np.array([0.5, 0.5], dtype=np.complex128)  # norm 0.5
    assert O.measure_u(short, 0.9) == 1  # falls back to size - 1


def pauli_matrix(p: str):
    m = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.diag([1, -1])}
    out = np.array([[1]], dtype=np.complex128)
    for ch in p.upper():  # leftmost = highest qubit
        out = np.kron(out, m[ch])
    return out


def test_expval_eigenstates_and_bruteforce():
    c = Circuit(1); s = O.run(c); assert O.expval(s, "Z") == 1.0
    c.x(0); s = O.run(c); assert O.expval(s, "Z") == -1.0
    c = Circuit(1); c.h(0); s = O.run(c); assert abs(O.expval(s, "X") - 1) < 1e-12
    c = Circuit(1); c.h(0); c.s(0); s = O.run(c); assert abs(O.expval(s, "Y") - 1) < 1e-12
    c = Circuit(2); c.h(0); c.cx(0, 1); s = O.run(c)
    for p, v in [("ZZ", 1), ("XX", 1), ("YY", -1)]:
        assert abs(O.expval(s, p) - v) < 1e-12
    c = Circuit(2); c.x(0); s = O.run(c)
    assert O.expval(s, "IZ") == -1.0 and O.expval(s, "ZI") == 1.0  # rightmost = qubit 0
    with pytest.raises(ValueError):
        O.expval(s, "Z")
    with pytest.raises(ValueError):
        O.expval(s, "ZQ")
    rng = np.random.default_rng(4)
    s = O.run(W.random_mixed(6, 6, 8))
    for _ in range(20):
        p = "".join(rng.choice(list("IXYZ"), 6))
        want = np.vdot(s, pauli_matrix(p) @ s).real
        assert abs(O.expval(s, p) - want) <= 1e-12, p
