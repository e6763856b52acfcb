"""Seeded circuit generators for the benchmark and parity workloads.

Shapes follow the reference benchmark generators
(.github/scripts/benchmark/circuits.py:33-43 `qft`, 46-60 `random_circuit`).
Angles come from numpy's PCG64 generator (`np.random.default_rng(seed)`),
drawn in gate order; the choices the reference leaves open are frozen here:

* brickwork(n, layers): each layer is U(theta, phi, lambda) on every qubit,
  then CX(i, i+1) for i = layer % 2, layer % 2 + 2, ... (i + 1 < n).
* qft_u(n): one U layer, then H(j) and CP(pi / 2^(k-j)) on (k, j) for k > j,
  then SWAP(i, n-1-i).
* brickwork_mixed(n, depth): U on every qubit, then per staggered pair a coin
  flip between CX and CZ; on every 4th layer (layer % 4 == 3) also
  floor(n / 3) CCX gates on disjoint random triples (a random permutation cut
  into triples, the first two as controls).
* hea(n, layers): Ry then Rz on every qubit with angles U[-pi, pi], then a CZ
  ring (i, (i+1) % n); the observable is the open chain sum_i Z_i Z_{i+1}.
"""
from __future__ import annotations

import math

import numpy as np

from .gates import Circuit

TWO_PI = 2.0 * math.pi


def brickwork(n: int, layers: int, seed: int = 0, measure: bool = True) -> Circuit:
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for layer in range(layers):
        for q in range(n):
            th, ph, la = rng.uniform(0.0, TWO_PI, 3)
            c.u(q, th, ph, la)
        for a in range(layer % 2, n - 1, 2):
            c.cx(a, a + 1)
    if measure:
        c.measure_all()
    return c


def qft_u(n: int, seed: int = 0) -> Circuit:
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for q in range(n):
        th, ph, la = rng.uniform(0.0, TWO_PI, 3)
        c.u(q, th, ph, la)
    for j in range(n):
        c.h(j)
        for k in range(j + 1, n):
            c.cp(k, j, math.pi / 2 ** (k - j))
    for i in range(n // 2):
        c.swap(i, n - 1 - i)
    return c


def brickwork_mixed(n: int, depth: int, seed: int = 0) -> Circuit:
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for layer in range(depth):
        for q in range(n):
            th, ph, la = rng.uniform(0.0, TWO_PI, 3)
            c.u(q, th, ph, la)
        for a in range(layer % 2, n - 1, 2):
            if rng.random() < 0.5:
                c.cx(a, a + 1)
            else:
                c.cz(a, a + 1)
        if layer % 4 == 3:
            perm = rng.permutation(n)
            for t in range(n // 3):
                a, b, tgt = (int(x) for x in perm[3 * t: 3 * t + 3])
                c.ccx(a, b, tgt)
    return c


def hea(n: int, layers: int = 8, seed: int = 0) -> Circuit:
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for _ in range(layers):
        for q in range(n):
            c.ry(q, rng.uniform(-math.pi, math.pi))
            c.rz(q, rng.uniform(-math.pi, math.pi))
        for q in range(n):
            if (q + 1) % n != q and not (n == 2 and q == 1):
                c.cz(q, (q + 1) % n)
    return c


def zz_chain(n: int) -> list[str]:
    """Pauli strings of sum_i Z_i Z_{i+1}, i = 0..n-2 (rightmost char = qubit 0)."""
    out = []
    for i in range(n - 1):
        s = ["I"] * n
        s[n - 1 - i] = "Z"
        s[n - 2 - i] = "Z"
        out.append("".join(s))
    return out


GATE_SET_13 = ["h", "s", "t", "rx", "ry", "rz", "u", "cx", "cz", "cp", "swap",
               "ccx", "cswap"]


def random_mixed(n: int, depth: int, seed: int) -> Circuit:
    """`depth` layers of one gate per free qubit slot, gates uniform over the
    13-gate set {H,S,T,Rx,Ry,Rz,U,CX,CZ,CP,SWAP,CCX,CSWAP}, angles U[0, 2pi).
    Within a layer, qubits are visited in a random order and a multi-qubit gate
    takes the next free qubits of that order."""
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for _ in range(depth):
        order = [int(x) for x in rng.permutation(n)]
        i = 0
        while i < n:
            name = GATE_SET_13[int(rng.integers(len(GATE_SET_13)))]
            k = {"cx": 2, "cz": 2, "cp": 2, "swap": 2, "ccx": 3, "cswap": 3}.get(name, 1)
            if i + k > n:
                name, k = "h", 1
            qs = order[i: i + k]
            i += k
            if name in ("rx", "ry", "rz", "cp"):
                getattr(c, name)(*qs, rng.uniform(0.0, TWO_PI))
            elif name == "u":
                c.u(qs[0], *rng.uniform(0.0, TWO_PI, 3))
            else:
                getattr(c, name)(*qs)
    return c


def diag_mix(n: int, depth: int, seed: int) -> Circuit:
    """Dense gates interleaved with every diagonal gate type (Z, S, Sdg, T,
    Tdg, Rz, CZ, CP, CRz) on random, often distant, qubit pairs."""
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for _ in range(depth):
        for q in range(n):
            c.u(q, *rng.uniform(0.0, TWO_PI, 3))
        for _ in range(n):
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            kind = int(rng.integers(9))
            if kind == 0: c.cz(a, b)
            elif kind == 1: c.cp(a, b, rng.uniform(0.0, TWO_PI))
            elif kind == 2: c.crz(a, b, rng.uniform(0.0, TWO_PI))
            elif kind == 3: c.rz(a, rng.uniform(0.0, TWO_PI))
            elif kind == 4: c.z(a)
            elif kind == 5: c.s(a)
            elif kind == 6: c.sdg(a)
            elif kind == 7: c.t(a)
            else: c.tdg(a)
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        c.cx(a, b)
        c.h(int(rng.integers(n)))
    return c


CLIFFORD_1Q = ("h", "s", "sdg", "x", "y", "z")
CLIFFORD_2Q = ("cx", "cy", "cz", "swap")


def random_clifford(n: int, depth: int, seed: int, measures: int = 0) -> Circuit:
    """Random circuit over ket's Clifford set (H, S, Sdg, X, Y, Z, CX, CY, CZ,
    SWAP), with `measures` measurements of random qubits into random clbits
    at random positions (mid-circuit and repeated qubits included)."""
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    slots = sorted(rng.choice(depth + 1, size=measures, replace=True)) if measures else []
    k = 0
    for d in range(depth):
        while k < len(slots) and slots[k] == d:
            c.measure(int(rng.integers(n)), int(rng.integers(max(1, min(n, 64)))))
            k += 1
        if n > 1 and rng.random() < 0.4:
            a, b = rng.choice(n, 2, replace=False)
            getattr(c, CLIFFORD_2Q[rng.integers(len(CLIFFORD_2Q))])(int(a), int(b))
        else:
            getattr(c, CLIFFORD_1Q[rng.integers(len(CLIFFORD_1Q))])(int(rng.integers(n)))
    while k < len(slots):
        c.measure(int(rng.integers(n)), int(rng.integers(max(1, min(n, 64)))))
        k += 1
    return c


_SELF_INVERSE = {"h", "x", "y", "z", "cx", "cy", "cz", "ch", "swap", "ccx", "cswap"}
_SWAP_NAME = {"s": "sdg", "sdg": "s", "t": "tdg", "tdg": "t"}


def inverse(c: Circuit) -> Circuit:
    """The unitary part of `c` reversed and adjointed gate by gate:
    U(t, p, l)^+ = U(-t, -l, -p), rotations and phases negate their angle."""
    out = Circuit(c.n_qubits)
    for name, qs, ps in reversed(c.ops):
        if name in ("measure", "barrier"):
            continue
        if name in _SELF_INVERSE:
            getattr(out, name)(*qs)
        elif name in _SWAP_NAME:
            getattr(out, _SWAP_NAME[name])(*qs)
        elif name in ("u", "cu"):
            t, p, l = ps
            getattr(out, name)(*qs, -t, -l, -p)
        else:  # rx, ry, rz, cp, crx, cry, crz
            getattr(out, name)(*qs, *[-x for x in ps])
    return out
