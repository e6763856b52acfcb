"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Run in a container where the reference tree is present and oracle/_ref is
built (`make -C oracle ref`). Every vector here comes out of the reference's
own ket::run / Simulation::sample / Simulation::expval through
tests/cpp/ket_driver.cpp; nothing is computed by this repository's code.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ket_amd import workloads as W  # noqa: E402
from ket_amd.gates import Circuit  # noqa: E402
from oracle import oracle as O  # noqa: E402


def closed_form_circuits():
    """Small circuits covering every gate type and operand order."""
    out = {}
    c = Circuit(4)
    for g in ["h", "x", "y", "z", "s", "sdg", "t", "tdg"]:
        getattr(c, g)(1)
        c.h(0)
    c.rx(2, 0.3); c.ry(3, 1.1); c.rz(0, 2.2); c.u(1, 0.1, 0.2, 0.3)
    c.ch(0, 1); c.cx(1, 2); c.cy(2, 3); c.cz(3, 0); c.crx(0, 2, 0.5)
    c.cry(1, 3, 0.7); c.crz(2, 0, 0.9); c.cu(3, 1, 0.1, 0.4, 0.8)
    c.cp(0, 3, 1.3); c.swap(1, 2); c.ccx(0, 1, 3); c.cswap(2, 0, 1)
    c.measure_all()
    out["allgates4"] = c
    # Same-qubit runs that ket's Fuser merges (one- and two-qubit blocks).
    c = Circuit(5)
    for q in range(5):
        c.h(q); c.t(q); c.rz(q, 0.1 * (q + 1)); c.ry(q, 0.2 * (q + 2))
    for q in range(4):
        c.cx(q, q + 1); c.rz(q + 1, 0.3); c.cx(q, q + 1); c.cz(q, q + 1); c.cp(q + 1, q, 0.7)
    c.ccx(0, 1, 2); c.cswap(0, 1, 2); c.ccx(0, 1, 2)
    out["fuser_runs5"] = c
    # Many random angles through every parametrised gate: pins the exact
    # libm path ket's matrices take (sin/cos pairs via sincos).
    rng = np.random.default_rng(31)
    c = Circuit(3)
    for _ in range(400):
        a, b, d = (float(x) for x in rng.uniform(0, 2 * math.pi, 3))
        q = [int(x) for x in rng.permutation(3)]
        c.u(q[0], a, b, d); c.rx(q[1], a); c.ry(q[2], b); c.rz(q[0], d)
        c.cu(q[0], q[1], a, b, d); c.crx(q[1], q[2], b); c.cry(q[2], q[0], d)
        c.crz(q[0], q[2], a); c.cp(q[1], q[0], d)
    out["angles3"] = c
    out["qft10"] = W.qft_u(10, seed=3)
    out["brick10"] = W.brickwork(10, 8, seed=4)
    out["mixed10"] = W.random_mixed(10, 12, seed=5)
    out["bwm10"] = W.brickwork_mixed(10, 12, seed=6)
    out["hea8"] = W.hea(8, 8, seed=7)
    return out


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/ket_driver is not built (make -C oracle ref)")
    small = closed_form_circuits()
    arrays = {}
    for name, c in small.items():
        arrays[f"{name}__text"] = np.frombuffer(c.to_text().encode(), dtype=np.uint8)
        arrays[f"{name}__state"] = O.ref_run(c)
    paulis = ["ZZZZZZZZZZ", "XYZIXYZIXY", "IIIIIIIIYY", "xzxzxzxzxz"]
    bk, ev = O.ref_expval(small["mixed10"], paulis)
    arrays["mixed10__paulis"] = np.frombuffer("\n".join(paulis).encode(), dtype=np.uint8)
    arrays["mixed10__expval"] = np.array(ev)
    # Config 1: 20-qubit brickwork, 20 layers, seed 0; 10^4 shots, mt19937{1}.
    c1 = W.brickwork(20, 20, seed=0)
    s1 = O.ref_run(c1)
    rng = np.random.default_rng(2024)
    idx = np.sort(rng.choice(1 << 20, 4096, replace=False))
    arrays["brick20__text"] = np.frombuffer(c1.to_text().encode(), dtype=np.uint8)
    arrays["brick20__idx"] = idx.astype(np.int64)
    arrays["brick20__amps"] = s1[idx]
    p = s1.real ** 2 + s1.imag ** 2
    arrays["brick20__block_sums"] = p.reshape(-1, 1 << 14).sum(axis=1)
    arrays["brick20__norm"] = np.array([p.sum()])
    shots = O.ref_sample(c1, 1, 10_000)
    arrays["brick20__shots_seed1"] = shots.astype(np.uint32)
    u = O.ref_uniforms(1, 20_000)
    arrays["uniforms_seed1_first256"] = u[:256]
    arrays["uniforms_seed1_sum"] = np.array([math.fsum(u)])
    path = os.path.join(HERE, "golden.npz")
    np.savez_compressed(path, **arrays)
    print(f"wrote {path} ({os.path.getsize(path)} bytes), backend={bk}")


if __name__ == "__main__":
    main()
