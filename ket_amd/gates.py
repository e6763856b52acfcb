"""Flat gate records and a small circuit builder for the C ABI.

The record layout is `ket_gate` in include/ket_amd.h (48 bytes): type code in
the order of ket::GateType (reference include/ket/gate.hpp:16-29), arity, up to
three qubits, up to three parameters. `Circuit` mirrors the gate methods of
ket::Circuit (reference include/ket/circuit.hpp:29-131) so workloads read the
same way in Python and C++; it validates operands like Circuit::add
(reference src/circuit.cpp:20-40).
"""
from __future__ import annotations

import numpy as np

NAMES = [
    "h", "x", "y", "z", "s", "sdg", "t", "tdg", "rx", "ry", "rz", "u",
    "ch", "cx", "cy", "cz", "crx", "cry", "crz", "cu", "cp", "swap",
    "ccx", "cswap", "measure", "barrier", "composite", "probe",
]
CODE = {n: i for i, n in enumerate(NAMES)}
ARITY = {n: (1 if i <= CODE["u"] else 2 if i <= CODE["swap"] else 3)
         for i, n in enumerate(NAMES[:CODE["cswap"] + 1])}
NPARAMS = {"rx": 1, "ry": 1, "rz": 1, "u": 3, "crx": 1, "cry": 1, "crz": 1,
           "cu": 3, "cp": 1}
CLIFFORD = {"h", "x", "y", "z", "s", "sdg", "cx", "cy", "cz", "swap"}
STRUCTURAL = {"measure", "barrier", "probe"}

GATE_DTYPE = np.dtype([
    ("type", "<i4"), ("nq", "<i4"), ("q", "<i4", (3,)), ("pad", "<i4"),
    ("p", "<f8", (3,)),
])
assert GATE_DTYPE.itemsize == 48


class Circuit:
    """An ordered gate list over `n_qubits` qubits."""

    def __init__(self, n_qubits: int):
        self.n_qubits = int(n_qubits)
        self.n_clbits = 0
        self.ops: list[tuple[str, tuple[int, ...], tuple[float, ...]]] = []

    def _add(self, name, qubits, params=()):
        for q in qubits:
            if not 0 <= q < self.n_qubits:
                raise IndexError(f"qubit index {q} is out of range for a "
                                 f"{self.n_qubits}-qubit circuit")
        if len(set(qubits)) != len(qubits):
            raise ValueError("a gate cannot act on a qubit more than once")
        self.ops.append((name, tuple(int(q) for q in qubits),
                         tuple(float(p) for p in params)))

    def __getattr__(self, name):
        if name in ARITY:
            k, npar = ARITY[name], NPARAMS.get(name, 0)

            def add(*args):
                if len(args) != k + npar:
                    raise TypeError(f"{name} takes {k} qubits and {npar} params")
                self._add(name, args[:k], args[k:])
            return add
        raise AttributeError(name)

    def measure(self, qubit, clbit):
        self._add("measure", (qubit,))
        self.ops[-1] = ("measure", (int(qubit), int(clbit)), ())
        self.n_clbits = max(self.n_clbits, int(clbit) + 1)

    def measure_all(self):
        for q in range(self.n_qubits):
            self.measure(q, q)

    def barrier(self):
        self.ops.append(("barrier", (), ()))

    def n_unitary(self) -> int:
        return sum(1 for op in self.ops if op[0] in ARITY)

    def is_clifford(self) -> bool:
        return all(n in CLIFFORD or n in STRUCTURAL for n, _, _ in self.ops)

    def records(self) -> np.ndarray:
        rec = np.zeros(len(self.ops), dtype=GATE_DTYPE)
        for i, (name, qs, ps) in enumerate(self.ops):
            rec[i]["type"] = CODE[name]
            if name == "measure":
                rec[i]["nq"] = 1
                rec[i]["q"][0] = qs[0]
                rec[i]["q"][1] = qs[1]
            elif name in ARITY:
                rec[i]["nq"] = len(qs)
                rec[i]["q"][: len(qs)] = qs
            rec[i]["p"][: len(ps)] = ps
        return rec

    def to_text(self) -> str:
        lines = [str(self.n_qubits)]
        for name, qs, ps in self.ops:
            parts = [name] + [str(q) for q in qs] + [repr(float(p)) for p in ps]
            lines.append(" ".join(parts))
        return "\n".join(lines) + "\n"

    def prefix(self, k: int) -> "Circuit":
        """The first k unitary gates (structural records dropped)."""
        c = Circuit(self.n_qubits)
        for op in self.ops:
            if op[0] in ARITY and c.n_unitary() < k:
                c.ops.append(op)
        return c
