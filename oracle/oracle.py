"""ctypes loader for the CPU oracle (liboracle.so) and the reference driver.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the CPU baseline, never
as the thing measured. The GPU product path never imports this module.
"""
from __future__ import annotations

import ctypes
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ket_driver")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, u64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_double
        L.oracle_set_threads.argtypes = [i32]
        L.oracle_num_threads.restype = i32
        L.oracle_ground_state.argtypes = [i32, vp]
        L.oracle_apply_gate.argtypes = [i32, vp, vp]
        L.oracle_apply_gate.restype = i32
        L.oracle_apply_matrix.argtypes = [i32, vp, vp, i32, vp]
        L.oracle_run.argtypes = [i32, vp, vp, ctypes.c_size_t, i32]
        L.oracle_run.restype = i32
        L.oracle_block_matrix.argtypes = [vp, ctypes.c_size_t, vp, i32, vp]
        L.oracle_block_matrix.restype = i32
        L.oracle_measure_u.argtypes = [vp, u64, dbl]
        L.oracle_measure_u.restype = u64
        L.oracle_mt_seed.argtypes = [vp, ctypes.c_uint32]
        L.oracle_mt_next.argtypes = [vp]
        L.oracle_mt_next.restype = ctypes.c_uint32
        L.oracle_canonical.argtypes = [vp]
        L.oracle_canonical.restype = dbl
        L.oracle_sample.argtypes = [vp, u64, ctypes.c_uint32, u64, vp]
        L.oracle_expval.argtypes = [vp, i32, ctypes.c_char_p, ctypes.POINTER(i32)]
        L.oracle_expval.restype = dbl
        L.oracle_pauli_masks.argtypes = [ctypes.c_char_p, i32, ctypes.POINTER(u64),
                                         ctypes.POINTER(u64), ctypes.POINTER(i32)]
        L.oracle_pauli_masks.restype = i32
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(t: int) -> None:
    lib().oracle_set_threads(int(t))


def ground_state(n: int) -> np.ndarray:
    s = np.zeros(1 << n, dtype=np.complex128)
    s[0] = 1.0
    return s


def run(circuit, fuse: bool = True, state: np.ndarray | None = None) -> np.ndarray:
    """Evolve |0..0> (or a copy of `state`) through the circuit."""
    rec = circuit.records()
    s = ground_state(circuit.n_qubits) if state is None else np.array(state, dtype=np.complex128)
    rc = lib().oracle_run(circuit.n_qubits, _p(s), _p(rec), len(rec), 1 if fuse else 0)
    if rc != 0:
        raise ValueError("oracle_run rejected the circuit")
    return s


def apply_matrix(state: np.ndarray, qubits, mat: np.ndarray) -> np.ndarray:
    s = np.array(state, dtype=np.complex128)
    n = int(s.size).bit_length() - 1
    q = np.asarray(qubits, dtype=np.int32)
    m = np.ascontiguousarray(mat, dtype=np.complex128)
    lib().oracle_apply_matrix(n, _p(s), _p(q), len(q), _p(m))
    return s


def block_matrix(circuit, qubits) -> np.ndarray:
    rec = circuit.records()
    q = np.asarray(qubits, dtype=np.int32)
    d = 1 << len(q)
    out = np.zeros((d, d), dtype=np.complex128)
    lib().oracle_block_matrix(_p(rec), len(rec), _p(q), len(q), _p(out))
    return out


def measure_u(state: np.ndarray, u: float) -> int:
    s = np.ascontiguousarray(state, dtype=np.complex128)
    return int(lib().oracle_measure_u(_p(s), s.size, float(u)))


class MT19937:
    def __init__(self, seed: int):
        self._buf = (ctypes.c_uint8 * (624 * 4 + 8))()
        lib().oracle_mt_seed(ctypes.cast(self._buf, ctypes.c_void_p), seed)

    def next(self) -> int:
        return int(lib().oracle_mt_next(ctypes.cast(self._buf, ctypes.c_void_p)))

    def canonical(self) -> float:
        return float(lib().oracle_canonical(ctypes.cast(self._buf, ctypes.c_void_p)))


def uniforms(seed: int, count: int) -> np.ndarray:
    r = MT19937(seed)
    return np.array([r.canonical() for _ in range(count)], dtype=np.float64)


def sample(state: np.ndarray, seed: int, shots: int) -> np.ndarray:
    s = np.ascontiguousarray(state, dtype=np.complex128)
    out = np.zeros(shots, dtype=np.uint64)
    lib().oracle_sample(_p(s), s.size, seed, shots, _p(out))
    return out


def expval(state: np.ndarray, pauli: str) -> float:
    s = np.ascontiguousarray(state, dtype=np.complex128)
    n = int(s.size).bit_length() - 1
    err = ctypes.c_int(0)
    v = lib().oracle_expval(_p(s), n, pauli.encode(), ctypes.byref(err))
    if err.value:
        raise ValueError(f"expval: bad Pauli string {pauli!r}")
    return float(v)


def pauli_masks(pauli: str, n: int):
    f, s, y = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_int(0)
    rc = lib().oracle_pauli_masks(pauli.encode(), n, ctypes.byref(f), ctypes.byref(s),
                                  ctypes.byref(y))
    if rc:
        raise ValueError(pauli)
    return int(f.value), int(s.value), int(y.value)


# ---- the compiled reference (oracle/_ref), when it was built ----

def ref_available() -> bool:
    return os.path.exists(REF_DRIVER)


def _ref(args, **kw):
    return subprocess.run([REF_DRIVER] + [str(a) for a in args], check=True,
                          capture_output=True, text=True, **kw)


def ref_run(circuit, threads: int = 1) -> np.ndarray:
    with tempfile.TemporaryDirectory() as d:
        cpath, opath = os.path.join(d, "c.txt"), os.path.join(d, "s.bin")
        with open(cpath, "w") as f:
            f.write(circuit.to_text())
        _ref(["run", cpath, opath, threads])
        return np.fromfile(opath, dtype=np.complex128)


def ref_sample(circuit, seed: int, shots: int, per_call: bool = False) -> np.ndarray:
    with tempfile.TemporaryDirectory() as d:
        cpath, opath = os.path.join(d, "c.txt"), os.path.join(d, "o.bin")
        with open(cpath, "w") as f:
            f.write(circuit.to_text())
        _ref(["sample-calls" if per_call else "sample", cpath, seed, shots, opath])
        return np.fromfile(opath, dtype=np.uint64)


def ref_expval(circuit, paulis) -> tuple[str, list[float]]:
    with tempfile.TemporaryDirectory() as d:
        cpath = os.path.join(d, "c.txt")
        with open(cpath, "w") as f:
            f.write(circuit.to_text())
        out = _ref(["expval", cpath] + list(paulis)).stdout.split("\n")
        return out[0], [float(x) for x in out[1:] if x.strip()]


def ref_uniforms(seed: int, count: int) -> np.ndarray:
    with tempfile.TemporaryDirectory() as d:
        opath = os.path.join(d, "u.bin")
        _ref(["uniforms", seed, count, opath])
        return np.fromfile(opath, dtype=np.float64)


def ref_time(circuit, threads: int, reps: int) -> dict:
    with tempfile.TemporaryDirectory() as d:
        cpath = os.path.join(d, "c.txt")
        with open(cpath, "w") as f:
            f.write(circuit.to_text())
        return json.loads(_ref(["time", cpath, threads, reps]).stdout)


def ref_step_time(circuit, threads: int, reps: int) -> dict:
    """The reference's Stepper over `circuit` on one resident state: setup
    seconds and per-step seconds (reps lists), tests/cpp/ket_driver.cpp."""
    with tempfile.TemporaryDirectory() as d:
        cpath = os.path.join(d, "c.txt")
        with open(cpath, "w") as f:
            f.write(circuit.to_text())
        return json.loads(_ref(["step-time", cpath, threads, reps]).stdout)
