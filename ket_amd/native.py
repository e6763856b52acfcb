"""ctypes binding of the C ABI in include/ket_amd.h (libket_amd.so).

The library is built in-tree by `make -C ket_amd`; importing this module never
builds anything and raises if the library is missing, so a GPU run without
the HIP extension fails loudly instead of falling back to anything else.
"""
from __future__ import annotations

import contextlib
import ctypes
import os

import numpy as np

from .gates import GATE_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libket_amd.so")

C128, C64 = 0, 1
FUSE_NONE, FUSE_REFERENCE, FUSE_BLOCKS, FUSE_LEARNED, FUSE_BLOCKS_ALL, FUSE_EXACT, FUSE_RUNS = (
    0, 1, 2, 3, 4, 5, 6)

STATUS = {1: ValueError, 2: IndexError, 3: RuntimeError, 4: MemoryError,
          5: NotImplementedError}


MLP_INPUTS, MLP_HIDDEN = 16, 64
MLP_PARAMS = MLP_HIDDEN * MLP_INPUTS + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + 2 * MLP_HIDDEN + 1


class LaunchRecord(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("k", ctypes.c_int32), ("n_gates", ctypes.c_int32),
                ("n_terms", ctypes.c_int32), ("qubit_mask", ctypes.c_uint64),
                ("alg_bytes", ctypes.c_double), ("flops", ctypes.c_double),
                ("ms", ctypes.c_double), ("features", ctypes.c_float * MLP_INPUTS)]


LAUNCH_KINDS = ["dense1q", "ch", "permute", "phase", "matrix", "block", "run"]


class BatchStats(ctypes.Structure):
    _fields_ = [("instances_lds", ctypes.c_uint64), ("instances_hbm", ctypes.c_uint64),
                ("launches", ctypes.c_uint64), ("gates", ctypes.c_uint64),
                ("alg_bytes", ctypes.c_double), ("plan_seconds", ctypes.c_double),
                ("device_ms", ctypes.c_double), ("graph_replayed", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("pack_seconds", ctypes.c_double),
                ("upload_seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class RunStats(ctypes.Structure):
    _fields_ = [("gates", ctypes.c_uint64), ("launches", ctypes.c_uint64),
                ("blocks", ctypes.c_uint64), ("alg_bytes", ctypes.c_double),
                ("flops", ctypes.c_double), ("plan_seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C ket_amd` "
                "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, u64, dbl, sz = (ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64,
                                 ctypes.c_double, ctypes.c_size_t)
        P = ctypes.POINTER
        sigs = {
            "ket_last_error": ([], ctypes.c_char_p),
            "ket_device_info": ([ctypes.c_char_p, sz, P(i32), P(u64)], i32),
            "ket_sv_create": ([i32, i32, P(vp)], i32),
            "ket_sv_destroy": ([vp], i32),
            "ket_sv_info": ([vp, P(i32), P(i32)], i32),
            "ket_sv_set_basis": ([vp, u64], i32),
            "ket_sv_copy_in": ([vp, u64, u64, vp], i32),
            "ket_sv_copy_out": ([vp, u64, u64, vp], i32),
            "ket_sv_copy": ([vp, vp], i32),
            "ket_sv_sync": ([vp], i32),
            "ket_sv_apply_1q": ([vp, i32, u64, vp], i32),
            "ket_sv_apply_phase": ([vp, u64, dbl, dbl], i32),
            "ket_sv_apply_x": ([vp, i32, u64], i32),
            "ket_sv_apply_swap": ([vp, i32, i32, u64], i32),
            "ket_sv_apply_gate": ([vp, vp], i32),
            "ket_sv_apply_matrix": ([vp, i32, vp, vp], i32),
            "ket_sv_apply_circuit": ([vp, vp, sz, i32, i32, P(RunStats)], i32),
            "ket_sv_norm": ([vp, P(dbl)], i32),
            "ket_sv_renormalize": ([vp], i32),
            "ket_sv_sample": ([vp, vp, u64, vp], i32),
            "ket_sv_prob_prefix": ([vp, vp, u64, vp], i32),
            "ket_sv_expval_pauli": ([vp, u64, u64, i32, P(dbl)], i32),
            "ket_sv_expval_pauli_sum": ([vp, sz, vp, vp, vp, vp, P(dbl)], i32),
            "ket_sv_fidelity": ([vp, vp, P(dbl), P(dbl)], i32),
            "ket_expval_pauli_sum_noalloc": ([vp, i32, i32, sz, vp, vp, vp, vp, vp, P(dbl)], i32),
            "ket_sv_profile_circuit": ([vp, vp, sz, i32, i32, vp, sz, P(sz)], i32),
            "ket_set_tuning": ([vp], i32),
            "ket_set_tuning_for": ([i32, vp], i32),
            "ket_get_tuning_for": ([i32, vp], i32),
            "ket_get_tuning": ([vp], i32),
            "ket_calibrate_copy": ([u64, i32, i32, P(dbl)], i32),
            "ket_batch_create": ([sz, vp, i32, P(vp)], i32),
            "ket_batch_destroy": ([vp], i32),
            "ket_batch_run": ([vp, vp, vp, vp, vp, vp, i32, i32, vp, P(BatchStats)], i32),
            "ket_mlp_create": ([u64, P(vp)], i32),
            "ket_mlp_destroy": ([vp], i32),
            "ket_mlp_set_params": ([vp, vp, sz], i32),
            "ket_mlp_get_params": ([vp, vp, sz], i32),
            "ket_mlp_forward": ([vp, vp, sz, vp], i32),
            "ket_mlp_gradients": ([vp, vp, vp, sz, vp, vp], i32),
            "ket_mlp_train_step": ([vp, vp, vp, sz, ctypes.c_float, vp], i32),
            "ket_mlp_use_for_planning": ([vp], i32),
            "ket_sv_event_record": ([vp, i32], i32),
            "ket_sv_event_elapsed": ([vp, i32, i32, P(ctypes.c_float)], i32),
        }
        for name, (args, res) in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().ket_last_error().decode(errors="replace")
        raise STATUS.get(rc, RuntimeError)(msg)


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def device_info() -> dict:
    name = ctypes.create_string_buffer(256)
    cu, hbm = ctypes.c_int(0), ctypes.c_uint64(0)
    check(lib().ket_device_info(name, 256, ctypes.byref(cu), ctypes.byref(hbm)))
    return {"name": name.value.decode(), "compute_units": cu.value, "hbm_bytes": hbm.value}


class Tuning(ctypes.Structure):
    _fields_ = [("items_per_thread", ctypes.c_int32), ("nontemporal", ctypes.c_int32),
                ("blocks_per_cu", ctypes.c_int32), ("block_tiles", ctypes.c_int32),
                ("block_bpc", ctypes.c_int32), ("block_stage", ctypes.c_int32),
                ("pair_exchange", ctypes.c_int32), ("block_k", ctypes.c_int32),
                ("exact_runs", ctypes.c_int32), ("run_low_bits", ctypes.c_int32),
                ("device_build", ctypes.c_int32), ("batch_form", ctypes.c_int32)]


TUNING_FIELDS = ("items_per_thread", "nontemporal", "blocks_per_cu", "block_tiles", "block_bpc",
                 "block_stage", "pair_exchange", "block_k", "exact_runs", "run_low_bits",
                 "device_build", "batch_form")


def get_tuning(precision: int = C128) -> dict:
    t = Tuning()
    check(lib().ket_get_tuning_for(int(precision), ctypes.byref(t)))
    return {f: getattr(t, f) for f in TUNING_FIELDS}


def set_tuning(precision: int | None = None, **changes):
    """Sets the named ket_tuning fields for one precision (C128 or C64) or,
    with precision None, for both; the other fields keep their values."""
    unknown = set(changes) - set(TUNING_FIELDS)
    if unknown:
        raise TypeError(f"set_tuning: unknown fields {sorted(unknown)}")
    for p in ((C128, C64) if precision is None else (precision,)):
        cur = get_tuning(p)
        cur.update(changes)
        t = Tuning(*(int(cur[f]) for f in TUNING_FIELDS))
        check(lib().ket_set_tuning_for(int(p), ctypes.byref(t)))


@contextlib.contextmanager
def tuning(precision: int | None = None, **changes):
    """Applies the changes for the duration of a with-block, then restores
    both precisions' tunings."""
    saved = {p: get_tuning(p) for p in (C128, C64)}
    try:
        set_tuning(precision, **changes)
        yield
    finally:
        for p, t in saved.items():
            set_tuning(p, **t)


def calibrate_copy(nbytes: int = 4 << 30, reps: int = 10, nontemporal: int = 0) -> float:
    v = ctypes.c_double(0)
    check(lib().ket_calibrate_copy(int(nbytes), int(reps), int(nontemporal), ctypes.byref(v)))
    return v.value


def pauli_masks(pauli: str, n: int):
    """flip/sign masks and the Y count ket's expval derives from a Pauli string
    (reference src/observable.cpp:34-60); rightmost character is qubit 0."""
    if len(pauli) != n:
        raise ValueError("expval: Pauli string length must equal the number of qubits")
    flip = sign = n_y = 0
    for i in range(n):
        p = pauli[n - 1 - i].upper()
        b = 1 << i
        if p == "I":
            continue
        if p == "X":
            flip |= b
        elif p == "Y":
            flip |= b
            sign |= b
            n_y += 1
        elif p == "Z":
            sign |= b
        else:
            raise ValueError(f"expval: invalid Pauli character '{pauli[n - 1 - i]}'")
    return flip, sign, n_y


class StateVector:
    """A 2^n state resident on the GPU (complex128 or complex64)."""

    def __init__(self, n_qubits: int, precision: int = C128):
        h = ctypes.c_void_p()
        check(lib().ket_sv_create(int(n_qubits), int(precision), ctypes.byref(h)))
        self._h = h
        self.n = int(n_qubits)
        self.precision = int(precision)
        self.last_stats: dict | None = None

    def close(self):
        if getattr(self, "_h", None):
            lib().ket_sv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_basis(self, index: int = 0):
        check(lib().ket_sv_set_basis(self._h, int(index)))

    def load(self, state: np.ndarray):
        a = np.ascontiguousarray(state, dtype=np.complex128)
        check(lib().ket_sv_copy_in(self._h, 0, a.size, _p(a)))

    def to_numpy(self, offset: int = 0, count: int | None = None) -> np.ndarray:
        count = (1 << self.n) - offset if count is None else count
        out = np.empty(count, dtype=np.complex128)
        check(lib().ket_sv_copy_out(self._h, offset, count, _p(out)))
        return out

    def copy_from(self, other: "StateVector"):
        check(lib().ket_sv_copy(self._h, other._h))

    def sync(self):
        check(lib().ket_sv_sync(self._h))

    def apply_gate_record(self, rec):
        r = np.ascontiguousarray(rec, dtype=GATE_DTYPE).reshape(1)
        check(lib().ket_sv_apply_gate(self._h, _p(r)))

    def apply_1q(self, target, m, ctrl_mask=0):
        mm = np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(4))
        check(lib().ket_sv_apply_1q(self._h, int(target), int(ctrl_mask), _p(mm)))

    def apply_phase(self, mask, phase):
        check(lib().ket_sv_apply_phase(self._h, int(mask), float(np.real(phase)),
                                       float(np.imag(phase))))

    def apply_x(self, target, ctrl_mask=0):
        check(lib().ket_sv_apply_x(self._h, int(target), int(ctrl_mask)))

    def apply_swap(self, a, b, ctrl_mask=0):
        check(lib().ket_sv_apply_swap(self._h, int(a), int(b), int(ctrl_mask)))

    def apply_matrix(self, qubits, mat):
        q = np.ascontiguousarray(qubits, dtype=np.int32)
        m = np.ascontiguousarray(mat, dtype=np.complex128)
        check(lib().ket_sv_apply_matrix(self._h, len(q), _p(q), _p(m)))

    def apply_records(self, rec: np.ndarray, fusion: int = FUSE_REFERENCE,
                      block_k: int = 0) -> dict:
        r = np.ascontiguousarray(rec, dtype=GATE_DTYPE)
        st = RunStats()
        check(lib().ket_sv_apply_circuit(self._h, _p(r), len(r), int(fusion), int(block_k),
                                         ctypes.byref(st)))
        self.last_stats = st.as_dict()
        return self.last_stats

    def apply_circuit(self, circuit, fusion: int = FUSE_REFERENCE, block_k: int = 0) -> dict:
        return self.apply_records(circuit.records(), fusion, block_k)

    def profile_circuit(self, circuit, fusion: int = FUSE_NONE, block_k: int = 0):
        """Runs the circuit, timing each launch; returns a list of dicts."""
        r = np.ascontiguousarray(circuit.records(), dtype=GATE_DTYPE)
        cap = 4 * len(r) + 16
        buf = (LaunchRecord * cap)()
        nl = ctypes.c_size_t(0)
        check(lib().ket_sv_profile_circuit(self._h, _p(r), len(r), int(fusion), int(block_k),
                                           ctypes.cast(buf, ctypes.c_void_p), cap,
                                           ctypes.byref(nl)))
        out = []
        for i in range(min(nl.value, cap)):
            x = buf[i]
            out.append({"kind": LAUNCH_KINDS[x.kind], "k": x.k, "n_gates": x.n_gates,
                        "n_terms": x.n_terms, "qubit_mask": int(x.qubit_mask),
                        "alg_bytes": x.alg_bytes, "flops": x.flops, "ms": x.ms,
                        "features": np.ctypeslib.as_array(x.features).copy()})
        return out

    def norm(self) -> float:
        v = ctypes.c_double(0)
        check(lib().ket_sv_norm(self._h, ctypes.byref(v)))
        return v.value

    def renormalize(self):
        check(lib().ket_sv_renormalize(self._h))

    def sample(self, u: np.ndarray) -> np.ndarray:
        uu = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(uu.size, dtype=np.uint64)
        check(lib().ket_sv_sample(self._h, _p(uu), uu.size, _p(out)))
        return out

    def prob_prefix(self, idx) -> np.ndarray:
        """ket's sequential running sum of |a|^2 through each index."""
        ii = np.ascontiguousarray(idx, dtype=np.uint64)
        out = np.zeros(ii.size, dtype=np.float64)
        check(lib().ket_sv_prob_prefix(self._h, _p(ii), ii.size, _p(out)))
        return out

    def expval(self, pauli: str) -> float:
        flip, sign, n_y = pauli_masks(pauli, self.n)
        v = ctypes.c_double(0)
        check(lib().ket_sv_expval_pauli(self._h, flip, sign, n_y, ctypes.byref(v)))
        return v.value

    def expval_sum(self, terms) -> float:
        """terms: iterable of (coeff, pauli)."""
        terms = list(terms)
        fl = np.zeros(len(terms), np.uint64)
        sg = np.zeros(len(terms), np.uint64)
        ny = np.zeros(len(terms), np.int32)
        cf = np.zeros(len(terms), np.float64)
        for i, (c, p) in enumerate(terms):
            fl[i], sg[i], ny[i] = pauli_masks(p, self.n)
            cf[i] = c
        v = ctypes.c_double(0)
        check(lib().ket_sv_expval_pauli_sum(self._h, len(terms), _p(fl), _p(sg), _p(ny),
                                            _p(cf), ctypes.byref(v)))
        return v.value

    def fidelity(self, other: "StateVector"):
        f, d = ctypes.c_double(0), ctypes.c_double(0)
        check(lib().ket_sv_fidelity(self._h, other._h, ctypes.byref(f), ctypes.byref(d)))
        return f.value, d.value

    def event_record(self, slot: int):
        check(lib().ket_sv_event_record(self._h, slot))

    def event_elapsed(self, a: int, b: int) -> float:
        ms = ctypes.c_float(0)
        check(lib().ket_sv_event_elapsed(self._h, a, b, ctypes.byref(ms)))
        return ms.value


class Batch:
    """Many circuits of uneven widths evaluated in one sweep (ket_batch_*)."""

    def __init__(self, n_qubits, precision: int = C128):
        nq = np.ascontiguousarray(n_qubits, dtype=np.int32)
        h = ctypes.c_void_p()
        check(lib().ket_batch_create(nq.size, _p(nq), int(precision), ctypes.byref(h)))
        self._h = h
        self.n_qubits = nq
        self.precision = int(precision)

    def close(self):
        if getattr(self, "_h", None):
            lib().ket_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @staticmethod
    def pack(circuits, observables):
        """Flattens circuits and diagonal observables into the arrays
        ket_batch_run takes: (gates, gate_off, zmask, coeff, term_off)."""
        recs = [c.records() for c in circuits]
        goff = np.zeros(len(recs) + 1, np.uint64)
        goff[1:] = np.cumsum([len(r) for r in recs])
        gates = np.concatenate(recs) if recs else np.zeros(0, GATE_DTYPE)
        zm, cf, toff = [], [], [0]
        for c, obs in zip(circuits, observables):
            for coeff, pauli in obs:
                flip, sign, _ = pauli_masks(pauli, c.n_qubits)
                if flip:
                    raise ValueError("batched observables must be diagonal (I/Z only)")
                zm.append(sign)
                cf.append(coeff)
            toff.append(len(zm))
        zm = np.ascontiguousarray(zm, np.uint64)
        cf = np.ascontiguousarray(cf, np.float64)
        toff = np.ascontiguousarray(toff, np.uint64)
        return gates, goff, zm, cf, toff

    def run(self, circuits, observables, block_k: int = 0, use_graph: bool = True):
        """circuits[i]: ket_amd.gates.Circuit; observables[i]: list of
        (coeff, pauli) with only I/Z characters. Returns (values, stats)."""
        return self.run_packed(self.pack(circuits, observables), block_k, use_graph)

    def run_packed(self, packed, block_k: int = 0, use_graph: bool = True):
        gates, goff, zm, cf, toff = packed
        out = np.zeros(len(goff) - 1, np.float64)
        st = BatchStats()
        check(lib().ket_batch_run(self._h, _p(gates), _p(goff), _p(zm), _p(cf), _p(toff),
                                  int(block_k), 1 if use_graph else 0, _p(out),
                                  ctypes.byref(st)))
        return out, st.as_dict()


class MLP:
    """The planner's cost model on the GPU (ket_mlp_*)."""

    def __init__(self, seed: int = 0):
        h = ctypes.c_void_p()
        check(lib().ket_mlp_create(int(seed), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().ket_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(MLP_PARAMS, np.float32)
        check(lib().ket_mlp_get_params(self._h, _p(out), MLP_PARAMS))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(lib().ket_mlp_set_params(self._h, _p(p), p.size))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        y = np.zeros(len(x), np.float32)
        check(lib().ket_mlp_forward(self._h, _p(x), len(x), _p(y)))
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        check(lib().ket_mlp_gradients(self._h, _p(x), _p(t), len(x), _p(g), _p(loss)))
        return g, float(loss[0])

    def train_step(self, x, t, lr: float) -> float:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_INPUTS)
        t = np.ascontiguousarray(t, np.float32)
        loss = np.zeros(1, np.float32)
        check(lib().ket_mlp_train_step(self._h, _p(x), _p(t), len(x), ctypes.c_float(lr),
                                       _p(loss)))
        return float(loss[0])

    def use_for_planning(self, on: bool = True):
        check(lib().ket_mlp_use_for_planning(self._h if on else None))
