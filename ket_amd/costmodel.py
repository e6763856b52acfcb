"""Training and evaluation of the learned fusion planner's cost model.

Data: pass timings the simulator records itself (ket_sv_profile_circuit: one
HIP event pair around every launch, plus the pass's feature vector). Each
training circuit contributes a window of consecutive gates, profiled as
per-gate passes and as forced blocks of width 3, 4 and 5, so the model sees
both options of every decision the planner makes. Timing whole 28-32-qubit
circuits would take hours; timing windows of them samples the same passes.

Model: the 16-64-64-1 MLP of include/ket_amd.h, trained on the GPU with Adam
on log(milliseconds).

Evaluation: held-out circuits of the same distribution run whole with the
greedy packer (KET_FUSE_BLOCKS, k = 5) and with the learned planner
(KET_FUSE_LEARNED); wall time from HIP events, plus the model's error on the
held-out circuits' passes.
"""
from __future__ import annotations

import math
import os
import time

import numpy as np

from . import native as N
from . import workloads as W
from .gates import Circuit

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "cost_model.npz")


def window(c: Circuit, start: int, length: int) -> Circuit:
    w = Circuit(c.n_qubits)
    w.ops = [op for op in c.ops if op[0] not in ("measure", "barrier")][start:start + length]
    return w


class States:
    """One resident state per width, reused for every timing."""

    def __init__(self, precision: int):
        self.precision = precision
        self.sv = {}

    def get(self, n: int) -> N.StateVector:
        if n not in self.sv:
            self.sv[n] = N.StateVector(n, self.precision)
        return self.sv[n]

    def close(self):
        for s in self.sv.values():
            s.close()
        self.sv.clear()


def collect(circuits, states: States, window_len: int = 12, seed: int = 0, log=None):
    rng = np.random.default_rng(seed)
    X, y = [], []
    t0 = time.perf_counter()
    for i, c in enumerate(circuits):
        total = c.n_unitary()
        start = int(rng.integers(0, max(1, total - window_len)))
        sub = window(c, start, window_len)
        sv = states.get(c.n_qubits)
        for fusion, k in ((N.FUSE_NONE, 5), (N.FUSE_BLOCKS_ALL, 3), (N.FUSE_BLOCKS_ALL, 4),
                          (N.FUSE_BLOCKS_ALL, 5)):
            for r in sv.profile_circuit(sub, fusion, k):
                if r["ms"] > 0:
                    X.append(r["features"])
                    y.append(math.log(r["ms"]))
        if log and (i + 1) % 250 == 0:
            log(f"[costmodel] timed {i + 1}/{len(circuits)} windows, {len(y)} passes, "
                f"{time.perf_counter() - t0:.0f} s")
    return np.array(X, np.float32), np.array(y, np.float32)


def train(X, y, epochs: int = 60, batch: int = 1024, lr: float = 3e-3, seed: int = 0, log=None):
    mlp = N.MLP(seed)
    rng = np.random.default_rng(seed)
    n = len(X)
    hist = []
    for ep in range(epochs):
        perm = rng.permutation(n)
        losses = []
        step_lr = lr * (0.5 * (1 + math.cos(math.pi * ep / epochs)))
        for b in range(0, n, batch):
            idx = perm[b:b + batch]
            losses.append(mlp.train_step(X[idx], y[idx], step_lr))
        hist.append(float(np.mean(losses)))
        if log and (ep % 10 == 0 or ep == epochs - 1):
            log(f"[costmodel] epoch {ep}: loss {hist[-1]:.4f}")
    return mlp, hist


def save(mlp: N.MLP, path: str = MODEL_PATH, **meta):
    np.savez(path, params=mlp.params, **{k: np.asarray(v) for k, v in meta.items()})


def load(path: str = MODEL_PATH) -> N.MLP | None:
    if not os.path.exists(path):
        return None
    data = np.load(path, allow_pickle=False)
    mlp = N.MLP(0)
    mlp.params = data["params"]
    return mlp


def time_circuit(sv: N.StateVector, c: Circuit, fusion: int, k: int) -> float:
    rec = c.records()
    sv.set_basis(0)
    sv.sync()
    sv.event_record(0)
    sv.apply_records(rec, fusion, k)
    sv.event_record(1)
    sv.sync()
    return sv.event_elapsed(0, 1)


def evaluate(mlp: N.MLP, circuits, states: States, log=None, ref_subset: int = 0) -> dict:
    """Learned planner vs the greedy packer (k = 5, bytes rule) on whole
    circuits, the model's relative error on the learned plans' passes, and,
    on the first `ref_subset` circuits, learned vs ket's own greedy Fuser
    (KET_FUSE_REFERENCE, the reference's fusion rule), run as the reference
    runs it (one pass per fused group) and as exact runs on LDS tiles."""
    tg, tl, errs, tr, trr, tlr = [], [], [], [], [], []
    for i, c in enumerate(circuits):
        sv = states.get(c.n_qubits)
        mlp.use_for_planning(False)
        time_circuit(sv, c, N.FUSE_BLOCKS, 5)  # warm
        tg.append(time_circuit(sv, c, N.FUSE_BLOCKS, 5))
        if i < ref_subset:
            with N.tuning(sv.precision, exact_runs=0):
                time_circuit(sv, c, N.FUSE_REFERENCE, 5)  # warm
                tr.append(time_circuit(sv, c, N.FUSE_REFERENCE, 5))
            with N.tuning(sv.precision, exact_runs=1):
                time_circuit(sv, c, N.FUSE_REFERENCE, 5)
                trr.append(time_circuit(sv, c, N.FUSE_REFERENCE, 5))
        mlp.use_for_planning(True)
        tl.append(time_circuit(sv, c, N.FUSE_LEARNED, 5))
        if i < ref_subset:
            tlr.append(tl[-1])
        sv.set_basis(0)
        recs = sv.profile_circuit(c, N.FUSE_LEARNED, 5)
        feats = np.array([r["features"] for r in recs], np.float32)
        pred = np.exp(mlp.forward(feats))
        meas = np.array([r["ms"] for r in recs])
        errs.append(float(abs(pred.sum() - meas.sum()) / meas.sum()))
        if log and (i + 1) % 25 == 0:
            log(f"[costmodel] evaluated {i + 1}/{len(circuits)}: greedy {sum(tg):.0f} ms, "
                f"learned {sum(tl):.0f} ms")
    mlp.use_for_planning(False)
    tg, tl = np.array(tg), np.array(tl)
    out = {"circuits": len(circuits), "greedy_ms": float(tg.sum()), "learned_ms": float(tl.sum()),
           "speedup": float(tg.sum() / tl.sum()),
           "per_circuit_speedup_median": float(np.median(tg / tl)),
           "learned_faster_fraction": float(np.mean(tl < tg)),
           "model_total_time_rel_error_median": float(np.median(errs))}
    if tr:
        out.update({"reference_fuser_circuits": len(tr), "reference_fuser_ms": float(sum(tr)),
                    "reference_fuser_runs_ms": float(sum(trr)),
                    "learned_ms_same_circuits": float(sum(tlr)),
                    "speedup_vs_reference_fuser": float(sum(tr) / sum(tlr)),
                    "speedup_vs_reference_fuser_runs": float(sum(trr) / sum(tlr))})
    return out


def config5_circuits(count: int, seed: int, lo: int = 28, hi: int = 32, depth: int = 40):
    rng = np.random.default_rng(seed)
    return [W.random_mixed(int(rng.integers(lo, hi + 1)), depth, int(rng.integers(1 << 30)))
            for _ in range(count)]
