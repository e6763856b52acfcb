"""Training and evaluation of the learned fusion planner's cost model.

Data: pass timings the simulator records itself (ket_sv_profile_circuit: one
HIP event pair around every launch, plus the pass's feature vector). Each
training circuit contributes two windows of consecutive gates from one
random start: 8 gates profiled as per-gate passes, as per-gate exact runs
and as forced blocks of width 3, 4 and 5, and 48 gates (the learned
planner's smallest region) profiled in the run forms it weighs, the run
packer (KET_FUSE_RUNS) and ket's Fuser as exact runs. So the model sees
every kind of pass the planner's candidate plans hold. Timing whole
28-32-qubit circuits would take hours; timing windows of them samples the
same passes.

Optionally (`bench.py --workload planner --ladder-circuits N`), QFT-like
circuits contribute too, their long window (by default the whole circuit)
timed as blocks of width 3, 4 and 5 as well as runs: blocks with dozens of
trailing phase terms and runs of hundreds of phase passes, which the random
circuits' windows do not hold.

Model: the 16-64-64-1 MLP of include/ket_amd.h, trained on the GPU with Adam
on log(milliseconds), from a seeded start or from a previous model's
parameters (fine-tuning on new timings).

Evaluation: held-out circuits of the same distribution run whole in every
form: the greedy block packer (KET_FUSE_BLOCKS, k = 5), ket's Fuser as exact
runs (KET_FUSE_REFERENCE, the fallback the learned planner is judged
against), the run packer (KET_FUSE_RUNS) and the learned planner
(KET_FUSE_LEARNED). Each form runs once cold (planning included; its
planning time is recorded) and then `reps` times with its plan cached; a
circuit's time in a form is the median of those. Wall time from HIP events.
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

# name -> (fusion mode, block_k); "learned" plans with the model
FORMS = {
    "greedy_blocks": (N.FUSE_BLOCKS, 5),
    "fuser_runs": (N.FUSE_REFERENCE, 5),
    "packed_runs": (N.FUSE_RUNS, 5),
    "learned": (N.FUSE_LEARNED, 5),
}
FALLBACK = "fuser_runs"


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


SHORT_FORMS = ((N.FUSE_NONE, 5), (N.FUSE_EXACT, 5), (N.FUSE_BLOCKS_ALL, 3),
               (N.FUSE_BLOCKS_ALL, 4), (N.FUSE_BLOCKS_ALL, 5))
LONG_FORMS = ((N.FUSE_RUNS, 5), (N.FUSE_REFERENCE, 5))
# Phase ladders (QFT-like circuits): their blocks carry dozens of trailing
# phase terms, which 8-gate windows of the random circuits never produce,
# so their long windows are timed as blocks too.
LADDER_FORMS = LONG_FORMS + ((N.FUSE_BLOCKS, 3), (N.FUSE_BLOCKS, 4), (N.FUSE_BLOCKS, 5))


def collect(circuits, states: States, short_len: int = 8, long_len: int = 48, seed: int = 0,
            log=None, long_forms=LONG_FORMS):
    """Feature rows and log(ms) of every pass of both windows of each
    circuit, in every form of SHORT_FORMS / long_forms."""
    rng = np.random.default_rng(seed)
    X, y = [], []
    t0 = time.perf_counter()
    for i, c in enumerate(circuits):
        total = c.n_unitary()
        start = int(rng.integers(0, max(1, total - long_len)))
        sv = states.get(c.n_qubits)
        for length, forms in ((short_len, SHORT_FORMS), (long_len, long_forms)):
            sub = window(c, start, length)
            for fusion, k in forms:
                for r in sv.profile_circuit(sub, fusion, k):
                    if r["ms"] > 0:
                        X.append(r["features"])
                        y.append(math.log(r["ms"]))
        if log and (i + 1) % 250 == 0:
            log(f"[costmodel] timed {i + 1}/{len(circuits)} windows, {len(y)} passes, "
                f"{time.perf_counter() - t0:.0f} s")
    return np.array(X, np.float32).reshape(-1, N.MLP_INPUTS), np.array(y, np.float32)


def train(X, y, epochs: int = 60, batch: int = 1024, lr: float = 3e-3, seed: int = 0, log=None,
          init=None):
    """Adam with a cosine learning rate; init: parameters to start from (a
    model fine-tuned on new timings) instead of a seeded initialisation."""
    mlp = N.MLP(seed)
    if init is not None:
        mlp.params = init
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


def time_forms(mlp: N.MLP, circuits, states: States, reps: int = 3, log=None) -> list:
    """Per circuit and form: the cold time (planning included), its planning
    time, the median and spread of `reps` cached repeats, launches and
    blocks; for the learned form also the model's relative error on the
    total of its plan's passes."""
    recs = []
    mlp.use_for_planning(True)
    try:
        for i, c in enumerate(circuits):
            sv = states.get(c.n_qubits)
            row = {"n": c.n_qubits, "gates": c.n_unitary()}
            for name, (fusion, k) in FORMS.items():
                cold = time_circuit(sv, c, fusion, k)
                st = dict(sv.last_stats)
                ts = [time_circuit(sv, c, fusion, k) for _ in range(reps)]
                med = float(np.median(ts))
                row[name] = {"cold_ms": cold, "plan_ms": st["plan_seconds"] * 1e3, "ms": med,
                             "spread": float((max(ts) - min(ts)) / med),
                             "launches": int(st["launches"]), "blocks": int(st["blocks"])}
            sv.set_basis(0)
            prof = sv.profile_circuit(c, N.FUSE_LEARNED, 5)
            feats = np.array([r["features"] for r in prof], np.float32)
            meas = np.array([r["ms"] for r in prof])
            pred = np.exp(mlp.forward(feats))
            row["learned"]["model_rel_error"] = float(abs(pred.sum() - meas.sum()) / meas.sum())
            recs.append(row)
            if log and (i + 1) % 10 == 0:
                tot = {f: sum(r[f]["ms"] for r in recs) for f in FORMS}
                log(f"[costmodel] timed {i + 1}/{len(circuits)} circuits: " +
                    ", ".join(f"{f} {v:.0f} ms" for f, v in tot.items()))
    finally:
        mlp.use_for_planning(False)
    return recs


def summarize(recs: list) -> dict:
    """Totals per form, the learned planner against the best non-learned
    form (the one with the lowest total), the fallback and the greedy
    packer, and the timing noise (90th percentile of the repeats' relative
    spread)."""
    fixed = [f for f in FORMS if f != "learned"]
    tot = {f: float(sum(r[f]["ms"] for r in recs)) for f in FORMS}
    cold = {f: float(sum(r[f]["cold_ms"] for r in recs)) for f in FORMS}
    plan = {f: float(np.mean([r[f]["plan_ms"] for r in recs])) for f in FORMS}
    best = min(fixed, key=tot.get)
    per = np.array([min(r[f]["ms"] for f in fixed) / r["learned"]["ms"] for r in recs])
    spreads = [r[f]["spread"] for r in recs for f in FORMS]
    return {
        "circuits": len(recs), "total_ms": tot, "cold_total_ms": cold, "plan_ms_mean": plan,
        "best_fixed_form": best,
        "speedup_vs_best_fixed": tot[best] / tot["learned"],
        "speedup_vs_fallback": tot[FALLBACK] / tot["learned"],
        "speedup_vs_greedy_blocks": tot["greedy_blocks"] / tot["learned"],
        "cold_speedup_vs_fallback": cold[FALLBACK] / cold["learned"],
        "per_circuit_vs_fastest_fixed_median": float(np.median(per)),
        "learned_fastest_fraction": float(np.mean(per >= 1.0)),
        "noise": float(np.percentile(spreads, 90)),
        "learned_launches_mean": float(np.mean([r["learned"]["launches"] for r in recs])),
        "learned_blocks_mean": float(np.mean([r["learned"]["blocks"] for r in recs])),
        "model_total_time_rel_error_median":
            float(np.median([r["learned"]["model_rel_error"] for r in recs])),
    }


def time_cases(mlp: N.MLP, cases, states: States, reps: int = 3, log=None) -> dict:
    """Named circuits in every form (median of `reps` cached runs each) and
    the form the learned planner's plan takes (its launch and block counts)."""
    out = {}
    mlp.use_for_planning(True)
    try:
        for name, c in cases:
            sv = states.get(c.n_qubits)
            row = {}
            for form, (fusion, k) in FORMS.items():
                time_circuit(sv, c, fusion, k)  # plan and warm
                row[form] = {"ms": float(np.median([time_circuit(sv, c, fusion, k)
                                                    for _ in range(reps)])),
                             "launches": sv.last_stats["launches"],
                             "blocks": sv.last_stats["blocks"]}
            out[name] = row
            if log:
                log(f"[costmodel] {name}: " + ", ".join(
                    f"{f} {v['ms']:.1f} ms ({v['blocks']} blocks)" for f, v in row.items()))
    finally:
        mlp.use_for_planning(False)
    return out


def evaluate(mlp: N.MLP, circuits, states: States, reps: int = 3, log=None) -> dict:
    return summarize(time_forms(mlp, circuits, states, reps, log))


def ladder_circuits(count: int, seed: int, lo: int = 28, hi: int = 32):
    """QFT-like circuits: a random U layer, then the QFT's H + controlled-phase
    ladders and closing SWAPs, on 28-32 qubits."""
    rng = np.random.default_rng(seed)
    return [W.qft_u(int(rng.integers(lo, hi + 1)), int(rng.integers(1 << 30)))
            for _ in range(count)]


def config5_circuits(count: int, seed: int, lo: int = 28, hi: int = 32, depth: int = 40):
    rng = np.random.default_rng(seed)
    return [W.random_mixed(int(rng.integers(lo, hi + 1)), depth, int(rng.integers(1 << 30)))
            for _ in range(count)]
