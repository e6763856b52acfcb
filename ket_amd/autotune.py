"""Autotuner for launch shapes and fusion width.

Three sweeps, each timed on a resident state with HIP events (per launch) or
wall clocks (whole runs), keeping the fastest point:

* per-gate passes: a probe circuit with dense, controlled, diagonal and
  permuting gates on targets 0, 1, 2, 3, 5, 9, 17 and n - 1 (every access
  pattern of the plain form) over items per thread, nontemporal mode,
  persistent grid and the lane-pair exchange threshold;
* fused blocks: blocks on low, middle and high qubits with trailing phase
  terms over LDS staging and workgroups per CU;
* fusion width: whole runs of a brickwork and a QFT circuit and a batched
  sweep of hardware-efficient-ansatz instances of uneven widths, for block
  widths 3, 4 and 5; the score is the geometric mean over the three of the
  time relative to the best width for that workload;
* exact runs: whole KET_FUSE_EXACT runs of a brickwork, a QFT and a random
  13-gate-set circuit for tile low-bit counts 3..6, scored the same way.

Results are stored per precision in a JSON table that `apply_table` loads.

    python -m ket_amd.autotune --qubits 30 --out table.json
"""
from __future__ import annotations

import argparse
import itertools
import json
import math
import os
import sys
import time

import numpy as np

from . import native as N
from . import workloads as W
from .gates import Circuit

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")

SPACE = {"items_per_thread": [1, 2, 4], "nontemporal": [0, 2, 3], "blocks_per_cu": [0, 32],
         "pair_exchange": [0, 4, 16]}


def probe_circuit(n: int) -> Circuit:
    """One pass per gate: U, CX with the control on qubit 0 (1 for target 0)
    and on the next qubit up, CP to a high qubit, T, SWAP with a partner 7
    above, on each target; then a CCX and a controlled RY across the state."""
    if n < 20:
        raise ValueError("the probe needs at least 20 qubits")
    c = Circuit(n)
    for t in (0, 1, 2, 3, 5, 9, 17, n - 1):
        c.u(t, 0.3, 0.2, 0.1)
        c.cx(0 if t else 1, t)
        c.cx((t + 1) % n, t)
        c.cp(t, n - 2 if t != n - 2 else 0, 0.3)
        c.t(t)
        c.swap(t, (t + 7) % n)
    c.ccx(0, n // 2, n - 1)
    c.cry(n - 1, n // 2, 0.4)
    return c


BLOCK_SPACE = {"block_stage": [0, 3, 64], "block_bpc": [8, 16, 32, 64]}


def block_probe(n: int, k: int) -> Circuit:
    """Fused blocks on low, middle and high qubits, each once without and
    once with trailing controlled phases to qubits outside it (the launch
    rule treats the two differently)."""
    c = Circuit(n)
    for lo in (0, n // 2 - k // 2, n - k):
        qs = list(range(lo, lo + k))
        for phases in (False, True):
            for q in qs:
                c.u(q, 0.3, 0.2, 0.1)
            for a, b in zip(qs, qs[1:]):
                c.cx(a, b)
            if phases:
                for q in qs:
                    c.cp(q, (q + k + 3) % n if (q + k + 3) % n not in qs else (q + 2 * k) % n, 0.4)
            c.barrier()
    return c


def tune_blocks(n: int = 30, precision: int = N.C64, reps: int = 3, log=print) -> dict:
    k = 5 if precision == N.C64 else 4
    circ = block_probe(n, k)
    results = []
    with N.StateVector(n, precision) as sv:
        for point in itertools.product(*BLOCK_SPACE.values()):
            cfg = dict(zip(BLOCK_SPACE, point))
            N.set_tuning(precision, **cfg)
            sv.profile_circuit(circ, N.FUSE_BLOCKS_ALL, k)
            ms = []
            for _ in range(reps):
                recs = [r for r in sv.profile_circuit(circ, N.FUSE_BLOCKS_ALL, k)
                        if r["kind"] == "block"]
                ms.append(sum(r["ms"] for r in recs))
            bytes_ = sum(r["alg_bytes"] for r in recs)
            m = {**cfg, "ms": float(np.median(ms)),
                 "GBps": bytes_ / (float(np.median(ms)) / 1e3) / 1e9,
                 "per_block_GBps": {f"q{(r['qubit_mask'] & -r['qubit_mask']).bit_length() - 1}"
                                    + ("+phases" if r["n_terms"] else ""):
                                    round(r["alg_bytes"] / (r["ms"] / 1e3) / 1e9)
                                    for r in recs}}
            results.append(m)
            log(f"[tune] blocks k={k} {cfg}: {m['ms']:.2f} ms, {m['GBps']:.0f} GB/s "
                f"by lowest qubit {m['per_block_GBps']}")
    best = min(results, key=lambda r: r["ms"])
    N.set_tuning(precision, **{kk: best[kk] for kk in BLOCK_SPACE})
    return {"k": k, "best": {kk: best[kk] for kk in BLOCK_SPACE}, "best_GBps": best["GBps"],
            "all": results}


def measure(sv, circ, reps: int) -> dict:
    sv.profile_circuit(circ, N.FUSE_NONE)  # warm
    tot_ms, tot_bytes = 0.0, 0.0
    by = {}
    for _ in range(reps):
        for r in sv.profile_circuit(circ, N.FUSE_NONE):
            tot_ms += r["ms"]
            tot_bytes += r["alg_bytes"]
            d = by.setdefault(r["kind"], [0.0, 0.0])
            d[0] += r["ms"]
            d[1] += r["alg_bytes"]
    return {"ms": tot_ms / reps, "GBps": tot_bytes / (tot_ms / 1e3) / 1e9,
            "by_kind_GBps": {k: v[1] / (v[0] / 1e3) / 1e9 for k, v in by.items()}}


def tune(n: int = 30, precision: int = N.C128, reps: int = 3, log=print) -> dict:
    circ = probe_circuit(n)
    results = []
    with N.StateVector(n, precision) as sv:
        for point in itertools.product(*SPACE.values()):
            cfg = dict(zip(SPACE, point))
            N.set_tuning(precision, **cfg)
            m = measure(sv, circ, reps)
            m.update(cfg)
            results.append(m)
            log(f"[tune] {cfg}: {m['ms']:.2f} ms, {m['GBps']:.0f} GB/s "
                + " ".join(f"{k}={v:.0f}" for k, v in m["by_kind_GBps"].items()))
    best = min(results, key=lambda r: r["ms"])
    N.set_tuning(precision, **{k: best[k] for k in SPACE})
    return {"device": N.device_info()["name"], "precision": precision, "qubits": n,
            "best": {k: best[k] for k in SPACE}, "best_GBps": best["GBps"], "all": results}


WIDTHS = [3, 4, 5]


def hea_instances(count: int, lo: int = 10, hi: int = 22, seed: int = 0):
    rng = np.random.default_rng(seed)
    ns = [int(x) for x in rng.integers(lo, hi + 1, count)]
    circs = [W.hea(n, 8, seed=i) for i, n in enumerate(ns)]
    obs = [[(1.0, p) for p in W.zz_chain(n)] for n in ns]
    return ns, circs, obs


def _median_wall(fn, reps: int) -> float:
    fn()  # warm (first plan, graph capture)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def tune_width(n: int = 30, precision: int = N.C128, reps: int = 3, instances: int = 512,
               log=print) -> dict:
    """Whole-run wall time (host planning included) of FUSE_BLOCKS at each
    width on a brickwork circuit, a QFT and a batched sweep."""
    singles = {"brick": W.brickwork(n, 6, seed=3, measure=False), "qft": W.qft_u(n, seed=3)}
    ns, bc, bo = hea_instances(instances)
    packed = N.Batch.pack(bc, bo)
    rows = []
    with N.StateVector(n, precision) as sv, N.Batch(ns, precision) as b:
        for k in WIDTHS:
            row = {"block_k": k}
            for name, c in singles.items():
                rec = c.records()

                def run(rec=rec):
                    sv.set_basis(0)
                    sv.apply_records(rec, N.FUSE_BLOCKS, k)
                    sv.sync()
                row[name + "_ms"] = _median_wall(run, reps) * 1e3
            row["hea_ms"] = _median_wall(lambda: b.run_packed(packed, block_k=k), reps) * 1e3
            rows.append(row)
            log(f"[tune] width k={k}: " + " ".join(f"{kk}={v:.1f}" for kk, v in row.items()
                                                     if kk != "block_k"))
    keys = [kk for kk in rows[0] if kk.endswith("_ms")]
    best_of = {kk: min(r[kk] for r in rows) for kk in keys}
    for r in rows:
        r["score"] = math.exp(sum(math.log(r[kk] / best_of[kk]) for kk in keys) / len(keys))
    best = min(rows, key=lambda r: r["score"])
    N.set_tuning(precision, block_k=best["block_k"])
    return {"qubits": n, "instances": instances, "best": {"block_k": best["block_k"]},
            "all": rows}


RUN_LOWS = [3, 4, 5, 6]


def tune_runs(n: int = 30, precision: int = N.C128, reps: int = 3, log=print) -> dict:
    """Whole-run wall time of KET_FUSE_EXACT (consecutive per-gate passes on
    LDS tiles) for each number of low bits the tiles keep."""
    circs = {"brick": W.brickwork(n, 6, seed=3, measure=False), "qft": W.qft_u(n, seed=3),
             "mixed": W.random_mixed(n, 12, seed=3)}
    rows = []
    with N.StateVector(n, precision) as sv:
        for low in RUN_LOWS:
            N.set_tuning(precision, run_low_bits=low)
            row = {"run_low_bits": low}
            for name, c in circs.items():
                rec = c.records()

                def run(rec=rec):
                    sv.set_basis(0)
                    sv.apply_records(rec, N.FUSE_EXACT, 0)
                    sv.sync()
                row[name + "_ms"] = _median_wall(run, reps) * 1e3
            rows.append(row)
            log(f"[tune] runs low={low}: " + " ".join(f"{kk}={v:.1f}" for kk, v in row.items()
                                                       if kk != "run_low_bits"))
    keys = [kk for kk in rows[0] if kk.endswith("_ms")]
    best_of = {kk: min(r[kk] for r in rows) for kk in keys}
    for r in rows:
        r["score"] = math.exp(sum(math.log(r[kk] / best_of[kk]) for kk in keys) / len(keys))
    best = min(rows, key=lambda r: r["score"])
    N.set_tuning(precision, run_low_bits=best["run_low_bits"])
    return {"qubits": n, "best": {"run_low_bits": best["run_low_bits"]}, "all": rows}


def apply_table(path: str = TABLE, precision: int = N.C128) -> dict | None:
    """Set the tuning recorded for this precision (gate passes, fused blocks
    and fusion width), if the table has it; returns the fields set."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    key = "c128" if precision == N.C128 else "c64"
    applied = {}
    for entry in (table.get(key), table.get(key + "_blocks"), table.get(key + "_width"),
                  table.get(key + "_runs")):
        if entry:
            applied.update({k: v for k, v in entry["best"].items() if k in N.TUNING_FIELDS})
    if not applied:
        return None
    N.set_tuning(precision, **applied)
    return applied


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=None)
    ap.add_argument("--blocks-only", action="store_true")
    ap.add_argument("--instances", type=int, default=512, help="width sweep: batch size")
    args = ap.parse_args()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    for nt in (0, 1):
        log(f"[tune] copy kernel nt={nt}: {N.calibrate_copy(4 << 30, 10, nt):.0f} GB/s")
    table = {}
    for prec, key in ((N.C128, "c128"), (N.C64, "c64")):
        if not args.blocks_only:
            table[key] = tune(args.qubits, prec, args.reps, log)
            log(f"[tune] {key} best {table[key]['best']} {table[key]['best_GBps']:.0f} GB/s")
        table[key + "_blocks"] = tune_blocks(args.qubits, prec, args.reps, log)
        log(f"[tune] {key} blocks best {table[key + '_blocks']['best']}")
        table[key + "_width"] = tune_width(args.qubits, prec, args.reps, args.instances, log)
        log(f"[tune] {key} width best {table[key + '_width']['best']}")
        table[key + "_runs"] = tune_runs(args.qubits, prec, args.reps, log)
        log(f"[tune] {key} runs best {table[key + '_runs']['best']}")
    out = json.dumps(table, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(out)
    else:
        print(out)


if __name__ == "__main__":
    main()
