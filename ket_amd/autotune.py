"""Autotuner for the launch shape of the per-gate passes.

Times a fixed probe circuit (dense, controlled, diagonal and permuting gates
on low, middle and high qubits, so every access pattern of the plain form is
represented) on a resident state for every point of the search space the C
ABI exposes (ket_tuning: items per thread, nontemporal access, persistent
grid), with per-launch HIP events, and keeps the fastest. Results are stored
per device and precision in a JSON table that `apply_table` loads.

    python -m ket_amd.autotune --qubits 30 --out table.json
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys

import numpy as np

from . import native as N
from .gates import Circuit

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")

SPACE = {"items_per_thread": [1, 2, 4], "nontemporal": [0, 2, 3], "blocks_per_cu": [0, 32],
         "pair_exchange": [0, 4, 16]}


def probe_circuit(n: int) -> Circuit:
    c = Circuit(n)
    lo, mid, hi = 0, n // 2, n - 1
    for q in (lo, 2, mid, hi):
        c.u(q, 0.3, 0.2, 0.1)          # dense 2x2
    c.h(1)
    c.cry(hi, mid, 0.4)                 # controlled dense
    c.cry(lo, mid, 0.4)
    for a, b in ((mid, hi), (lo, hi), (5, mid)):
        c.cp(a, b, 0.7)                 # two-bit diagonal
    c.t(mid)                            # one-bit diagonal
    c.x(lo)                             # permutes
    c.cx(mid, hi)
    c.swap(3, hi)
    c.ccx(lo, mid, hi - 1)
    return c


BLOCK_SPACE = {"block_stage": [0, 3, 64], "block_bpc": [8, 16, 32]}


def block_probe(n: int, k: int) -> Circuit:
    """Fused blocks on low, middle and high qubits, each with trailing
    controlled phases to qubits outside it."""
    c = Circuit(n)
    for lo in (0, n // 2 - k // 2, n - k):
        qs = list(range(lo, lo + k))
        for q in qs:
            c.u(q, 0.3, 0.2, 0.1)
        for a, b in zip(qs, qs[1:]):
            c.cx(a, b)
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
                 "per_block_GBps": {f"q{(r['qubit_mask'] & -r['qubit_mask']).bit_length() - 1}":
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


def apply_table(path: str = TABLE, precision: int = N.C128) -> dict | None:
    """Set the tuning recorded for this precision, if the table has one."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    entry = table.get("c128" if precision == N.C128 else "c64")
    if entry:
        b = entry["best"]
        N.set_tuning(precision, **{k: b[k] for k in SPACE if k in b})
        return b
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=None)
    ap.add_argument("--blocks-only", action="store_true")
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
    out = json.dumps(table, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(out)
    else:
        print(out)


if __name__ == "__main__":
    main()
