"""Benchmark of the dense state-vector hot path on one MI355X.

Workload (default): the 30-qubit QFT preceded by one random U layer
(510 gates; 16 GiB complex128 state resident in HBM). One step = reset to
|0...0> and evolve the whole circuit with the plain per-gate complex128
kernels. `value` is gates per second of that form.

Besides the timed steps the run reports:
  * the same circuit with fused dense blocks (complex128, k = 4) and in
    complex64 with fused blocks (k = 4, 5), with 1-F and max |amplitude error|
    of the complex64 state against the plain complex128 state, both resident;
  * `roofline`: the dominant kernel of the plain form, its algorithmic bytes
    per launch over its average launch time from HIP events on the state's
    stream, against the 8 TB/s HBM peak;
  * `cpu_baseline`: ket's own simulator (the reference library built from its
    sources, oracle/_ref) timed on this host on a bounded sample.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--workload qft30]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ket_amd import native as N  # noqa: E402
from ket_amd import workloads as W  # noqa: E402

METRIC = ("gates/s & HBM GB/s at 30-33q; fused FP32 vs plain FP64 GPU at "
          "1-F<=1e-6")
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    if name == "qft30":
        return W.qft_u(30, seed=0), "qft30_u: 30-qubit QFT after one U layer (510 gates)"
    if name == "brick33":
        return (W.brickwork_mixed(33, 24, seed=0),
                "brick33: 33-qubit random brickwork, depth 24, {U,CX,CZ,CCX}")
    if name.startswith("qft"):
        n = int(name[3:])
        return W.qft_u(n, seed=0), f"qft{n}_u: {n}-qubit QFT after one U layer"
    raise SystemExit(f"unknown workload {name}")


def time_steps(sv, circ, fusion, k, steps, warmup):
    rec = circ.records()
    for _ in range(warmup):
        sv.set_basis(0)
        sv.apply_records(rec, fusion, k)
    sv.sync()
    sv.event_record(0)
    t0 = time.perf_counter()
    for _ in range(steps):
        sv.set_basis(0)
        sv.apply_records(rec, fusion, k)
    sv.event_record(1)
    sv.sync()
    wall = time.perf_counter() - t0
    ev = sv.event_elapsed(0, 1) / 1e3
    return wall, ev, dict(sv.last_stats)


def dominant_kernel(records):
    by = {}
    for r in records:
        d = by.setdefault(r["kind"], {"ms": 0.0, "n": 0, "bytes": 0.0, "flops": 0.0})
        d["ms"] += r["ms"]
        d["n"] += 1
        d["bytes"] += r["alg_bytes"]
        d["flops"] += r["flops"]
    kind = max(by, key=lambda k: by[k]["ms"])
    d = by[kind]
    avg_ms = d["ms"] / d["n"]
    avg_bytes = d["bytes"] / d["n"]
    return kind, d, avg_ms, avg_bytes, by


def cpu_baseline(circ, threads, budget_s=20.0):
    """ket's own run() (reference library, oracle/_ref) on prefixes of the
    workload at its real size: rate = extra gates / extra time between a short
    and a longer prefix, so the 2^n allocation is not charged per gate."""
    from oracle import oracle as O
    if not O.ref_available():
        return None
    short, long_ = 4, 24
    t_s = min(O.ref_time(circ.prefix(short), threads, 1)["seconds"])
    t_l = min(O.ref_time(circ.prefix(long_), threads, 1)["seconds"])
    rate = (long_ - short) / max(t_l - t_s, 1e-9)
    return {"value": rate, "unit": "gates/s", "cores": threads, "kind": "reference",
            "sample": (f"ket::run (reference built from source, set_num_threads({threads})) "
                       f"on the first {short} and first {long_} gates of the same "
                       f"{circ.n_qubits}-qubit circuit; rate = {long_ - short} gates / "
                       f"({t_l:.2f}s - {t_s:.2f}s)"),
            "seconds": [t_s, t_l]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="qft30")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fused", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this backend has no multi-GPU path "
              "(the sharding design is described in DESIGN.md); use --gpus 1.",
              file=sys.stderr)
        return 2

    circ, desc = workload(args.workload)
    n = circ.n_qubits
    gates = circ.n_unitary()
    info = N.device_info()
    log(f"[bench] device {info['name']} CUs={info['compute_units']} "
        f"HBM={info['hbm_bytes'] / 2**30:.0f} GiB; workload {desc}")

    out = {"metric": METRIC}
    plain = N.StateVector(n, N.C128)
    log(f"[bench] plain complex128: warmup {args.warmup}, steps {args.steps}")
    wall, ev_s, stats = time_steps(plain, circ, N.FUSE_NONE, 5, args.steps, args.warmup)
    per_step = wall / args.steps
    value = gates / per_step
    gbps = stats["alg_bytes"] / per_step / 1e9
    log(f"[bench] plain: {per_step * 1e3:.1f} ms/step, {value:.1f} gates/s, "
        f"{gbps:.0f} GB/s algorithmic (events {ev_s / args.steps * 1e3:.1f} ms/step)")

    # Roofline of the dominant kernel, from per-launch HIP events.
    plain.set_basis(0)
    recs = plain.profile_circuit(circ, N.FUSE_NONE)
    kind, d, avg_ms, avg_bytes, by = dominant_kernel(recs)
    achieved = avg_bytes / (avg_ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                "kernel": kind, "launches": d["n"], "avg_launch_ms": avg_ms,
                "alg_bytes_per_launch": avg_bytes,
                "by_kind": {k: {"launches": v["n"], "ms": v["ms"],
                                "GBps": v["bytes"] / (v["ms"] / 1e3) / 1e9 if v["ms"] else None}
                            for k, v in by.items()}}
    log(f"[bench] dominant kernel {kind}: {d['n']} launches, avg {avg_ms:.3f} ms, "
        f"{achieved:.0f} GB/s ({achieved / HBM_PEAK_GBPS:.1%} of 8 TB/s)")

    fused = {}
    if not args.no_fused:
        plain.set_basis(0)
        plain.apply_circuit(circ, N.FUSE_NONE)
        plain.sync()
        for label, prec, k in [("c128_k4", N.C128, 4), ("c64_k4", N.C64, 4), ("c64_k5", N.C64, 5)]:
            sv = N.StateVector(n, prec)
            w, e, st = time_steps(sv, circ, N.FUSE_BLOCKS, k, args.steps, args.warmup)
            f, dmax = sv.fidelity(plain)
            ps = w / args.steps
            fused[label] = {"ms_per_step": ps * 1e3, "gates_per_s": gates / ps,
                            "alg_GBps": st["alg_bytes"] / ps / 1e9, "launches": st["launches"],
                            "blocks": st["blocks"], "one_minus_F": 1.0 - f,
                            "max_abs_err": dmax, "speedup_vs_plain": per_step / ps}
            log(f"[bench] fused {label}: {ps * 1e3:.1f} ms/step, {gates / ps:.1f} gates/s, "
                f"{st['launches']} launches, 1-F={1 - f:.2e}, max|d|={dmax:.2e}, "
                f"x{per_step / ps:.2f} vs plain")
            sv.close()
    plain.close()

    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline with {args.cpu_threads} threads")
        try:
            cpu = cpu_baseline(circ, args.cpu_threads)
        except (subprocess.CalledProcessError, OSError) as e:
            log(f"[bench] cpu baseline failed: {e}")
        if cpu:
            log(f"[bench] cpu: {cpu['value']:.3f} gates/s ({cpu['sample']})")

    out.update({
        "value": value, "unit": "gates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": per_step * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c128", "data": "synthetic",
        "config": {"workload": desc, "n_qubits": n, "gates": gates, "state_bytes": 16 << n,
                   "form": "plain per-gate complex128 kernels", "parallelism": "none (1 GPU)"},
        "hbm_GBps_algorithmic": gbps,
        "roofline": roofline,
        "fused": fused,
        "cpu_baseline": cpu,
        "device": info["name"],
    })
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
