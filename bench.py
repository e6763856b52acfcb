"""Benchmark of the dense state-vector hot path on one MI355X.

Workload (default): the 30-qubit QFT preceded by one random U layer
(510 gates; 16 GiB complex128 state resident in HBM). One step = reset to
|0...0> and evolve the whole circuit with the plain per-gate complex128
kernels. `value` is gates per second of that form.

Besides the timed steps the run reports:
  * the same circuit as exact runs (consecutive per-gate passes applied on
    LDS tiles, equal to the plain form bit for bit) and with ket's Fuser
    rule plus exact runs (the drop-in default, equal to ket::run);
  * the same circuit with fused dense blocks (complex128, k = 4) and in
    complex64 with fused blocks (k = 4, 5), with 1-F and max |amplitude error|
    of the complex64 state against the plain complex128 state, both resident;
  * `roofline`: the dominant kernel of the plain form, its algorithmic bytes
    per launch over its average launch time from HIP events on the state's
    stream, against the 8 TB/s HBM peak;
  * `brick33`: the 33-qubit random brickwork (depth 24, {U, CX, CZ, CCX},
    1242 gates) at its real size: plain complex128 (128 GiB) for one step,
    fused complex64 k = 5 blocks (64 GiB, resident beside it) for
    min(steps, 5) steps, 1-F and max |amplitude error| between them, and
    exact runs in complex128 for one step (--no-brick33 skips it);
  * `cpu_baseline`: ket's own simulator (the reference library built from its
    sources, oracle/_ref) timed on this host on a bounded sample of the
    workload's gates at its real size, weighted by its gate mix, on every CPU
    this process may use; `host` names the CPU model and counts.

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


def hea_batch(count=4096, lo=10, hi=22, seed=0, angle_seed=0):
    """Config 4: `count` hardware-efficient-ansatz instances, widths uniform
    in [lo, hi], 8 layers each; observable sum_i Z_i Z_{i+1} (open chain)."""
    rng = np.random.default_rng(seed)
    ns = [int(x) for x in rng.integers(lo, hi + 1, count)]
    circs = [W.hea(n, 8, seed=angle_seed * 1_000_003 + i) for i, n in enumerate(ns)]
    obs = [[(1.0, p) for p in W.zz_chain(n)] for n in ns]
    return ns, circs, obs


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


def traffic_from_profile(kind):
    """HBM bytes per launch of the dominant kernel as measured by the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs committed under profiles/ (scaled
    by the same counters on the copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "qft30_plain_c128_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"traffic_bytes_per_launch {kind} "):
            return float(line.split()[2])
    return None


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


def host_cpus():
    """This host's CPU model, the logical CPUs it reports, and how many of them
    this process may use (its affinity mask, capped by a cgroup CPU quota)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    allowed = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()
            if q != "max":
                quota = max(1, int(-(-int(q) // int(period))))
    except (OSError, ValueError):
        pass
    usable = min(allowed, quota) if quota else allowed
    return {"model": model, "reported": os.cpu_count(), "affinity": allowed,
            "cgroup_quota": quota, "usable": usable}


def cpu_baseline_mix(circ, threads, per_kind=6, reps=3, seed=0):
    """ket's own simulator (reference library, oracle/_ref) at the workload's
    real size, weighted by the workload's gate mix: per_kind gates of each
    kind drawn from the circuit, each applied `reps` times on one resident
    state by ket::Stepper (the kernels ket::run applies, without
    reallocating the 2^n state per gate), median per gate; estimated run time
    = sum over kinds of (gates of that kind) x (mean median time of that
    kind). ket's Fuser merges nothing in this circuit (no gate's qubits are a
    subset of the previous gate's), so run() applies the same passes."""
    from ket_amd.gates import ARITY, Circuit
    from oracle import oracle as O
    if not O.ref_available():
        return None
    kinds = {}
    for op in circ.ops:
        if op[0] in ARITY:
            kinds.setdefault(op[0], []).append(op)
    rng = np.random.default_rng(seed)
    chosen = []
    for name in sorted(kinds):
        ops = kinds[name]
        idx = rng.choice(len(ops), min(per_kind, len(ops)), replace=False)
        chosen += [ops[i] for i in sorted(idx)]
    sample = Circuit(circ.n_qubits)
    for _ in range(reps):
        sample.ops += chosen
    r = O.ref_step_time(sample, threads, 1)
    t = np.array(r["steps"][0]).reshape(reps, len(chosen))
    med = np.median(t, axis=0)
    per_kind_s = {}
    for name in sorted(kinds):
        sel = [i for i, op in enumerate(chosen) if op[0] == name]
        per_kind_s[name] = float(np.mean(med[sel]))
    est = sum(len(kinds[k]) * per_kind_s[k] for k in kinds)
    gates = circ.n_unitary()
    counts = {k: len(v) for k, v in sorted(kinds.items())}
    return {"value": gates / est, "unit": "gates/s", "cores": r["threads"], "kind": "reference",
            "sample": (f"ket::Stepper of the reference built from source "
                       f"(set_num_threads({threads})) on the {circ.n_qubits}-qubit state: "
                       f"{len(chosen)} gates of the circuit ({per_kind} per kind), each "
                       f"applied {reps}x, median per gate; run time estimated as "
                       f"sum_kind count x mean median = {est:.1f} s for {gates} gates "
                       f"({counts})"),
            "estimated_run_s": est, "per_gate_s": per_kind_s, "gate_counts": counts,
            "setup_s": r["setup"], "measured_s": float(t.sum()), "threads": r["threads"]}


def emit(out):
    print(json.dumps(out), flush=True)
    return 0


def brick33_forms(circ, fused_steps, fused_warmup):
    """Config 3 at its real size: plain complex128 (128 GiB, one step), fused
    complex64 k = 5 blocks on MFMA (64 GiB, resident beside it), 1-F and
    max |amplitude error| between the two, and the same complex128
    arithmetic as exact runs (one step, equal to the plain form bit for bit).
    Each form's state is reset to |0...0> inside its timed steps."""
    n, gates = circ.n_qubits, circ.n_unitary()
    res = {"workload": "brick33: 33-qubit random brickwork, depth 24, {U,CX,CZ,CCX}",
           "n_qubits": n, "gates": gates}
    plain = N.StateVector(n, N.C128)
    log(f"[bench] brick33 plain complex128 (one step, {gates} gates)")
    wp, ep, stp = time_steps(plain, circ, N.FUSE_NONE, 5, 1, 0)
    gbps = stp["alg_bytes"] / wp / 1e9
    log(f"[bench] brick33 plain: {wp:.2f} s, {gates / wp:.1f} gates/s, {gbps:.0f} GB/s "
        f"({gbps / HBM_PEAK_GBPS:.1%} of 8 TB/s)")
    res["plain_c128"] = {"s_per_step": wp, "gates_per_s": gates / wp, "alg_GBps": gbps,
                         "roofline_frac": gbps / HBM_PEAK_GBPS, "launches": stp["launches"],
                         "alg_bytes": stp["alg_bytes"]}
    fused = N.StateVector(n, N.C64)
    log(f"[bench] brick33 fused complex64 k=5: warmup {fused_warmup}, steps {fused_steps}")
    w, e, st = time_steps(fused, circ, N.FUSE_BLOCKS, 5, fused_steps, fused_warmup)
    ps_f = w / fused_steps
    f, dmax = fused.fidelity(plain)
    log(f"[bench] brick33 fused: {ps_f:.2f} s/step, {gates / ps_f:.1f} gates/s, "
        f"{st['launches']} launches, {st['alg_bytes'] / ps_f / 1e9:.0f} GB/s, x{wp / ps_f:.2f} vs "
        f"plain; 1-F={1 - f:.2e} max|d|={dmax:.2e}")
    res["fused_c64_k5"] = {"s_per_step": ps_f, "steps": fused_steps, "warmup": fused_warmup,
                           "gates_per_s": gates / ps_f, "alg_GBps": st["alg_bytes"] / ps_f / 1e9,
                           "launches": st["launches"], "blocks": st["blocks"],
                           "plan_s": st["plan_seconds"], "one_minus_F": 1 - f,
                           "max_abs_err": dmax, "speedup_vs_plain_c128": wp / ps_f}
    fused.close()
    exact = N.StateVector(n, N.C128)
    log("[bench] brick33 exact runs complex128 (one step)")
    wx, ex, stx = time_steps(exact, circ, N.FUSE_EXACT, 0, 1, 0)
    fx, dx = exact.fidelity(plain)
    log(f"[bench] brick33 exact runs: {wx:.2f} s, {gates / wx:.1f} gates/s, {stx['launches']} "
        f"launches, x{wp / wx:.2f} vs plain, max|d| vs plain {dx:.1e}")
    res["exact_runs_c128"] = {"s_per_step": wx, "gates_per_s": gates / wx,
                              "launches": stx["launches"], "max_abs_diff_vs_plain": dx,
                              "speedup_vs_plain": wp / wx}
    exact.close()
    plain.close()
    return res


def bench_brick33(args, circ, desc):
    """Config 3 alone: the brick33 forms with --steps fused steps."""
    res = brick33_forms(circ, args.steps, args.warmup)
    p = res["plain_c128"]
    out = {"metric": METRIC, "value": p["gates_per_s"], "unit": "gates/s", "n_gpus": 1,
           "steps": 1, "warmup": 0, "ms_per_step": p["s_per_step"] * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
           "data": "synthetic",
           "config": {"workload": desc, "n_qubits": circ.n_qubits, "gates": circ.n_unitary(),
                      "form": "plain per-gate complex128"},
           "brick33": res, "device": N.device_info()["name"]}
    if not args.no_cpu:
        threads = host_cpus()["usable"]
        log(f"[bench] cpu baseline with {threads} threads (33-qubit state)")
        out["cpu_baseline"] = cpu_baseline_mix(circ, threads, per_kind=3, reps=3)
        out["host"] = host_cpus()
    return emit(out)


FORMS = {  # name: (precision, fusion, block width)
    "plain": (N.C128, N.FUSE_NONE, 0),
    "c128_exact_runs": (N.C128, N.FUSE_EXACT, 0),
    "c128_ket_fuser_runs": (N.C128, N.FUSE_REFERENCE, 0),
    "c128_k4": (N.C128, N.FUSE_BLOCKS, 4),
    "c64_exact_runs": (N.C64, N.FUSE_EXACT, 0),
    "c64_k4": (N.C64, N.FUSE_BLOCKS, 4),
    "c64_k5": (N.C64, N.FUSE_BLOCKS, 5),
}


def bench_form(args, circ, desc):
    """One form of a workload on its own (for profiling): --steps timed steps
    after --warmup, then one pass with a HIP event pair around every launch."""
    prec, fusion, k = FORMS[args.form]
    n, gates = circ.n_qubits, circ.n_unitary()
    sv = N.StateVector(n, prec)
    w, e, st = time_steps(sv, circ, fusion, k, args.steps, args.warmup)
    ps = w / args.steps
    sv.set_basis(0)
    recs = sv.profile_circuit(circ, fusion, k)
    sv.close()
    launches = [{"kind": r["kind"], "k": r["k"], "n_gates": r["n_gates"], "ms": r["ms"],
                 "alg_bytes": r["alg_bytes"]} for r in recs]
    log(f"[bench] {args.form}: {ps * 1e3:.1f} ms/step, {gates / ps:.1f} gates/s, "
        f"{st['launches']} launches")
    for i, r in enumerate(launches):
        log(f"[bench]   launch {i}: {r['kind']} k={r['k']} gates={r['n_gates']} {r['ms']:.3f} ms "
            f"{r['alg_bytes'] / (r['ms'] / 1e3) / 1e9:.0f} GB/s")
    return emit({"metric": METRIC, "value": gates / ps, "unit": "gates/s", "n_gpus": 1,
                 "steps": args.steps, "warmup": args.warmup, "ms_per_step": ps * 1e3,
                 "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                 "dtype": "c128" if prec == N.C128 else "c64", "data": "synthetic",
                 "config": {"workload": desc, "form": args.form, "n_qubits": n, "gates": gates},
                 "launches": launches, "device": N.device_info()["name"]})


def bench_hea(args):
    """Config 4: 4096 instances, widths U{10..22}, one sweep per step with a
    new parameter binding (two angle sets alternate), complex128 and complex64,
    HBM instances as fused blocks and as packed exact runs (ket_tuning
    batch_form), hipGraph replay; `value` is complex128 in the library's
    default form. CPU baseline: ket::run + expval per instance."""
    sets = [hea_batch(angle_seed=a) for a in (0, 1)]
    ns = sets[0][0]
    gates = sum(c.n_unitary() for c in sets[0][1])
    packed = [N.Batch.pack(s[1], s[2]) for s in sets]  # records built once, outside timing
    log(f"[bench] hea4096: {len(ns)} instances, {gates} gates per sweep")
    out = {"metric": METRIC}
    res = {}
    forms = [(f"{p}_{f}", prec, bf) for p, prec in (("c128", N.C128), ("c64", N.C64))
             for f, bf in (("blocks", 0), ("runs", 1))]
    default_form = {prec: N.get_tuning(prec)["batch_form"] for prec in (N.C128, N.C64)}
    for label, prec, bf in forms:
        N.set_tuning(prec, batch_form=bf)
        with N.Batch(ns, prec) as b:
            for w in range(max(args.warmup, 1)):
                b.run_packed(packed[w % 2], block_k=0, use_graph=True)
            t0 = time.perf_counter()
            dev_ms, plan_s, pack_s, up_s = 0.0, 0.0, 0.0, 0.0
            for k in range(args.steps):
                vals, st = b.run_packed(packed[k % 2], block_k=0, use_graph=True)
                dev_ms += st["device_ms"]
                plan_s += st["plan_seconds"]
                pack_s += st["pack_seconds"]
                up_s += st["upload_seconds"]
            wall = (time.perf_counter() - t0) / args.steps
            res[label] = {"s_per_sweep": wall, "instances_per_s": len(ns) / wall,
                          "gates_per_s": gates / wall, "device_ms_per_sweep": dev_ms / args.steps,
                          "plan_s_per_sweep": plan_s / args.steps,
                          "pack_s_per_sweep": pack_s / args.steps,
                          "upload_s_per_sweep": up_s / args.steps,
                          "alg_GBps_device": st["alg_bytes"] / (st["device_ms"] / 1e3) / 1e9,
                          "launches": st["launches"],
                          "instances_lds": st["instances_lds"], "graph_replayed": st["graph_replayed"],
                          "batch_form": "runs" if bf else "blocks",
                          "block_k": N.get_tuning(prec)["block_k"] or (4 if prec == N.C128 else 5)}
            log(f"[bench] {label}: {wall:.3f} s/sweep ({dev_ms / args.steps:.1f} ms device, "
                f"{plan_s / args.steps:.3f} s planning of which {pack_s / args.steps:.3f} s "
                f"packing, {up_s / args.steps:.3f} s upload), {len(ns) / wall:.0f} instances/s, "
                f"{st['launches']} launches")
        N.set_tuning(prec, batch_form=default_form[prec])
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        if O.ref_available():
            rng = np.random.default_rng(7)
            pick = rng.choice(len(ns), 32, replace=False)
            t0 = time.perf_counter()
            for i in pick:
                O.ref_expval(sets[0][1][i], [p for _, p in sets[0][2][i]])
            dt = time.perf_counter() - t0
            frac_gates = sum(sets[0][1][i].n_unitary() for i in pick) / gates
            cpu = {"value": len(pick) / dt, "unit": "instances/s", "cores": 1, "kind": "reference",
                   "sample": f"ket::run + expval(PauliSum) on 32 random instances of the sweep "
                             f"({frac_gates:.2%} of its gates), one process per instance, 1 thread"}
            log(f"[bench] cpu: {cpu['value']:.1f} instances/s")
    main = res["c128_runs" if default_form[N.C128] else "c128_blocks"]
    out.update({
        "value": main["instances_per_s"], "unit": "instances/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": main["s_per_sweep"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic",
        "config": {"workload": "hea4096: 4096 HEA instances (8 layers Ry,Rz + CZ ring), "
                               "n ~ U{10..22}, sum Z_i Z_i+1", "gates_per_sweep": gates},
        "forms": res, "cpu_baseline": cpu, "device": N.device_info()["name"]})
    return emit(out)


def bench_brick20(args):
    """Config 1: 20-qubit brickwork, 20 layers (590 gates); final state via
    ket's Fuser rule in complex128 (exact), 10^4 shots with mt19937{1}."""
    from oracle import oracle as O
    circ = W.brickwork(20, 20, seed=0)
    gates = circ.n_unitary()
    sv = N.StateVector(20, N.C128)
    w, e, st = time_steps(sv, circ, N.FUSE_REFERENCE, 5, args.steps, args.warmup)
    u = O.uniforms(1, 10_000)
    for _ in range(max(args.warmup, 1)):
        # the process's first sample() loads the sampler's kernels and sets up
        # the runtime's staging for host copies of this size
        sv.set_basis(0)
        sv.apply_circuit(circ, N.FUSE_REFERENCE)
        sv.sample(u)
    sv.set_basis(0)
    sv.apply_circuit(circ, N.FUSE_REFERENCE)
    sv.sync()
    t0 = time.perf_counter()
    shots = sv.sample(u)
    t_draw = time.perf_counter() - t0
    t0 = time.perf_counter()
    shots2 = sv.sample(u)  # sampler tables reused
    t_draw2 = time.perf_counter() - t0
    run_s = w / args.steps
    log(f"[bench] brick20: run {run_s * 1e3:.2f} ms ({gates / run_s:.0f} gates/s); 10^4 shots "
        f"{t_draw * 1e3:.2f} ms (tables built) / {t_draw2 * 1e3:.2f} ms (reused)")
    cpu = None
    if not args.no_cpu and O.ref_available():
        t = O.ref_time(circ, 1, 3)["seconds"]
        t0 = time.perf_counter()
        O.ref_sample(circ, 1, 10_000)
        t_s = time.perf_counter() - t0
        cpu = {"value": gates / min(t), "unit": "gates/s", "cores": 1, "kind": "reference",
               "sample": f"ket::run of the full circuit (1 thread, best of 3: {min(t) * 1e3:.1f} ms); "
                         f"simulate + 10^4 Simulation::sample: {t_s:.2f} s incl. process start"}
    out = {"metric": METRIC, "value": gates / run_s, "unit": "gates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": run_s * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
           "data": "synthetic",
           "config": {"workload": "brick20: 20-qubit brickwork, 20 layers, 590 gates; 10^4 shots",
                      "n_qubits": 20, "gates": gates},
           "shots_10k_ms": t_draw * 1e3, "shots_10k_ms_reused": t_draw2 * 1e3,
           "shots_reproducible": bool(np.array_equal(shots, shots2)),
           "cpu_baseline": cpu, "device": N.device_info()["name"]}
    sv.close()
    return emit(out)


def bench_planner(args):
    """Config 5: the learned fusion planner, in stages that can run in
    separate processes (--shard i --shards S splits the circuits; each
    stage's files go to --data-dir):
      collect: time windows of the 5000 training circuits (28-32 qubits,
               depth 40, 13-gate set) in every pass form -> data_<i>.npz;
      train:   train the cost model on the GPU on every data_*.npz and
               write it to --model-out;
      eval:    200 held-out circuits of the same distribution, whole, in
               every form (greedy blocks, ket's Fuser as exact runs, the
               run packer, learned), complex64 -> eval_<i>.json;
      report:  merge the eval_*.json into the result line;
      cases:   QFT-30, 30-qubit brickwork and one 30-qubit random circuit
               in every form and learned (where the forms trade places);
      all:     collect, train, eval and report in this process."""
    from ket_amd import costmodel as CM
    out = {"metric": METRIC}
    ddir = os.path.abspath(args.data_dir)
    os.makedirs(ddir, exist_ok=True)
    shard = lambda xs: xs[args.shard::args.shards]
    stages = ("collect", "train", "eval", "report", "cases") if args.stage == "all" else (args.stage,)
    states = CM.States(N.C64)
    mlp = None
    try:
        if "collect" in stages:
            circs = shard(CM.config5_circuits(args.train_circuits, seed=1234))
            log(f"[bench] planner: timing windows of {len(circs)} training circuits "
                f"(shard {args.shard}/{args.shards})")
            t0 = time.perf_counter()
            X, y = CM.collect(circs, states, seed=1 + args.shard, log=log)
            if args.ladder_circuits:
                lad = shard(CM.ladder_circuits(args.ladder_circuits, seed=4321))
                log(f"[bench] planner: timing windows of {len(lad)} QFT-like circuits")
                X2, y2 = CM.collect(lad, states, long_len=args.ladder_window, seed=101 + args.shard,
                                    log=log, long_forms=CM.LADDER_FORMS)
                X, y = np.concatenate([X, X2]), np.concatenate([y, y2])
                circs = circs + lad
            np.savez(os.path.join(ddir, f"data_{args.shard}.npz"), X=X, y=y,
                     circuits=len(circs), seconds=time.perf_counter() - t0)
            log(f"[bench] {len(y)} timed passes in {time.perf_counter() - t0:.0f} s")
            out["collect"] = {"circuits": len(circs), "passes": int(len(y)),
                              "collect_s": time.perf_counter() - t0}
        if "train" in stages:
            files = sorted(f for f in os.listdir(ddir) if f.startswith("data_"))
            parts = [np.load(os.path.join(ddir, f), allow_pickle=False) for f in files]
            X = np.concatenate([p["X"] for p in parts])
            y = np.concatenate([p["y"] for p in parts])
            n_circ = int(sum(int(p["circuits"]) for p in parts))
            log(f"[bench] planner: training on {len(y)} passes of {n_circ} circuits")
            t0 = time.perf_counter()
            init = None
            if args.init_model:
                init = np.load(args.init_model, allow_pickle=False)["params"]
                log(f"[bench] planner: fine-tuning {args.init_model}")
            mlp, hist = CM.train(X, y, epochs=args.epochs, lr=args.lr, log=log, init=init)
            t_train = time.perf_counter() - t0
            dest = os.path.abspath(args.model_out)
            os.makedirs(os.path.dirname(dest), exist_ok=True)
            CM.save(mlp, dest, train_passes=len(y), train_circuits=n_circ,
                    final_loss=hist[-1], epochs=len(hist))
            out["train"] = {"circuits": n_circ, "passes": int(len(y)),
                            "collect_s": float(sum(float(p["seconds"]) for p in parts)),
                            "train_s": t_train, "loss_first": hist[0], "loss_final": hist[-1]}
        if "eval" in stages:
            if mlp is None:  # the shipped model
                mlp = CM.load()
            if mlp is None:
                raise SystemExit("no ket_amd/cost_model.npz: run --stage train first")
            held = shard(CM.config5_circuits(args.eval_circuits, seed=777))
            log(f"[bench] planner: timing {len(held)} held-out circuits in every form")
            recs = CM.time_forms(mlp, held, states, reps=3, log=log)
            with open(os.path.join(ddir, f"eval_{args.shard}.json"), "w") as f:
                json.dump(recs, f)
        if "cases" in stages:
            # named circuits where the forms trade places: QFT-30 (blocks
            # fastest), 30-qubit brickwork and a config-5 circuit (runs)
            if mlp is None:
                mlp = CM.load()
            cases = [("qft30", W.qft_u(30, seed=0)), ("brick30", W.brickwork_mixed(30, 24, seed=0)),
                     ("random30", W.random_mixed(30, 40, seed=5))]
            out["cases"] = CM.time_cases(mlp, cases, states, log=log)
        if "report" in stages:
            recs = []
            for name in sorted(os.listdir(ddir)):
                if name.startswith("eval_"):
                    with open(os.path.join(ddir, name)) as f:
                        recs += json.load(f)
            res = CM.summarize(recs)
            out["eval"] = res
            log(f"[bench] {res['circuits']} circuits: " +
                ", ".join(f"{k} {v:.0f} ms" for k, v in res["total_ms"].items()) +
                f"; learned x{res['speedup_vs_best_fixed']:.3f} vs the best fixed form "
                f"({res['best_fixed_form']}), x{res['speedup_vs_fallback']:.3f} vs ket's Fuser "
                f"as exact runs, noise {res['noise']:.3f}")
    finally:
        states.close()
    ev = out.get("eval", {})
    out.update({
        "value": ev.get("speedup_vs_fallback"),
        "unit": "x (ket's Fuser as exact runs wall time / learned wall time)",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": ev.get("total_ms", {}).get("learned"),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": "planner: random 13-gate-set circuits, 28-32 qubits, depth 40"},
        "device": N.device_info()["name"]})
    return emit(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="qft30",
                    help="qft30 (default), brick33, hea4096, brick20, qftNN")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stage", default="all",
                    help="planner workload: collect | train | eval | report | cases | all")
    ap.add_argument("--shard", type=int, default=0, help="planner collect/eval: this shard")
    ap.add_argument("--shards", type=int, default=1, help="planner collect/eval: shard count")
    ap.add_argument("--data-dir", default="gpurun_out/planner",
                    help="planner: where collect/eval write and train/report read")
    ap.add_argument("--train-circuits", type=int, default=5000)
    ap.add_argument("--eval-circuits", type=int, default=200)
    ap.add_argument("--ladder-circuits", type=int, default=0,
                    help="planner collect: QFT-like circuits added to the training windows")
    ap.add_argument("--ladder-window", type=int, default=1000,
                    help="planner collect: gates in a QFT-like circuit's long window "
                         "(the default takes the whole circuit)")
    ap.add_argument("--init-model", default="",
                    help="planner train: start from this model's parameters (fine-tuning)")
    ap.add_argument("--epochs", type=int, default=60, help="planner train: epochs")
    ap.add_argument("--lr", type=float, default=3e-3, help="planner train: peak learning rate")
    ap.add_argument("--model-out", default="cost_model_trained.npz",
                    help="planner train: where the trained cost model is written "
                         "(the shipped one is ket_amd/cost_model.npz)")
    ap.add_argument("--no-fused", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (0: every CPU this process may use)")
    ap.add_argument("--form", choices=sorted(FORMS),
                    help="run only this form of the workload (profiling)")
    ap.add_argument("--no-brick33", action="store_true",
                    help="default run: skip the 33-qubit forms (about 90 s)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this backend has no multi-GPU path "
              "(a state larger than one GPU's HBM would be split by its top qubits "
              "across GPUs, which is not built); use --gpus 1.",
              file=sys.stderr)
        return 2

    if args.workload == "hea4096":
        return bench_hea(args)
    if args.workload == "planner":
        return bench_planner(args)
    if args.workload == "brick20":
        return bench_brick20(args)
    circ, desc = workload(args.workload)
    n = circ.n_qubits
    gates = circ.n_unitary()
    info = N.device_info()
    log(f"[bench] device {info['name']} CUs={info['compute_units']} "
        f"HBM={info['hbm_bytes'] / 2**30:.0f} GiB; workload {desc}")

    out = {"metric": METRIC}
    if args.form:
        return bench_form(args, circ, desc)
    if args.workload == "brick33":
        return bench_brick33(args, circ, desc)
    copy_gbps = max(N.calibrate_copy(4 << 30, 5, nt) for nt in (0, 1))
    log(f"[bench] copy kernel (4 GiB, read+write): {copy_gbps:.0f} GB/s")
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
                "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic_from_profile(kind),
                "copy_kernel_GBps": copy_gbps, "frac_of_copy": achieved / copy_gbps,
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
        # exact runs: per-gate arithmetic on LDS tiles (equal to the plain
        # form bit for bit), and ket's Fuser with runs (the drop-in default)
        for label, prec, fusion, k in [("c128_exact_runs", N.C128, N.FUSE_EXACT, 0),
                                       ("c128_ket_fuser_runs", N.C128, N.FUSE_REFERENCE, 0),
                                       ("c128_k4", N.C128, N.FUSE_BLOCKS, 4),
                                       ("c64_k4", N.C64, N.FUSE_BLOCKS, 4),
                                       ("c64_k5", N.C64, N.FUSE_BLOCKS, 5)]:
            sv = N.StateVector(n, prec)
            w, e, st = time_steps(sv, circ, fusion, k, args.steps, args.warmup)
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

    brick = None
    if not args.no_brick33 and args.workload == "qft30":
        brick = brick33_forms(W.brickwork_mixed(33, 24, seed=0), min(args.steps, 5), 1)

    cpu = None
    host = host_cpus()
    if not args.no_cpu:
        threads = args.cpu_threads or host["usable"]
        log(f"[bench] cpu baseline with {threads} threads ({host['model']}, "
            f"{host['reported']} logical CPUs reported, {host['usable']} usable)")
        try:
            cpu = cpu_baseline_mix(circ, threads)
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
        "brick33": brick,
        "cpu_baseline": cpu,
        "host": host,
        "device": info["name"],
    })
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
