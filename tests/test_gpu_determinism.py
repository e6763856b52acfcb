"""Repeatability: the same circuit gives the same bits every time, on fresh
states and on a state that reuses its cached plan (ket_sv_apply_circuit keeps
the last plan per state and skips replanning and the pool upload for the
same records), for fused complex64 blocks on MFMA, ket's Fuser as exact runs
and batched sweeps with and without graph replay; a launch-tuning change
makes the state plan afresh."""
import numpy as np
import pytest

from ket_amd import native as N
from ket_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,fusion,k", [(N.C64, N.FUSE_BLOCKS, 5),
                                                (N.C64, N.FUSE_BLOCKS, 4),
                                                (N.C128, N.FUSE_REFERENCE, 0),
                                                (N.C128, N.FUSE_BLOCKS, 4)])
def test_same_circuit_same_bits(precision, fusion, k):
    c = W.random_mixed(20, 12, seed=17)
    rec = c.records()
    with N.StateVector(20, precision) as a, N.StateVector(20, precision) as b:
        s1 = a.apply_records(rec, fusion, k)
        first = a.to_numpy()
        b.apply_records(rec, fusion, k)
        assert np.array_equal(b.to_numpy(), first)
        for _ in range(3):  # the cached plan, its pool still on the device
            a.set_basis(0)
            s2 = a.apply_records(rec, fusion, k)
            assert np.array_equal(a.to_numpy(), first)
        assert s2["launches"] == s1["launches"] and s2["plan_seconds"] < s1["plan_seconds"]
        b.profile_circuit(c, N.FUSE_NONE)  # another upload into b's pool
        b.set_basis(0)
        b.apply_records(rec, fusion, k)  # cached plan, pool uploaded again
        assert np.array_equal(b.to_numpy(), first)


def test_tuning_change_replans():
    c = W.qft_u(16, seed=3)
    rec = c.records()
    saved = N.get_tuning(N.C128)
    try:
        with N.StateVector(16, N.C128) as sv:
            sv.apply_records(rec, N.FUSE_EXACT)
            want = sv.to_numpy()
            sv.set_basis(0)
            cached = sv.apply_records(rec, N.FUSE_EXACT)["plan_seconds"]
            N.set_tuning(N.C128, **dict(saved, run_low_bits=6))
            sv.set_basis(0)
            fresh = sv.apply_records(rec, N.FUSE_EXACT)["plan_seconds"]
            assert fresh > 5 * cached  # planned again, with the new tile shape
            assert np.array_equal(sv.to_numpy(), want)  # exact runs: same bits
    finally:
        N.set_tuning(N.C128, **saved)


@pytest.mark.parametrize("precision", [N.C128, N.C64])
def test_batched_sweep_repeats(precision):
    rng = np.random.default_rng(4)
    ns = [int(x) for x in rng.integers(8, 19, 24)]
    circs = [W.hea(n, 8, seed=60 + i) for i, n in enumerate(ns)]
    obs = [[(1.0, p) for p in W.zz_chain(n)] for n in ns]
    with N.Batch(ns, precision) as b:
        v0, _ = b.run(circs, obs, block_k=0, use_graph=False)
        v1, _ = b.run(circs, obs, block_k=0, use_graph=False)
        v2, _ = b.run(circs, obs, block_k=0, use_graph=True)
        v3, s3 = b.run(circs, obs, block_k=0, use_graph=True)
    assert s3["graph_replayed"] == 1
    for v in (v1, v2, v3):
        assert np.array_equal(v, v0)
