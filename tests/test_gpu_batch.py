"""Batched sweeps (ket_batch_*): every instance's expectation value against
the oracle's run + expval, for LDS-resident and HBM instances, uniform and
non-uniform structures, with and without hipGraph replay."""
import numpy as np
import pytest

from ket_amd import native as N
from ket_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_value(c, obs):
    s = O.run(c)
    return sum(cf * O.expval(s, p) for cf, p in obs)


def zz_obs(n):
    return [(1.0, p) for p in W.zz_chain(n)] if n > 1 else [(1.0, "Z")]


@pytest.fixture
def restore_tuning():
    saved = {p: N.get_tuning(p) for p in (N.C128, N.C64)}
    yield
    for p, t in saved.items():
        N.set_tuning(p, **t)


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("precision,tol", [(N.C128, 1e-10), (N.C64, 2e-4)])
def test_hea_sweep_matches_oracle(precision, tol, form, restore_tuning):
    N.set_tuning(precision, batch_form=form)  # HBM instances as blocks / as runs
    rng = np.random.default_rng(0)
    ns = [int(x) for x in rng.integers(2, 17, 40)] + [1, 12, 13, 14]
    circs = [W.hea(n, 8, seed=100 + i) for i, n in enumerate(ns)]
    obs = [zz_obs(n) for n in ns]
    with N.Batch(ns, precision) as b:
        vals, st = b.run(circs, obs, block_k=4, use_graph=False)
        assert st["instances_lds"] + st["instances_hbm"] == len(ns)
        assert st["instances_hbm"] > 0 and st["instances_lds"] > 0
        want = np.array([oracle_value(c, o) for c, o in zip(circs, obs)])
        assert np.abs(vals - want).max() <= tol
        # New angles, same structure: the captured graph is replayed.
        circs2 = [W.hea(n, 8, seed=500 + i) for i, n in enumerate(ns)]
        want2 = np.array([oracle_value(c, o) for c, o in zip(circs2, obs)])
        v1, s1 = b.run(circs2, obs, block_k=4, use_graph=True)
        v2, s2 = b.run(circs2, obs, block_k=4, use_graph=True)
        v3, s3 = b.run(circs, obs, block_k=4, use_graph=True)
        assert s1["graph_replayed"] == 0 and s2["graph_replayed"] == 1 and s3["graph_replayed"] == 1
        assert np.abs(v1 - want2).max() <= tol
        assert np.abs(v2 - want2).max() <= tol
        assert np.abs(v3 - want).max() <= tol


@pytest.mark.parametrize("precision,tol", [(N.C128, 1e-10), (N.C64, 2e-4)])
def test_hea_widths_17_to_22_in_the_arena(precision, tol):
    """The top of the sweep's width range (config 4 draws widths 10..22):
    instances of 17..22 qubits in the HBM arena, with graph capture and
    replay, against the oracle's run + expval."""
    ns = [17, 18, 19, 20, 21, 22, 17, 22]
    circs = [W.hea(n, 8, seed=300 + i) for i, n in enumerate(ns)]
    obs = [zz_obs(n) for n in ns]
    want = np.array([oracle_value(c, o) for c, o in zip(circs, obs)])
    with N.Batch(ns, precision) as b:
        v1, s1 = b.run(circs, obs, block_k=0, use_graph=True)
        v2, s2 = b.run(circs, obs, block_k=0, use_graph=True)
    assert s1["instances_hbm"] == len(ns) and s2["graph_replayed"] == 1
    assert np.abs(v1 - want).max() <= tol and np.abs(v2 - want).max() <= tol


def test_non_uniform_structures_and_all_gates():
    rng = np.random.default_rng(3)
    ns = [int(x) for x in rng.integers(3, 16, 24)]
    circs = [W.random_mixed(n, 8, seed=40 + i) for i, n in enumerate(ns)]
    obs = []
    for n in ns:
        terms = []
        for _ in range(3):
            p = "".join(rng.choice(list("IZ"), n))
            terms.append((float(rng.uniform(-1, 1)), p))
        obs.append(terms)
    with N.Batch(ns, N.C128) as b:
        for k in (3, 5):
            vals, st = b.run(circs, obs, block_k=k, use_graph=(k == 5))
            want = np.array([oracle_value(c, o) for c, o in zip(circs, obs)])
            assert np.abs(vals - want).max() <= 1e-10, k


def blocks_on(n, groups, seed):
    """U gates on each listed group of qubits in turn: with block width 4 and
    every group emitted as a block, one block per group."""
    rng = np.random.default_rng(seed)
    c = W.Circuit(n)
    for g in groups:
        for q in g:
            c.u(q, *rng.uniform(0, 2 * np.pi, 3))
    return c


def test_graph_replay_follows_instance_boundaries(restore_tuning):
    """Two same-width instances whose passes split 3 + 2 and then 2 + 3: the
    launch counts, pool size and per-pass shapes agree, but each launch
    addresses its own instance's state, so the second sweep must not replay
    the first sweep's graph."""
    N.set_tuning(N.C128, batch_form=0)  # fused blocks
    n = 14
    three = [range(0, 4), range(4, 8), range(8, 12)]
    two = [range(0, 4), range(6, 10)]
    obs = [zz_obs(n), zz_obs(n)]
    with N.Batch([n, n], N.C128) as b:
        first = [blocks_on(n, three, 1), blocks_on(n, two, 2)]
        v1, s1 = b.run(first, obs, block_k=4, use_graph=True)
        second = [blocks_on(n, two, 3), blocks_on(n, three, 4)]
        v2, s2 = b.run(second, obs, block_k=4, use_graph=True)
        v3, s3 = b.run(second, obs, block_k=4, use_graph=True)
        assert s1["instances_hbm"] == 2 and s1["launches"] == s2["launches"]
        assert s2["graph_replayed"] == 0 and s3["graph_replayed"] == 1
        for vals, circs in ((v1, first), (v2, second), (v3, second)):
            want = np.array([oracle_value(c, o) for c, o in zip(circs, obs)])
            assert np.abs(vals - want).max() <= 1e-10


def test_batch_rejects_bad_input():
    with pytest.raises(ValueError):
        N.Batch([3, 31])
    with N.Batch([3]) as b:
        c = W.hea(3, 1)
        with pytest.raises(ValueError):
            b.run([c], [[(1.0, "XII")]])


@pytest.mark.parametrize("precision", [N.C128, N.C64])
@pytest.mark.parametrize("k", [3, 4, 5])
def test_batched_lds_tiles_equal_direct(precision, k, restore_tuning):
    """HBM instances through the batched LDS-tile kernels (block_stage 64)
    against the direct batched kernels (block_stage 0): the same bits where
    both use the vector FMA path; complex64 k = 4, 5 tiles run on MFMA while
    the direct batched kernel uses FP32 vector FMAs, so there both are held
    to the oracle's tolerance only."""
    ns = [13, 14, 15, 16, 17, 14, 15] if precision == N.C128 else [14, 15, 16, 17, 15, 16]
    circs = [W.hea(n, 4, seed=900 + i) for i, n in enumerate(ns)]
    obs = [zz_obs(n) for n in ns]
    out = {}
    N.set_tuning(precision, batch_form=0)  # fused blocks
    for stage in (0, 64):
        N.set_tuning(block_stage=stage)
        with N.Batch(ns, precision) as b:
            vals, st = b.run(circs, obs, block_k=k, use_graph=False)
            assert st["instances_hbm"] == len(ns)
            out[stage] = vals
    want = np.array([oracle_value(c, o) for c, o in zip(circs, obs)])
    tol = 1e-10 if precision == N.C128 else 2e-4
    if precision == N.C128 or k == 3:
        assert np.array_equal(out[0], out[64])
    assert np.abs(out[0] - want).max() <= tol
    assert np.abs(out[64] - want).max() <= tol


@pytest.mark.parametrize("precision,tol", [(N.C128, 1e-10), (N.C64, 2e-4)])
def test_batched_runs_match_oracle(precision, tol, restore_tuning):
    """HBM instances as packed exact runs (ket_tuning.batch_form = 1): one
    batched launch per run for instances of a width, each workgroup reading
    its own instance's record; widths 14..22 (above the LDS-resident widths
    of both precisions) and a non-uniform group (other gate lists, so other
    run counts: one launch per instance and run), with graph capture and
    replay on new angles."""
    N.set_tuning(precision, batch_form=1)
    ns = [14, 15, 16, 17, 18, 19, 20, 21, 22, 15, 22]
    circs = [W.hea(n, 8, seed=700 + i) for i, n in enumerate(ns)]
    circs[9] = W.random_mixed(15, 8, seed=3)  # width 15: instances plan differently
    obs = [zz_obs(n) for n in ns]
    want = np.array([oracle_value(c, o) for c, o in zip(circs, obs)])
    with N.Batch(ns, precision) as b:
        v1, s1 = b.run(circs, obs, block_k=0, use_graph=True)
        assert s1["instances_hbm"] == len(ns) and s1["graph_replayed"] == 0
        assert np.abs(v1 - want).max() <= tol
        circs2 = [W.hea(n, 8, seed=900 + i) if i != 9 else circs[9] for i, n in enumerate(ns)]
        want2 = np.array([oracle_value(c, o) for c, o in zip(circs2, obs)])
        v2, s2 = b.run(circs2, obs, block_k=0, use_graph=True)
        v3, s3 = b.run(circs2, obs, block_k=0, use_graph=False)
    assert s2["graph_replayed"] == 1
    assert np.abs(v2 - want2).max() <= tol and np.abs(v3 - want2).max() <= tol
    assert np.array_equal(v2, v3)  # the replay gives the eager launches' bits
