"""GPU parity: the HIP path through the C ABI against the CPU oracle.

complex128 per-gate passes and ket's Fuser rule must equal the oracle value
for value (IEEE ==, which lets only the sign of a zero differ); dense fused
blocks (fused multiply-adds) within 1e-12; complex64 within 1-F <= 1e-6.
"""
import itertools

import numpy as np
import pytest

from conftest import random_state
from ket_amd import native as N
from ket_amd import workloads as W
from ket_amd.gates import ARITY, NPARAMS, Circuit
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def gpu_run(c, fusion, precision=N.C128, k=5, state=None):
    with N.StateVector(c.n_qubits, precision) as sv:
        if state is not None:
            sv.load(state)
        sv.apply_circuit(c, fusion, k)
        return sv.to_numpy()


def fid(a, b):
    return abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


GATES = [g for g in ARITY]


@pytest.mark.parametrize("name", GATES)
def test_each_gate_every_placement_exact(name):
    n = 7
    k = ARITY[name]
    rng = np.random.default_rng(hash(name) % 1000)
    placements = [p for p in itertools.permutations(range(n), k)]
    rng.shuffle(placements)
    placements = placements[:24] + [tuple(range(k)), tuple(range(n - k, n))]
    s0 = random_state(n, 3)
    for qs in placements:
        c = Circuit(n)
        params = list(rng.uniform(0, 2 * np.pi, NPARAMS.get(name, 0)))
        getattr(c, name)(*qs, *params)
        want = O.run(c, fuse=False, state=s0)
        got = gpu_run(c, N.FUSE_NONE, state=s0)
        assert np.array_equal(got, want), (name, qs, np.abs(got - want).max())


@pytest.mark.parametrize("name", GATES)
def test_pair_exchange_form_exact(name, restore_tuning):
    """Pair passes with one amplitude per thread (lane-pair swap) give ket's
    values in complex128 and the same bits as the two-amplitude form in
    complex64, at every tested placement."""
    n = 7
    k = ARITY[name]
    rng = np.random.default_rng(hash(name) % 1000 + 7)
    placements = [p for p in itertools.permutations(range(n), k)]
    rng.shuffle(placements)
    placements = placements[:12] + [tuple(range(k)), tuple(range(n - k, n))]
    s0 = random_state(n, 5)
    for qs in placements:
        c = Circuit(n)
        params = list(rng.uniform(0, 2 * np.pi, NPARAMS.get(name, 0)))
        getattr(c, name)(*qs, *params)
        want = O.run(c, fuse=False, state=s0)
        for nt in (0, 1, 2, 3):
            N.set_tuning(pair_exchange=64, nontemporal=nt)
            got = gpu_run(c, N.FUSE_NONE, state=s0)
            got64 = gpu_run(c, N.FUSE_NONE, N.C64, state=s0.astype(np.complex64))
            N.set_tuning(pair_exchange=0, nontemporal=nt)
            ref = gpu_run(c, N.FUSE_NONE, state=s0)
            ref64 = gpu_run(c, N.FUSE_NONE, N.C64, state=s0.astype(np.complex64))
            assert np.array_equal(got, want), (name, qs, nt, np.abs(got - want).max())
            assert np.array_equal(ref, want), (name, qs, nt)
            assert np.array_equal(got64, ref64), (name, qs, nt)


CIRCUITS = {
    "mixed12": lambda: W.random_mixed(12, 14, 7),
    "brick14": lambda: W.brickwork(14, 10, seed=2),
    "qft12": lambda: W.qft_u(12, seed=3),
    "hea10": lambda: W.hea(10, 8, seed=4),
    "bwm13": lambda: W.brickwork_mixed(13, 12, seed=5),
    "diag12": lambda: W.diag_mix(12, 6, seed=6),
    "qft14": lambda: W.qft_u(14, seed=8),
}


@pytest.mark.parametrize("cname", sorted(CIRCUITS))
@pytest.mark.parametrize("fusion", [N.FUSE_NONE, N.FUSE_REFERENCE, N.FUSE_EXACT])
def test_circuit_exact(cname, fusion):
    c = CIRCUITS[cname]()
    want = O.run(c, fuse=(fusion == N.FUSE_REFERENCE))
    got = gpu_run(c, fusion)
    assert np.array_equal(got, want), np.abs(got - want).max()


def test_golden_states_from_reference_exact():
    """Every small golden circuit, run by the GPU with ket's Fuser rule in
    complex128, equals the state the reference library wrote."""
    import os
    from test_oracle import GOLD, circuit_from_text
    g = np.load(GOLD, allow_pickle=False)
    names = sorted({k.split("__")[0] for k in g.files if k.endswith("__state")})
    for name in names:
        c = circuit_from_text(g[f"{name}__text"].tobytes().decode())
        got = gpu_run(c, N.FUSE_REFERENCE)
        assert np.array_equal(got, g[f"{name}__state"]), name


def test_golden_config1_from_reference():
    from test_oracle import GOLD, circuit_from_text
    g = np.load(GOLD, allow_pickle=False)
    c = circuit_from_text(g["brick20__text"].tobytes().decode())
    with N.StateVector(20) as sv:
        sv.apply_circuit(c, N.FUSE_REFERENCE)
        s = sv.to_numpy()
        shots = sv.sample(O.uniforms(1, 10_000))
    assert np.array_equal(s[g["brick20__idx"]], g["brick20__amps"])
    assert np.array_equal(shots.astype(np.uint32), g["brick20__shots_seed1"])


@pytest.mark.parametrize("cname", sorted(CIRCUITS))
@pytest.mark.parametrize("k", [2, 3, 4, 5, 6])
def test_blocks_complex128(cname, k):
    c = CIRCUITS[cname]()
    want = O.run(c)
    got = gpu_run(c, N.FUSE_BLOCKS, k=k)
    assert np.abs(got - want).max() <= 1e-12
    assert 1 - fid(got, want) <= 1e-12


STAGED = {
    "qft11": lambda: W.qft_u(11, seed=9),
    "brick11": lambda: W.brickwork(11, 8, seed=10),
    "mixed12": CIRCUITS["mixed12"],
    "qft12": CIRCUITS["qft12"],
    "bwm13": CIRCUITS["bwm13"],
    "diag12": CIRCUITS["diag12"],
    "qft14": CIRCUITS["qft14"],
    "brick14": CIRCUITS["brick14"],
}


@pytest.fixture
def restore_tuning():
    saved = {p: N.get_tuning(p) for p in (N.C128, N.C64)}
    yield
    for p, t in saved.items():
        N.set_tuning(p, **t)


@pytest.mark.parametrize("cname", sorted(STAGED))
@pytest.mark.parametrize("precision,k", [(N.C128, 3), (N.C128, 4), (N.C128, 5),
                                         (N.C64, 3), (N.C64, 4), (N.C64, 5)])
def test_blocks_lds_tiles_equal_direct(cname, precision, k, restore_tuning):
    """Blocks moved through LDS tiles (every block, block_stage 64) give the
    same bits as the direct kernels (block_stage 0): same products in the
    same order, only the memory path differs."""
    c = STAGED[cname]()
    N.set_tuning(block_stage=0)
    direct = gpu_run(c, N.FUSE_BLOCKS, precision, k)
    N.set_tuning(block_stage=64)
    staged = gpu_run(c, N.FUSE_BLOCKS, precision, k)
    assert np.array_equal(staged, direct), np.abs(staged - direct).max()
    want = O.run(c)
    tol = 1e-12 if precision == N.C128 else 1e-6
    assert 1 - fid(staged, want) <= tol


@pytest.mark.parametrize("cname", sorted(CIRCUITS))
@pytest.mark.parametrize("fusion,k", [(N.FUSE_NONE, 5), (N.FUSE_REFERENCE, 5),
                                      (N.FUSE_BLOCKS, 4), (N.FUSE_BLOCKS, 5)])
def test_complex64(cname, fusion, k):
    c = CIRCUITS[cname]()
    want = O.run(c)
    got = gpu_run(c, fusion, N.C64, k)
    assert 1 - fid(got, want) <= 1e-6
    assert np.abs(got - want).max() <= 1e-5


def test_apply_matrix_exact():
    n = 9
    rng = np.random.default_rng(11)
    s0 = random_state(n, 12)
    for k in (1, 2, 3, 4, 5, 6):
        q = [int(x) for x in rng.permutation(n)[:k]]
        a = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
        m, _ = np.linalg.qr(a)
        want = O.apply_matrix(s0, q, m)
        with N.StateVector(n) as sv:
            sv.load(s0)
            sv.apply_matrix(q, m)
            got = sv.to_numpy()
        assert np.array_equal(got, want), (k, q)


def test_sampling_matches_sequential_scan_config1():
    c = W.brickwork(20, 20, seed=0)
    state = O.run(c)
    u = O.uniforms(1, 10_000)
    want = np.array([O.measure_u(state, x) for x in u[:2000]], dtype=np.uint64)
    with N.StateVector(20) as sv:
        sv.apply_circuit(c, N.FUSE_REFERENCE)
        got = sv.sample(u)
    assert np.array_equal(got[:2000], want)
    want_all = O.sample(state, 1, 10_000)
    assert np.array_equal(got, want_all)


@pytest.mark.parametrize("n", [0, 1, 3, 8, 12, 13])
def test_sampling_small_and_edge(n):
    rng = np.random.default_rng(n)
    s = random_state(n, 100 + n) if n else np.ones(1, np.complex128)
    u = np.concatenate([rng.uniform(size=500), [0.0, np.nextafter(1.0, 0.0), 0.999999999999]])
    with N.StateVector(n) as sv:
        sv.load(s)
        got = sv.sample(u)
    want = np.array([O.measure_u(s, x) for x in u], dtype=np.uint64)
    assert np.array_equal(got, want)


def test_sampling_boundaries_exact():
    # u drawn exactly at (and one ulp around) the sequential cumulative sums.
    n = 14
    s = random_state(n, 77)
    p = s.real ** 2 + s.imag ** 2
    cum = np.empty_like(p)
    acc = 0.0
    for i, x in enumerate(p):
        acc += x
        cum[i] = acc
    rng = np.random.default_rng(5)
    idx = rng.integers(0, (1 << n) - 1, 300)
    u = np.concatenate([cum[idx], np.nextafter(cum[idx], 0), np.nextafter(cum[idx], 2)])
    u = u[u < 1.0]
    with N.StateVector(n) as sv:
        sv.load(s)
        got = sv.sample(u)
    want = np.array([O.measure_u(s, x) for x in u], dtype=np.uint64)
    assert np.array_equal(got, want)


def test_sampling_complex64_uses_fp64_probabilities():
    c = W.brickwork(16, 6, seed=9)
    with N.StateVector(16, N.C64) as sv:
        sv.apply_circuit(c, N.FUSE_BLOCKS, 5)
        s = sv.to_numpy()  # float amplitudes widened exactly
        u = O.uniforms(3, 3000)
        got = sv.sample(u)
    want = np.array([O.measure_u(s, x) for x in u], dtype=np.uint64)
    assert np.array_equal(got, want)


def test_expval_matches_oracle():
    c = W.random_mixed(11, 10, 21)
    s = O.run(c)
    rng = np.random.default_rng(1)
    paulis = ["".join(rng.choice(list("IXYZ"), 11)) for _ in range(12)]
    paulis += ["Z" * 11, "I" * 11, "xyzxyzxyzxy"]
    with N.StateVector(11) as sv:
        sv.apply_circuit(c)
        for p in paulis:
            assert abs(sv.expval(p) - O.expval(s, p)) <= 1e-12, p
        terms = [(0.5, p) for p in paulis] + [(0.25, q) for q in W.zz_chain(11)]
        want = sum(cf * O.expval(s, p) for cf, p in terms)
        assert abs(sv.expval_sum(terms) - want) <= 1e-12


def test_norm_fidelity_renormalize():
    n = 15
    s = random_state(n, 8) * 1.7
    with N.StateVector(n) as a, N.StateVector(n, N.C64) as b:
        a.load(s)
        assert abs(a.norm() - np.vdot(s, s).real) <= 1e-12 * 3
        b.copy_from(a)
        f, d = a.fidelity(b)
        assert 1 - f <= 1e-12 and d <= 1e-6
        a.renormalize()
        assert abs(a.norm() - 1.0) <= 1e-14


def test_errors_are_statuses():
    with N.StateVector(4) as sv:
        with pytest.raises(IndexError):
            sv.apply_x(4)
        with pytest.raises(ValueError):
            sv.apply_swap(1, 1)
        c = Circuit(6)
        c.h(5)
        with pytest.raises(IndexError):
            sv.apply_circuit(c)
    with pytest.raises(MemoryError):
        N.StateVector(40)
    with pytest.raises(ValueError):
        N.StateVector(41)


@pytest.mark.parametrize("precision", [N.C128, N.C64])
def test_tuned_block_width_applies_when_call_passes_zero(precision, restore_tuning):
    """ket_tuning.block_k is the fusion width of calls that pass block_k = 0,
    and the result equals the same width passed explicitly."""
    c = W.brickwork(12, 6, seed=21, measure=False)
    with N.StateVector(12, precision) as sv:
        for k in (3, 4, 5):
            N.set_tuning(precision, block_k=k)
            recs = sv.profile_circuit(c, N.FUSE_BLOCKS_ALL, 0)
            ks = [r["k"] for r in recs if r["kind"] == "block"]
            assert ks and max(ks) == k
            sv.set_basis(0)
            sv.apply_circuit(c, N.FUSE_BLOCKS, 0)
            auto = sv.to_numpy()
            sv.set_basis(0)
            sv.apply_circuit(c, N.FUSE_BLOCKS, k)
            assert np.array_equal(auto, sv.to_numpy())


@pytest.mark.parametrize("precision", [N.C128, N.C64])
@pytest.mark.parametrize("cname", ["mixed12", "qft14", "bwm13", "diag12", "brick14"])
def test_exact_runs_equal_per_gate_passes(cname, precision, restore_tuning):
    """Consecutive passes applied as exact runs on LDS tiles (KET_FUSE_EXACT;
    ket's Fuser and the block packer's leftover passes with exact_runs on)
    give the bits of one pass per gate, from a dense random state, in both
    precisions."""
    c = CIRCUITS[cname]()
    n = c.n_qubits
    start = random_state(n, 40 + n)
    with N.StateVector(n, precision) as sv:
        out = {}
        for label, fusion, runs in (("none", N.FUSE_NONE, 1), ("exact", N.FUSE_EXACT, 1),
                                    ("ref_off", N.FUSE_REFERENCE, 0), ("ref_on", N.FUSE_REFERENCE, 1),
                                    ("blocks_off", N.FUSE_BLOCKS, 0), ("blocks_on", N.FUSE_BLOCKS, 1)):
            N.set_tuning(precision, exact_runs=runs)
            sv.load(start)
            sv.apply_circuit(c, fusion)
            out[label] = sv.to_numpy()
        assert np.array_equal(out["exact"], out["none"])
        assert np.array_equal(out["ref_on"], out["ref_off"])
        # the passes the block packer leaves, as runs
        assert np.array_equal(out["blocks_on"], out["blocks_off"])
        if n >= (11 if precision == N.C128 else 12):
            kinds = [r["kind"] for r in sv.profile_circuit(c, N.FUSE_EXACT)]
            assert kinds.count("run") >= 1


@pytest.mark.parametrize("precision", [N.C128, N.C64])
@pytest.mark.parametrize("cname", sorted(CIRCUITS))
def test_packed_runs(cname, precision):
    """KET_FUSE_RUNS: per-gate passes with commuting gates pulled forward
    into fewer exact runs; the rounding order changes, so the bar is that of
    the fused modes (complex128 within 1e-12, complex64 1-F <= 1e-6), and
    the plan has no more launches than KET_FUSE_EXACT's."""
    c = CIRCUITS[cname]()
    want = O.run(c)
    got = gpu_run(c, N.FUSE_RUNS, precision)
    if precision == N.C128:
        assert np.abs(got - want).max() <= 1e-12
        assert 1 - fid(got, want) <= 1e-12
    else:
        assert 1 - fid(got, want) <= 1e-6
        assert np.abs(got - want).max() <= 1e-5
    with N.StateVector(c.n_qubits, precision) as sv:
        packed = sv.apply_circuit(c, N.FUSE_RUNS)["launches"]
        exact = sv.apply_circuit(c, N.FUSE_EXACT)["launches"]
    assert packed <= exact


@pytest.mark.parametrize("precision", [N.C128, N.C64])
@pytest.mark.parametrize("n", [3, 12, 15])
def test_prob_prefix_is_kets_sequential_sum(n, precision):
    """ket_sv_prob_prefix equals the FP64 running sum measure() builds
    (numpy's add.accumulate is the same sequential sum) bit for bit, at
    random indices, chunk boundaries and the last index; sampling agrees
    with it (the outcome for u is the first index whose prefix exceeds u)."""
    c = W.random_mixed(n, 6, seed=n)
    with N.StateVector(n, precision) as sv:
        sv.apply_circuit(c, N.FUSE_NONE)
        a = sv.to_numpy()
        if precision == N.C64:  # the state as stored, widened exactly
            a = a.astype(np.complex64).astype(np.complex128)
        p = a.real * a.real + a.imag * a.imag
        cum = np.add.accumulate(p)
        rng = np.random.default_rng(n)
        dim = 1 << n
        idx = np.unique(np.concatenate([rng.integers(0, dim, 300), [0, dim - 1],
                                        np.arange(0, dim, 4096), np.arange(4095, dim, 4096)]))
        got = sv.prob_prefix(idx)
        assert np.array_equal(got, cum[idx])
        u = O.uniforms(n, 200)
        shots = sv.sample(u)
        pre = sv.prob_prefix(shots)
        inside = u < cum[-1]  # u past the total: ket returns the last index
        assert np.all(u[inside] < pre[inside]) and np.all(shots[~inside] == dim - 1)
        prev = sv.prob_prefix(np.maximum(shots.astype(np.int64) - 1, 0).astype(np.uint64))
        assert np.all((shots == 0) | (u >= prev))
        with pytest.raises(IndexError):
            sv.prob_prefix([dim])
