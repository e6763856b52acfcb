"""The exact Born sampler and ket_sv_prob_prefix at scale, on dense states.

ket's measure() returns the first index whose sequential FP64 running sum of
|a_i|^2 exceeds u, else the last index (reference src/measurement.cpp:12-21).
The host side here computes that rule directly: |a|^2 as re*re + im*im
(separately rounded, as std::norm), np.add.accumulate for the running sum (a
sequential loop, rounding once per addition), and searchsorted(side="right")
for the first index above u. The device must agree bit for bit at uniforms
placed exactly on running-sum values and one ulp either side (where an
off-by-one or a differently rounded sum changes the outcome), at chunk
boundaries of the sampler's 4096-amplitude chunks and at the indices where
the sum crosses a power of two, and at seeded uniforms.
"""
import numpy as np
import pytest

from ket_amd import native as N
from ket_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu
CHUNK = 4096


def rule(cum, u):
    return np.minimum(np.searchsorted(cum, u, side="right"), len(cum) - 1).astype(np.uint64)


def probe_indices(cum, rng, count):
    n = len(cum)
    idx = set(int(i) for i in rng.integers(0, n, count))
    chunks = n // CHUNK
    for k in range(1, chunks, max(1, chunks // 400)):
        idx |= {k * CHUNK - 1, k * CHUNK, k * CHUNK + 1}
    for e in range(1, 64):  # where the running sum crosses 2^-e
        j = int(np.searchsorted(cum, 2.0 ** -e))
        idx |= {max(j - 1, 0), min(j, n - 1)}
    idx |= {0, 1, n - 2, n - 1}
    return np.array(sorted(idx), dtype=np.uint64)


def probes(cum, idx):
    c = cum[idx.astype(np.int64)]
    u = np.concatenate([c, np.nextafter(c, -1.0), np.nextafter(c, 2.0)])
    return u[(u >= 0.0) & (u < 1.0)]


def check_state(sv, a, rng, count):
    p = a.real * a.real + a.imag * a.imag
    cum = np.add.accumulate(p)
    idx = probe_indices(cum, rng, count)
    u = np.concatenate([probes(cum, idx), O.uniforms(7, 10_000)])
    print(f"{len(a)} amplitudes, {len(idx)} probe indices, {len(u)} draws", flush=True)
    got = sv.sample(u)
    want = rule(cum, u)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, (bad[:5], u[bad[:5]], got[bad[:5]], want[bad[:5]])
    assert np.array_equal(sv.prob_prefix(idx), cum[idx.astype(np.int64)])
    return cum


@pytest.mark.parametrize("precision", [N.C128, N.C64])
def test_dense_26_qubit_state(precision):
    n = 26
    c = W.brickwork(n, 4, seed=11, measure=False)
    with N.StateVector(n, precision) as sv:
        sv.apply_circuit(c, N.FUSE_BLOCKS)
        a = sv.to_numpy()  # complex64 widens exactly
        assert np.count_nonzero(a) == a.size
        cum = check_state(sv, a, np.random.default_rng(5 + precision), 3000)
        # the sum crossed many binades on the way (dense, spread-out state)
        assert np.unique(np.floor(np.log2(cum[cum > 0]))).size >= 20


def amp_for(p):
    """An amplitude with |a|^2 == p exactly, for p a power of two."""
    e = -int(np.log2(p))
    return 2.0 ** (-e / 2) if e % 2 == 0 else (1 + 1j) * 2.0 ** (-(e + 1) / 2)


def tie_state(n, seed):
    """Amplitudes whose running sum crosses 44 binades (growing steps 2^-45
    .. 2^-2) and, between the steps, adds half an ulp of the current sum (an
    exact tie, resolved to even), a whole ulp (exact) or nothing."""
    rng = np.random.default_rng(seed)
    size = 1 << n
    steps = set(int(i) for i in np.linspace(size // 64, size - size // 64, 44))
    a = np.zeros(size, dtype=np.complex128)
    cum, k = 0.0, 0
    for i in range(size):
        if i in steps:
            p = 2.0 ** -(45 - k)
            k += 1
        elif cum == 0.0:
            continue
        else:
            r = rng.integers(0, 3)
            if r == 2:
                continue
            p = float(np.spacing(cum)) / (2 if r == 0 else 1)
        a[i] = amp_for(p)
        cum += abs(a[i].real) ** 2 + abs(a[i].imag) ** 2
    return a


def test_running_sum_ties_and_binade_crossings():
    """Every |a|^2 is a power of two, so the running sum's roundings are
    exact ties at chosen places (the sampler's tie path) in every binade it
    crosses (its binade-crossing path)."""
    n = 16
    a = tie_state(n, 21)
    with N.StateVector(n, N.C128) as sv:
        sv.load(a)
        cum = check_state(sv, a, np.random.default_rng(8), 2000)
    # The premise, checked on the host: additions that were exact ties, in
    # both directions (x86 long double holds each cum + p exactly here), and
    # the binades the sum crossed.
    p = a.real * a.real + a.imag * a.imag
    prev = np.concatenate([[0.0], cum[:-1]]).astype(np.longdouble)
    exact = prev + p.astype(np.longdouble)
    half = np.spacing(cum).astype(np.longdouble) / 2
    diff = cum.astype(np.longdouble) - exact
    assert np.count_nonzero(diff == half) > 100 and np.count_nonzero(diff == -half) > 100
    assert np.unique(np.floor(np.log2(cum[cum > 0]))).size >= 40
