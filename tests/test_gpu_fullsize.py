"""Parity at the benchmark sizes (30 and 33 qubits), where no CPU oracle run
fits in a test: properties that hold whatever the size. A circuit followed by
its exact inverse returns |0...0> (per-gate complex128 passes and fused
blocks); the norm stays 1; exact sampling of a state with known support draws
only from it, each outcome as the sequential scan over its probabilities
places the uniform."""
import numpy as np
import pytest

from ket_amd import native as N
from ket_amd import workloads as W
from ket_amd.gates import Circuit
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def back_to_zero(sv, n, tol):
    a0 = sv.to_numpy(0, 1)[0]
    assert abs(a0 - 1.0) <= tol, a0
    assert abs(sv.norm() - 1.0) <= tol


@pytest.mark.parametrize("fusion,k", [(N.FUSE_NONE, 0), (N.FUSE_REFERENCE, 0), (N.FUSE_BLOCKS, 4)])
def test_qft30_then_inverse_complex128(fusion, k):
    c = W.qft_u(30, seed=11)
    inv = W.inverse(c)
    with N.StateVector(30, N.C128) as sv:
        sv.apply_circuit(c, fusion, k)
        assert abs(sv.norm() - 1.0) <= 1e-12
        sv.apply_circuit(inv, fusion, k)
        back_to_zero(sv, 30, 1e-12)


def test_brick33_then_inverse_complex64_fused():
    """64 GiB complex64 state, fused k = 5 blocks on MFMA (LDS tiles)."""
    c = W.brickwork(33, 6, seed=5, measure=False)
    inv = W.inverse(c)
    with N.StateVector(33, N.C64) as sv:
        sv.apply_circuit(c, N.FUSE_BLOCKS, 5)
        assert abs(sv.norm() - 1.0) <= 1e-5
        sv.apply_circuit(inv, N.FUSE_BLOCKS, 5)
        back_to_zero(sv, 33, 1e-5)


def test_exact_sampling_30q_known_support():
    n, qs = 30, (0, 7, 15, 29)
    c = Circuit(n)
    for q in qs:
        c.h(q)
    support = np.array(sorted(sum(1 << q for i, q in enumerate(qs) if (m >> i) & 1)
                              for m in range(16)), dtype=np.uint64)
    u = O.uniforms(3, 10_000)
    with N.StateVector(n, N.C128) as sv:
        sv.apply_circuit(c, N.FUSE_NONE)
        shots = sv.sample(u)
        amps = np.array([sv.to_numpy(int(i), 1)[0] for i in support])
    # The sequential running sum over the nonzero probabilities (zeros add
    # nothing exactly) and ket's rule: the first index whose sum exceeds u.
    cum = np.cumsum(amps.real * amps.real + amps.imag * amps.imag)
    j = np.searchsorted(cum, u, side="right")
    ext = np.append(support, np.uint64((1 << n) - 1))  # u past the total: last index
    want = ext[j]
    assert np.array_equal(shots, want)
    counts = np.array([np.sum(shots == s) for s in support])
    assert counts.min() > 625 - 5 * 25 and counts.max() < 625 + 5 * 25


def test_qft30_exact_runs_equal_per_gate_passes():
    """QFT-30 as exact runs on LDS tiles equals one pass per gate bit for bit
    over the whole 16 GiB state (max |difference| exactly 0)."""
    c = W.qft_u(30, seed=12)
    with N.StateVector(30, N.C128) as a, N.StateVector(30, N.C128) as b:
        a.apply_circuit(c, N.FUSE_NONE)
        st = b.apply_circuit(c, N.FUSE_EXACT)
        assert st["launches"] < c.n_unitary() // 4
        f, dmax = a.fidelity(b)
        assert dmax == 0.0 and abs(f - 1.0) <= 1e-12
