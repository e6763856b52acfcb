"""Block matrices built on the GPU (ket_tuning.device_build = 1, the default)
against the same plans with the matrices built on the host: the build kernel
uses the host builder's products and sums, so every state must be equal bit
for bit, for fused blocks (complex128 and complex64, widths 2..6, folded
phase terms), ket's Fuser (exact matrices, also copied into exact runs) and
the learned planner; and the device-built result must still meet the oracle
bar (ket's Fuser value for value, fused complex128 within 1e-12)."""
import numpy as np
import pytest

from ket_amd import native as N
from ket_amd import workloads as W
from oracle import oracle as O
from test_runs_cpu import every_gate_circuit

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_tuning():
    saved = {p: N.get_tuning(p) for p in (N.C128, N.C64)}
    yield
    for p, t in saved.items():
        N.set_tuning(p, **t)


def run(c, precision, fusion, k, device_build):
    N.set_tuning(precision, device_build=device_build)
    with N.StateVector(c.n_qubits, precision) as sv:
        sv.apply_circuit(c, fusion, k)
        st = dict(sv.last_stats)
        return sv.to_numpy(), st


CIRCUITS = {
    "every13": lambda: every_gate_circuit(13, 6, 21),
    "qft14": lambda: W.qft_u(14, seed=2),
    "mixed15": lambda: W.random_mixed(15, 12, seed=3),
    "hea14": lambda: W.hea(14, 8, seed=4),
}


@pytest.mark.parametrize("name", sorted(CIRCUITS))
@pytest.mark.parametrize("precision,fusion,k", [
    (N.C128, N.FUSE_REFERENCE, 0), (N.C128, N.FUSE_BLOCKS, 3), (N.C128, N.FUSE_BLOCKS, 4),
    (N.C128, N.FUSE_BLOCKS, 6), (N.C64, N.FUSE_BLOCKS, 4), (N.C64, N.FUSE_BLOCKS, 5),
    (N.C64, N.FUSE_BLOCKS_ALL, 5), (N.C64, N.FUSE_LEARNED, 0), (N.C128, N.FUSE_LEARNED, 0)])
def test_device_built_matrices_equal_host_built(name, precision, fusion, k, restore_tuning):
    c = CIRCUITS[name]()
    host, _ = run(c, precision, fusion, k, 0)
    dev, st = run(c, precision, fusion, k, 1)
    assert np.array_equal(dev.view(np.uint8), host.view(np.uint8)), \
        f"max |diff| {np.abs(dev - host).max():.3e}"
    if fusion == N.FUSE_REFERENCE:
        assert np.array_equal(dev, O.run(c))  # ket::run's values
    elif precision == N.C128 and fusion == N.FUSE_BLOCKS:
        assert np.abs(dev - O.run(c)).max() <= 1e-12


def test_device_build_repeats_and_cached_plan(restore_tuning):
    """The same circuit twice on one state reuses the uploaded pool, whose
    matrices the first call built; a third call after a tuning change plans
    and builds again. All three equal bit for bit."""
    c = W.random_mixed(16, 10, seed=9)
    N.set_tuning(N.C64, device_build=1)
    with N.StateVector(16, N.C64) as sv:
        outs = []
        for i in range(3):
            if i == 2:
                N.set_tuning(N.C64, block_stage=N.get_tuning(N.C64)["block_stage"])
            sv.set_basis(0)
            sv.apply_circuit(c, N.FUSE_BLOCKS, 5)
            outs.append(sv.to_numpy())
    assert all(np.array_equal(o.view(np.uint8), outs[0].view(np.uint8)) for o in outs)
