"""The largest states one MI355X holds (288 GB of HBM3E): 34 qubits in
complex128 and 35 in complex64, 256 GiB each, so every index, grid-stride
loop, tile count and sampler chunk table runs past 2^33. No CPU oracle run
fits at these sizes, so the checks are properties that hold whatever the
size: a circuit followed by its exact inverse returns |0...0> and keeps the
norm at 1 (per-gate passes and fused blocks, all 13 gate types); exact
sampling of a state with known support returns the sequential rule's
outcome for every uniform; Pauli expectation values of a known product
state. One state past the device's memory fails with KET_ERR_ALLOC."""
import numpy as np
import pytest

from ket_amd import native as N
from ket_amd import workloads as W
from ket_amd.gates import Circuit
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GiB = 1 << 30


def need_bytes(nbytes):
    hbm = N.device_info()["hbm_bytes"]
    if hbm < nbytes + 8 * GiB:
        pytest.skip(f"device has {hbm / GiB:.0f} GiB, the state needs {nbytes / GiB:.0f} GiB")


def round_trip(sv, c, fusion, k, tol):
    sv.set_basis(0)
    sv.apply_circuit(c, fusion, k)
    assert abs(sv.norm() - 1.0) <= tol
    sv.apply_circuit(W.inverse(c), fusion, k)
    a0 = sv.to_numpy(0, 1)[0]
    assert abs(a0 - 1.0) <= tol, a0
    # a few amplitudes spread over the upper half of the index range
    rest = [sv.to_numpy((1 << (sv.n - 1)) + (7 << j), 1)[0] for j in (0, 11, 23, sv.n - 4)]
    assert max(abs(x) for x in rest) <= tol


def test_34_qubits_complex128():
    n = 34
    need_bytes(16 << n)
    c = W.random_mixed(n, 2, seed=34)
    with N.StateVector(n, N.C128) as sv:
        round_trip(sv, c, N.FUSE_NONE, 0, 1e-12)
        round_trip(sv, c, N.FUSE_BLOCKS, 4, 1e-12)


def test_35_qubits_complex64_fused():
    n = 35
    need_bytes(8 << n)
    c = W.random_mixed(n, 2, seed=35)
    with N.StateVector(n, N.C64) as sv:
        round_trip(sv, c, N.FUSE_BLOCKS, 5, 1e-5)
        round_trip(sv, c, N.FUSE_NONE, 0, 1e-5)


def test_34_qubits_sampling_and_expval():
    n, qs = 34, (0, 17, 33)
    need_bytes(16 << n)
    c = Circuit(n)
    for q in qs:
        c.h(q)
    c.x(20)
    support = np.array(sorted(sum(1 << q for i, q in enumerate(qs) if (m >> i) & 1) | (1 << 20)
                              for m in range(8)), dtype=np.uint64)
    u = O.uniforms(5, 4000)
    with N.StateVector(n, N.C128) as sv:
        sv.apply_circuit(c, N.FUSE_NONE)
        shots = sv.sample(u)
        amps = np.array([sv.to_numpy(int(i), 1)[0] for i in support])
        assert abs(np.sum(np.abs(amps) ** 2) - 1.0) <= 1e-12

        def pauli(**at):
            s = ["I"] * n
            for q, p in at.items():
                s[n - 1 - int(q[1:])] = p
            return "".join(s)

        assert abs(sv.expval(pauli(q20="Z")) + 1.0) <= 1e-12
        assert abs(sv.expval(pauli(q0="X", q33="X")) - 1.0) <= 1e-12
        assert abs(sv.expval(pauli(q0="Z"))) <= 1e-12
        assert abs(sv.expval(pauli(q17="Y"))) <= 1e-12
    cum = np.cumsum(amps.real * amps.real + amps.imag * amps.imag)
    j = np.searchsorted(cum, u, side="right")
    ext = np.append(support, np.uint64((1 << n) - 1))  # u past the total: last index
    assert np.array_equal(shots, ext[j])
    counts = np.array([np.sum(shots == s) for s in support])
    assert counts.min() > 500 - 5 * 21 and counts.max() < 500 + 5 * 21


def test_state_beyond_device_memory_is_an_allocation_error():
    hbm = N.device_info()["hbm_bytes"]
    n = 34
    while (16 << n) <= hbm:
        n += 1
    with pytest.raises(MemoryError, match="cannot allocate"):  # KET_ERR_ALLOC
        N.StateVector(n, N.C128)
    with N.StateVector(4, N.C128) as sv:  # the device is still usable
        sv.apply_circuit(W.random_mixed(4, 3, seed=1), N.FUSE_NONE)
        assert abs(sv.norm() - 1.0) <= 1e-12
