"""ket_expval_pauli_sum_noalloc (the entry point that allocates nothing)
against ket_sv_expval_pauli_sum on the same amplitudes, through the C ABI:
equal bit for bit, on two calls in a row, the second on a smaller state, in
both precisions. The amplitudes are copied into device memory this test
allocates itself with the HIP runtime (the entry point takes the caller's
buffer)."""
import ctypes

import numpy as np
import pytest

from ket_amd import native as N
from ket_amd import workloads as W

pytestmark = pytest.mark.gpu

hip = ctypes.CDLL("libamdhip64.so")
hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
hip.hipFree.argtypes = [ctypes.c_void_p]
H2D = 1


class DeviceCopy:
    """Device memory holding a copy of a host array (owned by the test)."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), host.nbytes) == 0
        assert hip.hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, H2D) == 0

    def __enter__(self):
        return self.ptr

    def __exit__(self, *exc):
        assert hip.hipFree(self.ptr) == 0


def masks(terms, n):
    fl = np.zeros(len(terms), np.uint64)
    sg = np.zeros(len(terms), np.uint64)
    ny = np.zeros(len(terms), np.int32)
    cf = np.zeros(len(terms), np.float64)
    for i, (c, p) in enumerate(terms):
        fl[i], sg[i], ny[i] = N.pauli_masks(p, n)
        cf[i] = c
    return fl, sg, ny, cf


def noalloc(ptr, n, precision, terms):
    fl, sg, ny, cf = masks(terms, n)
    v = ctypes.c_double(0)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    N.check(N.lib().ket_expval_pauli_sum_noalloc(ptr, n, precision, len(terms), p(fl), p(sg),
                                                  p(ny), p(cf), None, ctypes.byref(v)))
    return v.value


def random_terms(n, count, rng):
    out = []
    for _ in range(count):
        s = "".join(rng.choice(list("IXYZ"), n))
        out.append((float(rng.uniform(-2, 2)), s))
    out.append((0.5, "Z" * n))  # a diagonal term for sure
    out.append((-0.25, "I" * (n - 1) + "Z"))
    return out


@pytest.mark.parametrize("precision", [N.C128, N.C64])
def test_noalloc_equals_expval_sum_twice_shrinking(precision):
    rng = np.random.default_rng(7 + precision)
    for n, depth, count in [(20, 6, 40), (11, 4, 9)]:  # the second call is smaller
        c = W.random_mixed(n, depth, 100 + n)
        terms = random_terms(n, count, rng)
        with N.StateVector(n, precision) as sv:
            sv.apply_circuit(c, N.FUSE_BLOCKS)
            want = sv.expval_sum(terms)
            host = sv.to_numpy()
        if precision == N.C64:
            host = host.astype(np.complex64)  # exact: the values are complex64
        with DeviceCopy(host) as ptr:
            got = noalloc(ptr, n, precision, terms)
            again = noalloc(ptr, n, precision, terms)
        assert got == want, (n, got, want)
        assert again == want


def test_noalloc_rejects_what_it_cannot_hold():
    n = 4
    host = np.zeros(1 << n, np.complex128)
    host[0] = 1
    with DeviceCopy(host) as ptr:
        assert noalloc(ptr, n, N.C128, [(1.0, "ZIII")]) == 1.0
        with pytest.raises(ValueError, match="at most 64 terms"):
            noalloc(ptr, n, N.C128, [(1.0, "ZIII")] * 65)
        fl, sg, ny, cf = masks([(1.0, "ZIII")], n)
        v = ctypes.c_double(0)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        rc = N.lib().ket_expval_pauli_sum_noalloc(ptr, 3, N.C128, 1, p(fl), p(sg), p(ny), p(cf),
                                                  None, ctypes.byref(v))
        assert rc == 2  # the term's mask reaches past a 3-qubit state
        rc = N.lib().ket_expval_pauli_sum_noalloc(None, n, N.C128, 1, p(fl), p(sg), p(ny), p(cf),
                                                  None, ctypes.byref(v))
        assert rc == 1
