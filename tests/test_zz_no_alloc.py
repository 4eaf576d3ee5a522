"""qpe_qift_noalloc (the allocation-free QIFT + P(k)) against qpe_qift +
qpe_identify_phase_amplitudes on identical registers: the same kernels, so
the QIFT output and P(k) agree bit for bit.  Two calls in a row, the second on
a smaller register; complex128 and complex64 (with pending renormalisation
factors); and a host pointer for P(k) is refused.  The test allocates the
device buffer for P(k) itself (hipMalloc through ctypes): the entry point
allocates nothing."""
import ctypes
import math

import numpy as np
import pytest

import qpesim
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Caller-owned device memory for P(k), through the HIP runtime."""

    _hip = None

    def __init__(self, nbytes):
        if DeviceBuffer._hip is None:
            DeviceBuffer._hip = ctypes.CDLL("libamdhip64.so")
        self.nbytes = nbytes
        self.ptr = ctypes.c_void_p()
        assert self._hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes)) == 0

    def to_host(self, count):
        out = np.zeros(count, np.float64)
        assert self._hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(count * 8),
                                   2) == 0   # hipMemcpyDeviceToHost
        return out

    def free(self):
        self._hip.hipFree(self.ptr)


def _pair(n, e, m, seed, dtype):
    """two registers holding the same evolved state"""
    one, two = qpesim.synthetic_integrals(n, seed)
    h, _ = qpesim.build_hamiltonian(n, e, one, two)
    psi = qpesim.initial_state(n, e, seed, 0.9)
    out = []
    for _ in range(2):
        sv = qpesim.StateVector(m, n, e, dtype)
        sv.set(0, psi / math.sqrt(1 << m))
        qpesim.time_evolve_over_terms(sv, h)     # complex64: leaves renormalisation factors pending
        out.append(sv)
    return out


@pytest.mark.parametrize("dtype", [qpesim.COMPLEX128, qpesim.COMPLEX64])
def test_qift_noalloc_matches_qift_two_calls(dtype):
    buf = DeviceBuffer(8 << 10)
    try:
        for n, e, m, seed in [(6, 3, 6, 3), (5, 2, 4, 4)]:     # the second call on a smaller register
            say(f"qift_noalloc dtype={dtype} n={n} e={e} m={m}")
            a, b = _pair(n, e, m, seed, dtype)
            a.qift()
            want_P = a.identify_phase_amplitudes()
            want_y = a.get()
            b.qift_noalloc(buf.ptr.value)
            got_P = buf.to_host(1 << m)
            assert np.array_equal(got_P, want_P)
            assert np.array_equal(b.get(), want_y)
            # the register itself still gives P(k) through the ordinary entry point (from the
            # stored amplitudes: complex64 rounds the QIFT output to FP32)
            tol = 1e-12 if dtype == qpesim.COMPLEX128 else 1e-6
            assert np.allclose(b.identify_phase_amplitudes(), want_P, rtol=0, atol=tol)
            a.close()
            b.close()
    finally:
        buf.free()


def test_qift_noalloc_refuses_host_memory():
    sv = qpesim.StateVector(3, 4, 2)
    host = np.zeros(8, np.float64)
    with pytest.raises(qpesim.QpeError) as ei:
        sv.qift_noalloc(host.ctypes.data)
    assert ei.value.code == qpesim.ERR_ARG
    # the refused call leaves no error behind for the next call
    sv.qift()
    assert sv.identify_phase_amplitudes().shape == (8,)
    sv.close()


def test_qift_noalloc_above_1024_outcomes():
    """Up to 4096 outcomes the allocation-free QIFT takes the radix-2 kernel (the register
    kernel would need a second register there); above that it is refused."""
    M = 1 << 12
    buf = DeviceBuffer(8 * M)
    try:
        sv = qpesim.StateVector(12, 4, 2, qpesim.COMPLEX128)
        rng = np.random.default_rng(3)
        X = (rng.normal(size=(M, 36)) + 1j * rng.normal(size=(M, 36))) / math.sqrt(M * 36)
        sv.set(0, X.reshape(-1))
        sv.qift_noalloc(buf.ptr.value)
        Pr = np.sum(np.abs(np.fft.fft(X, axis=0) / math.sqrt(M)) ** 2, axis=1)
        dp = float(np.max(np.abs(buf.to_host(M) - Pr)))
        say(f"no-alloc QIFT 2^12 outcomes: max|dP| vs numpy {dp:.2e}")
        assert dp < 1e-12
        sv.close()
        sv = qpesim.StateVector(13, 4, 2, qpesim.COMPLEX128)
        with pytest.raises(qpesim.QpeError) as ei:
            sv.qift_noalloc(buf.ptr.value)
        assert ei.value.code == qpesim.ERR_ARG
        sv.close()
    finally:
        buf.free()
