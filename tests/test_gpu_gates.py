"""Ancilla gates on the GPU (qpe_state_apply_gate / _apply_controlled_gate,
cpp/state_vector.hpp:51-113) against the oracle's restatement, the
reference's own gate tests, and a gate-level inverse QFT circuit against the
product's qift kernel (the gate-level QIFT validation)."""
import math

import numpy as np
import pytest

import gates_oracle as go
import qpesim
from test_gates import inverse_qft_circuit
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def state_with(m, n, e, x, dtype=qpesim.COMPLEX128):
    sv = qpesim.StateVector(m, n, e, dtype)
    sv.set(0, x.ravel())
    return sv


def test_reference_gate_cases_on_the_device():
    """cpp/state_vector.hpp:383-455 with StateVector(m, 2, 1): four amplitudes per block."""
    say("gates: reference cases")
    X = qpesim.pauli_x()
    for m, start, ops, want in [
        (1, 0, [(X, None, 0)], 4),
        (2, 0, [(X, None, 0)], 8),
        (2, 0, [(X, 1, 0)], 0),
        (2, 4, [(X, 1, 0)], 12),
        (2, 4, [(X, 1, 0), (X, 0, 1)], 8),
        (3, 4, [(X, 2, 0)], 20),
        (3, 4, [(X, 2, 0), (X, 0, 2)], 16),
    ]:
        sv = qpesim.StateVector(m, 2, 1)
        x = np.zeros(sv.size(), np.complex128)
        x[start] = 1.0
        sv.set(0, x)
        for g, c, q in ops:
            if c is None:
                sv.apply_gate(g, q)
            else:
                sv.apply_controlled_gate(g, c, q)
        y = sv.get()
        sv.close()
        assert y[want] == 1.0 and np.count_nonzero(y) == 1
    sv = qpesim.StateVector(2, 2, 1)
    x = np.zeros(sv.size(), np.complex128)
    x[0] = 1.0
    sv.set(0, x)
    sv.apply_gate(qpesim.hadamard(), 0)
    y = sv.get()
    sv.close()
    assert y[0] == 1.0 / 2 ** 0.5 and y[8] == 1.0 / 2 ** 0.5


@pytest.mark.parametrize("m,n,e", [(1, 3, 1), (3, 4, 2), (5, 5, 2), (6, 6, 3)])
def test_random_gates_match_oracle(m, n, e):
    say(f"gates: random m={m}")
    rng = np.random.default_rng(m * 10 + n)
    D = math.comb(n, e) ** 2
    x = rng.normal(size=(1 << m, D)) + 1j * rng.normal(size=(1 << m, D))
    sv = state_with(m, n, e, x)
    ref = x.copy()
    for _ in range(12):
        g = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        q = int(rng.integers(0, m))
        if m >= 2 and rng.integers(0, 2):
            c = int(rng.choice([k for k in range(m) if k != q]))
            sv.apply_controlled_gate(g, c, q)
            go.apply_gate(ref, m, g, q, control=c)
        else:
            sv.apply_gate(g, q)
            go.apply_gate(ref, m, g, q)
    y = sv.get().reshape(1 << m, D)
    sv.close()
    assert np.max(np.abs(y - ref)) / np.max(np.abs(ref)) < 1e-14


@pytest.mark.parametrize("m", [1, 2, 4, 7])
def test_gate_level_inverse_qft_equals_qift(m):
    """H and controlled InvRotate gates plus swaps against qpe_qift on the same
    register (both natural order, 1/sqrt(M))."""
    say(f"gates: circuit vs qift m={m}")
    n, e = 5, 2
    D = math.comb(n, e) ** 2
    rng = np.random.default_rng(100 + m)
    x = rng.normal(size=(1 << m, D)) + 1j * rng.normal(size=(1 << m, D))
    a = state_with(m, n, e, x)
    inverse_qft_circuit(lambda g, q, c: a.apply_gate(g, q) if c is None else a.apply_controlled_gate(g, c, q), m)
    ya = a.get()
    a.close()
    b = state_with(m, n, e, x)
    b.qift()
    yb = b.get()
    b.close()
    assert np.max(np.abs(ya - yb)) / np.max(np.abs(yb)) < 1e-13


def test_complex64_gates_track_complex128():
    m, n, e = 4, 5, 2
    D = math.comb(n, e) ** 2
    rng = np.random.default_rng(7)
    x = rng.normal(size=(1 << m, D)) + 1j * rng.normal(size=(1 << m, D))
    s64, s32 = state_with(m, n, e, x), state_with(m, n, e, x, qpesim.COMPLEX64)
    for sv in (s64, s32):
        sv.apply_gate(qpesim.hadamard(), 1)
        sv.apply_controlled_gate(qpesim.rotate(3), 0, 2)
    y64, y32 = s64.get(), s32.get()
    s64.close()
    s32.close()
    assert np.max(np.abs(y32 - y64)) / np.max(np.abs(y64)) < 3e-7


def test_gate_arguments_are_checked():
    sv = qpesim.StateVector(2, 3, 1)
    for call in (lambda: sv.apply_gate(qpesim.hadamard(), 2),
                 lambda: sv.apply_controlled_gate(qpesim.pauli_x(), 1, 1),
                 lambda: sv.apply_controlled_gate(qpesim.pauli_x(), 2, 0),
                 lambda: sv.apply_gate(np.array([[np.nan, 0], [0, 1]]), 0)):
        with pytest.raises(qpesim.QpeError) as ei:
            call()
        assert ei.value.code == qpesim.ERR_ARG
    sv.close()
