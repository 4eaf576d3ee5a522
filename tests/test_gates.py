"""Ancilla gates on the CPU: the oracle's restatement of apply_gate /
apply_controlled_gate (cpp/state_vector.hpp:51-113) against the reference's own
gate tests (cpp/state_vector.hpp:383-455, restated) and against the
mathematics (a gate circuit equals the discrete Fourier transform)."""
import math

import numpy as np
import pytest

import gates_oracle as go
import qpesim


def basis_state(m, D, index):
    s = np.zeros((1 << m) * D, np.complex128)
    s[index] = 1.0
    return s.reshape(1 << m, D)


def test_reference_x_and_hadamard_cases():
    """cpp/state_vector.hpp:383-418: StateVector(m, 2, 1) has D = 4 amplitudes per block."""
    X = qpesim.pauli_x()
    s = go.apply_gate(basis_state(1, 4, 0), 1, X, 0)
    assert s.ravel()[4] == 1.0
    s = go.apply_gate(basis_state(2, 4, 0), 2, X, 0)
    assert s.ravel()[8] == 1.0
    s = go.apply_gate(basis_state(1, 4, 0), 1, qpesim.hadamard(), 0).ravel()
    assert s[0] == 1.0 / 2 ** 0.5 and s[4] == 1.0 / 2 ** 0.5
    s = go.apply_gate(basis_state(2, 4, 0), 2, qpesim.hadamard(), 0).ravel()
    assert s[0] == 1.0 / 2 ** 0.5 and s[8] == 1.0 / 2 ** 0.5


def test_reference_controlled_cases():
    """cpp/state_vector.hpp:420-455."""
    X = qpesim.pauli_x()
    s = go.apply_gate(basis_state(2, 4, 0), 2, X, 0, control=1)
    assert s.ravel()[0] == 1.0
    s = go.apply_gate(basis_state(2, 4, 4), 2, X, 0, control=1)
    assert s.ravel()[12] == 1.0
    s = go.apply_gate(s, 2, X, 1, control=0)
    assert s.ravel()[8] == 1.0
    s = go.apply_gate(basis_state(3, 4, 4), 3, X, 0, control=2)
    assert s.ravel()[20] == 1.0
    s = go.apply_gate(s, 3, X, 2, control=0)
    assert s.ravel()[16] == 1.0


def inverse_qft_circuit(apply, m):
    """The textbook QFT circuit with conjugated phases: for q = 0..m-1 (qubit 0
    the most significant bit of the block index), H on q, then controlled
    InvRotate(r - q + 1) from each r > q; then the qubit order reversed by
    swaps of three controlled X.  It maps block j to (1/sqrt M) sum_k
    e^{-2 pi i jk/M} block k."""
    X = qpesim.pauli_x()
    for q in range(m):
        apply(qpesim.hadamard(), q, None)
        for r in range(q + 1, m):
            apply(qpesim.inv_rotate(r - q + 1), q, r)
    for q in range(m // 2):
        a, b = q, m - 1 - q
        apply(X, b, a)
        apply(X, a, b)
        apply(X, b, a)


@pytest.mark.parametrize("m", [1, 2, 3, 5])
def test_gate_circuit_is_the_dft(m):
    rng = np.random.default_rng(m)
    D = 3
    x = rng.normal(size=(1 << m, D)) + 1j * rng.normal(size=(1 << m, D))
    s = x.copy()
    inverse_qft_circuit(lambda g, q, c: go.apply_gate(s, m, g, q, control=c), m)
    want = np.fft.fft(x, axis=0) / math.sqrt(1 << m)     # e^{-2 pi i jk/M}, natural order
    assert np.max(np.abs(s - want)) < 1e-13


def test_hadamard_squared_and_rotations():
    m, D = 3, 2
    rng = np.random.default_rng(9)
    x = rng.normal(size=(1 << m, D)) + 1j * rng.normal(size=(1 << m, D))
    s = x.copy()
    for q in range(m):
        go.apply_gate(s, m, qpesim.hadamard(), q)
        go.apply_gate(s, m, qpesim.hadamard(), q)
    assert np.max(np.abs(s - x)) < 1e-14
    go.apply_gate(s, m, qpesim.rotate(3), 1, control=0)
    go.apply_gate(s, m, qpesim.inv_rotate(3), 1, control=0)
    assert np.max(np.abs(s - x)) < 1e-14
