"""Resumable pieces of the evolution chain (qpe_time_evolve_blocks) on the GPU."""
import math

import numpy as np
import pytest

import qpesim
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [qpesim.COMPLEX128, qpesim.COMPLEX64])
def test_pieces_equal_the_whole_chain(dtype):
    n, e, seed, m = 7, 3, 21, 5
    say(f"chain pieces dtype={dtype}")
    one, two = qpesim.synthetic_integrals(n, seed)
    h, _ = qpesim.build_hamiltonian(n, e, one, two)
    psi = qpesim.initial_state(n, e, seed, 0.9) / math.sqrt(1 << m)
    sv = qpesim.StateVector(m, n, e, dtype)
    sv.set(0, psi)
    qpesim.time_evolve_over_terms(sv, h)
    whole = sv.get()
    sv.set(0, psi)
    for a, b in ((1, 4), (4, 5), (5, 17), (17, 32)):
        qpesim.time_evolve_blocks(sv, h, a, b)
    assert np.array_equal(sv.get(), whole)
    sv.qift()
    P = sv.identify_phase_amplitudes()
    assert abs(P.sum() - 1.0) < 1e-6
    sv.close()


def test_piece_must_continue_the_chain():
    n, e, seed, m = 5, 2, 22, 3
    one, two = qpesim.synthetic_integrals(n, seed)
    h, _ = qpesim.build_hamiltonian(n, e, one, two)
    sv = qpesim.StateVector(m, n, e)
    sv.set(0, qpesim.initial_state(n, e, seed, 1.0))
    with pytest.raises(qpesim.QpeError) as ei:
        qpesim.time_evolve_blocks(sv, h, 3, 5)          # blocks 1, 2 never evolved
    assert ei.value.code == qpesim.ERR_STATE
    qpesim.time_evolve_blocks(sv, h, 1, 3)
    with pytest.raises(qpesim.QpeError) as ei:
        qpesim.time_evolve_blocks(sv, h, 4, 6)          # gap at block 3
    assert ei.value.code == qpesim.ERR_STATE
    with pytest.raises(qpesim.QpeError) as ei:
        qpesim.time_evolve_blocks(sv, h, 3, 9)          # beyond 2^m
    assert ei.value.code == qpesim.ERR_ARG
    qpesim.time_evolve_blocks(sv, h, 3, 8)
    sv.close()
