"""The dataflow evolution (one launch per call; items wait for the items whose
writes they read) against the pass-by-pass launches: both apply the same ops
to the same tiles in the same order per column with the same rounding dither,
so the registers, block norms and P(k) must be identical bit for bit, in both
precisions, both Trotter orders and both term orders, for a whole chain and for
a chain split over several calls.  A wait limit of zero makes the first
unsatisfied wait stop the launch (QPE_ERR_STALLED), after which the state
evolves normally again.  The parity of either path with the oracle is covered
by the other GPU tests, which run the default (dataflow where the plan allows
it)."""
import math
import sys

import numpy as np
import pytest

import qpesim

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(n, e, m, seed, dtype, execution, order=1, term_order=0, split=None, tile=0, wait=None):
    one, two = qpesim.synthetic_integrals(n, seed)
    h, tau = qpesim.build_hamiltonian(n, e, one, two)
    psi0 = qpesim.initial_state(n, e, seed, 0.9)
    M = 1 << m
    sv = qpesim.StateVector(m, n, e, dtype)
    sv.set_option(qpesim.OPT_EXECUTION, execution)
    if order != 1:
        sv.set_option(qpesim.OPT_TROTTER_ORDER, order)
    if term_order:
        sv.set_option(qpesim.OPT_TERM_ORDER, term_order)
    if tile:
        sv.set_option(qpesim.OPT_MAX_TILE_BYTES, tile)
    if wait is not None:
        sv.set_option(qpesim.OPT_WAIT_LIMIT_MS, wait)
    sv.set(0, psi0)
    sv.scale_block(0, 1.0 / math.sqrt(M))
    if split is None:
        qpesim.time_evolve_over_terms(sv, h)
    else:
        qpesim.time_evolve_blocks(sv, h, 1, split)
        qpesim.time_evolve_blocks(sv, h, split, M)
    launches = sv.timings()["launches_per_step"]
    norms = sv.block_norms()
    sv.qift()
    y = sv.get()
    P = sv.identify_phase_amplitudes()
    sv.close()
    return y, norms, P, launches


CASES = [
    # n, e, m, dtype, Trotter order, term order, tile budget
    (6, 3, 4, qpesim.COMPLEX128, 1, 0, 0),
    (6, 3, 4, qpesim.COMPLEX64, 1, 0, 0),
    (8, 3, 3, qpesim.COMPLEX128, 2, 0, 0),
    (8, 4, 3, qpesim.COMPLEX64, 1, 1, 0),
    (10, 4, 2, qpesim.COMPLEX64, 1, 0, 6 * math.comb(10, 4) * 8),   # narrow tiles: many windows
    (16, 4, 2, qpesim.COMPLEX64, 1, 0, 0),    # the 32 spin-orbital / 8 electron block
    (14, 5, 2, qpesim.COMPLEX128, 1, 0, 0),   # the 28 spin-orbital / 10 electron block
]


@pytest.mark.parametrize("n,e,m,dtype,order,term_order,tile", CASES)
def test_dataflow_equals_pass_launches(n, e, m, dtype, order, term_order, tile):
    say(f"dataflow vs passes n={n} e={e} m={m} dtype={dtype} order={order} term_order={term_order} tile={tile}")
    y0, n0, p0, l0 = run(n, e, m, 3 + n, dtype, qpesim.EXEC_PASSES, order, term_order, tile=tile)
    y1, n1, p1, l1 = run(n, e, m, 3 + n, dtype, qpesim.EXEC_DATAFLOW, order, term_order, tile=tile)
    assert l1 == 2 and l0 > 2, (l0, l1)
    assert np.array_equal(n0, n1)
    assert np.array_equal(y0, y1)
    assert np.array_equal(p0, p1)


def test_dataflow_chain_split_over_calls():
    say("dataflow chain split over calls n=8 e=3 m=4")
    y0, n0, p0, _ = run(8, 3, 4, 5, qpesim.COMPLEX64, qpesim.EXEC_DATAFLOW)
    y1, n1, p1, _ = run(8, 3, 4, 5, qpesim.COMPLEX64, qpesim.EXEC_DATAFLOW, split=7)
    y2, n2, p2, _ = run(8, 3, 4, 5, qpesim.COMPLEX64, qpesim.EXEC_PASSES, split=11)
    assert np.array_equal(n0, n1) and np.array_equal(n0, n2)
    assert np.array_equal(y0, y1) and np.array_equal(y0, y2)
    assert np.array_equal(p0, p1)


def test_dataflow_stall_stops_the_launch():
    say("dataflow stall: zero wait limit")
    n, e, m = 8, 3, 4
    one, two = qpesim.synthetic_integrals(n, 9)
    h, _ = qpesim.build_hamiltonian(n, e, one, two)
    psi0 = qpesim.initial_state(n, e, 9, 0.9)
    sv = qpesim.StateVector(m, n, e, qpesim.COMPLEX128)
    sv.set_option(qpesim.OPT_EXECUTION, qpesim.EXEC_DATAFLOW)
    sv.set_option(qpesim.OPT_WAIT_LIMIT_MS, 0)
    sv.set(0, psi0)
    with pytest.raises(qpesim.QpeError) as ei:
        qpesim.time_evolve_over_terms(sv, h)
    assert ei.value.code == qpesim.ERR_STALLED
    # the same state evolves normally afterwards, identical to a fresh run
    sv.set_option(qpesim.OPT_WAIT_LIMIT_MS, 10000)
    sv.set(0, psi0)
    qpesim.time_evolve_over_terms(sv, h)
    a = sv.get()
    sv.close()
    ref = qpesim.StateVector(m, n, e, qpesim.COMPLEX128)
    ref.set_option(qpesim.OPT_EXECUTION, qpesim.EXEC_PASSES)
    ref.set(0, psi0)
    qpesim.time_evolve_over_terms(ref, h)
    b = ref.get()
    ref.close()
    assert np.array_equal(a, b)


def test_dataflow_refused_for_global_windows():
    say("dataflow refused where a pass keeps its windows in global memory")
    n, e, m = 8, 3, 1
    one, two = qpesim.synthetic_integrals(n, 2)
    h, _ = qpesim.build_hamiltonian(n, e, one, two)
    sv = qpesim.StateVector(m, n, e, qpesim.COMPLEX128)
    # a tile of one column and the alpha groups' angle tables: the mixed windows
    # (two columns at least) rotate in global memory
    sv.set_option(qpesim.OPT_MAX_TILE_BYTES, math.comb(n, e) * 16 + math.comb(n, 2) * 16 + 64)
    sv.set_option(qpesim.OPT_EXECUTION, qpesim.EXEC_DATAFLOW)
    sv.set(0, qpesim.initial_state(n, e, 2, 1.0))
    with pytest.raises(qpesim.QpeError) as ei:
        qpesim.time_evolve_over_terms(sv, h)
    assert ei.value.code == qpesim.ERR_ARG
    sv.set_option(qpesim.OPT_EXECUTION, qpesim.EXEC_AUTO)   # falls back to pass launches
    qpesim.time_evolve_over_terms(sv, h)
    assert sv.timings()["launches_per_step"] > 2
    sv.close()
