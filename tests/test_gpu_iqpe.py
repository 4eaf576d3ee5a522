"""applyUnitary's term order and time_evolve_iqpe on the GPU through the C ABI
(cpp/time_evolution_over_terms.hpp:141-155, 174-372), against the oracle's
restatement, which the CPU tests pin to products of dense matrix exponentials
in that order.

applyUnitary differs from the reordered step time_evolve_over_terms runs: pp,
then every pq, then pqqp, then every pqqr term on its own (also those without a
pq term), then pqrs.  time_evolve_iqpe applies it k times in place to block 1."""
import math

import numpy as np
import pytest

import oracle
import qpesim
from test_gpu_parity import rel_l2, say

pytestmark = pytest.mark.gpu


def integrals(n, seed, drop_pairs=()):
    """Synthetic integrals; drop_pairs removes one-body h_pq so that the pqqr
    terms of that (p, q) have no pq partner (applied by applyUnitary only)."""
    one, two = qpesim.synthetic_integrals(n, seed)
    one = [x for x in one if (x[0], x[1]) not in drop_pairs and (x[1], x[0]) not in drop_pairs]
    return one, two


CHAIN_CASES = [
    # n, e, m, seed, dropped one-body pairs, tile columns (0: device maximum)
    (4, 2, 4, 31, ((3, 1),), 0),
    (5, 2, 4, 8, ((4, 2), (5, 1)), 0),
    (6, 3, 5, 0, (), 0),
    (7, 3, 4, 10, ((5, 2),), 0),
    (8, 1, 4, 11, ((6, 3),), 0),          # one electron per spin: same-spin pqqr terms act on nothing
    (7, 3, 4, 14, ((4, 1),), 3),          # small tiles: many column windows
]


@pytest.mark.parametrize("n,e,m,seed,drop,tile_cols", CHAIN_CASES)
def test_applyunitary_order_chain_matches_oracle(n, e, m, seed, drop, tile_cols):
    say(f"applyUnitary chain n={n} e={e} m={m} drop={drop}")
    one, two = integrals(n, seed, drop)
    h, tau = qpesim.build_hamiltonian(n, e, one, two)
    psi = qpesim.initial_state(n, e, seed, 0.9)
    C = math.comb(n, e)
    M = 1 << m
    sv = qpesim.StateVector(m, n, e)
    sv.set_option(qpesim.OPT_TERM_ORDER, 1)
    if tile_cols:
        sv.set_option(qpesim.OPT_MAX_TILE_BYTES, tile_cols * C * 16)
    sv.set(0, psi)
    sv.scale_block(0, 1 / math.sqrt(M))
    qpesim.time_evolve_over_terms(sv, h)
    amps = sv.get()
    sv.qift()
    P = sv.identify_phase_amplitudes()
    sv.close()
    T = oracle.compile_terms(n, one, two)
    st = np.zeros((M, C * C), np.complex128)
    st[0] = psi / math.sqrt(M)
    oracle.time_evolve(n, e, m, T, tau, st, original=True)
    ra = st.ravel().copy()
    oracle.qift(m, st)
    rP = oracle.phase_probabilities(m, st)
    assert rel_l2(amps, ra) < 1e-12
    assert np.max(np.abs(P - rP)) < 1e-12
    # and it is not the reordered product
    st2 = np.zeros((M, C * C), np.complex128)
    st2[0] = psi / math.sqrt(M)
    oracle.time_evolve(n, e, m, T, tau, st2)
    assert rel_l2(st2.ravel(), ra) > 1e-9


@pytest.mark.parametrize("n,e,k,seed,drop", [(4, 2, 5, 31, ((3, 1),)), (6, 3, 3, 2, ((5, 3),)), (7, 2, 4, 3, ())])
def test_iqpe_matches_oracle_in_place_on_block_1(n, e, k, seed, drop):
    say(f"iqpe n={n} e={e} k={k}")
    one, two = integrals(n, seed, drop)
    h, tau = qpesim.build_hamiltonian(n, e, one, two)
    t = 8 * tau
    h.set_timestep(t)
    C = math.comb(n, e)
    D = C * C
    rng = np.random.default_rng(seed)
    psi = rng.normal(size=D) + 1j * rng.normal(size=D)
    psi /= np.linalg.norm(psi)
    other = rng.normal(size=D) + 1j * rng.normal(size=D)
    m = 2
    sv = qpesim.StateVector(m, n, e)
    sv.set(0, other)
    sv.set(D, psi)
    sv.set(2 * D, 2 * other)
    qpesim.time_evolve_iqpe(sv, h, k)
    out = sv.get()
    norms = sv.block_norms()
    sv.close()
    ref = psi.copy()
    oracle.time_evolve_iqpe(n, e, oracle.compile_terms(n, one, two), t, ref, k)
    assert rel_l2(out[D:2 * D], ref) < 1e-12
    assert np.array_equal(out[:D], other) and np.array_equal(out[2 * D:3 * D], 2 * other)
    assert np.all(out[3 * D:] == 0)
    assert abs(norms[1] - 1.0) < 1e-13


def test_iqpe_complex64_and_zero_applications():
    n, e, k, seed = 6, 3, 6, 4
    say("iqpe complex64")
    one, two = integrals(n, seed, ((4, 2),))
    h, tau = qpesim.build_hamiltonian(n, e, one, two)
    h.set_timestep(5 * tau)
    D = math.comb(n, e) ** 2
    psi = qpesim.initial_state(n, e, seed, 0.7)
    sv = qpesim.StateVector(1, n, e, qpesim.COMPLEX64)
    sv.set(D, psi)
    qpesim.time_evolve_iqpe(sv, h, 0)
    assert rel_l2(sv.get(D, D), psi) < 1e-7           # zero applications: only the FP32 rounding of psi
    qpesim.time_evolve_iqpe(sv, h, k)
    got = sv.get(D, D)
    sv.close()
    ref = psi.copy()
    oracle.time_evolve_iqpe(n, e, oracle.compile_terms(n, one, two), 5 * tau, ref, k)
    assert rel_l2(got, ref) < 2e-6


def test_iqpe_ends_a_chain_and_options_are_checked():
    n, e = 5, 2
    one, two = qpesim.synthetic_integrals(n, 5)
    h, _ = qpesim.build_hamiltonian(n, e, one, two)
    sv = qpesim.StateVector(2, n, e)
    sv.set(0, qpesim.initial_state(n, e))
    qpesim.time_evolve_blocks(sv, h, 1, 2)
    qpesim.time_evolve_iqpe(sv, h, 1)
    with pytest.raises(qpesim.QpeError) as ei:
        qpesim.time_evolve_blocks(sv, h, 2, 4)       # block 1 no longer continues the chain
    assert ei.value.code == qpesim.ERR_STATE
    with pytest.raises(qpesim.QpeError) as ei:
        sv.set_option(qpesim.OPT_TERM_ORDER, 2)
    assert ei.value.code == qpesim.ERR_ARG
    sv.close()


def test_applyunitary_order_batch_equals_single_runs():
    say("applyUnitary order batched")
    n, e, m = 6, 3, 3
    states, hams, singles = [], [], []
    for seed in (40, 41, 42):
        one, two = integrals(n, seed, ((5, 2),))
        h, _ = qpesim.build_hamiltonian(n, e, one, two)
        psi = qpesim.initial_state(n, e, seed, 0.8)
        a = qpesim.StateVector(m, n, e)
        b = qpesim.StateVector(m, n, e)
        for sv in (a, b):
            sv.set_option(qpesim.OPT_TERM_ORDER, 1)
            sv.set(0, psi)
        qpesim.time_evolve_over_terms(a, h)
        singles.append(a.get())
        a.close()
        states.append(b)
        hams.append(h)
    qpesim.batch_time_evolve(states, hams)
    for sv, ref in zip(states, singles):
        assert np.array_equal(sv.get(), ref)
        sv.close()


@pytest.mark.parametrize("n,e,seed", [(6, 3, 5)])
def test_applyunitary_order_second_order_is_time_reversible(n, e, seed):
    say("applyUnitary order, second order reversibility")
    one, two = integrals(n, seed, ((4, 1),))
    h, tau = qpesim.build_hamiltonian(n, e, one, two)
    D = math.comb(n, e) ** 2
    psi = qpesim.initial_state(n, e, seed, 0.6)

    def step(t, x):
        h.set_timestep(t)
        sv = qpesim.StateVector(1, n, e)
        sv.set_option(qpesim.OPT_TERM_ORDER, 1)
        sv.set_option(qpesim.OPT_TROTTER_ORDER, 2)
        sv.set(0, x)
        qpesim.time_evolve_over_terms(sv, h)
        y = sv.get(D, D)
        sv.close()
        return y

    back = step(-40 * tau, step(40 * tau, psi))
    assert np.linalg.norm(back - psi) < 1e-12
