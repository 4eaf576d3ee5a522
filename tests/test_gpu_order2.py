"""Symmetric second-order Trotter steps and global-memory windows on the GPU.

The second-order step is no part of the reference (whose product is first
order); it is pinned by what the mathematics fixes: the symmetric product is
time-reversible, U2(-t) U2(t) = 1 (the first-order product is not), and its
error against exp(-iHT) falls like 1/K^2 in the number of steps K (first
order: 1/K), with H the textbook Jordan-Wigner Hamiltonian."""
import math

import numpy as np
import pytest
import scipy.linalg

import jw
import qpesim
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _ham(n, e, seed, t=None):
    one, two = qpesim.synthetic_integrals(n, seed)
    h, tau = qpesim.build_hamiltonian(n, e, one, two)
    if t is not None:
        h.set_timestep(t)
    return h, tau, one, two


def _step(h, n, e, psi, order, dtype=qpesim.COMPLEX128, tile_bytes=0):
    """block 1 = U psi (one Trotter step)"""
    sv = qpesim.StateVector(1, n, e, dtype)
    sv.set_option(qpesim.OPT_TROTTER_ORDER, order)
    sv.set_option(qpesim.OPT_RENORMALIZE, 0)
    if tile_bytes:
        sv.set_option(qpesim.OPT_MAX_TILE_BYTES, tile_bytes)
    sv.set(0, psi)
    qpesim.time_evolve_over_terms(sv, h)
    D = psi.size
    out = sv.get(D, D)
    sv.close()
    return out


@pytest.mark.parametrize("n,e,seed", [(6, 3, 5), (7, 2, 6)])
def test_second_order_step_is_time_reversible(n, e, seed):
    say(f"order2 reversibility n={n} e={e}")
    h, tau, _, _ = _ham(n, e, seed)
    t = 40 * tau                                     # a large step: first-order asymmetry is visible
    hf, _, _, _ = _ham(n, e, seed, t)
    hb, _, _, _ = _ham(n, e, seed, -t)
    psi = qpesim.initial_state(n, e, seed, 0.6)
    psi = psi / np.linalg.norm(psi)
    back2 = _step(hb, n, e, _step(hf, n, e, psi, 2), 2)
    assert np.linalg.norm(back2 - psi) < 1e-12
    back1 = _step(hb, n, e, _step(hf, n, e, psi, 1), 1)
    assert np.linalg.norm(back1 - psi) > 1e-6       # the check can fail: the first-order product is not symmetric


def test_second_order_converges_quadratically():
    n, e, seed = 4, 2, 7
    say("order2 convergence")
    one, two = qpesim.synthetic_integrals(n, seed)
    H = jw.textbook_hamiltonian(n, e, one, two)
    psi = qpesim.initial_state(n, e, seed, 0.7)
    psi = psi / np.linalg.norm(psi)
    T = 2.0
    exact = scipy.linalg.expm(-1j * T * H) @ psi
    err = {}
    for order in (1, 2):
        for K in (8, 16, 32):
            h, _, _, _ = _ham(n, e, seed, T / K)
            m = 6
            sv = qpesim.StateVector(m, n, e)
            sv.set_option(qpesim.OPT_TROTTER_ORDER, order)
            sv.set(0, psi)
            qpesim.time_evolve_over_terms(sv, h)
            D = psi.size
            err[(order, K)] = np.linalg.norm(sv.get(K * D, D) - exact)
            sv.close()
    r1 = [err[(1, K)] / err[(1, 2 * K)] for K in (8, 16)]
    r2 = [err[(2, K)] / err[(2, 2 * K)] for K in (8, 16)]
    say(f"first-order ratios {r1}, second-order ratios {r2}")
    assert all(1.7 < r < 2.3 for r in r1)
    assert all(3.5 < r < 4.5 for r in r2)
    assert err[(2, 32)] < err[(1, 32)]


@pytest.mark.parametrize("order", [1, 2])
def test_global_windows_match_lds_windows(order):
    """complex128 with a tile budget of one column: mixed double excitations
    run in global memory; every amplitude sees the same rotations in the same
    order, so the result equals the LDS-tiled plan's bit for bit."""
    n, e, seed, m = 8, 4, 9, 3
    say(f"global windows order={order}")
    h, _, _, _ = _ham(n, e, seed)
    psi = qpesim.initial_state(n, e, seed, 0.9)
    C = math.comb(n, e)
    res = []
    # one column plus the staged angles of the n(n-1)/2 excitation groups of a spin
    for budget in (0, C * 16 + (n * (n - 1) // 2) * 16 + 64):
        sv = qpesim.StateVector(m, n, e)
        sv.set_option(qpesim.OPT_TROTTER_ORDER, order)
        if budget:
            sv.set_option(qpesim.OPT_MAX_TILE_BYTES, budget)
        sv.set(0, psi / math.sqrt(1 << m))
        qpesim.time_evolve_over_terms(sv, h)
        res.append(sv.get())
        sv.close()
    assert np.array_equal(res[0], res[1])


def test_second_order_complex64_tracks_complex128():
    n, e, seed, m = 8, 3, 12, 5
    say("order2 c64 vs c128")
    h, _, _, _ = _ham(n, e, seed)
    psi = qpesim.initial_state(n, e, seed, 0.95)
    P = []
    for dt in (qpesim.COMPLEX128, qpesim.COMPLEX64):
        sv = qpesim.StateVector(m, n, e, dt)
        sv.set_option(qpesim.OPT_TROTTER_ORDER, 2)
        sv.set(0, psi / math.sqrt(1 << m))
        qpesim.time_evolve_over_terms(sv, h)
        sv.qift()
        P.append(sv.identify_phase_amplitudes())
        sv.close()
    assert np.max(np.abs(P[0] - P[1])) < 1e-6
