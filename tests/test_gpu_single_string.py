"""Single-string (interleaved Jordan-Wigner) register on the GPU.

* one Trotter step pinned from first principles: the product of exp(-i t G)
  over the reference's term sequence (applyUnitary_reordered,
  cpp/time_evolution_over_terms.hpp:374-562), every generator G built in the
  interleaved mode order on the single-string basis (tests/jw.py), FP64 within
  1e-12 relative L2, and P(k) within 1e-12 of the oracle's;
* against the alpha x beta register on the same state: the evolved amplitudes
  are sign * (alpha x beta amplitudes) under qpesim.single_string_map bit for
  bit, the Sz != 0 entries stay exactly zero, and P(k) agrees within 1e-12,
  for both dtypes, at n=16 / N=8 and at the n=32 / N=8 workload's full size;
* a block 0 with amplitudes outside the (e alpha, e beta) sector is rejected."""
import math

import numpy as np
import pytest
from scipy.linalg import expm

import jw
import oracle
import qpesim
from test_gpu_parity import PK_TOL, oracle_run, rel_l2, say

pytestmark = pytest.mark.gpu


def _reference_order(T):
    """(cls, idx, coeff) in applyUnitary_reordered order (TEOT:374-562)."""
    L = T.as_lists()
    seq = [(0, k, c) for k, c in L[0]] + [(2, k, c) for k, c in L[2]]
    for (p, q), c in L[1]:
        seq.append((1, (p, q), c))
        for k, c2 in L[3]:
            if k[0] == q and k[3] == p:
                seq.append((3, k, c2))
    seq += [(4, k, c) for k, c in L[4]]
    return seq


def _to_single(psi_ab, n, e):
    idx, sg = qpesim.single_string_map(n, e)
    out = np.zeros(math.comb(2 * n, 2 * e), np.complex128)
    out[idx] = sg * psi_ab
    return out


def _run(n, e, m, seed, dtype, layout, hf=1.0, split=None, opts=()):
    one, two = qpesim.synthetic_integrals(n, seed)
    h, tau = qpesim.build_hamiltonian(n, e, one, two)
    psi = qpesim.initial_state(n, e, seed, hf)
    if layout == qpesim.LAYOUT_SINGLE_STRING:
        psi = _to_single(psi, n, e)
    M = 1 << m
    sv = qpesim.StateVector(m, n, e, dtype, layout)
    for opt, value in opts:
        sv.set_option(opt, value)
    sv.set(0, psi / math.sqrt(M))
    if split:
        qpesim.time_evolve_blocks(sv, h, 1, split)
        qpesim.time_evolve_blocks(sv, h, split, M)
    else:
        qpesim.time_evolve_over_terms(sv, h)
    amps = sv.get().reshape(M, -1)
    sv.qift()
    P = sv.identify_phase_amplitudes()
    sv.close()
    return amps, P, tau


def test_single_string_step_from_first_principles():
    n, e, m, seed = 4, 2, 3, 11
    say(f"single-string n={2 * n} N={2 * e} m={m}: dense interleaved exponentials")
    amps, P, tau = _run(n, e, m, seed, qpesim.COMPLEX128, qpesim.LAYOUT_SINGLE_STRING, hf=0.6)
    one, two = qpesim.synthetic_integrals(n, seed)
    T = oracle.compile_terms(n, one, two)
    dim = math.comb(2 * n, 2 * e)
    U = np.eye(dim, dtype=np.complex128)
    for cls, idx, c in _reference_order(T):
        U = expm(-1j * tau * jw.term_generator(cls, list(idx), c, n, e, interleaved=True)) @ U
    v = amps[0].copy()
    for j in range(1, 1 << m):
        v = U @ v
        err = rel_l2(amps[j], v)
        assert err < 1e-12, (j, err)
    _, _, rP = oracle_run(n, e, m, seed, hf=0.6)
    dP = float(np.max(np.abs(P - rP)))
    say(f"  max|dP| vs oracle {dP:.3e}")
    assert dP < PK_TOL


def _compare_layouts(n, e, m, seed, dtype, hf, split=None, opts=(), exact=True):
    a2, P2, _ = _run(n, e, m, seed, dtype, qpesim.LAYOUT_ALPHA_BETA, hf, opts=opts)
    a1, P1, _ = _run(n, e, m, seed, dtype, qpesim.LAYOUT_SINGLE_STRING, hf, split, opts=opts)
    idx, sg = qpesim.single_string_map(n, e)
    if exact:
        assert np.array_equal(a1[:, idx], sg[None, :] * a2)
    else:
        assert rel_l2(a1[:, idx], sg[None, :] * a2) < 1e-14
    rest = np.ones(a1.shape[1], bool)
    rest[idx] = False
    assert not np.any(a1[:, rest])
    dP = float(np.max(np.abs(P1 - P2)))
    say(f"  max|dP| single-string vs alpha x beta {dP:.3e}")
    assert dP < PK_TOL


@pytest.mark.parametrize("dtype", [qpesim.COMPLEX128, qpesim.COMPLEX64])
def test_single_string_equals_alpha_beta(dtype):
    n, e, m, seed = 8, 4, 5, 12
    say(f"single-string vs alpha x beta n={2 * n} N={2 * e} m={m} dtype={dtype}")
    _compare_layouts(n, e, m, seed, dtype, 0.8, split=9)


@pytest.mark.parametrize("opts", [((qpesim.OPT_TROTTER_ORDER, 2),), ((qpesim.OPT_TERM_ORDER, 1),),
                                  ((qpesim.OPT_ENGINE, 1),), ((qpesim.OPT_RENORMALIZE, 1),)])
def test_single_string_with_step_options(opts):
    """Same passes, so the same amplitudes bit for bit; with renormalisation in
    complex128 block 0's norm is summed over the longer single-string block in another
    order (products of FP64 amplitudes are rounded), so there the bar is 1e-14."""
    n, e, m, seed = 7, 3, 4, 15
    say(f"single-string vs alpha x beta n={2 * n} N={2 * e} m={m} options {opts}")
    renorm = any(o == qpesim.OPT_RENORMALIZE for o, _ in opts)
    _compare_layouts(n, e, m, seed, qpesim.COMPLEX128, 0.7, opts=opts, exact=not renorm)


def test_single_string_n32_full_blocks():
    """The n=32 / N=8 workload's blocks (C(32,8) = 10,518,300 single-string
    amplitudes, C = 1820) in complex64, m = 3."""
    n, e, m, seed = 16, 4, 3, 2
    say(f"single-string vs alpha x beta n=32 N=8 complex64 m={m}")
    _compare_layouts(n, e, m, seed, qpesim.COMPLEX64, 1.0)


def test_single_string_rejects_other_sectors():
    n, e, m = 4, 2, 2
    one, two = qpesim.synthetic_integrals(n, 13)
    h, _ = qpesim.build_hamiltonian(n, e, one, two)
    sv = qpesim.StateVector(m, n, e, qpesim.COMPLEX128, qpesim.LAYOUT_SINGLE_STRING)
    idx, _ = qpesim.single_string_map(n, e)
    psi = np.zeros(math.comb(2 * n, 2 * e), np.complex128)
    psi[idx[0]] = 1.0
    outside = int(np.setdiff1d(np.arange(psi.size), idx)[0])
    psi[outside] = 0.5
    sv.set(0, psi)
    with pytest.raises(qpesim.QpeError) as ei:
        qpesim.time_evolve_over_terms(sv, h)
    assert ei.value.code == qpesim.ERR_ARG
    with pytest.raises(qpesim.QpeError) as ei:
        qpesim.time_evolve_iqpe(sv, h, 1)
    assert ei.value.code == qpesim.ERR_ARG
    psi[outside] = 0.0
    sv.set(0, psi)
    qpesim.time_evolve_over_terms(sv, h)
    sv.close()
