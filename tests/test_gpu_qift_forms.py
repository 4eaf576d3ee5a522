"""The QIFT kernels against numpy's FFT (the reference's DftiComputeForward with
scale 1/sqrt(2^m), cpp/state_vector.hpp:133-168) on the same random registers:

* in-register DFTs (2^m = 2..1024), the DFT as a matrix product on the matrix
  cores (2^m = 16..64) and radix-2 stages in LDS (any 2^m);
* block sizes with a partial last column tile;
* P(k) = sum over the block of |Y[k]|^2 from the same launch.

Bars: complex128 1e-12 relative L2 and 1e-12 absolute on P(k); complex64 (the
DFTs in FP32) 2e-6 relative L2 and 2e-6 absolute on P(k).  A QPE with
renormalisation (complex64) gives the same P(k) through every form."""
import math

import numpy as np
import pytest

import qpesim
from test_gpu_parity import oracle_run, rel_l2, say

pytestmark = pytest.mark.gpu

FORMS = {qpesim.QIFT_REGISTER: range(1, 13), qpesim.QIFT_MFMA: range(4, 7), qpesim.QIFT_LDS_RADIX2: range(1, 13)}


def _case(m, n, e, dtype, form, seed):
    rng = np.random.default_rng(seed)
    M = 1 << m
    C = math.comb(n, e)
    X = (rng.normal(size=(M, C * C)) + 1j * rng.normal(size=(M, C * C))) / math.sqrt(M * C * C)
    sv = qpesim.StateVector(m, n, e, dtype)
    sv.set_option(qpesim.OPT_QIFT_FORM, form)
    sv.set(0, X.reshape(-1))
    X = sv.get().reshape(M, -1)   # what the register holds (rounded to complex64 there)
    sv.qift()
    Y = sv.get().reshape(M, -1)
    P = sv.identify_phase_amplitudes()
    sv.close()
    Yr = np.fft.fft(X, axis=0) / math.sqrt(M)
    Pr = np.sum(np.abs(Yr) ** 2, axis=1)
    return rel_l2(Y, Yr), float(np.max(np.abs(P - Pr)))


@pytest.mark.parametrize("dtype", [qpesim.COMPLEX128, qpesim.COMPLEX64])
@pytest.mark.parametrize("form", sorted(FORMS))
def test_qift_forms_against_numpy(form, dtype):
    tol = 1e-12 if dtype == qpesim.COMPLEX128 else 2e-6
    worst = (0.0, 0.0)
    for m in FORMS[form]:
        for n, e in ((4, 2), (8, 3)):       # 36 and 3136 columns: partial last tiles
            l2, dp = _case(m, n, e, dtype, form, 100 * m + n)
            assert l2 < tol and dp < tol, (m, n, e, l2, dp)
            worst = (max(worst[0], l2), max(worst[1], dp))
    say(f"form {form} dtype {dtype}: worst L2 {worst[0]:.2e}, max|dP| {worst[1]:.2e}")


@pytest.mark.parametrize("dtype", [qpesim.COMPLEX128, qpesim.COMPLEX64])
def test_register_form_up_to_2_to_the_20(dtype):
    """2^m > 1024: two launches of the register kernel (four-step split) and a second
    register; 36 determinant columns so that numpy holds the register."""
    tol = 1e-12 if dtype == qpesim.COMPLEX128 else 2e-6
    for m in (13, 16, 17, 20):
        l2, dp = _case(m, 4, 2, dtype, qpesim.QIFT_AUTO, 7 * m)
        say(f"m={m} dtype {dtype}: L2 {l2:.2e}, max|dP| {dp:.2e}")
        assert l2 < tol and dp < tol, (m, l2, dp)


def test_large_m_qpe_against_oracle():
    """A whole FP64 QPE at m = 13 (8191 Trotter steps of a small instance): the QIFT and
    P(k) of the evolved register against numpy within 1e-12, the evolved register and
    P(k) against the oracle's within 1e-11 (8191 steps accumulate the rounding
    differences of the fused rotations against the oracle's per-term loop: measured
    1.5e-12 on P(k))."""
    n, e, m, seed = 4, 2, 13, 9
    M = 1 << m
    one, two = qpesim.synthetic_integrals(n, seed)
    h, _ = qpesim.build_hamiltonian(n, e, one, two)
    psi = qpesim.initial_state(n, e, seed, 0.8)
    sv = qpesim.StateVector(m, n, e, qpesim.COMPLEX128)
    sv.set(0, psi / math.sqrt(M))
    qpesim.time_evolve_over_terms(sv, h)
    X = sv.get().reshape(M, -1)
    sv.qift()
    Y = sv.get().reshape(M, -1)
    P = sv.identify_phase_amplitudes()
    sv.close()
    Yr = np.fft.fft(X, axis=0) / math.sqrt(M)
    assert rel_l2(Y, Yr) < 1e-12
    assert float(np.max(np.abs(P - np.sum(np.abs(Yr) ** 2, axis=1)))) < 1e-12
    ra, _, rP = oracle_run(n, e, m, seed, hf=0.8)
    l2 = rel_l2(X.reshape(-1), ra.reshape(-1))
    dp = float(np.max(np.abs(P - rP)))
    say(f"m=13 QPE: evolved register L2 vs oracle {l2:.2e}, max|dP| vs oracle {dp:.2e}")
    assert l2 < 1e-11 and dp < 1e-11


def test_matrix_core_form_in_a_renormalised_qpe():
    n, e, seed = 6, 3, 5
    one, two = qpesim.synthetic_integrals(n, seed)
    h, _ = qpesim.build_hamiltonian(n, e, one, two)
    psi = qpesim.initial_state(n, e, seed, 0.9)
    for m in (4, 5, 6):
        P = {}
        for form in (qpesim.QIFT_REGISTER, qpesim.QIFT_MFMA):
            sv = qpesim.StateVector(m, n, e, qpesim.COMPLEX64)
            sv.set_option(qpesim.OPT_QIFT_FORM, form)
            sv.set(0, psi / math.sqrt(1 << m))
            qpesim.time_evolve_over_terms(sv, h)
            sv.qift()
            P[form] = sv.identify_phase_amplitudes()
            sv.close()
        dp = float(np.max(np.abs(P[qpesim.QIFT_REGISTER] - P[qpesim.QIFT_MFMA])))
        say(f"m={m}: max|dP| matrix cores vs register DFTs {dp:.2e}")
        assert dp < 2e-6


def test_qift_form_must_cover_the_size():
    sv = qpesim.StateVector(3, 4, 2, qpesim.COMPLEX64)
    with pytest.raises(qpesim.QpeError) as ei:
        sv.set_option(qpesim.OPT_QIFT_FORM, qpesim.QIFT_MFMA)
    assert ei.value.code == qpesim.ERR_ARG
    with pytest.raises(qpesim.QpeError) as ei:
        sv.set_option(qpesim.OPT_QIFT_FORM, 7)
    assert ei.value.code == qpesim.ERR_ARG
    sv.close()
