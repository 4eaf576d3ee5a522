"""The benchmark workloads themselves on the GPU, at their real block sizes,
with the ancilla count cut so that each test stays within seconds to a minute.

* n=32 spin-orbitals / N=8 (C = 1820) in the FP32 precision form: against the
  FP64 oracle (m=3) and against the FP64 GPU path over 63 steps (m=6);
* the 48-instance batching form (n=12..24, N=4..8) at m=2: batched FP32 equals
  single-instance FP32 bit for bit, the smallest and largest instances' FP64
  P(k) equal the oracle's, and every instance's FP32 P(k) is within 1e-5 of its
  FP64 P(k);
* n=34 / N=10 (C = 6188) with second-order steps: FP32 (LDS windows) tracks the
  complex128 run (global-memory windows), and U2(-t) U2(t) = 1 in FP64.

Tolerances: FP64 P(k) 1e-12 absolute; FP32 form max|dP(k)| <= 1e-5 against FP64."""
import math

import numpy as np
import pytest

import qpesim
from test_gpu_parity import PK_TOL, PK_TOL_F32, gpu_run, oracle_run, rel_l2, say

pytestmark = pytest.mark.gpu


# ---------------------------------------------------------------- n=32, N=8 (C=1820), FP32 form
def test_n32_fp32_form_against_oracle():
    n, e, seed, m = 16, 4, 2, 3
    say("n=32 N=8 complex64 m=3 vs FP64 oracle")
    amps, y, P, norms, _ = gpu_run(n, e, m, seed, dtype=qpesim.COMPLEX64)
    ra, ry, rP = oracle_run(n, e, m, seed)
    dP = float(np.max(np.abs(P - rP)))
    l2 = rel_l2(y, ry)
    say(f"  max|dP| {dP:.3e}  L2(QIFT output) {l2:.3e}  L2(evolved) {rel_l2(amps, ra):.3e}")
    assert dP <= PK_TOL_F32
    assert l2 < 1e-5


def test_n32_fp32_form_against_fp64_gpu_63_steps():
    n, e, seed, m = 16, 4, 2, 6
    say("n=32 N=8 complex64 vs complex128, m=6 (63 Trotter steps)")
    _, y32, P32, _, _ = gpu_run(n, e, m, seed, dtype=qpesim.COMPLEX64)
    _, y64, P64, _, _ = gpu_run(n, e, m, seed, dtype=qpesim.COMPLEX128)
    dP = float(np.max(np.abs(P32 - P64)))
    say(f"  max|dP| {dP:.3e}  L2 {rel_l2(y32, y64):.3e}")
    assert dP <= PK_TOL_F32
    assert abs(P64.sum() - 1.0) < 1e-12


# ---------------------------------------------------------------- n=28, N=10 (C=2002), FP64, m=8
def test_n28_fp64_full_qpe_qift_and_norms():
    """The plain FP64 workload at its full size (m=8, 255 Trotter steps; per-step parity
    with the oracle is tested at this block size in test_gpu_parity): every block keeps
    block 0's norm, the QIFT of the evolved register equals numpy's FFT of the same
    amplitudes (64 determinants of every block) within 1e-12, and P(k) sums
    to 1."""
    n, e, m, seed = 14, 5, 8, 1
    M = 1 << m
    one, two = qpesim.synthetic_integrals(n, seed)
    h, _ = qpesim.build_hamiltonian(n, e, one, two)
    psi = qpesim.initial_state(n, e, seed, 0.9)
    say("n=28 N=10 complex128 m=8: full chain, QIFT against numpy on a slice")
    sv = qpesim.StateVector(m, n, e, qpesim.COMPLEX128)
    D = sv.D
    sv.set(0, psi / math.sqrt(M))
    qpesim.time_evolve_over_terms(sv, h)
    norms = sv.block_norms()
    assert np.max(np.abs(norms - norms[0])) < 1e-12
    c0 = D // 3
    X = np.stack([sv.get(j * D + c0, 64) for j in range(M)])
    sv.qift()
    Y = np.stack([sv.get(j * D + c0, 64) for j in range(M)])
    P = sv.identify_phase_amplitudes()
    sv.close()
    Yr = np.fft.fft(X, axis=0) / math.sqrt(M)
    l2 = rel_l2(Y, Yr)
    say(f"  norms spread {np.max(np.abs(norms - norms[0])):.1e}  QIFT slice L2 {l2:.2e}  sum P - 1 {P.sum() - 1:.1e}")
    assert l2 < 1e-12
    assert abs(P.sum() - 1.0) < 1e-12


# ---------------------------------------------------------------- 48-instance batching form
def _batch(specs, m, dtype):
    inst = []
    for sp in specs:
        one, two = qpesim.synthetic_integrals(sp["n"], sp["seed"])
        h, _ = qpesim.build_hamiltonian(sp["n"], sp["e"], one, two)
        psi = qpesim.initial_state(sp["n"], sp["e"], sp["seed"], sp["hf"])
        sv = qpesim.StateVector(m, sp["n"], sp["e"], dtype)
        sv.set(0, psi / math.sqrt(1 << m))
        inst.append((sv, h, psi))
    return inst


def _probs(sv):
    sv.qift()
    return sv.identify_phase_amplitudes()


def test_batch_of_48_instances():
    m = 2
    specs = qpesim.batch_instances()
    sizes = [math.comb(sp["n"], sp["e"]) for sp in specs]
    say(f"48-instance batch, m={m}, C from {min(sizes)} to {max(sizes)}")
    single32 = _batch(specs, m, qpesim.COMPLEX64)
    for sv, h, _ in single32:
        qpesim.time_evolve_over_terms(sv, h)
    ref = [sv.get() for sv, _, _ in single32]
    for sv, _, psi in single32:
        sv.set(0, psi / math.sqrt(1 << m))
    qpesim.batch_time_evolve([i[0] for i in single32], [i[1] for i in single32])
    for k, (sv, _, _) in enumerate(single32):
        assert np.array_equal(sv.get(), ref[k]), f"instance {k}"
    P32 = [_probs(sv) for sv, _, _ in single32]
    for sv, _, _ in single32:
        sv.close()

    b64 = _batch(specs, m, qpesim.COMPLEX128)
    qpesim.batch_time_evolve([i[0] for i in b64], [i[1] for i in b64])
    P64 = [_probs(sv) for sv, _, _ in b64]
    for sv, _, _ in b64:
        sv.close()
    dps = [float(np.max(np.abs(a - b))) for a, b in zip(P32, P64)]
    say(f"  per-instance max|dP| FP32 vs FP64: max {max(dps):.3e}")
    assert max(dps) <= PK_TOL_F32
    for k in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        sp = specs[k]
        _, _, rP = oracle_run(sp["n"], sp["e"], m, sp["seed"], hf=sp["hf"])
        assert np.max(np.abs(P64[k] - rP)) < PK_TOL, f"instance {k} (C={sizes[k]})"


# ---------------------------------------------------------------- n=34, N=10 (C=6188), second order
def _order2_run(n, e, m, h, psi, dtype, renorm=None):
    sv = qpesim.StateVector(m, n, e, dtype)
    sv.set_option(qpesim.OPT_TROTTER_ORDER, 2)
    if renorm is not None:
        sv.set_option(qpesim.OPT_RENORMALIZE, renorm)
    sv.set(0, psi)
    qpesim.time_evolve_over_terms(sv, h)
    return sv


def test_n34_second_order_fp32_tracks_fp64():
    n, e, seed, m = 17, 5, 4, 2
    say("n=34 N=10 second order, complex64 vs complex128 (global windows), m=2")
    one, two = qpesim.synthetic_integrals(n, seed)
    h, _ = qpesim.build_hamiltonian(n, e, one, two)
    psi = qpesim.initial_state(n, e, seed, 0.95) / math.sqrt(1 << m)
    P = []
    for dt in (qpesim.COMPLEX128, qpesim.COMPLEX64):
        sv = _order2_run(n, e, m, h, psi, dt)
        P.append(_probs(sv))
        sv.close()
    dP = float(np.max(np.abs(P[0] - P[1])))
    say(f"  max|dP| {dP:.3e}")
    assert dP <= 1e-6
    assert abs(P[0].sum() - 1.0) < 1e-12


def test_n34_second_order_is_time_reversible():
    n, e, seed = 17, 5, 4
    say("n=34 N=10 U2(-t) U2(t) = 1 in FP64")
    one, two = qpesim.synthetic_integrals(n, seed)
    hf, tau = qpesim.build_hamiltonian(n, e, one, two)
    hb, _ = qpesim.build_hamiltonian(n, e, one, two)
    hb.set_timestep(-tau)
    psi = qpesim.initial_state(n, e, seed, 0.95)
    psi = psi / np.linalg.norm(psi)
    D = psi.size
    sv = _order2_run(n, e, 1, hf, psi, qpesim.COMPLEX128, renorm=0)
    fwd = sv.get(D, D)
    sv.close()
    sv = _order2_run(n, e, 1, hb, fwd, qpesim.COMPLEX128, renorm=0)
    back = sv.get(D, D)
    sv.close()
    err = float(np.linalg.norm(back - psi))
    say(f"  |U2(-t)U2(t) psi - psi| = {err:.3e}; |U2(t) psi - psi| = {np.linalg.norm(fwd - psi):.3e}")
    assert err < 1e-12
    assert np.linalg.norm(fwd - psi) > 1e-3
