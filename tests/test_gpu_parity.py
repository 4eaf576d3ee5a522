"""GPU parity: libqpesim.so (through its C ABI) against the CPU oracle on the
same seeded synthetic Hamiltonians.

Bar (FP64): determinant indices and signs bit-exact; amplitudes within 1e-12
relative L2; P(k) within 1e-12 absolute.  FP32 precision form: max |dP(k)|
<= 1e-5 against the FP64 oracle."""
import math
import sys

import numpy as np
import pytest

import oracle
import qpesim

pytestmark = pytest.mark.gpu

AMP_TOL = 1e-12     # relative L2, FP64
PK_TOL = 1e-12      # max |dP(k)|, FP64
PK_TOL_F32 = 1e-5   # max |dP(k)|, FP32 state and rotations


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def gpu_run(n, e, m, seed, dtype=qpesim.COMPLEX128, hf=1.0, tile=0, psi0=None, want_amps=True):
    one, two = qpesim.synthetic_integrals(n, seed)
    h, tau = qpesim.build_hamiltonian(n, e, one, two)
    if psi0 is None:
        psi0 = qpesim.initial_state(n, e, seed, hf)
    M = 1 << m
    sv = qpesim.StateVector(m, n, e, dtype)
    if tile:
        sv.set_option(qpesim.OPT_MAX_TILE_BYTES, tile)
    sv.set(0, psi0)
    sv.scale_block(0, 1.0 / math.sqrt(M))
    qpesim.time_evolve_over_terms(sv, h)
    amps = sv.get() if want_amps else None
    norms = sv.block_norms()
    sv.qift()
    y = sv.get() if want_amps else None
    P = sv.identify_phase_amplitudes()
    info = sv.timings()
    sv.close()
    return amps, y, P, norms, info


def oracle_run(n, e, m, seed, hf=1.0, psi0=None):
    one, two = qpesim.synthetic_integrals(n, seed)
    T = oracle.compile_terms(n, one, two)
    h, tau = qpesim.build_hamiltonian(n, e, one, two)   # tau exactly as the product uses it
    if psi0 is None:
        psi0 = qpesim.initial_state(n, e, seed, hf)
    C = math.comb(n, e)
    M = 1 << m
    st = np.zeros((M, C * C), np.complex128)
    st[0] = psi0 / math.sqrt(M)
    oracle.time_evolve(n, e, m, T, tau, st)
    amps = st.ravel().copy()
    oracle.qift(m, st)
    P = oracle.phase_probabilities(m, st)
    return amps, st.ravel(), P


# ------------------------------------------------------------------ indices
@pytest.mark.parametrize("n,e", [(6, 3), (8, 3), (10, 5), (14, 5), (16, 4), (17, 5)])
def test_device_pairs_bit_exact(n, e):
    say(f"device pairs n={n} e={e}")
    cases = []
    for p in range(n):
        for r in range(n):
            if p != r:
                cases.append((2, (p, r)))
    rng = np.random.default_rng(n * 100 + e)
    for _ in range(40):
        p, q, r, s = (int(x) for x in rng.choice(n, 4, replace=False))
        cases.append((4, (p, q, r, s)))
        a, b, c = (int(x) for x in rng.choice(n, 3, replace=False))
        lo, hi = min(a, c), max(a, c)
        cases.append((3, (lo, b, hi)))
    kinds = {2: oracle.KIND_CA, 3: oracle.KIND_CNA, 4: oracle.KIND_CCAA}
    for kind, args in cases:
        if kind in (3, 4) and e < 2:
            continue
        g1, g2, gp = qpesim.debug_pairs(kind, n, e, args)
        got = sorted(zip(g1.tolist(), g2.tolist(), gp.tolist()))
        ref = sorted(oracle.expand_pairs(kinds[kind], n, e, args))
        assert got == ref, (kind, args)


# ------------------------------------------------------------------ FP64 parity
PARITY_CASES = [
    # n, e, m, seed, hf
    (2, 1, 3, 5, 1.0),
    (3, 1, 3, 6, 0.9),
    (4, 2, 4, 7, 0.9),
    (5, 2, 4, 8, 0.9),
    (5, 4, 3, 9, 0.9),     # e = n - 1
    (6, 3, 6, 0, 1.0),     # the n=12 spin-orbital, N=6, m=6 workload (seed 0, HF)
    (7, 3, 5, 10, 0.9),
    (8, 2, 4, 11, 0.9),
    (8, 4, 3, 12, 0.9),
    (9, 3, 3, 13, 0.9),
]


@pytest.mark.parametrize("n,e,m,seed,hf", PARITY_CASES)
def test_fp64_parity_with_oracle(n, e, m, seed, hf):
    say(f"fp64 parity n={n} e={e} m={m} seed={seed}")
    amps, y, P, norms, _ = gpu_run(n, e, m, seed, hf=hf)
    ra, ry, rP = oracle_run(n, e, m, seed, hf=hf)
    assert rel_l2(amps, ra) < AMP_TOL
    assert rel_l2(y, ry) < AMP_TOL
    assert np.max(np.abs(P - rP)) < PK_TOL
    assert abs(P.sum() - 1.0) < 1e-12                      # unitarity
    assert np.max(np.abs(norms - 1.0 / (1 << m))) < 1e-14  # every block keeps its norm


@pytest.mark.slow
@pytest.mark.parametrize("n,e,seed,hf", [(14, 5, 1, 0.9), (16, 4, 2, 1.0)])
def test_fp64_parity_with_oracle_at_full_block_size(n, e, seed, hf):
    """The block sizes of the n=28/N=10 and n=32/N=8 workloads (C = 2002, 1820),
    three Trotter steps: amplitudes and P(k) against the oracle (about 20 s of
    host time each)."""
    m = 2
    say(f"fp64 parity at full block size n={n} e={e}")
    amps, y, P, norms, _ = gpu_run(n, e, m, seed, hf=hf)
    ra, ry, rP = oracle_run(n, e, m, seed, hf=hf)
    assert rel_l2(amps, ra) < AMP_TOL
    assert rel_l2(y, ry) < AMP_TOL
    assert np.max(np.abs(P - rP)) < PK_TOL


def test_global_windows_at_full_block_size():
    """n=34/N=10 (C = 6188): complex128 runs its mixed windows in global memory
    (a column is 99 KB, the LDS tile holds one), complex64 in LDS tiles of three
    columns.  Two independent code paths; one Trotter step must agree to the
    complex64 precision."""
    n, e, m, seed = 17, 5, 1, 4
    say("global windows vs LDS windows at C=6188")
    amps64, _, P64, _, _ = gpu_run(n, e, m, seed, hf=0.95)
    amps32, _, P32, _, _ = gpu_run(n, e, m, seed, hf=0.95, dtype=qpesim.COMPLEX64)
    assert rel_l2(amps32, amps64) < 1e-6
    assert np.max(np.abs(P32 - P64)) < 1e-7


@pytest.mark.parametrize("tile_cols", [2, 3, 5])
def test_fp64_parity_small_tiles(tile_cols):
    """Small LDS budgets split the mixed double excitations into many windows."""
    n, e, m, seed = 7, 3, 4, 14
    C = math.comb(n, e)
    say(f"fp64 small tiles cols={tile_cols}")
    amps, y, P, _, info = gpu_run(n, e, m, seed, hf=0.9, tile=tile_cols * C * 16)
    ra, ry, rP = oracle_run(n, e, m, seed, hf=0.9)
    assert rel_l2(amps, ra) < AMP_TOL
    assert np.max(np.abs(P - rP)) < PK_TOL


def test_random_initial_state_parity():
    n, e, m = 6, 3, 5
    C = math.comb(n, e)
    rng = np.random.default_rng(77)
    psi = rng.normal(size=C * C) + 1j * rng.normal(size=C * C)
    psi /= np.linalg.norm(psi)
    say("fp64 random psi0")
    amps, y, P, _, _ = gpu_run(n, e, m, 3, psi0=psi)
    ra, ry, rP = oracle_run(n, e, m, 3, psi0=psi)
    assert rel_l2(amps, ra) < AMP_TOL
    assert np.max(np.abs(P - rP)) < PK_TOL


# ------------------------------------------------------------------ FP32 form
@pytest.mark.parametrize("n,e,m,seed", [(6, 3, 6, 0), (8, 4, 5, 21), (10, 3, 6, 22)])
def test_fp32_precision_form(n, e, m, seed):
    say(f"fp32 n={n} e={e} m={m}")
    amps, y, P, norms, _ = gpu_run(n, e, m, seed, dtype=qpesim.COMPLEX64, hf=0.9)
    ra, ry, rP = oracle_run(n, e, m, seed, hf=0.9)
    assert np.max(np.abs(P - rP)) <= PK_TOL_F32
    assert rel_l2(amps, ra) < 1e-4
    # every block is renormalised to block 0's norm, i.e. to |psi0|^2 after rounding to FP32;
    # the factor is applied in FP32 (relative rounding 2^-24, so 2^-23 on a squared norm)
    assert abs(P.sum() - norms[0] * (1 << m)) < 2.0 ** -23
    assert abs(P.sum() - 1.0) < 1e-6


# ------------------------------------------------------------------ failure handling
def test_oom_is_reported_before_allocation():
    with pytest.raises(qpesim.QpeError) as ei:
        qpesim.StateVector(20, 16, 8, qpesim.COMPLEX128)    # ~2.8 PB
    assert ei.value.code == qpesim.ERR_OOM


def test_non_finite_input_is_detected():
    n, e, m = 4, 2, 3
    one, two = qpesim.synthetic_integrals(n, 1)
    h, tau = qpesim.build_hamiltonian(n, e, one, two)
    sv = qpesim.StateVector(m, n, e)
    psi = qpesim.initial_state(n, e)
    psi[5] = np.nan
    sv.set(0, psi)
    with pytest.raises(qpesim.QpeError) as ei:
        qpesim.time_evolve_over_terms(sv, h)
    assert ei.value.code == qpesim.ERR_NUMERIC
    sv.close()


def test_mismatched_hamiltonian_is_rejected():
    one, two = qpesim.synthetic_integrals(4, 1)
    h, _ = qpesim.build_hamiltonian(4, 2, one, two)
    sv = qpesim.StateVector(2, 5, 2)
    with pytest.raises(qpesim.QpeError) as ei:
        qpesim.time_evolve_over_terms(sv, h)
    assert ei.value.code == qpesim.ERR_ARG
    sv.close()


def test_set_get_roundtrip_reference_order():
    n, e, m = 5, 2, 2
    sv = qpesim.StateVector(m, n, e)
    rng = np.random.default_rng(2)
    x = rng.normal(size=sv.size()) + 1j * rng.normal(size=sv.size())
    sv.set(0, x)
    assert np.array_equal(sv.get(), x)
    sv.set(7, x[:11] * 2)
    y = sv.get(5, 20)
    assert np.array_equal(y[2:13], x[:11] * 2) and np.array_equal(y[:2], x[5:7])
    sv.close()
