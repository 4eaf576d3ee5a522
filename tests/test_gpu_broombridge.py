"""The reference program's call sequence (cpp/main.cc:99-167) on real molecules
read from Broombridge files: qpe_state_ingest_broombridge -> IndexTerms ->
SetTimestep(1/steps) -> 1/sqrt(2^m) -> time_evolve_over_terms -> qift ->
IdentifyPhaseAmplitudes -> energy estimate, on the GPU through the C ABI,
against the oracle (PyYAML ingest + the C oracle) on the same files.

Bars as for the synthetic workloads: FP64 amplitudes within 1e-12 relative L2
and P(k) within 1e-12; FP32 max |dP(k)| <= 1e-5.  The physics check: the P(k)
peak sits at the exact ground-state phase -E0 t / 2 pi (mod 1) of the textbook
Hamiltonian, E0 from dense diagonalisation, within one bin."""
import math
import os
import sys

import numpy as np
import pytest

import broombridge_oracle as bo
import jw
import oracle
import qpesim

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = os.path.join(HERE, "golden", "broombridge")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def product_run(path, m, steps, dtype=qpesim.COMPLEX128, t=None):
    h = qpesim.Hamiltonian()
    sv = h.ingest_broombridge(path, m, dtype)
    try:
        h.index_terms()
        if t is None:
            t = 1.0 / steps
        h.set_timestep(t)
        sv.scale_block(0, 1.0 / math.sqrt(1 << m))
        qpesim.time_evolve_over_terms(sv, h)
        amps = sv.get()
        sv.qift()
        P = sv.identify_phase_amplitudes()
        return amps, P, h.identity_term()
    finally:
        sv.close()


def oracle_run(path, m, steps, t=None):
    r = bo.ingest_file(path)
    n, e = r["n"], r["e"]
    C = math.comb(n, e)
    M = 1 << m
    st = np.zeros((M, C * C), np.complex128)
    for a, c in r["initial"]:
        st[0, a] = c
    st[0] *= 1.0 / math.sqrt(M)
    oracle.time_evolve(n, e, m, oracle.compile_terms(n, r["one"], r["two"]), 1.0 / steps if t is None else t, st)
    amps = st.ravel().copy()
    oracle.qift(m, st)
    return r, amps, oracle.phase_probabilities(m, st)


CASES = [
    # file, ancillas, steps (t = 1/steps)
    ("PQ.yml", 4, 10),
    ("PQQR.yml", 4, 10),
    ("PQRS.yml", 4, 10),
    ("H2_2_sto-6g.yml", 6, 10),
    ("h2_ccpvtz_ccsd_1_4008au.yaml", 6, 10),
    ("h4_sto6g_0.100.nw.out.yaml", 5, 10),
    ("Be-ccpVDZ-5orb.yaml", 6, 10),
    ("h2o_oh0.9_sto3g.nw.out.yaml", 6, 10),
    ("N2-10orb.yaml", 5, 10),          # three determinants, C(10,7) = 120
    ("H2O-10orb.yaml", 4, 10),         # C(10,5) = 252
]


@pytest.mark.parametrize("name,m,steps", CASES)
def test_broombridge_qpe_fp64_matches_oracle(name, m, steps):
    path = os.path.join(FIXTURES, name)
    say(f"broombridge fp64 {name} m={m}")
    amps, P, ident = product_run(path, m, steps)
    r, ra, rP = oracle_run(path, m, steps)
    assert ident == r["identity"]
    assert rel_l2(amps, ra) < 1e-12
    assert np.max(np.abs(P - rP)) < 1e-12
    est = qpesim.energy_estimate(P, m, ident, 1.0 / steps)
    phase = float(np.dot(np.arange(1 << m), rP)) / (1 << m)
    assert abs(est["energy"] - (-2 * math.pi * phase + r["identity"])) < 1e-10


@pytest.mark.slow
def test_broombridge_qpe_at_full_block_size():
    """H2O in 14 orbitals with 5+5 electrons: C = 2002, the block size of the
    n=28 / N=10 workload, from a real molecule (three Trotter steps), at the
    synthetic workloads' step t = 0.5 / sum|c| and at the reference program's
    t = 1/steps = 0.1, where one step turns sum|c| t = 55 rad.

    The amplitudes agree within the FP64 bar at both steps.  At t = 0.1 the
    reference's P(k) accumulation (a plain double sum over the 4.0M
    determinants in index order, cpp/state_vector.hpp:341-372, restated by the
    oracle) is itself up to 2.9e-12 off the correctly rounded sum of the same
    |y_k(d)|^2; the product's compensated sum is checked against the correctly
    rounded sum (math.fsum of the oracle's QIFT output) at 1e-12, and against the
    plain sum within that sum's own rounding error."""
    path = os.path.join(FIXTURES, "H2O-14orb.yaml")
    m = 2
    h = qpesim.Hamiltonian()
    h.load_broombridge(path=path)
    h.index_terms()
    tau = 0.5 / h.sum_abs_coeff()
    say("broombridge fp64 H2O-14orb m=2 t=0.5/sum|c|")
    amps, P, _ = product_run(path, m, 0, t=tau)
    _, ra, rP = oracle_run(path, m, 0, t=tau)
    assert rel_l2(amps, ra) < 1e-12
    assert np.max(np.abs(P - rP)) < 1e-12
    say("broombridge fp64 H2O-14orb m=2 t=0.1")
    amps, P, _ = product_run(path, m, 10)
    _, ra, rP = oracle_run(path, m, 10)
    assert rel_l2(amps, ra) < 1e-12
    M = 1 << m
    y = np.fft.fft(ra.reshape(M, -1), axis=0) / math.sqrt(M)     # the oracle's QIFT (pinned to numpy.fft)
    exact = np.array([math.fsum((y[k].real ** 2 + y[k].imag ** 2).tolist()) for k in range(M)])
    say(f"  |P - exact| {np.max(np.abs(P - exact)):.2e}; plain-sum oracle off exact by {np.max(np.abs(rP - exact)):.2e}")
    assert np.max(np.abs(P - exact)) < 1e-12
    assert np.max(np.abs(P - rP)) < 1e-12 + 2 * np.max(np.abs(rP - exact))


@pytest.mark.parametrize("name", ["Be-ccpVDZ-5orb.yaml", "h2o_oh0.9_sto3g.nw.out.yaml", "H2O-10orb.yaml"])
def test_broombridge_qpe_fp32_matches_oracle(name):
    path = os.path.join(FIXTURES, name)
    m, steps = 6, 10
    say(f"broombridge fp32 {name}")
    _, P, _ = product_run(path, m, steps, qpesim.COMPLEX64)
    _, _, rP = oracle_run(path, m, steps)
    assert np.max(np.abs(P - rP)) <= 1e-5


@pytest.mark.parametrize("name", ["H2_2_sto-6g.yml", "Be-ccpVDZ-5orb.yaml", "h2o_oh0.9_sto3g.nw.out.yaml"])
def test_qpe_peak_sits_at_the_ground_state_phase(name):
    path = os.path.join(FIXTURES, name)
    m, steps = 8, 10
    M = 1 << m
    say(f"broombridge ground-state phase {name}")
    res = qpesim.run_main(path, m, steps)
    r = bo.ingest_file(path)
    E0 = np.linalg.eigvalsh(jw.textbook_hamiltonian(r["n"], r["e"], r["one"], r["two"]))[0]
    want = (-E0 * (1.0 / steps) / (2 * math.pi)) % 1.0 * M
    k = int(np.argmax(res["probs"]))
    assert min(abs(k - want), M - abs(k - want)) <= 1.0
    assert res["probs"][k] > 0.4
    assert abs(res["probs"].sum() - 1.0) < 1e-12
    assert res["moe"] == pytest.approx(2 * math.pi / ((1.0 / steps) * M))


def test_repeated_determinant_keeps_the_last_coefficient(tmp_path):
    from test_broombridge import DUPLICATE
    path = tmp_path / "dup.yml"
    path.write_text(DUPLICATE)
    h = qpesim.Hamiltonian()
    sv = h.ingest_broombridge(str(path), 2)
    block0 = sv.get(0, sv.D)
    sv.close()
    want = np.zeros(block0.size, np.complex128)
    for a, c in bo.ingest(DUPLICATE)["initial"]:
        want[a] = c                               # later entries overwrite earlier ones
    assert block0[0] == -0.5
    assert np.array_equal(block0, want)


def test_ingest_errors_leave_no_state():
    h = qpesim.Hamiltonian()
    with pytest.raises(qpesim.QpeError) as ei:
        h.ingest_broombridge(os.path.join(FIXTURES, "PP.yml"), 3)    # one orbital: below the register's minimum
    assert ei.value.code == qpesim.ERR_ARG
    assert h.sizes() == (0, 0)


def test_batch_of_molecules_equals_single_runs():
    """The batching form on real molecules of different sizes (C = 6 .. 252, one
    to three determinants): one launch sequence for all of them gives every
    register bit for bit as its own time_evolve_over_terms."""
    names = ["H2_2_sto-6g.yml", "h2_ccpvtz_ccsd_1_4008au.yaml", "Be-ccpVDZ-5orb.yaml",
             "h2o_oh0.9_sto3g.nw.out.yaml", "N2-10orb.yaml", "H2O-10orb.yaml"]
    m, t = 4, 0.1
    say("broombridge batch of molecules")
    for dtype in (qpesim.COMPLEX128, qpesim.COMPLEX64):
        singles, states, hams = [], [], []
        for name in names:
            path = os.path.join(FIXTURES, name)
            for batch in (False, True):
                h = qpesim.Hamiltonian()
                sv = h.ingest_broombridge(path, m, dtype)
                h.index_terms()
                h.set_timestep(t)
                sv.scale_block(0, 1.0 / math.sqrt(1 << m))
                if batch:
                    states.append(sv)
                    hams.append(h)
                else:
                    qpesim.time_evolve_over_terms(sv, h)
                    singles.append(sv.get())
                    sv.close()
        qpesim.batch_time_evolve(states, hams)
        for sv, ref in zip(states, singles):
            assert np.array_equal(sv.get(), ref)
            sv.close()
