"""The committed n=12 spin-orbital / N=6 / m=6 fixture (tests/golden/cfg0_golden.json,
written by tests/golden/make_cfg0_golden.py from the oracle): the oracle still
reproduces it, and the product's term compiler gives its canonical term list
bit for bit (no GPU needed)."""
import json
import math
import os

import numpy as np

import oracle
import qpesim

HERE = os.path.dirname(os.path.abspath(__file__))
G = json.load(open(os.path.join(HERE, "golden", "cfg0_golden.json")))


def test_fixture_shape():
    M = 1 << G["ancillas"]
    assert len(G["terms"]) == 948            # the term count of this workload in the reference's canonical forms
    assert len(G["P"]) == M and len(G["block_norms"]) == M
    assert abs(sum(G["P"]) - 1.0) < 1e-12
    assert max(abs(x - 1.0 / M) for x in G["block_norms"]) < 1e-14


def test_oracle_reproduces_the_fixture():
    n, e, m, seed = G["n_spatial"], G["electrons_per_spin"], G["ancillas"], G["seed"]
    one, two = qpesim.synthetic_integrals(n, seed)
    T = oracle.compile_terms(n, one, two)
    assert [[int(t.cls), [int(v) for v in t.idx], float(t.coeff)] for t in T.arr] == G["terms"]
    assert 0.5 / T.sum_abs() == G["tau"]
    C = math.comb(n, e)
    M = 1 << m
    st = np.zeros((M, C * C), np.complex128)
    st[0] = qpesim.initial_state(n, e, seed) / math.sqrt(M)
    oracle.time_evolve(n, e, m, T, G["tau"], st)
    flat = st.ravel()
    for i, re, im in G["samples"]:
        assert abs(flat[i] - complex(re, im)) < 1e-15
    oracle.qift(m, st)
    assert np.max(np.abs(oracle.phase_probabilities(m, st) - np.array(G["P"]))) < 1e-14


def test_product_term_compiler_gives_the_fixture_terms():
    n, e = G["n_spatial"], G["electrons_per_spin"]
    one, two = qpesim.synthetic_integrals(n, G["seed"])
    h, tau = qpesim.build_hamiltonian(n, e, one, two)
    assert abs(tau - G["tau"]) < 1e-15 * G["tau"] * 8   # the same sum of |coefficients|, another summation order
    got = []
    for cls in range(5):
        idx, co = h.terms(cls)
        for k, c in zip(idx, co):
            got.append([cls, [int(v) for v in k], float(c)])
    # one-body keys carry two indices; the fixture pads them with the oracle's fill
    want = [[c, ix[:2] if c in (0, 1) else ix, v] for c, ix, v in G["terms"]]
    got = [[c, ix[:2] if c in (0, 1) else ix, v] for c, ix, v in got]
    assert got == want
