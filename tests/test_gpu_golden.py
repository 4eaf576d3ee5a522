"""The product on the GPU against the committed n=12 spin-orbital / N=6 / m=6
fixture (tests/golden/cfg0_golden.json): P(k), block norms and sampled
amplitudes in the reference's index order, FP64 bar (1e-12)."""
import json
import math
import os

import numpy as np
import pytest

import qpesim
from test_gpu_parity import say

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
G = json.load(open(os.path.join(HERE, "golden", "cfg0_golden.json")))


def test_product_reproduces_the_fixture():
    n, e, m, seed = G["n_spatial"], G["electrons_per_spin"], G["ancillas"], G["seed"]
    say("golden cfg0 fixture")
    one, two = qpesim.synthetic_integrals(n, seed)
    h, _ = qpesim.build_hamiltonian(n, e, one, two)
    h.set_timestep(G["tau"])
    M = 1 << m
    sv = qpesim.StateVector(m, n, e)
    sv.set(0, qpesim.initial_state(n, e, seed))
    sv.scale_block(0, 1.0 / math.sqrt(M))
    qpesim.time_evolve_over_terms(sv, h)
    flat = sv.get()
    norms = sv.block_norms()
    sv.qift()
    P = sv.identify_phase_amplitudes()
    sv.close()
    assert np.max(np.abs(P - np.array(G["P"]))) < 1e-12
    assert np.max(np.abs(norms - np.array(G["block_norms"]))) < 1e-14
    for i, re, im in G["samples"]:
        assert abs(flat[i] - complex(re, im)) < 1e-12
