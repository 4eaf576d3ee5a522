"""The library's caches must never confuse a new object with a freed one that
lived at the same address: a state's device plan is keyed on the Hamiltonian's
generation number, and the cached batch graph on state and plan generations,
and is dropped when a member state is destroyed."""
import math

import numpy as np
import pytest

import qpesim
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def _ham(n, e, seed):
    one, two = qpesim.synthetic_integrals(n, seed)
    return qpesim.build_hamiltonian(n, e, one, two)[0]


def _evolve(sv, h, psi, m):
    sv.set(0, psi / math.sqrt(1 << m))
    qpesim.time_evolve_over_terms(sv, h)
    return sv.get()


def test_plan_is_rebuilt_for_a_new_hamiltonian_at_a_reused_address():
    n, e, m = 6, 3, 3
    say("plan cache vs reused Hamiltonian handles")
    psi = qpesim.initial_state(n, e, 1, 0.9)
    sv = qpesim.StateVector(m, n, e)
    h1 = _ham(n, e, 0)
    _evolve(sv, h1, psi, m)
    del h1                                  # frees the handle; the next one may take its address
    for seed in (1, 2, 3):
        h2 = _ham(n, e, seed)
        got = _evolve(sv, h2, psi, m)
        fresh = qpesim.StateVector(m, n, e)
        want = _evolve(fresh, h2, psi, m)
        fresh.close()
        assert np.array_equal(got, want), f"seed {seed}"
        del h2
    sv.close()


def test_batch_graph_is_dropped_with_a_destroyed_member():
    m = 3
    say("batch cache vs destroyed members")
    shapes = [(6, 3, 11), (7, 2, 12)]
    hams = [_ham(n, e, s) for n, e, s in shapes]
    psis = [qpesim.initial_state(n, e, s, 0.9) for n, e, s in shapes]

    def make():
        out = []
        for (n, e, _), psi in zip(shapes, psis):
            sv = qpesim.StateVector(m, n, e, qpesim.COMPLEX64)
            sv.set(0, psi / math.sqrt(1 << m))
            out.append(sv)
        return out

    svs = make()
    qpesim.batch_time_evolve(svs, hams)
    ref = [sv.get() for sv in svs]
    for sv in svs:
        sv.close()                           # destroys every member of the cached batch
    for rep in range(3):                     # new states, possibly at the old addresses
        svs = make()
        qpesim.batch_time_evolve(svs, hams)
        for k, sv in enumerate(svs):
            assert np.array_equal(sv.get(), ref[k]), f"rep {rep} instance {k}"
        for sv in svs:
            sv.close()
