"""Single-string (interleaved Jordan-Wigner) register, host side.

The index and sign map from the reference's alpha x beta order
(qpesim.single_string_map) is pinned by building the same fermion operators in
both mode orders from creation/annihilation operators (tests/jw.py): every
term's matrix in the interleaved order, restricted to the (e alpha, e beta)
sector, must equal sign * (its blocked-order matrix) * sign under the map, and
must not couple the sector to the other Sz sectors the register carries."""
import itertools
import math

import numpy as np
import pytest

import jw
import oracle
import qpesim


@pytest.mark.parametrize("n,e", [(3, 1), (4, 2), (5, 2), (6, 3)])
def test_map_is_a_bijection_onto_the_sector(n, e):
    idx, sg = qpesim.single_string_map(n, e)
    C = math.comb(n, e)
    assert idx.shape == (C * C,) and set(np.unique(sg)) <= {-1.0, 1.0}
    basis, _ = jw.interleaved_basis(n, e)
    assert len(basis) == math.comb(2 * n, 2 * e)
    even = sum(1 << (2 * p) for p in range(n))
    sector = [i for i, x in enumerate(basis) if bin(x & even).count("1") == e]
    assert sorted(idx.tolist()) == sector
    # strings in the reference's order: alpha rank major, beta rank minor
    strings = sorted(sum(1 << b for b in c) for c in itertools.combinations(range(n), e))
    for k in range(0, C * C, max(1, C * C // 50)):
        a, b = strings[k // C], strings[k % C]
        x = sum(((a >> p & 1) << (2 * p)) | ((b >> p & 1) << (2 * p + 1)) for p in range(n))
        assert basis[idx[k]] == x


@pytest.mark.parametrize("n,e,seed", [(3, 1, 5), (4, 2, 6), (5, 2, 7)])
def test_map_carries_the_hamiltonian(n, e, seed):
    one, two = qpesim.synthetic_integrals(n, seed)
    Hb = jw.textbook_hamiltonian(n, e, one, two)
    Hi = jw.textbook_hamiltonian(n, e, one, two, interleaved=True)
    idx, sg = qpesim.single_string_map(n, e)
    assert np.max(np.abs(Hi[np.ix_(idx, idx)] - sg[:, None] * Hb * sg[None, :])) < 1e-14
    out = np.setdiff1d(np.arange(Hi.shape[0]), idx)
    assert not np.any(Hi[np.ix_(out, idx)]) and not np.any(Hi[np.ix_(idx, out)])
    # the sign is needed: without it the two matrices differ
    assert np.max(np.abs(Hi[np.ix_(idx, idx)] - Hb)) > 1e-3


def test_map_carries_every_term():
    """Term by term (no cancellation between terms can hide a wrong sign)."""
    n, e = 4, 2
    one, two = qpesim.synthetic_integrals(n, 8)
    T = oracle.compile_terms(n, one, two)
    idx, sg = qpesim.single_string_map(n, e)
    L = T.as_lists()
    checked = 0
    for cls in range(5):
        for k, c in L[cls]:
            Gb = jw.term_generator(cls, list(k), c, n, e)
            Gi = jw.term_generator(cls, list(k), c, n, e, interleaved=True)
            assert np.max(np.abs(Gi[np.ix_(idx, idx)] - sg[:, None] * Gb * sg[None, :])) < 1e-14, (cls, k)
            checked += 1
    assert checked == sum(len(x) for x in L)


def test_layout_argument_errors_without_gpu():
    for args in [(4, 4, 2, qpesim.COMPLEX128, 2), (4, 18, 5, qpesim.COMPLEX128, qpesim.LAYOUT_SINGLE_STRING)]:
        with pytest.raises(qpesim.QpeError) as ei:
            qpesim.StateVector(*args)
        assert ei.value.code == qpesim.ERR_ARG
