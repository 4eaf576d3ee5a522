"""Pins the CPU oracle (oracle/qpe_oracle.c) before it is trusted as the checker.

* combinatorics: exhaustive against brute-force enumeration;
* term compiler: the reference's own KATs (cpp/hamiltonian.hpp:192-313 on the
  qsharp-verify/broombridge/{PP,PQ,PQQP,PQQR,PQRS}.yml integrals), and the sum
  of the term generators against the textbook second-quantised Hamiltonian
  built independently in tests/jw.py;
* run tables: every pair and sign against textbook Jordan-Wigner matrices;
* evolution: one Trotter step against the ordered product of exact matrix
  exponentials of the dense term generators;
* QIFT / P(k): numpy.fft (same sign convention) and the definition.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest
from scipy.linalg import expm

import jw
import oracle
import qpesim

HERE = os.path.dirname(os.path.abspath(__file__))


def popcount(x):
    return bin(x).count("1")


# --------------------------------------------------------------- combinatorics
def test_nchoosek_matches_math_comb():
    for n in range(0, 35):
        for k in range(0, n + 1):
            assert oracle.nchoosek(n, k) == math.comb(n, k)
    assert oracle.nchoosek(3, 5) == 0


@pytest.mark.parametrize("n", [1, 4, 7, 10, 13, 16])
def test_rank_is_colex_index_exhaustive(n):
    for e in range(0, n + 1):
        masks = sorted(x for x in range(1 << n) if popcount(x) == e)
        for i, x in enumerate(masks):
            assert oracle.rank(x) == i


def test_rank_large_strings_sampled():
    rng = np.random.default_rng(5)
    for n, e in [(34, 10), (32, 8), (28, 10)]:
        for _ in range(200):
            bits = rng.choice(n, size=e, replace=False)
            x = int(sum(1 << int(b) for b in bits))
            # colex rank = number of e-subsets that are numerically smaller
            expect = sum(math.comb(int(p), k + 1) for k, p in enumerate(sorted(bits)))
            assert oracle.rank(x) == expect


def test_next_comb_walks_colex_order():
    for n, e in [(8, 3), (10, 5), (12, 1)]:
        masks = sorted(x for x in range(1 << n) if popcount(x) == e)
        for a, b in zip(masks, masks[1:]):
            assert oracle.next_comb(a) == b


def test_parity_is_popcount_mod_2():
    rng = np.random.default_rng(1)
    for x in list(rng.integers(0, 2**62, size=500)) + [0, 1, 2**63 - 1]:
        assert oracle.parity(int(x)) == popcount(int(x)) % 2


# --------------------------------------------------------------- term compiler
def _kat(name):
    fx = json.load(open(os.path.join(HERE, "golden", "kat_integrals.json")))[name]
    T = oracle.compile_terms(fx["n_orb"], [tuple(x) for x in fx["one"]], [tuple(x) for x in fx["two"]])
    return dict((k, v) for lst in T.as_lists() for k, v in lst), T


def test_kat_pp():
    # cpp/hamiltonian.hpp:192-219: pp_term(0) == pp_term(1) == exp(-i) at t = 1
    d, T = _kat("PP")
    assert d == {(0, 0): 1.0, (1, 1): 1.0}
    assert T.counts == [2, 0, 0, 0, 0]
    # second half: AddOneBodyTerm((2,2), 1.0) with max_number_of_orbitals = 0 leaves index 0 empty
    T2 = oracle.compile_terms(0, [(2, 2, 1.0)], [])
    assert all(t.idx[0] != 0 for t in T2.arr)


def test_kat_pq():
    # cpp/hamiltonian.hpp:225-239: pq_term(1,0,0) == pq_term(3,2,0) == {cos 1, -i sin 1}
    d, T = _kat("PQ")
    assert d == {(1, 0): 1.0, (3, 2): 1.0}


def test_kat_pqqp():
    # cpp/hamiltonian.hpp:241-258: pqqp_term (0,1), (1,2), (0,3), (2,3) == exp(-i)
    d, T = _kat("PQQP")
    assert d == {(0, 1, 1, 0): 1.0, (1, 2, 2, 1): 1.0, (0, 3, 3, 0): 1.0, (2, 3, 3, 2): 1.0}


def test_kat_pqqr():
    # cpp/hamiltonian.hpp:260-280 queries (2,1,1,0), (5,1,1,3), (2,4,4,0), (5,4,4,3) with
    # coefficient 1; pqqr_term asserts r > p (cpp/hamiltonian.cc:734), and the compiler
    # stores the same four terms with p < r (cpp/hamiltonian.cc:538-542).
    d, T = _kat("PQQR")
    assert d == {(0, 1, 1, 2): 1.0, (3, 1, 1, 5): 1.0, (0, 4, 4, 2): 1.0, (3, 4, 4, 5): 1.0}


def test_kat_pqrs():
    # cpp/hamiltonian.hpp:282-313: eight terms with coefficient -1.  The same-spin keys are
    # stored as queried; the mixed keys are queried in (a, b, a, b) spin order, which the
    # compiler stores as (a, b, b, a) with r and s swapped, flipping the sign.
    d, T = _kat("PQRS")
    for key in [(0, 1, 3, 2), (0, 2, 3, 1), (4, 5, 7, 6), (4, 6, 7, 5)]:
        assert d[key] == -1.0
    for p, q, r, s in [(0, 5, 3, 6), (0, 6, 3, 5), (4, 1, 7, 2), (4, 2, 7, 1)]:
        pass  # mixed keys re-expressed below
    assert d[(0, 5, 6, 3)] == 1.0 and d[(0, 6, 5, 3)] == 1.0
    assert d[(1, 4, 7, 2)] == 1.0 and d[(1, 7, 4, 2)] == 1.0
    assert len(d) == 8


def _random_integrals(n, seed, dense=True):
    return qpesim.synthetic_integrals(n, seed)


@pytest.mark.parametrize("n,e,seed", [(2, 1, 11), (3, 1, 12), (3, 2, 13), (4, 2, 14), (4, 1, 15), (5, 2, 16)])
def test_terms_sum_to_textbook_hamiltonian(n, e, seed):
    one, two = _random_integrals(n, seed)
    T = oracle.compile_terms(n, one, two)
    H_terms = np.zeros((math.comb(n, e) ** 2,) * 2, np.complex128)
    for t in T.arr:
        idx = [int(v) for v in t.idx]
        H_terms += jw.term_generator(int(t.cls), idx, float(t.coeff), n, e)
    H_book = jw.textbook_hamiltonian(n, e, one, two)
    assert np.allclose(H_terms, H_book, atol=1e-12, rtol=0)
    assert np.allclose(H_book, H_book.conj().T)


def test_term_counts_match_closed_forms():
    # dense synthetic integrals: every canonical key is present (survey Appendix C)
    n = 6
    one, two = qpesim.synthetic_integrals(n, 0)
    T = oracle.compile_terms(n, one, two)
    pp, pq, pqqp, pqqr, pqrs = T.counts
    assert pp == 2 * n
    assert pq == 2 * math.comb(n, 2)
    assert pqqp == 2 * math.comb(n, 2) + n * n
    assert pqqr == 2 * (math.comb(n, 2) * (n - 2) + math.comb(n, 2) * n)
    assert pqrs == 2 * 3 * math.comb(n, 4) + 2 * math.comb(n, 2) ** 2


# --------------------------------------------------------------- run tables
def _pairs_textbook(ops_create, ops_annih, n, e):
    """All (rank_src, rank_dst, sign) with dst = a+... a... src inside one spin string."""
    strings = sorted(x for x in range(1 << n) if popcount(x) == e)
    rank = {s: i for i, s in enumerate(strings)}
    ops = [(m, True) for m in ops_create] + [(m, False) for m in ops_annih]
    out = {}
    for x in strings:
        s, y = jw.apply_ops(ops, x)
        if s:
            out[(rank[x], rank[y])] = s
    return out


@pytest.mark.parametrize("n,e", [(5, 2), (6, 3), (7, 3), (8, 4)])
def test_run_tables_ca_match_textbook_signs(n, e):
    for p, r in itertools.permutations(range(n), 2):
        pairs = oracle.expand_pairs(oracle.KIND_CA, n, e, (p, r))
        assert len(pairs) == math.comb(n - 2, e - 1)
        book = _pairs_textbook([p], [r], n, e)   # a+_p a_r : r-occupied -> p-occupied
        assert len(book) == len(pairs)
        for i01, i10, par in pairs:
            assert book[(i01, i10)] == (-1) ** par


@pytest.mark.parametrize("n,e", [(6, 2), (7, 3), (8, 3)])
def test_run_tables_ccaa_match_textbook_signs(n, e):
    for p, q, r, s in itertools.permutations(range(n), 4):
        if not (p < q and s < r and p < s):
            continue
        pairs = oracle.expand_pairs(oracle.KIND_CCAA, n, e, (p, q, r, s))
        assert len(pairs) == math.comb(n - 4, e - 2)
        book = _pairs_textbook([p, q], [r, s], n, e)   # a+_p a+_q a_r a_s
        for i01, i10, par in pairs:
            assert book[(i01, i10)] == (-1) ** par


@pytest.mark.parametrize("n,e", [(6, 2), (7, 3)])
def test_run_tables_cna_match_textbook_signs(n, e):
    for p, q, r in itertools.permutations(range(n), 3):
        if not p < r:
            continue
        pairs = oracle.expand_pairs(oracle.KIND_CNA, n, e, (p, q, r))
        book = _pairs_textbook([p, q], [q, r], n, e)   # a+_p a+_q a_q a_r
        assert len(book) == len(pairs)
        for i01, i10, par in pairs:
            assert book[(i01, i10)] == (-1) ** par


@pytest.mark.parametrize("n,e", [(6, 2), (7, 3)])
def test_run_tables_n_nn_cover_occupied_strings(n, e):
    strings = sorted(x for x in range(1 << n) if popcount(x) == e)
    for q in range(n):
        got = sorted(i for i, _, _ in oracle.expand_pairs(oracle.KIND_N, n, e, (q,)))
        assert got == [i for i, s in enumerate(strings) if s >> q & 1]
    for p, r in itertools.combinations(range(n), 2):
        got = sorted(i for i, _, _ in oracle.expand_pairs(oracle.KIND_NN, n, e, (p, r)))
        assert got == [i for i, s in enumerate(strings) if (s >> p & 1) and (s >> r & 1)]


# --------------------------------------------------------------- evolution
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


@pytest.mark.parametrize("n,e,seed", [(3, 1, 21), (4, 2, 22), (4, 1, 23), (5, 2, 24)])
def test_one_trotter_step_matches_dense_exponentials(n, e, seed):
    one, two = qpesim.synthetic_integrals(n, seed)
    T = oracle.compile_terms(n, one, two)
    t = 0.5 / T.sum_abs()
    dim = math.comb(n, e) ** 2
    U = np.eye(dim, dtype=np.complex128)
    for cls, idx, c in _reference_order(T):
        G = jw.term_generator(cls, list(idx), c, n, e)
        U = expm(-1j * t * G) @ U
    U_or = np.zeros((dim, dim), np.complex128)
    for col in range(dim):
        v = np.zeros(dim, np.complex128)
        v[col] = 1.0
        oracle.apply_unitary(n, e, T, t, v)
        U_or[:, col] = v
    assert np.max(np.abs(U_or - U)) < 1e-13


def _original_order(T):
    """(cls, idx, coeff) in applyUnitary order (TEOT:174-372): pp, every pq,
    pqqp, every pqqr (not gated by a pq term), pqrs."""
    L = T.as_lists()
    return ([(0, k, c) for k, c in L[0]] + [(1, k, c) for k, c in L[1]] + [(2, k, c) for k, c in L[2]] +
            [(3, k, c) for k, c in L[3]] + [(4, k, c) for k, c in L[4]])


def _dense_step(T, n, e, t, seq):
    dim = math.comb(n, e) ** 2
    U = np.eye(dim, dtype=np.complex128)
    for cls, idx, c in seq:
        U = expm(-1j * t * jw.term_generator(cls, list(idx), c, n, e)) @ U
    return U


@pytest.mark.parametrize("n,e,seed", [(3, 1, 21), (4, 2, 22), (4, 1, 23), (5, 2, 24)])
def test_original_order_step_matches_dense_exponentials(n, e, seed):
    one, two = qpesim.synthetic_integrals(n, seed)
    T = oracle.compile_terms(n, one, two)
    t = 0.5 / T.sum_abs()
    U = _dense_step(T, n, e, t, _original_order(T))
    dim = U.shape[0]
    U_or = np.zeros((dim, dim), np.complex128)
    for col in range(dim):
        v = np.zeros(dim, np.complex128)
        v[col] = 1.0
        oracle.apply_unitary_original(n, e, T, t, v)
        U_or[:, col] = v
    assert np.max(np.abs(U_or - U)) < 1e-13
    # the two orders are different products
    U_re = _dense_step(T, n, e, t, _reference_order(T))
    assert np.max(np.abs(U_re - U)) > 1e-6


def test_iqpe_is_the_original_step_repeated():
    """time_evolve_iqpe (TEOT:141-155): applyUnitary k times in place; also the
    pqqr terms without a pq partner act (unlike the reordered step)."""
    n, e = 4, 2
    one, two = qpesim.synthetic_integrals(n, 31)
    one = [x for x in one if not (x[0] == 3 and x[1] == 1)]   # drop h_31: some pqqr lose their pq
    T = oracle.compile_terms(n, one, two)
    t = 0.07
    U = _dense_step(T, n, e, t, _original_order(T))
    rng = np.random.default_rng(3)
    v = rng.normal(size=U.shape[0]) + 1j * rng.normal(size=U.shape[0])
    w = v.copy()
    oracle.time_evolve_iqpe(n, e, T, t, w, 3)
    assert np.max(np.abs(w - U @ U @ U @ v)) < 1e-13
    # chained form: block j = applyUnitary(block j-1)
    m = 2
    st = np.zeros((1 << m, U.shape[0]), np.complex128)
    st[0] = v
    oracle.time_evolve(n, e, m, T, t, st, original=True)
    for j in range(1, 1 << m):
        assert np.max(np.abs(st[j] - U @ st[j - 1])) < 1e-13


def test_pq_gated_pqqr_quirk():
    """A pqqr term whose outer pair has no pq term is never applied (TEOT:453-465)."""
    n, e = 4, 2
    one, two = qpesim.synthetic_integrals(n, 31)
    one = [x for x in one if not (x[0] == 3 and x[1] == 1)]   # drop h_31
    T = oracle.compile_terms(n, one, two)
    t = 0.05
    dim = math.comb(n, e) ** 2
    seq = _reference_order(T)
    U = np.eye(dim, dtype=np.complex128)
    for cls, idx, c in seq:
        U = expm(-1j * t * jw.term_generator(cls, list(idx), c, n, e)) @ U
    rng = np.random.default_rng(0)
    v = rng.normal(size=dim) + 1j * rng.normal(size=dim)
    w = v.copy()
    oracle.apply_unitary(n, e, T, t, w)
    assert np.max(np.abs(w - U @ v)) < 1e-13
    # and the skipped terms do matter: applying every pqqr would differ
    L = T.as_lists()
    skipped = [k for k, _ in L[3] if not any(k[0] == q and k[3] == p for (p, q), _ in L[1])]
    assert skipped


def test_full_qpe_matches_numpy():
    n, e, m = 4, 2, 4
    one, two = qpesim.synthetic_integrals(n, 41)
    T = oracle.compile_terms(n, one, two)
    t = 0.5 / T.sum_abs()
    C = math.comb(n, e)
    D, M = C * C, 1 << m
    rng = np.random.default_rng(3)
    psi = rng.normal(size=D) + 1j * rng.normal(size=D)
    psi /= np.linalg.norm(psi)
    st = np.zeros((M, D), np.complex128)
    st[0] = psi / math.sqrt(M)
    oracle.time_evolve(n, e, m, T, t, st)
    # independent: U from the oracle's single step, powers by matrix products
    U = np.zeros((D, D), np.complex128)
    for col in range(D):
        v = np.zeros(D, np.complex128)
        v[col] = 1
        oracle.apply_unitary(n, e, T, t, v)
        U[:, col] = v
    ref = np.zeros_like(st)
    ref[0] = psi / math.sqrt(M)
    for j in range(1, M):
        ref[j] = U @ ref[j - 1]
    assert np.max(np.abs(st - ref)) < 1e-14
    y_ref = np.fft.fft(ref, axis=0) / math.sqrt(M)
    oracle.qift(m, st)
    assert np.max(np.abs(st - y_ref)) < 1e-14
    P = oracle.phase_probabilities(m, st)
    assert np.allclose(P, (np.abs(y_ref) ** 2).sum(axis=1), atol=1e-15)
    assert abs(P.sum() - 1.0) < 1e-13


def test_qift_of_plane_wave_is_a_peak():
    # forward DFT: a_j = e^{-2 pi i j k0 / M}/sqrt(M) -> ... ; e^{+2 pi i j k0/M} peaks at k0
    for m in range(1, 8):
        M = 1 << m
        for k0 in range(M):
            st = (np.exp(2j * np.pi * np.arange(M) * k0 / M) / math.sqrt(M)).astype(np.complex128).reshape(M, 1).copy()
            oracle.qift(m, st)
            assert abs(st[k0, 0] - 1.0) < 1e-12
            assert np.sum(np.abs(st) ** 2) == pytest.approx(1.0, abs=1e-12)


def test_eigenstate_gives_fejer_kernel():
    n, e, m = 4, 2, 5
    one, two = qpesim.synthetic_integrals(n, 51)
    T = oracle.compile_terms(n, one, two)
    t = 0.5 / T.sum_abs()
    C = math.comb(n, e)
    D, M = C * C, 1 << m
    U = np.zeros((D, D), np.complex128)
    for col in range(D):
        v = np.zeros(D, np.complex128)
        v[col] = 1
        oracle.apply_unitary(n, e, T, t, v)
        U[:, col] = v
    w, V = np.linalg.eig(U)
    psi = V[:, 0] / np.linalg.norm(V[:, 0])
    phi = np.angle(w[0])          # U psi = e^{i phi} psi
    st = np.zeros((M, D), np.complex128)
    st[0] = psi / math.sqrt(M)
    oracle.time_evolve(n, e, m, T, t, st)
    oracle.qift(m, st)
    P = oracle.phase_probabilities(m, st)
    k = np.arange(M)
    delta = phi - 2 * np.pi * k / M            # forward DFT peaks where 2 pi k / M = phi
    with np.errstate(invalid="ignore", divide="ignore"):
        fe = np.where(np.abs(np.sin(delta / 2)) < 1e-15, 1.0,
                      np.sin(M * delta / 2) ** 2 / (M ** 2 * np.sin(delta / 2) ** 2))
    assert np.allclose(P, fe, atol=1e-10)


def test_update_counts_match_closed_forms():
    n, e = 6, 3
    one, two = qpesim.synthetic_integrals(n, 0)
    T = oracle.compile_terms(n, one, two)
    dg, pr = oracle.count_updates(n, e, T)
    C = math.comb(n, e)
    c = math.comb
    diag = n * c(n - 1, e - 1) * C * 2 + 2 * c(n, 2) * c(n - 2, e - 2) * C + n * n * c(n - 1, e - 1) ** 2
    pairs = (2 * c(n, 2) * c(n - 2, e - 1) * C
             + 2 * c(n, 2) * (n - 2) * c(n - 3, e - 2) * C
             + 2 * c(n, 2) * n * c(n - 2, e - 1) * c(n - 1, e - 1)
             + 2 * 3 * c(n, 4) * c(n - 4, e - 2) * C
             + 2 * c(n, 2) ** 2 * c(n - 2, e - 1) ** 2)
    assert dg == diag
    assert pr == pairs
