"""Textbook Jordan-Wigner fermions on the fixed (e alpha, e beta) sector, in the
reference's blocked ("HalfUp") mode order: alpha orbital p is mode p, beta
orbital p is mode n+p.  A basis state is a bit mask over the 2n modes; its
index in the register is rank(alpha bits) * C + rank(beta bits).

a+_j |x> = (-1)^{#occupied modes below j} |x + j>  (standard JW convention).

This module is an independent restatement used to pin the oracle: nothing in
it follows the reference's code paths (run tables, parity rules)."""
import itertools
import math

import numpy as np


def sector_basis(n, e):
    strings = sorted(sum(1 << b for b in c) for c in itertools.combinations(range(n), e))
    rank = {s: i for i, s in enumerate(strings)}
    C = len(strings)
    basis = []
    for a in strings:
        for b in strings:
            basis.append(a | (b << n))
    index = {x: i for i, x in enumerate(basis)}
    return strings, rank, basis, index, C


def interleaved_basis(n, e):
    """The single-string register's basis: every string of 2e electrons on 2n
    interleaved modes (alpha orbital p = mode 2p, beta orbital p = mode 2p+1),
    including the Sz != 0 ones, in colex order (ascending bit mask)."""
    basis = sorted(sum(1 << b for b in c) for c in itertools.combinations(range(2 * n), 2 * e))
    return basis, {x: i for i, x in enumerate(basis)}


def interleaved_mode(mode, n):
    """blocked mode (alpha p = p, beta p = n + p) -> interleaved mode"""
    return 2 * mode if mode < n else 2 * (mode - n) + 1


def apply_ops(ops, x):
    """ops: list of (mode, dagger) applied right-to-left; returns (sign, x') or (0, None)."""
    sign = 1
    for mode, dag in reversed(ops):
        bit = 1 << mode
        occ = bool(x & bit)
        if dag == occ:
            return 0, None
        if bin(x & (bit - 1)).count("1") & 1:
            sign = -sign
        x ^= bit
    return sign, x


def operator_matrix(terms, n, e, interleaved=False):
    """sum_k c_k * product_k on the sector; terms = [(c, [(mode, dagger), ...])]
    with blocked mode numbers.  interleaved: the same operators with the modes
    renumbered to the interleaved order, on the single-string basis."""
    if interleaved:
        basis, index = interleaved_basis(n, e)
        terms = [(c, [(interleaved_mode(md, n), dg) for md, dg in ops]) for c, ops in terms]
    else:
        _, _, basis, index, _ = sector_basis(n, e)
    dim = len(basis)
    Mx = np.zeros((dim, dim), dtype=np.complex128)
    for c, ops in terms:
        for col, x in enumerate(basis):
            s, y = apply_ops(ops, x)
            if s and y in index:
                Mx[index[y], col] += c * s
    return Mx


def term_generator(cls, idx, coeff, n, e, interleaved=False):
    """Hermitian generator of one indexed reference term."""
    cr = lambda m: (m, True)
    an = lambda m: (m, False)
    if cls == 0:      # pp: c n_p
        p = idx[0]
        ops = [(coeff, [cr(p), an(p)])]
    elif cls == 1:    # pq: c (a+_p a_q + a+_q a_p)
        p, q = idx[0], idx[1]
        ops = [(coeff, [cr(p), an(q)]), (coeff, [cr(q), an(p)])]
    elif cls == 2:    # pqqp: c a+_p a+_q a_q a_p
        p, q = idx[0], idx[1]
        ops = [(coeff, [cr(p), cr(q), an(q), an(p)])]
    else:             # pqqr / pqrs: c (a+_p a+_q a_r a_s + h.c.)
        p, q, r, s = idx
        ops = [(coeff, [cr(p), cr(q), an(r), an(s)]), (coeff, [cr(s), cr(r), an(q), an(p)])]
    return operator_matrix(ops, n, e, interleaved)


def textbook_hamiltonian(n, e, one_body, two_body, interleaved=False):
    """H = sum h_pq a+_{p s} a_{q s} + 1/2 sum (pq|rs) a+_{p s} a+_{r s'} a_{s s'} a_{q s}
    from Broombridge-style lists (1-based, Mulliken two-body, each 8-fold class once)."""
    h = np.zeros((n, n))
    for p, q, v in one_body:
        h[p - 1, q - 1] = v
        h[q - 1, p - 1] = v
    g = np.zeros((n, n, n, n))
    for a, b, c, d, v in two_body:
        p, q, r, s = a - 1, b - 1, c - 1, d - 1
        for (i, j, k, l) in {(p, q, r, s), (q, p, r, s), (p, q, s, r), (q, p, s, r),
                             (r, s, p, q), (s, r, p, q), (r, s, q, p), (s, r, q, p)}:
            g[i, j, k, l] = v
    terms = []
    for sp in range(2):
        for p in range(n):
            for q in range(n):
                if h[p, q] != 0.0:
                    terms.append((h[p, q], [(sp * n + p, True), (sp * n + q, False)]))
    for s1 in range(2):
        for s2 in range(2):
            for p, q, r, s in itertools.product(range(n), repeat=4):
                v = g[p, q, r, s]
                if v == 0.0:
                    continue
                terms.append((0.5 * v, [(s1 * n + p, True), (s2 * n + r, True), (s2 * n + s, False), (s1 * n + q, False)]))
    return operator_matrix(terms, n, e, interleaved)
