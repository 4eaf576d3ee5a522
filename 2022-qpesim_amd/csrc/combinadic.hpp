// Combinadic (colex) rank/unrank of fixed-popcount bit strings, host and device.
//
// rank(x) = sum_k C(pos_k, k) over the set bits of x in ascending position,
// k = 1, 2, ... (OffsetComputer::rank, christopherkang/2022-qpesim
// cpp/offset_computer.hpp:16-27).  Colex order of k-subsets equals the numeric
// order of their bit masks, so rank(x) is the index of x among all masks of
// the same popcount.  On the device the binomial table lives in LDS.
#pragma once

#include <cstdint>

#ifndef QPE_HD
#if defined(__HIPCC__)
#define QPE_HD __host__ __device__ inline
#else
#define QPE_HD inline
#endif
#endif

namespace qpe {

constexpr int kBinomRows = 35;  // C(i, k), i < 35 (strings of up to 34 orbitals)
constexpr int kBinomCols = 18;  // k < 18

// binomial table in row-major [i][k], filled by fill_binomials
QPE_HD void fill_binomials(uint32_t *tab) {
    for (int i = 0; i < kBinomRows; ++i)
        for (int k = 0; k < kBinomCols; ++k) {
            uint64_t v;
            if (k > i) v = 0;
            else if (k == 0 || k == i) v = 1;
            else v = (uint64_t)tab[(i - 1) * kBinomCols + k - 1] + tab[(i - 1) * kBinomCols + k];
            tab[i * kBinomCols + k] = (uint32_t)v;
        }
}

QPE_HD uint32_t binom(const uint32_t *tab, int i, int k) {
    return (k < 0 || i < 0 || k > i) ? 0u : tab[i * kBinomCols + k];
}

// colex rank of the set bits of x
QPE_HD uint32_t rank_of(uint32_t x, const uint32_t *tab) {
    uint32_t r = 0;
    int k = 1;
    while (x) {
        int pos = __builtin_ctz(x);
        r += binom(tab, pos, k);
        ++k;
        x &= x - 1;
    }
    return r;
}

// the `ones`-subset of [0, width) with colex rank r
QPE_HD uint32_t unrank_of(uint32_t r, int ones, int width, const uint32_t *tab) {
    uint32_t x = 0;
    int top = width - 1;
    for (int i = ones; i >= 1; --i) {
        while (binom(tab, top, i) > r) --top;  // C(i-1, i) = 0 stops the scan
        x |= 1u << top;
        r -= binom(tab, top, i);
        --top;
    }
    return x;
}

// insert a zero bit at position pos (bits at >= pos move up by one)
QPE_HD uint32_t open_slot(uint32_t f, int pos) {
    uint32_t low = f & ((1u << pos) - 1u);
    return ((f >> pos) << (pos + 1)) | low;
}

// mask of positions strictly between a and b
QPE_HD uint32_t between_mask(int a, int b) {
    int lo = a < b ? a : b, hi = a < b ? b : a;
    uint32_t below_hi = (1u << hi) - 1u;
    uint32_t upto_lo = (lo >= 31) ? 0xffffffffu : ((1u << (lo + 1)) - 1u);
    return below_hi & ~upto_lo;
}

QPE_HD int popcnt(uint32_t x) { return __builtin_popcount(x); }

// Single-string (interleaved) layout.  Spin-orbital 2p is alpha orbital p, 2p+1
// is beta orbital p; a determinant is one string of 2n bits, stored at its colex
// rank among the strings with the same number of electrons.  The reference's
// blocked order puts alpha orbitals before beta ones, so the same creation
// operators in ascending order differ by one swap per (alpha a, beta b) pair
// with b < a: amplitude(interleaved) = sign * amplitude(blocked).

// C(i, k) for any k: the table holds k < kBinomCols, and C(i, k) = C(i, i - k)
// brings every k of a row i < kBinomRows = 2 * kBinomCols - 1 into it
QPE_HD uint32_t binom_sym(const uint32_t *tab, int i, int k) {
    if (k > i - k) k = i - k;
    return binom(tab, i, k);
}

// colex rank of the interleaved string of (alpha string sa, beta string sb)
QPE_HD uint32_t interleaved_rank(uint32_t sa, uint32_t sb, int n, const uint32_t *tab) {
    uint32_t r = 0;
    int k = 1;
    for (int p = 0; p < n; ++p) {
        if ((sa >> p) & 1u) r += binom_sym(tab, 2 * p, k++);
        if ((sb >> p) & 1u) r += binom_sym(tab, 2 * p + 1, k++);
    }
    return r;
}

// parity of the pairs (alpha a, beta b) with b < a: the blocked-to-interleaved sign
QPE_HD uint32_t interleave_parity(uint32_t sa, uint32_t sb) {
    uint32_t par = 0;
    while (sa) {
        const int a = __builtin_ctz(sa);
        par += (uint32_t)popcnt(sb & ((1u << a) - 1u));
        sa &= sa - 1;
    }
    return par & 1u;
}

// the alpha and beta strings of the interleaved string with colex rank r
// (`ones` electrons on 2n spin-orbitals); false when the rank is out of range
QPE_HD bool interleaved_unrank(uint32_t r, int ones, int n, const uint32_t *tab, uint32_t *sa, uint32_t *sb) {
    uint32_t a = 0, b = 0;
    int top = 2 * n - 1;
    for (int i = ones; i >= 1; --i) {
        while (top >= 0 && binom_sym(tab, top, i) > r) --top;
        if (top < 0) return false;
        if (top & 1) b |= 1u << (top >> 1);
        else a |= 1u << (top >> 1);
        r -= binom_sym(tab, top, i);
        --top;
    }
    *sa = a;
    *sb = b;
    return r == 0;
}

// Pair k of a single excitation between orbitals x and y (discoverCA_BVs,
// cpp/time_evolution_over_terms.hpp:1276-1324): the free string f has e-1
// electrons on the other n-2 orbitals.  Returns rank(f|x) in *rx, rank(f|y) in
// *ry, and the number of occupied orbitals strictly between x and y modulo 2.
QPE_HD uint32_t ca_pair(uint32_t k, int x, int y, int n, int e, const uint32_t *tab,
                        uint32_t *rx, uint32_t *ry, uint32_t *par) {
    int lo = x < y ? x : y, hi = x < y ? y : x;
    uint32_t f = unrank_of(k, e - 1, n - 2, tab);
    f = open_slot(open_slot(f, lo), hi);
    *rx = rank_of(f | (1u << x), tab);
    *ry = rank_of(f | (1u << y), tab);
    *par = (uint32_t)popcnt(f & between_mask(x, y)) & 1u;
    return f;
}

// Pair k of a double excitation a+_p a+_q a_r a_s in one string
// (discoverCCAA_BVs, cpp/time_evolution_over_terms.hpp:1160-1210): f has e-2
// electrons on the other n-4 orbitals; *r01 = rank(f|r|s), *r10 = rank(f|p|q);
// parity = occupied between the two lowest + between the two highest of the
// sorted {p,q,r,s}.
QPE_HD void ccaa_pair(uint32_t k, int p, int q, int r, int s, int n, int e, const uint32_t *tab,
                      uint32_t *r01, uint32_t *r10, uint32_t *par) {
    int id[4] = {p, q, r, s};
    for (int a = 1; a < 4; ++a)
        for (int b = a; b > 0 && id[b] < id[b - 1]; --b) {
            int t = id[b]; id[b] = id[b - 1]; id[b - 1] = t;
        }
    uint32_t f = unrank_of(k, e - 2, n - 4, tab);
    for (int a = 0; a < 4; ++a) f = open_slot(f, id[a]);
    *r01 = rank_of(f | (1u << r) | (1u << s), tab);
    *r10 = rank_of(f | (1u << p) | (1u << q), tab);
    *par = (uint32_t)(popcnt(f & between_mask(id[0], id[1])) + popcnt(f & between_mask(id[2], id[3]))) & 1u;
}

// Pair k of a pqqr term inside one string (discoverCNA_BVs,
// cpp/time_evolution_over_terms.hpp:1212-1274): q stays occupied; *r01 =
// rank(f|q|r), *r10 = rank(f|q|p); parity = occupied strictly between p and r,
// q included when it lies between them.
QPE_HD void cna_pair(uint32_t k, int p, int q, int r, int n, int e, const uint32_t *tab,
                     uint32_t *r01, uint32_t *r10, uint32_t *par) {
    int id[3] = {p, q, r};
    for (int a = 1; a < 3; ++a)
        for (int b = a; b > 0 && id[b] < id[b - 1]; --b) {
            int t = id[b]; id[b] = id[b - 1]; id[b - 1] = t;
        }
    uint32_t f = unrank_of(k, e - 2, n - 3, tab);
    for (int a = 0; a < 3; ++a) f = open_slot(f, id[a]);
    uint32_t full = f | (1u << q);
    *r01 = rank_of(full | (1u << r), tab);
    *r10 = rank_of(full | (1u << p), tab);
    *par = (uint32_t)popcnt(full & between_mask(p, r)) & 1u;
}

}  // namespace qpe
