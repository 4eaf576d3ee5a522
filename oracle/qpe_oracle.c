/*
 * qpe_oracle.c -- CPU restatement of the reference QPE hot path (plain C99 +
 * OpenMP).  TEST INFRASTRUCTURE ONLY: loaded by tests/, smoke() and the
 * cpu_baseline leg of bench.py as the checker / CPU baseline, never by the
 * product.  Each function cites the reference file:line it follows.
 */
#include "qpe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* combinatorics: cpp/util.hpp:49-94, cpp/offset_computer.hpp:16-27          */
/* ------------------------------------------------------------------------ */

/* nChoosek, cpp/util.hpp:49-60 */
uint64_t orc_nchoosek(int n, int k) {
    if (k < 0 || n < 0 || k > n) return 0;
    if (k * 2 > n) k = n - k;
    if (k == 0) return 1;
    uint64_t r = (uint64_t)n;
    for (int i = 2; i <= k; ++i) {
        r *= (uint64_t)(n - i + 1);
        r /= (uint64_t)i;
    }
    return r;
}

/* OffsetComputer::rank, cpp/offset_computer.hpp:16-27: sum_k C(pos_k, k), k 1-based */
uint64_t orc_rank(uint64_t bv) {
    uint64_t r = 0;
    int i = 0, k = 1;
    for (; bv != 0; ++i, bv >>= 1) {
        if (bv & 1u) {
            r += orc_nchoosek(i, k);
            ++k;
        }
    }
    return r;
}

/* next_comb (Gosper), cpp/util.hpp:62-70 */
uint64_t orc_next_comb(uint64_t x) {
    uint64_t u = x & (0 - x);
    uint64_t v = u + x;
    return v + (((v ^ x) / u) >> 2);
}

/* parity, cpp/util.hpp:73-81 */
unsigned orc_parity(uint64_t x) {
    x ^= x >> 32; x ^= x >> 16; x ^= x >> 8; x ^= x >> 4; x ^= x >> 2; x ^= x >> 1;
    return (unsigned)(x & 1u);
}

static int popc(uint64_t x) { return __builtin_popcountll(x); }

/* ------------------------------------------------------------------------ */
/* term compiler: cpp/hamiltonian.cc:301-657, 160-197, 253-299                */
/* ------------------------------------------------------------------------ */

/* SpinOrbital (cpp/types.hpp:31-98): orbital (1-based as ingested), spin 0=Up 1=Down.
 * operator< orders by spin, then orbital. */
typedef struct { int orb; int spin; } so_t;

static int so_cmp(so_t a, so_t b) {
    if (a.spin != b.spin) return a.spin < b.spin ? -1 : 1;
    if (a.orb != b.orb) return a.orb < b.orb ? -1 : 1;
    return 0;
}
static int so_eq(so_t a, so_t b) { return a.orb == b.orb && a.spin == b.spin; }
static so_t so_min(so_t a, so_t b) { return so_cmp(b, a) < 0 ? b : a; } /* std::min */

typedef struct {
    int nso;        /* 2 (one-body) or 4 (two-body) */
    so_t k[4];
    double delta;
    int64_t seq;
} rec_t;

typedef struct { rec_t *v; int64_t n, cap; } recvec;

static int rv_push(recvec *rv, rec_t r) {
    if (rv->n == rv->cap) {
        int64_t nc = rv->cap ? rv->cap * 2 : 1024;
        rec_t *nv = (rec_t *)realloc(rv->v, (size_t)nc * sizeof(rec_t));
        if (!nv) return -1;
        rv->v = nv; rv->cap = nc;
    }
    r.seq = rv->n;
    rv->v[rv->n++] = r;
    return 0;
}

/* _UpdateDataModel_OneBody / _TwoBody, cpp/hamiltonian.cc:623-657: map[key] += delta,
 * skipped when delta == 0.  Recorded in arrival order; summed per key later. */
static int update_model(recvec *rv, int nso, const so_t *k, double delta) {
    if (delta == 0.0) return 0;
    rec_t r;
    memset(&r, 0, sizeof r);
    r.nso = nso;
    for (int i = 0; i < nso; ++i) r.k[i] = k[i];
    r.delta = delta;
    return rv_push(rv, r);
}

/* AddOneBodyTerm + _EnumerateSpinOrbitals_OneBody, cpp/hamiltonian.cc:301-323, 421-438 */
static int add_one_body(recvec *rv, int p, int q, double coeff) {
    if (coeff == 0.0) return 0;
    for (int spin = 0; spin < 2; ++spin) {
        so_t k[2] = {{p, spin}, {q, spin}};
        if (update_model(rv, 2, k, coeff)) return -1;
    }
    return 0;
}

/* _ReorderPQQPTerm, cpp/hamiltonian.cc:467-497 */
static double reorder_pqqp(so_t *t, double coeff) {
    so_t p = t[0], q = t[1], r = t[2], s = t[3], tmp;
    if (so_eq(r, s) || so_eq(p, q)) return 0.0;
    double c = coeff / 2.0;
    if (so_eq(p, r)) { tmp = r; r = s; s = tmp; c *= -1.0; }
    if (!(so_cmp(p, q) < 0)) { tmp = p; p = q; q = tmp; tmp = r; r = s; s = tmp; }
    t[0] = p; t[1] = q; t[2] = r; t[3] = s;
    return c;
}

/* _ReorderPQQRTerm, cpp/hamiltonian.cc:499-545 */
static double reorder_pqqr(so_t *t, double coeff, int *err) {
    so_t p = t[0], q = t[1], r = t[2], s = t[3], tmp;
    if (so_eq(p, q) || so_eq(r, s)) return 0.0;
    double c = coeff / 4.0;
    if (so_eq(q, r)) {
    } else if (so_eq(q, s)) {
        c *= -1.0; tmp = r; r = s; s = tmp;
    } else if (so_eq(p, r)) {
        c *= -1.0; tmp = p; p = q; q = tmp;
    } else if (so_eq(p, s)) {
        tmp = p; p = q; q = tmp; tmp = r; r = s; s = tmp;
    } else {
        *err = 1;
        return 0.0;
    }
    if (so_cmp(s, p) < 0) { tmp = p; p = s; s = tmp; }
    t[0] = p; t[1] = q; t[2] = r; t[3] = s;
    return c;
}

/* _ReorderPQRSTerm, cpp/hamiltonian.cc:547-619 */
static double reorder_pqrs(so_t *t, double coeff) {
    so_t p = t[0], q = t[1], r = t[2], s = t[3], tmp;
    int alphas = (p.spin == 0) + (q.spin == 0) + (r.spin == 0) + (s.spin == 0);
    double c = coeff / 4.0;
    if (so_cmp(so_min(p, q), so_min(r, s)) < 0) {
    } else {
        tmp = p; p = r; r = tmp; tmp = q; q = s; s = tmp;
    }
    if (alphas == 2) {
        if (q.spin == 0) { tmp = p; p = q; q = tmp; c *= -1.0; }
        if (r.spin == 0) { tmp = r; r = s; s = tmp; c *= -1.0; }
    } else {
        if (q.orb < p.orb) { tmp = p; p = q; q = tmp; c *= -1.0; }
        if (r.orb < s.orb) { tmp = r; r = s; s = tmp; c *= -1.0; }
    }
    t[0] = p; t[1] = q; t[2] = r; t[3] = s;
    return c;
}

static int tup_cmp4(const int *a, const int *b) {
    for (int i = 0; i < 4; ++i)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

/* AddTwoBodyTerm, cpp/hamiltonian.cc:325-398, with _EnumerateTwoBodyPermutations
 * (:400-417, a std::set, so iterated in sorted order) and
 * _EnumerateSpinOrbitals_TwoBody (:440-463: inner spin outer loop, outer spin inner loop). */
static int add_two_body(recvec *rv, int p, int q, int r, int s, double coeff) {
    if (coeff == 0.0) return 0;
    if (p == q && q == r && r == s) {
        so_t k[4] = {{p, 0}, {p, 1}, {p, 1}, {p, 0}};
        return update_model(rv, 4, k, coeff);
    }
    int i = p, j = q, k = r, l = s;
    int perms[8][4] = {{i, j, k, l}, {l, j, k, i}, {i, k, j, l}, {l, k, j, i},
                       {j, i, l, k}, {k, i, l, j}, {j, l, i, k}, {k, l, i, j}};
    /* sort + dedup == std::set iteration */
    for (int a = 1; a < 8; ++a)
        for (int b = a; b > 0 && tup_cmp4(perms[b], perms[b - 1]) < 0; --b) {
            int tmp[4];
            memcpy(tmp, perms[b], sizeof tmp);
            memcpy(perms[b], perms[b - 1], sizeof tmp);
            memcpy(perms[b - 1], tmp, sizeof tmp);
        }
    for (int a = 0; a < 8; ++a) {
        if (a > 0 && tup_cmp4(perms[a], perms[a - 1]) == 0) continue;
        const int *pm = perms[a];
        for (int inner = 0; inner < 2; ++inner) {
            for (int outer = 0; outer < 2; ++outer) {
                so_t t[4] = {{pm[0], outer}, {pm[1], inner}, {pm[2], inner}, {pm[3], outer}};
                /* number of distinct spin-orbitals (std::set<SpinOrbital>) */
                int uniq = 0;
                for (int x = 0; x < 4; ++x) {
                    int seen = 0;
                    for (int y = 0; y < x; ++y) seen |= so_eq(t[x], t[y]);
                    uniq += !seen;
                }
                double c;
                int err = 0;
                if (uniq == 2) c = reorder_pqqp(t, coeff);
                else if (uniq == 3) c = reorder_pqqr(t, coeff, &err);
                else if (uniq == 4) c = reorder_pqrs(t, coeff);
                else return -1; /* "Unknown type not caught" */
                if (err) return -1;
                if (update_model(rv, 4, t, c)) return -1;
            }
        }
    }
    return 0;
}

static int rec_cmp(const void *va, const void *vb) {
    const rec_t *a = (const rec_t *)va, *b = (const rec_t *)vb;
    if (a->nso != b->nso) return a->nso < b->nso ? -1 : 1;
    for (int i = 0; i < a->nso; ++i) {
        int c = so_cmp(a->k[i], b->k[i]);
        if (c) return c;
    }
    if (a->seq != b->seq) return a->seq < b->seq ? -1 : 1;
    return 0;
}

/* ConvertToInt(HalfUp), cpp/types.hpp:50-55 */
static int half_up(so_t s, int n_orb) { return s.spin * n_orb + s.orb - 1; }

int64_t orc_compile_terms(int n_orb,
                          int64_t n_one, const int32_t *one_pq, const double *one_val,
                          int64_t n_two, const int32_t *two_abcd, const double *two_val,
                          orc_term *out, int64_t cap, int64_t counts[5]) {
    recvec rv = {0, 0, 0};
    int64_t nout = -1;
    /* IngestBroombridge one-body loop, cpp/hamiltonian.cc:22-37: swap so p >= q */
    for (int64_t x = 0; x < n_one; ++x) {
        int p = one_pq[2 * x], q = one_pq[2 * x + 1];
        if (q > p) { int t = p; p = q; q = t; }
        if (add_one_body(&rv, p, q, one_val[x])) goto done;
    }
    /* two-body loop, cpp/hamiltonian.cc:40-52: [a,b,c,d] -> i=a,l=b,j=c,k=d -> (i,j,k,l) */
    for (int64_t x = 0; x < n_two; ++x) {
        const int32_t *e = two_abcd + 4 * x;
        int i = e[0], l = e[1], j = e[2], k = e[3];
        if (add_two_body(&rv, i, j, k, l, two_val[x])) goto done;
    }
    qsort(rv.v, (size_t)rv.n, sizeof(rec_t), rec_cmp);

    /* accumulate per key in arrival order (std::map operator[] starts at 0.0),
     * then IndexTerms(HalfUp) and segment (cpp/hamiltonian.cc:253-299).  The
     * indexed-map order equals the SpinOrbital-map order because HalfUp is
     * monotone in (spin, orbital). */
    int64_t ncls[5] = {0, 0, 0, 0, 0};
    orc_term *tmp = (orc_term *)malloc((size_t)(rv.n + 1) * sizeof(orc_term));
    int64_t nt = 0;
    if (!tmp) goto done;
    for (int64_t a = 0; a < rv.n;) {
        int64_t b = a;
        double sum = 0.0;
        while (b < rv.n && rv.v[b].nso == rv.v[a].nso) {
            int same = 1;
            for (int z = 0; z < rv.v[a].nso; ++z) same &= so_eq(rv.v[a].k[z], rv.v[b].k[z]);
            if (!same) break;
            sum = sum + rv.v[b].delta;
            ++b;
        }
        orc_term t;
        memset(&t, 0, sizeof t);
        t.coeff = sum;
        if (rv.v[a].nso == 2) {
            t.idx[0] = half_up(rv.v[a].k[0], n_orb);
            t.idx[1] = half_up(rv.v[a].k[1], n_orb);
            t.idx[2] = t.idx[3] = -1;
            t.cls = (t.idx[0] == t.idx[1]) ? ORC_PP : ORC_PQ;
        } else {
            for (int z = 0; z < 4; ++z) t.idx[z] = half_up(rv.v[a].k[z], n_orb);
            if (t.idx[1] == t.idx[2]) t.cls = (t.idx[0] == t.idx[3]) ? ORC_PQQP : ORC_PQQR;
            else t.cls = ORC_PQRS;
        }
        tmp[nt++] = t;
        ncls[t.cls]++;
        a = b;
    }
    /* stable partition by class (within a class the map order is kept) */
    int64_t w = 0;
    for (int c = 0; c < 5; ++c)
        for (int64_t x = 0; x < nt; ++x)
            if (tmp[x].cls == c) {
                if (w < cap) out[w] = tmp[x];
                ++w;
            }
    free(tmp);
    for (int c = 0; c < 5; ++c) counts[c] = ncls[c];
    nout = w;
done:
    free(rv.v);
    return nout;
}

/* ------------------------------------------------------------------------ */
/* run tables: cpp/time_evolution_over_terms.hpp:1160-1411                    */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t niter, off01, off10; unsigned mid_parity; } part_t;

static void sort_u(unsigned *v, int n) {
    for (int a = 1; a < n; ++a)
        for (int b = a; b > 0 && v[b] < v[b - 1]; --b) { unsigned t = v[b]; v[b] = v[b - 1]; v[b - 1] = t; }
}

static uint64_t lo_niter(uint64_t lo, unsigned bound) {
    return lo ? orc_nchoosek((int)bound, popc(lo)) : 1;
}

static void advance(uint64_t *bv, uint64_t niter) {
    if (*bv == 0) *bv = 1;
    for (uint64_t i = 0; i < niter; ++i) *bv = orc_next_comb(*bv);
}

/* discoverCCAA_BVs, TEOT:1160-1210 */
static int64_t disc_ccaa(part_t *pt, uint64_t size, unsigned p, unsigned q, unsigned r, unsigned s, unsigned e) {
    unsigned id[4] = {p, q, r, s};
    sort_u(id, 4);
    uint64_t ctr = 0, bv = (1ull << (e - 2)) - 1;
    int64_t np = 0;
    while (ctr < size) {
        uint64_t lo = bv & ((1ull << id[0]) - 1);
        uint64_t midlo = (bv & ((1ull << (id[1] - 1)) - 1)) >> id[0];
        uint64_t midmid = (bv & ((1ull << (id[2] - 2)) - 1)) >> (id[1] - 1);
        uint64_t midhi = (bv & ((1ull << (id[3] - 3)) - 1)) >> (id[2] - 2);
        uint64_t hi = bv >> (id[3] - 3);
        uint64_t bvb = lo | (midlo << (id[0] + 1)) | (midmid << (id[1] + 1)) | (midhi << (id[2] + 1)) | (hi << (id[3] + 1));
        pt[np].off01 = orc_rank(bvb | (1ull << r) | (1ull << s));
        pt[np].off10 = orc_rank(bvb | (1ull << q) | (1ull << p));
        pt[np].mid_parity = orc_parity(midlo) + orc_parity(midhi);
        uint64_t ni = lo_niter(lo, id[0]);
        pt[np].niter = ni;
        ++np;
        ctr += ni;
        advance(&bv, ni);
    }
    return np;
}

/* discoverCNA_BVs, TEOT:1212-1274 */
static int64_t disc_cna(part_t *pt, uint64_t size, unsigned p, unsigned q, unsigned r, unsigned e) {
    unsigned id[3] = {p, q, r};
    sort_u(id, 3);
    uint64_t ctr = 0, bv = (1ull << (e - 2)) - 1;
    int64_t np = 0;
    while (ctr < size) {
        uint64_t lo = bv & ((1ull << id[0]) - 1);
        uint64_t midlo = (bv & ((1ull << (id[1] - 1)) - 1)) >> id[0];
        uint64_t midhi = (bv & ((1ull << (id[2] - 2)) - 1)) >> (id[1] - 1);
        uint64_t hi = bv >> (id[2] - 2);
        uint64_t bvb = lo | (midlo << (id[0] + 1)) | (midhi << (id[1] + 1)) | (hi << (id[2] + 1));
        pt[np].off01 = orc_rank(bvb | (1ull << q) | (1ull << r));
        pt[np].off10 = orc_rank(bvb | (1ull << q) | (1ull << p));
        if (q > r) pt[np].mid_parity = orc_parity(midlo);
        else if (q < p) pt[np].mid_parity = orc_parity(midhi);
        else pt[np].mid_parity = orc_parity(midlo) + orc_parity(midhi) + 1;
        uint64_t ni = lo_niter(lo, id[0]);
        pt[np].niter = ni;
        ++np;
        ctr += ni;
        advance(&bv, ni);
    }
    return np;
}

/* discoverCA_BVs, TEOT:1276-1324 */
static int64_t disc_ca(part_t *pt, uint64_t size, unsigned p, unsigned r, unsigned e) {
    unsigned np_ = p < r ? p : r, nr = p < r ? r : p;
    uint64_t ctr = 0, bv = (1ull << (e - 1)) - 1;
    int64_t np = 0;
    while (ctr < size) {
        uint64_t lo = bv & ((1ull << np_) - 1);
        uint64_t mid = (bv & ((1ull << (nr - 1)) - 1)) >> np_;
        uint64_t hi = bv >> (nr - 1);
        uint64_t bvb = lo | (mid << (np_ + 1)) | (hi << (nr + 1));
        pt[np].off01 = orc_rank(bvb | (1ull << r));
        pt[np].off10 = orc_rank(bvb | (1ull << p));
        pt[np].mid_parity = orc_parity(mid);
        uint64_t ni = lo_niter(lo, np_);
        pt[np].niter = ni;
        ++np;
        ctr += ni;
        advance(&bv, ni);
    }
    return np;
}

/* discoverNN_BVs, TEOT:1326-1369 (off10 and mid_parity stay value-initialised 0) */
static int64_t disc_nn(part_t *pt, uint64_t size, unsigned p, unsigned r, unsigned e) {
    uint64_t ctr = 0, bv = (1ull << (e - 2)) - 1;
    int64_t np = 0;
    while (ctr < size) {
        uint64_t lo = bv & ((1ull << p) - 1);
        uint64_t mid = (bv & ((1ull << (r - 1)) - 1)) >> p;
        uint64_t hi = bv >> (r - 1);
        uint64_t bvb = lo | (mid << (p + 1)) | (hi << (r + 1));
        pt[np].off01 = orc_rank(bvb | (1ull << r) | (1ull << p));
        pt[np].off10 = 0;
        pt[np].mid_parity = 0;
        uint64_t ni = lo_niter(lo, p);
        pt[np].niter = ni;
        ++np;
        ctr += ni;
        advance(&bv, ni);
    }
    return np;
}

/* discoverN_BVs, TEOT:1371-1411 */
static int64_t disc_n(part_t *pt, uint64_t size, unsigned q, unsigned e) {
    uint64_t ctr = 0, bv = (1ull << (e - 1)) - 1;
    int64_t np = 0;
    while (ctr < size) {
        uint64_t lo = bv & ((1ull << q) - 1);
        uint64_t hi = bv >> q;
        uint64_t bvb = lo | (hi << (q + 1));
        uint64_t o = orc_rank(bvb | (1ull << q));
        pt[np].off01 = o;
        pt[np].off10 = o;
        pt[np].mid_parity = 0;
        uint64_t ni = lo_niter(lo, q);
        pt[np].niter = ni;
        ++np;
        ctr += ni;
        advance(&bv, ni);
    }
    return np;
}

static uint64_t table_size(int kind, int n, int e) {
    switch (kind) {
    case ORC_N: return orc_nchoosek(n - 1, e - 1);
    case ORC_NN: return orc_nchoosek(n - 2, e - 2);
    case ORC_CA: return orc_nchoosek(n - 2, e - 1);
    case ORC_CNA: return orc_nchoosek(n - 3, e - 2);
    case ORC_CCAA: return orc_nchoosek(n - 4, e - 2);
    }
    return 0;
}

static int64_t discover(int kind, int n, int e, const unsigned *a, part_t **out) {
    uint64_t size = table_size(kind, n, e);
    part_t *pt = (part_t *)calloc((size_t)(size ? size : 1), sizeof(part_t));
    *out = pt;
    if (!pt) return -1;
    if (size == 0) return 0;
    switch (kind) {
    case ORC_N: return disc_n(pt, size, a[0], (unsigned)e);
    case ORC_NN: return disc_nn(pt, size, a[0], a[1], (unsigned)e);
    case ORC_CA: return disc_ca(pt, size, a[0], a[1], (unsigned)e);
    case ORC_CNA: return disc_cna(pt, size, a[0], a[1], a[2], (unsigned)e);
    case ORC_CCAA: return disc_ccaa(pt, size, a[0], a[1], a[2], a[3], (unsigned)e);
    }
    return -1;
}

int64_t orc_run_table(int kind, int n, int e, const int32_t *args,
                      uint64_t *off01, uint64_t *off10, uint64_t *niter,
                      uint32_t *mid_parity, int64_t cap) {
    unsigned a[4] = {0, 0, 0, 0};
    int na = kind == ORC_N ? 1 : kind == ORC_CCAA ? 4 : kind == ORC_CNA ? 3 : 2;
    for (int i = 0; i < na; ++i) a[i] = (unsigned)args[i];
    part_t *pt = NULL;
    int64_t np = discover(kind, n, e, a, &pt);
    if (np < 0) { free(pt); return -1; }
    for (int64_t i = 0; i < np && i < cap; ++i) {
        off01[i] = pt[i].off01;
        off10[i] = pt[i].off10;
        niter[i] = pt[i].niter;
        mid_parity[i] = pt[i].mid_parity;
    }
    free(pt);
    return np;
}

/* ------------------------------------------------------------------------ */
/* evolution: cpp/time_evolution_over_terms.hpp:374-1158                      */
/* ------------------------------------------------------------------------ */

typedef struct { double re, im; } cplx;

static inline cplx cmul(cplx a, cplx b) {
    cplx r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}
static inline cplx cadd(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }

typedef struct {
    int n, e;
    uint64_t C;   /* e_of_n: strings per spin */
    cplx *blk;    /* one block, [alpha][beta] */
} ctx_t;

/* updateStateVector, TEOT:565-589 */
static inline void rot(cplx *sv, const cplx *c, uint64_t l, uint64_t r) {
    cplx vl = sv[l], vr = sv[r];
    sv[l] = cadd(cmul(vl, c[0]), cmul(vr, c[1]));
    sv[r] = cadd(cmul(vl, c[1]), cmul(vr, c[0]));
}

/* pp_term / pqqp_term: exp(-i c t), cpp/hamiltonian.cc:661-676, 708-729 */
static cplx phase_coeff(double c, double t) {
    double a = c * t;
    cplx r = {cos(a), -sin(a)};
    return r;
}

/* pq_term / pqqr_term / pqrs_term: {cos ct, -i(-1)^par sin ct, ., .},
 * cpp/hamiltonian.cc:678-706, 731-783 */
static void rot_coeffs(double c, double t, cplx even[2], cplx odd[2]) {
    double a = c * t;
    double cs = cos(a), sn = sin(a);
    even[0].re = cs; even[0].im = 0.0;
    even[1].re = 0.0; even[1].im = -sn;
    odd[0] = even[0];
    odd[1].re = 0.0; odd[1].im = sn;
}

/* _updateAAAA, TEOT:591-658 (pairs of whole beta rows) */
static void upd_aaaa(ctx_t *x, const part_t *pa, int64_t npa, const cplx *ev, const cplx *od) {
    uint64_t C = x->C;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t ip = 0; ip < npa; ++ip) {
        const cplx *c = (pa[ip].mid_parity & 1) ? od : ev;
        for (uint64_t na = 0; na < pa[ip].niter; ++na) {
            uint64_t s01 = (pa[ip].off01 + na) * C, s10 = (pa[ip].off10 + na) * C;
            for (uint64_t b = 0; b < C; ++b) rot(x->blk, c, s01 + b, s10 + b);
        }
    }
}

/* _updateBBBB, TEOT:660-729 (pairs inside every beta row) */
static void upd_bbbb(ctx_t *x, const part_t *pb, int64_t npb, const cplx *ev, const cplx *od) {
    uint64_t C = x->C;
    #pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < (int64_t)C; ++a) {
        uint64_t row = (uint64_t)a * C;
        for (int64_t ip = 0; ip < npb; ++ip) {
            const cplx *c = (pb[ip].mid_parity & 1) ? od : ev;
            for (uint64_t nb = 0; nb < pb[ip].niter; ++nb)
                rot(x->blk, c, row + pb[ip].off01 + nb, row + pb[ip].off10 + nb);
        }
    }
}

/* _updateABBA, TEOT:731-839 (parity = alpha + beta mid parity) */
static void upd_abba(ctx_t *x, const part_t *pa, int64_t npa, const part_t *pb, int64_t npb,
                     const cplx *ev, const cplx *od) {
    uint64_t C = x->C;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t ip = 0; ip < npa; ++ip) {
        for (uint64_t na = 0; na < pa[ip].niter; ++na) {
            uint64_t s01 = (pa[ip].off01 + na) * C, s10 = (pa[ip].off10 + na) * C;
            for (int64_t jb = 0; jb < npb; ++jb) {
                unsigned par = 1u & (pa[ip].mid_parity + pb[jb].mid_parity);
                const cplx *c = par ? od : ev;
                uint64_t i01 = s01 + pb[jb].off01, i10 = s10 + pb[jb].off10;
                for (uint64_t nb = 0; nb < pb[jb].niter; ++nb) rot(x->blk, c, i01 + nb, i10 + nb);
            }
        }
    }
}

/* updatePP_AA / updatePQQP_AA: diagonal phase on whole rows (TEOT:886-903, 988-1013) */
static void diag_rows(ctx_t *x, const part_t *pa, int64_t npa, cplx cf) {
    uint64_t C = x->C;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t ip = 0; ip < npa; ++ip)
        for (uint64_t na = 0; na < pa[ip].niter; ++na) {
            uint64_t row = (pa[ip].off01 + na) * C;
            for (uint64_t b = 0; b < C; ++b) x->blk[row + b] = cmul(x->blk[row + b], cf);
        }
}

/* updatePP_BB / updatePQQP_BB: diagonal phase on runs in every row (TEOT:904-925, 1014-1039) */
static void diag_cols(ctx_t *x, const part_t *pb, int64_t npb, cplx cf) {
    uint64_t C = x->C;
    #pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < (int64_t)C; ++a)
        for (int64_t ip = 0; ip < npb; ++ip)
            for (uint64_t nb = 0; nb < pb[ip].niter; ++nb) {
                uint64_t i = (uint64_t)a * C + pb[ip].off01 + nb;
                x->blk[i] = cmul(x->blk[i], cf);
            }
}

/* updatePQQP_AB, TEOT:1040-1071 */
static void diag_ab(ctx_t *x, const part_t *pa, int64_t npa, const part_t *pb, int64_t npb, cplx cf) {
    uint64_t C = x->C;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t ip = 0; ip < npa; ++ip)
        for (uint64_t na = 0; na < pa[ip].niter; ++na) {
            uint64_t row = (pa[ip].off01 + na) * C;
            for (int64_t jb = 0; jb < npb; ++jb)
                for (uint64_t nb = 0; nb < pb[jb].niter; ++nb) {
                    uint64_t i = row + pb[jb].off01 + nb;
                    x->blk[i] = cmul(x->blk[i], cf);
                }
        }
}

/* updatePQ_AA, TEOT:927-957 (serial in the reference) */
static void upd_pq_aa(ctx_t *x, const part_t *pa, int64_t npa, const cplx *ev, const cplx *od) {
    uint64_t C = x->C;
    for (int64_t ip = 0; ip < npa; ++ip) {
        const cplx *c = (pa[ip].mid_parity & 1) ? od : ev;
        for (uint64_t na = 0; na < pa[ip].niter; ++na) {
            uint64_t i01 = (pa[ip].off01 + na) * C, i10 = (pa[ip].off10 + na) * C;
            for (uint64_t b = 0; b < C; ++b) rot(x->blk, c, i01 + b, i10 + b);
        }
    }
}

/* updatePQ_BB, TEOT:958-986 (serial in the reference) */
static void upd_pq_bb(ctx_t *x, const part_t *pb, int64_t npb, const cplx *ev, const cplx *od) {
    uint64_t C = x->C;
    for (uint64_t a = 0; a < C; ++a) {
        uint64_t row = a * C;
        for (int64_t ip = 0; ip < npb; ++ip) {
            const cplx *c = (pb[ip].mid_parity & 1) ? od : ev;
            for (uint64_t nb = 0; nb < pb[ip].niter; ++nb)
                rot(x->blk, c, row + pb[ip].off01 + nb, row + pb[ip].off10 + nb);
        }
    }
}

/* updatePQQR_BAAB, TEOT:1115-1158: alpha N(q) rows x beta CA(p,r) pairs */
static void upd_pqqr_baab(ctx_t *x, const part_t *pa, int64_t npa, const part_t *pb, int64_t npb,
                          const cplx *ev, const cplx *od) {
    uint64_t C = x->C;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t ip = 0; ip < npa; ++ip)
        for (uint64_t na = 0; na < pa[ip].niter; ++na) {
            uint64_t row = (pa[ip].off01 + na) * C;
            for (int64_t jb = 0; jb < npb; ++jb) {
                const cplx *c = (pb[jb].mid_parity & 1) ? od : ev;
                for (uint64_t nb = 0; nb < pb[jb].niter; ++nb)
                    rot(x->blk, c, row + pb[jb].off01 + nb, row + pb[jb].off10 + nb);
            }
        }
}

/* one pqqr term, dispatch of TEOT:470-509 */
static int apply_pqqr(ctx_t *x, unsigned p_, unsigned q_, unsigned r_, double c, double t) {
    unsigned n = (unsigned)x->n, e = (unsigned)x->e;
    cplx ev[2], od[2];
    rot_coeffs(c, t, ev, od);
    part_t *pa = NULL, *pb = NULL;
    int64_t na = 0, nb = 0;
    if (p_ < n && q_ < n) {                     /* updatePQQR_AAAA, TEOT:1073-1086 */
        if (e < 2) return 0;
        unsigned a[3] = {p_, q_, r_};
        na = discover(ORC_CNA, x->n, x->e, a, &pa);
        upd_aaaa(x, pa, na, ev, od);
    } else if (p_ >= n && q_ >= n) {            /* updatePQQR_BBBB, TEOT:1087-1099 */
        if (e < 2) return 0;
        unsigned a[3] = {p_ - n, q_ - n, r_ - n};
        nb = discover(ORC_CNA, x->n, x->e, a, &pb);
        upd_bbbb(x, pb, nb, ev, od);
    } else if (q_ >= n) {                       /* updatePQQR_ABBA, TEOT:1100-1113 */
        unsigned a[2] = {p_, r_}, b[1] = {q_ - n};
        na = discover(ORC_CA, x->n, x->e, a, &pa);
        nb = discover(ORC_N, x->n, x->e, b, &pb);
        upd_abba(x, pa, na, pb, nb, ev, od);
    } else {                                    /* updatePQQR_BAAB, TEOT:1115-1158 */
        unsigned a[1] = {q_}, b[2] = {p_ - n, r_ - n};
        na = discover(ORC_N, x->n, x->e, a, &pa);
        nb = discover(ORC_CA, x->n, x->e, b, &pb);
        upd_pqqr_baab(x, pa, na, pb, nb, ev, od);
    }
    free(pa);
    free(pb);
    return 0;
}

/* terms are grouped class by class: [b0[c], b1[c]) holds class c */
static void class_bounds(const orc_term *terms, int64_t nterms, int64_t *b0, int64_t *b1) {
    for (int c = 0; c < 5; ++c) {
        int64_t s = -1, z = 0;
        for (int64_t i = 0; i < nterms; ++i)
            if (terms[i].cls == c) { if (s < 0) s = i; z = i + 1; }
        b0[c] = s < 0 ? 0 : s;
        b1[c] = s < 0 ? 0 : z;
    }
}

/* pp: updatePP_AA | updatePP_BB, TEOT:376-395 (TEOT:176-195 in applyUnitary) */
static void apply_pp_list(ctx_t *x, const orc_term *terms, int64_t i0, int64_t i1, double t) {
    unsigned n = (unsigned)x->n;
    for (int64_t i = i0; i < i1; ++i) {
        unsigned p = (unsigned)terms[i].idx[0];
        cplx cf = phase_coeff(terms[i].coeff, t);
        part_t *pt = NULL;
        unsigned a[1] = {p < n ? p : p - n};
        int64_t np = discover(ORC_N, x->n, x->e, a, &pt);
        if (p < n) diag_rows(x, pt, np, cf);
        else diag_cols(x, pt, np, cf);
        free(pt);
    }
}

/* pqqp: updatePQQP_AA | _BB | _AB, TEOT:399-424 (TEOT:225-250 in applyUnitary) */
static void apply_pqqp_list(ctx_t *x, const orc_term *terms, int64_t i0, int64_t i1, double t) {
    unsigned n = (unsigned)x->n, e = (unsigned)x->e;
    for (int64_t i = i0; i < i1; ++i) {
        unsigned p = (unsigned)terms[i].idx[0], q = (unsigned)terms[i].idx[1];
        cplx cf = phase_coeff(terms[i].coeff, t);
        if (p < n && q < n) {
            if (e < 2) continue;
            part_t *pt = NULL;
            unsigned a[2] = {p, q};
            int64_t np = discover(ORC_NN, x->n, x->e, a, &pt);
            diag_rows(x, pt, np, cf);
            free(pt);
        } else if (p >= n && q >= n) {
            if (e < 2) continue;
            part_t *pt = NULL;
            unsigned a[2] = {p - n, q - n};
            int64_t np = discover(ORC_NN, x->n, x->e, a, &pt);
            diag_cols(x, pt, np, cf);
            free(pt);
        } else {
            part_t *pa = NULL, *pb = NULL;
            unsigned a[1] = {p}, b[1] = {q - n};
            int64_t na = discover(ORC_N, x->n, x->e, a, &pa);
            int64_t nb = discover(ORC_N, x->n, x->e, b, &pb);
            diag_ab(x, pa, na, pb, nb, cf);
            free(pa);
            free(pb);
        }
    }
}

/* one pq term: updatePQ_AA | updatePQ_BB, TEOT:428-451 (TEOT:199-221 in applyUnitary) */
static void apply_pq(ctx_t *x, const orc_term *T, double t) {
    unsigned n = (unsigned)x->n;
    unsigned p = (unsigned)T->idx[0], q = (unsigned)T->idx[1];
    cplx ev[2], od[2];
    rot_coeffs(T->coeff, t, ev, od);
    part_t *pt = NULL;
    if (p < n && q < n) {
        unsigned a[2] = {p, q};
        int64_t np = discover(ORC_CA, x->n, x->e, a, &pt);
        upd_pq_aa(x, pt, np, ev, od);
    } else {
        unsigned a[2] = {p - n, q - n};
        int64_t np = discover(ORC_CA, x->n, x->e, a, &pt);
        upd_pq_bb(x, pt, np, ev, od);
    }
    free(pt);
}

/* pqrs: updatePQRS_AAAA | _BBBB | _ABBA, TEOT:519-560 (TEOT:313-357 in applyUnitary) */
static void apply_pqrs_list(ctx_t *x, const orc_term *terms, int64_t i0, int64_t i1, double t) {
    unsigned n = (unsigned)x->n, e = (unsigned)x->e;
    for (int64_t i = i0; i < i1; ++i) {
        unsigned p = (unsigned)terms[i].idx[0], q = (unsigned)terms[i].idx[1];
        unsigned r = (unsigned)terms[i].idx[2], s = (unsigned)terms[i].idx[3];
        cplx ev[2], od[2];
        rot_coeffs(terms[i].coeff, t, ev, od);
        if (p < n && q < n && r < n && s < n) {          /* TEOT:841-858 */
            if (e < 2) continue;
            part_t *pt = NULL;
            unsigned a[4] = {p, q, r, s};
            int64_t np = discover(ORC_CCAA, x->n, x->e, a, &pt);
            upd_aaaa(x, pt, np, ev, od);
            free(pt);
        } else if (p >= n && q >= n && r >= n && s >= n) { /* TEOT:860-871 */
            if (e < 2) continue;
            part_t *pt = NULL;
            unsigned a[4] = {p - n, q - n, r - n, s - n};
            int64_t np = discover(ORC_CCAA, x->n, x->e, a, &pt);
            upd_bbbb(x, pt, np, ev, od);
            free(pt);
        } else {                                           /* TEOT:873-884 */
            part_t *pa = NULL, *pb = NULL;
            unsigned a[2] = {p, s}, b[2] = {q - n, r - n};
            int64_t na = discover(ORC_CA, x->n, x->e, a, &pa);
            int64_t nb = discover(ORC_CA, x->n, x->e, b, &pb);
            upd_abba(x, pa, na, pb, nb, ev, od);
            free(pa);
            free(pb);
        }
    }
}

/* applyUnitary_reordered, TEOT:374-562: pp, pqqp, then each pq followed by the
 * pqqr terms whose (p_, r_) == (q, p) (TEOT:453-465), then pqrs */
static int apply_unitary(ctx_t *x, const orc_term *terms, int64_t nterms, double t) {
    int64_t b0[5], b1[5];
    class_bounds(terms, nterms, b0, b1);
    apply_pp_list(x, terms, b0[ORC_PP], b1[ORC_PP], t);
    apply_pqqp_list(x, terms, b0[ORC_PQQP], b1[ORC_PQQP], t);
    for (int64_t i = b0[ORC_PQ]; i < b1[ORC_PQ]; ++i) {
        unsigned p = (unsigned)terms[i].idx[0], q = (unsigned)terms[i].idx[1];
        apply_pq(x, &terms[i], t);
        for (int64_t k = b0[ORC_PQQR]; k < b1[ORC_PQQR]; ++k) {
            unsigned p_ = (unsigned)terms[k].idx[0], q_ = (unsigned)terms[k].idx[1];
            unsigned r_ = (unsigned)terms[k].idx[3];
            if (p_ == q && r_ == p) apply_pqqr(x, p_, q_, r_, terms[k].coeff, t);
        }
    }
    apply_pqrs_list(x, terms, b0[ORC_PQRS], b1[ORC_PQRS], t);
    return 0;
}

/* applyUnitary, TEOT:174-372: pp, every pq, pqqp, every pqqr (TEOT:254-301,
 * not gated by a pq term), pqrs */
static int apply_unitary_original(ctx_t *x, const orc_term *terms, int64_t nterms, double t) {
    int64_t b0[5], b1[5];
    class_bounds(terms, nterms, b0, b1);
    apply_pp_list(x, terms, b0[ORC_PP], b1[ORC_PP], t);
    for (int64_t i = b0[ORC_PQ]; i < b1[ORC_PQ]; ++i) apply_pq(x, &terms[i], t);
    apply_pqqp_list(x, terms, b0[ORC_PQQP], b1[ORC_PQQP], t);
    for (int64_t k = b0[ORC_PQQR]; k < b1[ORC_PQQR]; ++k)
        apply_pqqr(x, (unsigned)terms[k].idx[0], (unsigned)terms[k].idx[1], (unsigned)terms[k].idx[3],
                   terms[k].coeff, t);
    apply_pqrs_list(x, terms, b0[ORC_PQRS], b1[ORC_PQRS], t);
    return 0;
}

int orc_apply_unitary(int n, int e, const orc_term *terms, int64_t n_terms, double t, double *block) {
    ctx_t x;
    x.n = n; x.e = e; x.C = orc_nchoosek(n, e); x.blk = (cplx *)block;
    return apply_unitary(&x, terms, n_terms, t);
}

int orc_apply_unitary_original(int n, int e, const orc_term *terms, int64_t n_terms, double t, double *block) {
    ctx_t x;
    x.n = n; x.e = e; x.C = orc_nchoosek(n, e); x.blk = (cplx *)block;
    return apply_unitary_original(&x, terms, n_terms, t);
}

/* time_evolve_iqpe, TEOT:141-155: applyUnitary(..., exp = 1) num_of_applications
 * times, in place on block 1 (here: on the given block) */
int orc_time_evolve_iqpe(int n, int e, const orc_term *terms, int64_t n_terms, double t, double *block,
                         int64_t num_of_applications) {
    if (e < 1 || n < e || num_of_applications < 0) return -1;
    for (int64_t k = 0; k < num_of_applications; ++k)
        if (orc_apply_unitary_original(n, e, terms, n_terms, t, block)) return -1;
    return 0;
}

/* time_evolve_over_terms, TEOT:92-139: block j = copy(block j-1); applyUnitary_reordered
 * (original != 0: the applyUnitary call left commented out at TEOT:135) */
int orc_time_evolve_ordered(int n, int e, int m, const orc_term *terms, int64_t n_terms,
                            double t, double *state, int64_t j_begin, int64_t j_end, int original) {
    if (e < 1 || n < e || m < 1 || j_begin < 1) return -1;
    uint64_t C = orc_nchoosek(n, e), D = C * C, M = 1ull << m;
    if ((uint64_t)j_end > M) return -1;
    cplx *sv = (cplx *)state;
    for (int64_t j = j_begin; j < j_end; ++j) {
        memcpy(sv + (uint64_t)j * D, sv + (uint64_t)(j - 1) * D, (size_t)D * sizeof(cplx));
        ctx_t x;
        x.n = n; x.e = e; x.C = C; x.blk = sv + (uint64_t)j * D;
        if ((original ? apply_unitary_original : apply_unitary)(&x, terms, n_terms, t)) return -1;
    }
    return 0;
}

int orc_time_evolve(int n, int e, int m, const orc_term *terms, int64_t n_terms,
                    double t, double *state, int64_t j_begin, int64_t j_end) {
    return orc_time_evolve_ordered(n, e, m, terms, n_terms, t, state, j_begin, j_end, 0);
}

static uint64_t parts_total(const part_t *p, int64_t np) {
    uint64_t s = 0;
    for (int64_t i = 0; i < np; ++i) s += p[i].niter;
    return s;
}

/* det-term updates: one per updateStateVector call or diagonal multiply */
int orc_count_updates(int n, int e, const orc_term *terms, int64_t n_terms,
                      uint64_t *diag_updates, uint64_t *pair_updates) {
    unsigned un = (unsigned)n;
    uint64_t C = orc_nchoosek(n, e), dg = 0, pr = 0;
    for (int64_t i = 0; i < n_terms; ++i) {
        const orc_term *T = &terms[i];
        unsigned p = (unsigned)T->idx[0], q = (unsigned)T->idx[1];
        unsigned r = (unsigned)T->idx[2], s = (unsigned)T->idx[3];
        part_t *pa = NULL, *pb = NULL;
        int64_t na = 0, nb = 0;
        if (T->cls == ORC_PP) {
            unsigned a[1] = {p < un ? p : p - un};
            na = discover(ORC_N, n, e, a, &pa);
            dg += parts_total(pa, na) * C;
        } else if (T->cls == ORC_PQQP) {
            if (p < un && q < un) {
                if (e >= 2) { unsigned a[2] = {p, q}; na = discover(ORC_NN, n, e, a, &pa); dg += parts_total(pa, na) * C; }
            } else if (p >= un && q >= un) {
                if (e >= 2) { unsigned a[2] = {p - un, q - un}; na = discover(ORC_NN, n, e, a, &pa); dg += parts_total(pa, na) * C; }
            } else {
                unsigned a[1] = {p}, b[1] = {q - un};
                na = discover(ORC_N, n, e, a, &pa);
                nb = discover(ORC_N, n, e, b, &pb);
                dg += parts_total(pa, na) * parts_total(pb, nb);
            }
        } else if (T->cls == ORC_PQ) {
            unsigned a[2] = {p < un ? p : p - un, q < un ? q : q - un};
            na = discover(ORC_CA, n, e, a, &pa);
            pr += parts_total(pa, na) * C;
        } else if (T->cls == ORC_PQQR) {
            /* applied only when a pq term with key (r_, p_) exists (TEOT:453-465) */
            unsigned p_ = p, q_ = q, r_ = (unsigned)T->idx[3];
            int gated = 0;
            for (int64_t k = 0; k < n_terms; ++k)
                if (terms[k].cls == ORC_PQ && (unsigned)terms[k].idx[0] == r_ && (unsigned)terms[k].idx[1] == p_) gated = 1;
            if (gated) {
                if (p_ < un && q_ < un) {
                    if (e >= 2) { unsigned a[3] = {p_, q_, r_}; na = discover(ORC_CNA, n, e, a, &pa); pr += parts_total(pa, na) * C; }
                } else if (p_ >= un && q_ >= un) {
                    if (e >= 2) { unsigned a[3] = {p_ - un, q_ - un, r_ - un}; na = discover(ORC_CNA, n, e, a, &pa); pr += parts_total(pa, na) * C; }
                } else if (q_ >= un) {
                    unsigned a[2] = {p_, r_}, b[1] = {q_ - un};
                    na = discover(ORC_CA, n, e, a, &pa); nb = discover(ORC_N, n, e, b, &pb);
                    pr += parts_total(pa, na) * parts_total(pb, nb);
                } else {
                    unsigned a[1] = {q_}, b[2] = {p_ - un, r_ - un};
                    na = discover(ORC_N, n, e, a, &pa); nb = discover(ORC_CA, n, e, b, &pb);
                    pr += parts_total(pa, na) * parts_total(pb, nb);
                }
            }
        } else {
            if (p < un && q < un && r < un && s < un) {
                if (e >= 2) { unsigned a[4] = {p, q, r, s}; na = discover(ORC_CCAA, n, e, a, &pa); pr += parts_total(pa, na) * C; }
            } else if (p >= un && q >= un && r >= un && s >= un) {
                if (e >= 2) { unsigned a[4] = {p - un, q - un, r - un, s - un}; na = discover(ORC_CCAA, n, e, a, &pa); pr += parts_total(pa, na) * C; }
            } else {
                unsigned a[2] = {p, s}, b[2] = {q - un, r - un};
                na = discover(ORC_CA, n, e, a, &pa); nb = discover(ORC_CA, n, e, b, &pb);
                pr += parts_total(pa, na) * parts_total(pb, nb);
            }
        }
        free(pa);
        free(pb);
    }
    *diag_updates = dg;
    *pair_updates = pr;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* QIFT + P(k): cpp/state_vector.hpp:115-179, 341-372                         */
/* ------------------------------------------------------------------------ */

/* in-place iterative radix-2 DFT with sign -1 (MKL/FFTW forward), natural order */
static void fft_forward(cplx *x, uint64_t M, const cplx *tw) {
    for (uint64_t i = 1, j = 0; i < M; ++i) {
        uint64_t bit = M >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cplx t = x[i]; x[i] = x[j]; x[j] = t; }
    }
    for (uint64_t len = 2; len <= M; len <<= 1) {
        uint64_t half = len >> 1, step = M / len;
        for (uint64_t i = 0; i < M; i += len)
            for (uint64_t k = 0; k < half; ++k) {
                cplx w = tw[k * step];
                cplx u = x[i + k], v = cmul(x[i + k + half], w);
                x[i + k] = cadd(u, v);
                x[i + k + half].re = u.re - v.re;
                x[i + k + half].im = u.im - v.im;
            }
    }
}

int orc_qift(int m, uint64_t ld, double *state) {
    if (m < 1 || m > 24) return -1;
    uint64_t M = 1ull << m;
    cplx *tw = (cplx *)malloc((size_t)(M / 2 + 1) * sizeof(cplx));
    if (!tw) return -1;
    const double two_pi = 6.283185307179586476925286766559;
    for (uint64_t k = 0; k < M / 2 + 1; ++k) {
        double a = two_pi * (double)k / (double)M;
        tw[k].re = cos(a);
        tw[k].im = -sin(a);
    }
    cplx *sv = (cplx *)state;
    double norm = sqrt((double)M);
    int err = 0;
    #pragma omp parallel
    {
        cplx *col = (cplx *)malloc((size_t)M * sizeof(cplx));
        if (!col) err = 1;
        #pragma omp for schedule(static)
        for (int64_t d = 0; d < (int64_t)ld; ++d) {
            if (!col) continue;
            for (uint64_t j = 0; j < M; ++j) col[j] = sv[j * ld + (uint64_t)d];
            fft_forward(col, M, tw);
            for (uint64_t k = 0; k < M; ++k) {
                sv[k * ld + (uint64_t)d].re = col[k].re / norm;
                sv[k * ld + (uint64_t)d].im = col[k].im / norm;
            }
        }
        free(col);
    }
    free(tw);
    return err ? -1 : 0;
}

int orc_phase_probabilities(int m, uint64_t ld, const double *state, double *probs) {
    if (m < 1 || m > 24) return -1;
    int64_t M = (int64_t)1 << m;
    const cplx *sv = (const cplx *)state;
    #pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < M; ++k) {
        double s = 0.0;
        const cplx *row = sv + (uint64_t)k * ld;
        for (uint64_t d = 0; d < ld; ++d) s += row[d].re * row[d].re + row[d].im * row[d].im;
        probs[k] = s;
    }
    return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
