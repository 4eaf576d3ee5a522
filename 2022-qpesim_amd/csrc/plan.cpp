// Builds the tile-pass program of one Trotter step; see plan.hpp.
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>

#include "combinadic.hpp"

namespace qpe {

namespace {

struct Binom {
    uint32_t tab[kBinomRows * kBinomCols];
    Binom() { fill_binomials(tab); }
    uint64_t operator()(int i, int k) const { return binom(tab, i, k); }
};

// scatter the low bits of `pattern` onto the positions listed in tpos
uint32_t scatter(uint32_t pattern, const std::vector<int> &tpos) {
    uint32_t x = 0;
    for (size_t i = 0; i < tpos.size(); ++i)
        if (pattern & (1u << i)) x |= 1u << tpos[i];
    return x;
}

// expand a string over the n-|T| outside positions to n positions (zeros on T)
uint32_t expand_outside(uint32_t y, const std::vector<int> &tpos_sorted) {
    for (int pos : tpos_sorted) y = open_slot(y, pos);
    return y;
}

}  // namespace

Plan build_plan(const Hamiltonian &h, size_t elem_bytes, size_t tile_bytes) {
    if (!h.indexed()) throw std::runtime_error("IndexTerms has not been called");
    Plan P;
    const int n = static_cast<int>(h.num_orbitals());
    const int e = static_cast<int>(h.num_occupied());
    if (n < 2 || n > 32 || e < 1 || e >= n) throw std::runtime_error("unsupported (num_orbitals, num_occupied)");
    Binom B;
    P.n = n;
    P.e = e;
    P.C = static_cast<uint32_t>(B(n, e));
    P.ca_pairs = static_cast<uint32_t>(B(n - 2, e - 1));
    P.ccaa_pairs = (e >= 2 && n >= 4) ? static_cast<uint32_t>(B(n - 4, e - 2)) : 0u;
    P.elem_bytes = elem_bytes;
    P.tile_bytes = tile_bytes;
    const uint32_t C = P.C;
    const double t = h.timestep();
    const auto &seg = h.segmented();

    P.strings.resize(C);
    for (uint32_t i = 0; i < C; ++i) P.strings[i] = unrank_of(i, e, n, B.tab);

    if (C > kRankMask) throw std::runtime_error("C(num_orbitals, num_occupied) exceeds the packed-rank range");
    if (P.ca_pairs > kMaxPairs || P.ccaa_pairs > kMaxPairs)
        throw std::runtime_error("more determinant pairs per term than the tile engine unrolls");
    // single-excitation pair tables for every ordered orbital pair (x, y)
    P.ca_table.assign((size_t)n * n * P.ca_pairs, 0u);
    for (int x = 0; x < n; ++x)
        for (int y = 0; y < n; ++y) {
            if (x == y) continue;
            for (uint32_t k = 0; k < P.ca_pairs; ++k) {
                uint32_t rx, ry, par;
                ca_pair(k, x, y, n, e, B.tab, &rx, &ry, &par);
                P.ca_table[((size_t)x * n + y) * P.ca_pairs + k] = pack_pair(rx, ry, par);
            }
        }

    auto spin_of = [n](int g) { return g < n ? 0 : 1; };
    auto local = [n](int g) { return g < n ? g : g - n; };

    // ---- diagonal: pp (TEOT:376-395) and pqqp (TEOT:399-424) -------------
    std::vector<double> pp[2] = {std::vector<double>(n, 0.0), std::vector<double>(n, 0.0)};
    std::vector<double> vsame[2] = {std::vector<double>(n * n, 0.0), std::vector<double>(n * n, 0.0)};
    std::vector<double> vab(n * n, 0.0);
    for (const auto &T : seg[kPP]) {
        pp[spin_of(T.idx[0])][local(T.idx[0])] += T.coeff;
        P.diag_updates += (uint64_t)B(n - 1, e - 1) * C;
    }
    for (const auto &T : seg[kPQQP]) {
        int p = T.idx[0], q = T.idx[1];
        if (spin_of(p) == spin_of(q)) {
            if (e < 2) continue;
            vsame[spin_of(p)][local(p) * n + local(q)] += T.coeff;
            P.diag_updates += (uint64_t)B(n - 2, e - 2) * C;
        } else {
            vab[local(p) * n + local(q)] += T.coeff;   // p alpha, q beta (key order p < q)
            P.diag_updates += (uint64_t)B(n - 1, e - 1) * B(n - 1, e - 1);
        }
    }
    P.theta_a.assign(C, 0.0);
    P.theta_b.assign(C, 0.0);
    P.cross.assign((size_t)n * C, 0.0);
    for (uint32_t i = 0; i < C; ++i) {
        uint32_t s = P.strings[i];
        double ta = 0.0, tb = 0.0;
        for (int p = 0; p < n; ++p) {
            if (!(s & (1u << p))) continue;
            ta += pp[0][p];
            tb += pp[1][p];
            for (int q = p + 1; q < n; ++q)
                if (s & (1u << q)) {
                    ta += vsame[0][p * n + q];
                    tb += vsame[1][p * n + q];
                }
        }
        P.theta_a[i] = t * ta;
        P.theta_b[i] = t * tb;
        for (int p = 0; p < n; ++p) {
            double w = 0.0;
            for (int q = 0; q < n; ++q)
                if (s & (1u << q)) w += vab[p * n + q];
            P.cross[(size_t)p * C + i] = t * w;
        }
    }

    // ---- excitation groups: each pq term with the pqqr terms it gates (TEOT:428-513)
    std::map<std::pair<int, int>, std::vector<const IndexedTerm *>> gated;  // (p_, r_) -> pqqr terms
    for (const auto &T : seg[kPQQR]) gated[{T.idx[0], T.idx[3]}].push_back(&T);
    const uint32_t npa = P.ca_pairs;
    std::vector<int> group_spin;
    for (const auto &T : seg[kPQ]) {
        int gp = T.idx[0], gq = T.idx[1];   // gp > gq, same spin
        int sp = spin_of(gp);
        GroupDev g;
        g.x = local(gp);
        g.y = local(gq);
        g.spin = sp;
        g.pad = 0;
        std::vector<double> csame(n, 0.0), cother(n, 0.0);
        P.pair_updates += (uint64_t)npa * C;
        auto it = gated.find({gq, gp});
        if (it != gated.end()) {
            for (const IndexedTerm *Q : it->second) {
                int inner = Q->idx[1];
                if (spin_of(inner) == sp) {
                    if (e < 2) continue;   // updatePQQR_AAAA/BBBB return early (TEOT:1077-1080)
                    csame[local(inner)] += Q->coeff;
                    P.pair_updates += (uint64_t)B(n - 3, e - 2) * C;
                } else {
                    cother[local(inner)] += Q->coeff;
                    P.pair_updates += (uint64_t)npa * B(n - 1, e - 1);
                }
            }
        }
        // pair part: t*(h + sum of same-spin inner coefficients occupied in f)
        size_t base_a = P.grp_a.size();
        P.grp_a.resize(base_a + 2 * (size_t)npa);
        for (uint32_t k = 0; k < npa; ++k) {
            uint32_t r1, r2, par;
            uint32_t f = ca_pair(k, g.x, g.y, n, e, B.tab, &r1, &r2, &par);
            double ang = T.coeff;
            for (int x = 0; x < n; ++x)
                if (f & (1u << x)) ang += csame[x];
            ang *= t;
            P.grp_a[base_a + 2 * k] = std::cos(ang);
            P.grp_a[base_a + 2 * k + 1] = std::sin(ang);
        }
        // other-string part: t*(sum of opposite-spin inner coefficients occupied)
        size_t base_b = P.grp_b.size();
        P.grp_b.resize(base_b + 2 * (size_t)C);
        for (uint32_t i = 0; i < C; ++i) {
            double ang = 0.0;
            for (int y = 0; y < n; ++y)
                if (P.strings[i] & (1u << y)) ang += cother[y];
            ang *= t;
            P.grp_b[base_b + 2 * i] = std::cos(ang);
            P.grp_b[base_b + 2 * i + 1] = std::sin(ang);
        }
        P.groups.push_back(g);
    }

    // ---- pqrs terms (TEOT:519-560) --------------------------------------
    struct Rot { int cls; int index; int p, q, r, s; };
    std::vector<Rot> rots;
    std::vector<uint32_t> cc_offset;   // per term: offset in ccaa_table
    for (const auto &T : seg[kPQRS]) {
        int p = T.idx[0], q = T.idx[1], r = T.idx[2], s = T.idx[3];
        int cls;
        if (p < n && q < n && r < n && s < n) cls = OP_AAAA;
        else if (p >= n && q >= n && r >= n && s >= n) cls = OP_BBBB;
        else cls = OP_ABBA;
        if (cls != OP_ABBA && e < 2) continue;  // TEOT:845-848, 864-867
        TermDev d;
        d.p = local(p); d.q = local(q); d.r = local(r); d.s = local(s);
        double a = T.coeff * t;
        d.cs = std::cos(a);
        d.sn = std::sin(a);
        rots.push_back({cls, (int)P.terms.size(), d.p, d.q, d.r, d.s});
        if (cls != OP_ABBA) {
            uint32_t base = (uint32_t)P.ccaa_table.size();
            cc_offset.push_back(base);
            for (uint32_t k = 0; k < P.ccaa_pairs; ++k) {
                uint32_t r01, r10, par;
                ccaa_pair(k, d.p, d.q, d.r, d.s, n, e, B.tab, &r01, &r10, &par);
                P.ccaa_table.push_back(pack_pair(r01, r10, par));
            }
        } else {
            cc_offset.push_back(0);
        }
        P.terms.push_back(d);
        P.pair_updates += (cls == OP_ABBA) ? (uint64_t)npa * npa : (uint64_t)P.ccaa_pairs * C;
    }

    // ---- passes ----------------------------------------------------------
    const uint64_t line_bytes = (uint64_t)C * elem_bytes;
    // lines per workgroup for a pass that stages `ngroups` group angle rows
    auto lines_for = [&](int ngroups) {
        uint64_t per_line = line_bytes + (uint64_t)ngroups * 16u;
        return (int)std::min<uint64_t>(tile_bytes / per_line, 64);
    };
    const int W = lines_for(0);
    if (W < 1) throw std::runtime_error("a single column does not fit the tile budget");
    const uint32_t nca = P.ca_pairs;
    auto ca_off = [&](int x, int y) { return (uint32_t)(((size_t)x * n + y) * nca); };
    auto op_group = [&](int type, int g, int slot) {
        OpDev o;
        o.type = type;
        o.index = g;
        o.tab = ca_off(P.groups[g].x, P.groups[g].y);
        o.aux = (uint32_t)slot;
        o.cs = 1.0;
        o.sn = 0.0;
        return o;
    };
    auto op_term = [&](const Rot &R) {
        OpDev o;
        o.type = R.cls;
        o.index = R.index;
        if (R.cls == OP_ABBA) {
            o.tab = ca_off(R.s, R.p);        // alpha pair: rank(f|s), rank(f|p)
            o.aux = (uint32_t)R.q | ((uint32_t)R.r << 8);
        } else {
            o.tab = cc_offset[R.index];
            o.aux = 0;
        }
        o.cs = P.terms[R.index].cs;
        o.sn = P.terms[R.index].sn;
        return o;
    };

    auto add_pass = [&](int kind, int flags, int op_begin, std::vector<ItemDev> items, int max_lines,
                        int gfirst, int gcount) {
        PassDev pd;
        pd.kind = kind;
        pd.flags = flags;
        pd.op_begin = op_begin;
        pd.op_end = (int)P.ops.size();
        pd.item_begin = (int)P.items.size();
        pd.item_count = (int)items.size();
        pd.max_lines = max_lines;
        pd.group_first = gfirst;
        pd.group_count = gcount;
        pd.pad0 = pd.pad1 = pd.pad2 = 0;
        for (const auto &it : items) {
            P.items.push_back(it);
            P.tile_bytes_moved += 2ull * it.count * line_bytes;
        }
        P.max_items = std::max(P.max_items, pd.item_count);
        P.passes.push_back(pd);
    };
    auto chunk_columns = [&](int w) {
        std::vector<ItemDev> items;
        for (uint32_t c0 = 0; c0 < C; c0 += (uint32_t)w) {
            ItemDev it;
            it.begin = (uint32_t)P.col_list.size();
            it.count = std::min<uint32_t>((uint32_t)w, C - c0);
            for (uint32_t c = 0; c < it.count; ++c) P.col_list.push_back(c0 + c);
            items.push_back(it);
        }
        return items;
    };
    auto chunk_rows = [&](int w) {
        std::vector<ItemDev> items;
        for (uint32_t r0 = 0; r0 < C; r0 += (uint32_t)w) items.push_back({r0, std::min<uint32_t>((uint32_t)w, C - r0)});
        return items;
    };

    // groups are stored alpha first, then beta (pq map order), so each spin is a contiguous range
    int ga_first = -1, ga_count = 0, gb_first = -1, gb_count = 0;
    for (size_t g = 0; g < P.groups.size(); ++g) {
        if (P.groups[g].spin == 0) { if (ga_first < 0) ga_first = (int)g; ++ga_count; }
        else { if (gb_first < 0) gb_first = (int)g; ++gb_count; }
    }
    // pass 0: copy block j-1 -> j, diagonal phase, alpha groups (COL)
    {
        int w = lines_for(ga_count);
        if (w < 1) throw std::runtime_error("tile budget too small for the alpha excitation groups");
        int ob = (int)P.ops.size();
        OpDev d;
        d.type = OP_DIAG; d.index = 0; d.tab = 0; d.aux = 0; d.cs = 1.0; d.sn = 0.0;
        P.ops.push_back(d);
        for (int g = 0; g < ga_count; ++g) P.ops.push_back(op_group(OP_GA, ga_first + g, g));
        add_pass(PASS_COL, PF_FROM_PREV, ob, chunk_columns(w), w, ga_count ? ga_first : 0, ga_count);
    }
    // pass 1: beta groups (ROW)
    if (gb_count) {
        int w = lines_for(gb_count);
        if (w < 1) throw std::runtime_error("tile budget too small for the beta excitation groups");
        int ob = (int)P.ops.size();
        for (int g = 0; g < gb_count; ++g) P.ops.push_back(op_group(OP_GB, gb_first + g, g));
        add_pass(PASS_ROW, 0, ob, chunk_rows(w), w, gb_first, gb_count);
    }
    // alpha-leading double excitations, p by p: AAAA_p then ABBA_p (map order)
    auto max_orbit = [&](int tsize, int k0, int k1) {
        uint64_t m = 0;
        for (int k = k0; k <= k1; ++k) m = std::max<uint64_t>(m, B(tsize, k));
        return m;
    };
    auto k_range = [&](int tsize, bool all, int *k0, int *k1) {
        *k0 = std::max(all ? 0 : 1, e - (n - tsize));
        *k1 = std::min(e, all ? tsize : tsize - 1);
    };
    size_t ri = 0;
    while (ri < rots.size() && rots[ri].cls != OP_BBBB) {
        int p = rots[ri].p;
        std::vector<Rot> aaaa, abba;
        while (ri < rots.size() && rots[ri].cls != OP_BBBB && rots[ri].p == p) {
            (rots[ri].cls == OP_AAAA ? aaaa : abba).push_back(rots[ri]);
            ++ri;
        }
        bool pending_aaaa = !aaaa.empty();
        size_t ai = 0;
        while (ai < abba.size()) {
            // window: consecutive terms sharing q, over a run of r values
            int q = abba[ai].q;
            size_t aj = ai;
            uint32_t tmask = 1u << q;
            while (aj < abba.size() && abba[aj].q == q) {
                int r = abba[aj].r;
                uint32_t nm = tmask | (1u << r);
                if (nm != tmask) {
                    int ts = popcnt(nm);
                    int k0, k1;
                    k_range(ts, pending_aaaa, &k0, &k1);
                    if (max_orbit(ts, k0, k1) > (uint64_t)W) break;
                    tmask = nm;
                }
                ++aj;
            }
            if (aj == ai) throw std::runtime_error("tile budget too small for a mixed double excitation");
            std::vector<int> tpos;
            for (int b = 0; b < n; ++b)
                if (tmask & (1u << b)) tpos.push_back(b);
            int ts = (int)tpos.size();
            int k0, k1;
            k_range(ts, pending_aaaa, &k0, &k1);
            int ob = (int)P.ops.size();
            if (pending_aaaa)
                for (const auto &R : aaaa) P.ops.push_back(op_term(R));
            for (size_t x = ai; x < aj; ++x) P.ops.push_back(op_term(abba[x]));
            // orbits: outside string y (e-k electrons on n-ts positions) x all k-patterns of T
            std::vector<ItemDev> items;
            ItemDev cur{(uint32_t)P.col_list.size(), 0};
            for (int k = k0; k <= k1; ++k) {
                uint32_t ny = (uint32_t)B(n - ts, e - k), npat = (uint32_t)B(ts, k);
                for (uint32_t yi = 0; yi < ny; ++yi) {
                    uint32_t y = expand_outside(unrank_of(yi, e - k, n - ts, B.tab), tpos);
                    if (cur.count + npat > (uint32_t)W) {
                        items.push_back(cur);
                        cur = ItemDev{(uint32_t)P.col_list.size(), 0};
                    }
                    for (uint32_t pi = 0; pi < npat; ++pi) {
                        uint32_t str = y | scatter(unrank_of(pi, k, ts, B.tab), tpos);
                        P.col_list.push_back(rank_of(str, B.tab));
                        ++cur.count;
                    }
                }
            }
            if (cur.count) items.push_back(cur);
            add_pass(PASS_COL, 0, ob, items, W, 0, 0);
            pending_aaaa = false;
            ai = aj;
        }
        if (pending_aaaa) {
            int ob = (int)P.ops.size();
            for (const auto &R : aaaa) P.ops.push_back(op_term(R));
            add_pass(PASS_COL, 0, ob, chunk_columns(W), W, 0, 0);
        }
    }
    // pure-beta double excitations (ROW)
    {
        int ob = (int)P.ops.size();
        for (; ri < rots.size(); ++ri) P.ops.push_back(op_term(rots[ri]));
        if ((int)P.ops.size() > ob) add_pass(PASS_ROW, 0, ob, chunk_rows(W), W, 0, 0);
    }
    // the block norm is summed by the last pass; it must see every amplitude
    {
        const PassDev &last = P.passes.back();
        uint64_t covered = 0;
        for (int i = 0; i < last.item_count; ++i) covered += P.items[last.item_begin + i].count;
        if (covered != C) add_pass(PASS_COL, 0, (int)P.ops.size(), chunk_columns(W), W, 0, 0);
    }
    P.passes.back().flags |= PF_EMIT_NORM;
    return P;
}

}  // namespace qpe
