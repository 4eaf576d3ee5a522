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

Plan build_plan(const Hamiltonian &h, size_t elem_bytes, size_t tile_bytes, bool fuse, int order, int term_order) {
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
    if (order != 1 && order != 2) throw std::runtime_error("Trotter order must be 1 or 2");
    if (term_order != 0 && term_order != 1) throw std::runtime_error("term order must be 0 or 1");
    P.order = order;
    P.term_order = term_order;
    const bool original = term_order == 1;
    // second order: the step is the forward sequence at t/2 followed by the reversed sequence at t/2
    const double t = order == 2 ? 0.5 * h.timestep() : h.timestep();
    const auto &seg = h.segmented();

    P.strings.resize(C);
    for (uint32_t i = 0; i < C; ++i) P.strings[i] = unrank_of(i, e, n, B.tab);

    if (C > kSchedPad)
        throw std::runtime_error("C(num_orbitals, num_occupied) = " + std::to_string(C) + " exceeds " +
                                 std::to_string(kSchedPad) +
                                 ", the largest string count the 14-bit ranks of the pass schedules hold");
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
    // One table set per DIAG op: the reordered step applies pp and pqqp back to
    // back (one set); applyUnitary (TEOT:174-372) has the pq terms between them
    // (set 0: pp, set 1: pqqp).
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
    auto add_diag_set = [&](bool with_pp, bool with_pqqp) {
        const size_t a0 = P.theta_a.size(), c0 = P.cross.size();
        P.theta_a.resize(a0 + C, 0.0);
        P.theta_b.resize(a0 + C, 0.0);
        P.cross.resize(c0 + (size_t)n * C, 0.0);
        for (uint32_t i = 0; i < C; ++i) {
            uint32_t s = P.strings[i];
            double ta = 0.0, tb = 0.0;
            for (int p = 0; p < n; ++p) {
                if (!(s & (1u << p))) continue;
                if (with_pp) {
                    ta += pp[0][p];
                    tb += pp[1][p];
                }
                if (with_pqqp)
                    for (int q = p + 1; q < n; ++q)
                        if (s & (1u << q)) {
                            ta += vsame[0][p * n + q];
                            tb += vsame[1][p * n + q];
                        }
            }
            P.theta_a[a0 + i] = t * ta;
            P.theta_b[a0 + i] = t * tb;
            if (!with_pqqp) continue;
            for (int p = 0; p < n; ++p) {
                double w = 0.0;
                for (int q = 0; q < n; ++q)
                    if (s & (1u << q)) w += vab[p * n + q];
                P.cross[c0 + (size_t)p * C + i] = t * w;
            }
        }
    };
    if (original) {
        add_diag_set(true, false);
        add_diag_set(false, true);
    } else {
        add_diag_set(true, true);
    }

    // ---- excitation groups: each pq term with the pqqr terms it gates (TEOT:428-513)
    // (applyUnitary applies every pqqr term on its own, after the pqqp terms: no gating)
    std::map<std::pair<int, int>, std::vector<const IndexedTerm *>> gated;  // (p_, r_) -> pqqr terms
    if (!original)
        for (const auto &T : seg[kPQQR]) gated[{T.idx[0], T.idx[3]}].push_back(&T);
    const uint32_t npa = P.ca_pairs;
    std::vector<double> grp_pair;   // [G][npa][2] 1-cos, sin of the pair angle part, pair order k
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
        size_t base_a = grp_pair.size();
        grp_pair.resize(base_a + 2 * (size_t)npa);
        for (uint32_t k = 0; k < npa; ++k) {
            uint32_t r1, r2, par;
            uint32_t f = ca_pair(k, g.x, g.y, n, e, B.tab, &r1, &r2, &par);
            double ang = T.coeff;
            for (int x = 0; x < n; ++x)
                if (f & (1u << x)) ang += csame[x];
            ang *= t;
            grp_pair[base_a + 2 * k] = versine(ang);
            grp_pair[base_a + 2 * k + 1] = std::sin(ang);
        }
        // other-string part: t*(sum of opposite-spin inner coefficients occupied)
        size_t base_b = P.grp_b.size();
        P.grp_b.resize(base_b + 2 * (size_t)C);
        for (uint32_t i = 0; i < C; ++i) {
            double ang = 0.0;
            for (int y = 0; y < n; ++y)
                if (P.strings[i] & (1u << y)) ang += cother[y];
            ang *= t;
            P.grp_b[base_b + 2 * i] = versine(ang);
            P.grp_b[base_b + 2 * i + 1] = std::sin(ang);
        }
        P.groups.push_back(g);
    }

    // ---- applyUnitary's pqqr terms, each its own rotation (TEOT:254-301) ----
    // Term (p_, q_, q_, r_), p_ < r_ of one spin: the pairs (f|r_, f|p_) of that
    // spin's strings, with inner orbital q_ occupied: in f when q_ has the same
    // spin (the pairs are filtered), in the other string when not (a line gate).
    struct Single { int spin; uint32_t gate; int index; std::vector<uint32_t> entries; };
    std::vector<Single> singles;
    if (original) {
        for (const auto &T : seg[kPQQR]) {
            const int gp = T.idx[0], gi = T.idx[1], gr = T.idx[3];
            const int sp = spin_of(gp);
            const bool same = spin_of(gi) == sp;
            if (same && e < 2) continue;   // updatePQQR_AAAA/BBBB return early (TEOT:1077-1080)
            Single sg;
            sg.spin = sp;
            sg.gate = same ? 0u : (1u << local(gi));
            sg.index = (int)P.terms.size();
            const int x = local(gr), y = local(gp);
            for (uint32_t k = 0; k < npa; ++k) {
                uint32_t r1, r2, par;
                const uint32_t f = ca_pair(k, x, y, n, e, B.tab, &r1, &r2, &par);
                if (same && !(f & (1u << local(gi)))) continue;
                sg.entries.push_back(pack_sched(r1, r2, 0, par));
            }
            TermDev d;
            d.p = x; d.q = local(gi); d.r = y; d.s = y;
            const double a = T.coeff * t;
            d.cs = std::cos(a);
            d.sn = std::sin(a);
            P.terms.push_back(d);
            P.coef.push_back(versine(a));
            P.coef.push_back(d.sn);
            P.pair_updates += same ? (uint64_t)B(n - 3, e - 2) * C : (uint64_t)npa * B(n - 1, e - 1);
            singles.push_back(std::move(sg));
        }
    }

    // ---- pqrs terms (TEOT:519-560) --------------------------------------
    struct Rot { int cls; int index; int p, q, r, s; };
    std::vector<Rot> rots;
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
        P.terms.push_back(d);
        P.coef.push_back(versine(a));
        P.coef.push_back(d.sn);
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
    const int NS = kThreadsPlan;   // schedule slots: one per thread of the pass kernel
    const size_t max_group = fuse ? (size_t)kMaxGroup : 1u;
    auto blank = [](int type) {
        OpDev o;
        o.type = type; o.index = 0; o.tab = 0; o.aux = 0; o.lg = 0; o.ns = kThreadsPlan; o.g = 1; o.split = 0;
        return o;
    };
    const uint32_t pad_entry = pack_sched(kSchedPad, kSchedPad, 0, 0);
    // Appends rows [i0, i0+lg) of slot lists as a schedule: entry i of slot t at tab + i*ns + t.
    using Slots = std::vector<std::vector<uint32_t>>;
    auto emit_rows = [&](const Slots &slot, int i0, int lg, OpDev &o) {
        const int ns = (int)slot.size();
        o.tab = (uint32_t)P.sched.size();
        o.lg = lg;
        o.ns = ns;
        for (int i = i0; i < i0 + lg; ++i)
            for (int k = 0; k < ns; ++k)
                P.sched.push_back((size_t)i < slot[k].size() ? slot[k][i] : pad_entry);
    };
    auto slot_len = [](const Slots &slot) {
        size_t lg = 0;
        for (auto &sl : slot) lg = std::max(lg, sl.size());
        return (int)lg;
    };
    // excitation group g: its pairs k in slot k % ns, row k / ns, split into
    // ops of at most kSchedMax rows; the pair angles are laid out the same way
    auto push_group_ops = [&](int type, int g, int stage_slot, int units) {
        (void)units;
        const GroupDev &gd = P.groups[g];
        // the pair angles are read inside the op (one global latency per row): fewest rows
        const int ns = (int)std::min<uint32_t>(NS, npa);
        const int rows = (int)((npa + ns - 1) / ns);
        for (int i0 = 0; i0 < rows; i0 += kSchedMax) {
            const int lg = std::min(kSchedMax, rows - i0);
            OpDev o = blank(type);
            o.aux = (uint32_t)stage_slot;
            o.index = (int32_t)(P.grp_a.size() / 2);
            o.tab = (uint32_t)P.sched.size();
            o.lg = lg;
            o.ns = ns;
            for (int i = i0; i < i0 + lg; ++i)
                for (int k = 0; k < ns; ++k) {
                    const size_t kk = (size_t)i * ns + k;
                    if (kk < npa) {
                        const uint32_t e0 = P.ca_table[((size_t)gd.x * n + gd.y) * npa + kk];
                        P.sched.push_back(pack_sched(e0 & kRankMask, (e0 >> kRankBits) & kRankMask, 0, e0 >> 31));
                        P.grp_a.push_back(grp_pair[2 * ((size_t)g * npa + kk)]);
                        P.grp_a.push_back(grp_pair[2 * ((size_t)g * npa + kk) + 1]);
                    } else {
                        P.sched.push_back(pad_entry);
                        P.grp_a.push_back(0.0);
                        P.grp_a.push_back(0.0);
                    }
                }
            P.ops.push_back(o);
        }
    };
    // Packs per-orbit entry lists into ns slots, longest list first onto the
    // least-loaded slot (a min-heap over slot loads).  pack_rows returns only
    // the resulting schedule length; pack_slots builds the slot lists.
    using Orbits = std::vector<std::vector<uint32_t>>;
    auto lpt_order = [](const Orbits &orbits) {
        std::vector<uint32_t> idx(orbits.size());
        for (uint32_t i = 0; i < idx.size(); ++i) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(),
                         [&](uint32_t a, uint32_t b) { return orbits[a].size() > orbits[b].size(); });
        return idx;
    };
    auto lpt_assign = [](const Orbits &orbits, const std::vector<uint32_t> &order, int ns, std::vector<int> *where) {
        bool equal = true;
        for (const auto &o : orbits) equal = equal && o.size() == orbits[0].size();
        if (equal) {   // orbits of one length: round robin is the LPT packing (ties go to the lowest slot)
            if (where) {
                where->assign(orbits.size(), 0);
                for (size_t r = 0; r < order.size(); ++r) (*where)[order[r]] = (int)(r % ns);
            }
            return orbits.empty() ? 0 : (int)((orbits.size() + ns - 1) / ns * orbits[0].size());
        }
        std::vector<std::pair<size_t, int>> heap;   // (load, slot), min-heap
        heap.reserve(ns);
        for (int k = 0; k < ns; ++k) heap.push_back({0, k});
        auto cmp = [](const std::pair<size_t, int> &a, const std::pair<size_t, int> &b) { return a > b; };
        std::make_heap(heap.begin(), heap.end(), cmp);
        size_t lg = 0;
        if (where) where->assign(orbits.size(), 0);
        for (uint32_t oi : order) {
            std::pop_heap(heap.begin(), heap.end(), cmp);
            auto &top = heap.back();
            if (where) (*where)[oi] = top.second;
            top.first += orbits[oi].size();
            lg = std::max(lg, top.first);
            std::push_heap(heap.begin(), heap.end(), cmp);
        }
        return (int)lg;
    };
    auto pack_slots = [&](const Orbits &orbits, const std::vector<uint32_t> &order, int ns) {
        std::vector<int> where;
        lpt_assign(orbits, order, ns, &where);
        Slots slot(ns);
        for (uint32_t oi : order) slot[where[oi]].insert(slot[where[oi]].end(), orbits[oi].begin(), orbits[oi].end());
        return slot;
    };
    // orbits of a support set S (positions spos) inside one string: callback(y, kS)
    auto for_orbits = [&](const std::vector<int> &spos, auto &&fn) {
        const int ts = (int)spos.size();
        for (int k = std::max(0, e - (n - ts)); k <= std::min(e, ts); ++k) {
            uint32_t ny = (uint32_t)B(n - ts, e - k);
            for (uint32_t yi = 0; yi < ny; ++yi) fn(expand_outside(unrank_of(yi, e - k, n - ts, B.tab), spos), k);
        }
    };
    auto members = [&](const std::vector<int> &spos, int k) {
        std::vector<uint32_t> m;
        const int ts = (int)spos.size();
        for (uint32_t pi = 0; pi < (uint32_t)B(ts, k); ++pi) m.push_back(scatter(unrank_of(pi, k, ts, B.tab), spos));
        return m;
    };
    auto positions = [&](uint32_t mask) {
        std::vector<int> v;
        for (int b = 0; b < n; ++b)
            if (mask & (1u << b)) v.push_back(b);
        return v;
    };
    // Per-orbit entries of consecutive same-string double excitations T[0..g):
    // an orbit is the set of strings that agree outside the joint support, and
    // it is closed under every term of the group, so one thread applies the
    // group's rotations of an orbit in order without synchronising.
    auto cc_orbits = [&](const std::vector<Rot> &T) {
        uint32_t smask = 0;
        for (const auto &R : T) smask |= (1u << R.p) | (1u << R.q) | (1u << R.r) | (1u << R.s);
        std::vector<int> spos = positions(smask);
        Orbits orbits;
        std::vector<std::vector<uint32_t>> mem_k(spos.size() + 1);
        for (size_t k = 0; k <= spos.size(); ++k) mem_k[k] = members(spos, (int)k);
        for_orbits(spos, [&](uint32_t y, int k) {
            std::vector<uint32_t> ent;
            const auto &mem = mem_k[k];
            for (size_t ti = 0; ti < T.size(); ++ti) {
                const Rot &R = T[ti];
                const uint32_t from = (1u << R.r) | (1u << R.s), to = (1u << R.p) | (1u << R.q);
                int id[4] = {R.p, R.q, R.r, R.s};
                std::sort(id, id + 4);
                for (uint32_t m : mem) {
                    uint32_t a = y | m;
                    if ((a & from) != from || (a & to)) continue;
                    uint32_t b = a ^ from ^ to, f = a ^ from;
                    uint32_t par = (uint32_t)(popcnt(f & between_mask(id[0], id[1])) + popcnt(f & between_mask(id[2], id[3]))) & 1u;
                    ent.push_back(pack_sched(rank_of(a, B.tab), rank_of(b, B.tab), (uint32_t)ti, par));
                }
            }
            if (!ent.empty()) orbits.push_back(std::move(ent));
        });
        return orbits;
    };
    // consecutive ABBA terms (p, q, r, s_i): alpha orbits over {p} + {s_i}
    auto ab_orbits = [&](const std::vector<Rot> &T) {
        const int p = T[0].p;
        uint32_t smask = 1u << p;
        for (const auto &R : T) smask |= 1u << R.s;
        std::vector<int> spos = positions(smask);
        Orbits orbits;
        std::vector<std::vector<uint32_t>> mem_k(spos.size() + 1);
        for (size_t k = 0; k <= spos.size(); ++k) mem_k[k] = members(spos, (int)k);
        for_orbits(spos, [&](uint32_t y, int k) {
            std::vector<uint32_t> ent;
            const auto &mem = mem_k[k];
            for (size_t ti = 0; ti < T.size(); ++ti) {
                const int sv = T[ti].s;
                for (uint32_t m : mem) {
                    uint32_t a = y | m;
                    if (!((a >> sv) & 1u) || ((a >> p) & 1u)) continue;
                    uint32_t b = a ^ (1u << sv) ^ (1u << p), f = a ^ (1u << sv);
                    uint32_t par = (uint32_t)popcnt(f & between_mask(sv, p)) & 1u;
                    ent.push_back(pack_sched(rank_of(a, B.tab), rank_of(b, B.tab), (uint32_t)ti, par));
                }
            }
            if (!ent.empty()) orbits.push_back(std::move(ent));
        });
        return orbits;
    };
    // a single term: every determinant pair is its own orbit
    auto single_orbits = [&](const Rot &R, bool ab) {
        Orbits ob;
        if (ab) {
            const size_t base = ((size_t)R.s * n + R.p) * npa;   // pairs rank(f|s), rank(f|p)
            for (uint32_t k = 0; k < npa; ++k) {
                const uint32_t e0 = P.ca_table[base + k];
                ob.push_back({pack_sched(e0 & kRankMask, (e0 >> kRankBits) & kRankMask, 0, e0 >> 31)});
            }
        } else {
            for (uint32_t k = 0; k < P.ccaa_pairs; ++k) {
                uint32_t r01, r10, par;
                ccaa_pair(k, R.p, R.q, R.r, R.s, n, e, B.tab, &r01, &r10, &par);
                ob.push_back({pack_sched(r01, r10, 0, par)});
            }
        }
        return ob;
    };
    auto orbit_stats = [](const Orbits &ob, size_t &total, size_t &longest) {
        total = 0;
        longest = 0;
        for (auto &o : ob) {
            total += o.size();
            longest = std::max(longest, o.size());
        }
    };
    // Emits the ops of one term group from its orbits: the slot count with the
    // lowest cost among those whose schedule fits kSchedMax rows (a single term
    // may exceed it at any slot count: its rows are then split over ops).
    auto emit_orbits = [&](const Orbits &best, const std::vector<uint32_t> &order, int lg512, int type, int index,
                           int g, uint32_t aux, int units) {
        size_t total = 0, longest = 0;
        orbit_stats(best, total, longest);
        // Work split: ns schedule slots, the L = NS / ns threads of a slot share its lines
        // (or column pairs); threads past ns * L, and lanes past the units, idle.
        // With several units per slot, a VALU model: a wave runs lg rows of its busiest
        // thread's units (lane l of L takes units l, l + L, ...), the SIMDs (wave w on
        // SIMD w % 4) run their waves' instructions one after another, and the busiest
        // SIMD sets the op's time; the longest single chain breaks ties.  Candidates: NS / L
        // for each useful lane count, one orbit per slot, and the fewest slots for each row
        // count (no padded slots).  With one unit per slot (e.g. three-column windows) the
        // per-thread chain matters more than the padding (measured: fewer, fuller waves
        // were 4% slower there), so only NS / L slots compete, on rows x units per thread.
        const bool model = units >= 2;
        auto vcost = [&](int ns, int lg) {
            const int L = NS / ns;
            if (!model) return (double)lg * ((units + L - 1) / L) + 0.01 * lg;
            int simd[4] = {0, 0, 0, 0}, crit = 0;
            for (int w = 0; w < NS / 64; ++w) {
                const int lane = (w * 64) / ns;   // a wave's first thread has its busiest lane
                const int per = lane < L && lane < units ? (units - lane + L - 1) / L : 0;
                simd[w % 4] += per * lg;
                crit = std::max(crit, per * lg);
            }
            return (double)std::max(std::max(simd[0], simd[1]), std::max(simd[2], simd[3])) + 0.01 * crit;
        };
        int ns_best = NS;
        double bc = vcost(NS, std::min(lg512, kSchedMax)) * ((lg512 + kSchedMax - 1) / kSchedMax);
        std::vector<int> cands;
        for (int L = 2; L <= std::min(NS, units); ++L) cands.push_back(NS / L);
        if (model) {
            cands.push_back((int)best.size());
            for (size_t t = std::max<size_t>(longest, 1); t <= (size_t)kSchedMax; ++t)
                cands.push_back((int)((total + t - 1) / t));
        }
        for (int cand : cands) {
            if (cand < 1 || cand > NS) continue;
            if ((size_t)cand * kSchedMax < total || longest > (size_t)kSchedMax) continue;
            const int lower = (int)std::max<size_t>(longest, (total + cand - 1) / cand);
            if (vcost(cand, lower) >= bc - 1e-9) continue;
            const int lgc = lpt_assign(best, order, cand, nullptr);
            if (lgc > kSchedMax) continue;
            const double c = vcost(cand, lgc);
            if (c < bc - 1e-9) { bc = c; ns_best = cand; }
        }
        Slots slot = pack_slots(best, order, ns_best);
        const int rows = slot_len(slot);
        for (int i0 = 0; i0 < std::max(rows, 1); i0 += kSchedMax) {
            OpDev o = blank(type);
            o.index = index;
            o.g = g;
            o.aux = aux;
            emit_rows(slot, i0, std::min(kSchedMax, rows - i0), o);
            P.ops.push_back(o);
        }
    };
    // Greedy grouping of consecutive terms: a group grows while its joint
    // support stays within the limit, no orbit is longer than kSchedMax and
    // the entries fill at most 3/4 of the NS x kSchedMax schedule (so the
    // packing fits).  A single term's entries are independent: a term with more
    // than NS x kSchedMax of them is split over several ops.
    // Mixed-window runs (terms (p, q, r, s_i) sharing p, q, r): their grouping
    // and schedules depend only on p and the s_i (the alpha side), so every run
    // with the same alpha side reuses them, pointing at the same schedule rows.
    struct AbOp {
        uint32_t tab;
        int lg, ns, g;
        size_t first;   // the op's first term, as an offset into the run
    };
    std::map<std::vector<int>, std::vector<AbOp>> ab_cache;
    auto push_groups = [&](const std::vector<Rot> &list, bool ab, int units) {
        std::vector<int> key;
        bool consecutive = true;
        for (size_t k = 1; k < list.size(); ++k) consecutive = consecutive && list[k].index == list[0].index + (int)k;
        if (ab && consecutive && !list.empty()) {
            key.push_back(units);
            key.push_back(list[0].p);
            for (const auto &R : list) key.push_back(R.s);
            auto hit = ab_cache.find(key);
            if (hit != ab_cache.end()) {
                for (const AbOp &a : hit->second) {
                    OpDev o = blank(OP_FAB);
                    o.index = list[a.first].index;
                    o.g = a.g;
                    o.aux = (uint32_t)list[a.first].q | ((uint32_t)list[a.first].r << 8);
                    o.tab = a.tab;
                    o.lg = a.lg;
                    o.ns = a.ns;
                    P.ops.push_back(o);
                }
                return;
            }
        }
        const size_t op0 = P.ops.size();
        auto support = [&](const Rot &R) {
            return ab ? (1u << R.p) | (1u << R.s) : (1u << R.p) | (1u << R.q) | (1u << R.r) | (1u << R.s);
        };
        const int limit = ab ? kFuseSupportAB : kFuseSupportCC;
        const size_t cap = (size_t)NS * kSchedMax * 3 / 4;
        size_t i = 0;
        while (i < list.size()) {
            std::vector<Rot> grp{list[i]};
            Orbits best = single_orbits(list[i], ab);
            uint32_t sm = support(list[i]);
            size_t j = i + 1;
            // coefficients of a group are read as coef[index + slot]: term indices must be consecutive
            while (j < list.size() && grp.size() < max_group && list[j].index == grp[0].index + (int)grp.size()) {
                uint32_t nm = sm | support(list[j]);
                if (popcnt(nm) > limit) break;
                // closed forms: an orbit with k of the |S| support orbitals occupied holds, per
                // term, C(|S|-2, k-1) ABBA entries (s occupied, p empty) or C(|S|-4, k-2)
                // same-string entries (r, s occupied, p, q empty)
                const int ts = popcnt(nm), gg = (int)grp.size() + 1;
                uint64_t longest = 0;
                for (int k = std::max(0, e - (n - ts)); k <= std::min(e, ts); ++k) {
                    const uint64_t per = ab ? (k >= 1 ? B(ts - 2, k - 1) : 0) : (k >= 2 ? B(ts - 4, k - 2) : 0);
                    longest = std::max<uint64_t>(longest, per * gg);
                }
                const uint64_t total = (uint64_t)gg * (ab ? P.ca_pairs : P.ccaa_pairs);
                if (longest > (uint64_t)kSchedMax || total > cap) break;
                grp.push_back(list[j]);
                sm = nm;
                ++j;
            }
            if (grp.size() > 1) best = ab ? ab_orbits(grp) : cc_orbits(grp);
            // slot count with the lowest cost among those whose schedule fits kSchedMax rows
            // (a single term may exceed it at any slot count: its rows are then split over ops)
            std::vector<uint32_t> order = lpt_order(best);
            int lg512 = lpt_assign(best, order, NS, nullptr);
            while (grp.size() > 1 && lg512 > kSchedMax) {   // the packing did not fit: shrink
                grp.pop_back();
                --j;
                best = grp.size() > 1 ? (ab ? ab_orbits(grp) : cc_orbits(grp)) : single_orbits(grp[0], ab);
                order = lpt_order(best);
                lg512 = lpt_assign(best, order, NS, nullptr);
            }
            emit_orbits(best, order, lg512, ab ? OP_FAB : OP_FCC, grp[0].index, (int)grp.size(),
                        ab ? ((uint32_t)grp[0].q | ((uint32_t)grp[0].r << 8)) : 0u, units);
            i = j;
        }
        if (!key.empty()) {
            std::vector<AbOp> rec;
            for (size_t k = op0; k < P.ops.size(); ++k) {
                const OpDev &o = P.ops[k];
                rec.push_back({o.tab, o.lg, o.ns, o.g, (size_t)(o.index - list[0].index)});
            }
            ab_cache.emplace(std::move(key), std::move(rec));
        }
    };
    // one applyUnitary pqqr term: every entry its own orbit
    auto push_single = [&](const Single &sg, int units) {
        if (sg.entries.empty()) return;
        Orbits ob;
        ob.reserve(sg.entries.size());
        for (uint32_t en : sg.entries) ob.push_back({en});
        std::vector<uint32_t> order = lpt_order(ob);
        emit_orbits(ob, order, lpt_assign(ob, order, NS, nullptr), OP_PQQR, sg.index, 1, sg.gate, units);
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
        pd.pair_lists = 0;
        pd.pair_qr = 0;
        pd.pad2 = 0;
        // mixed windows: the distinct (q, r) of the pass's ops, in order, each op's index in aux
        {
            std::vector<uint32_t> qr;
            for (size_t o = (size_t)op_begin; o < P.ops.size(); ++o) {
                OpDev &op = P.ops[o];
                if (op.type != OP_FAB) continue;
                size_t k = 0;
                while (k < qr.size() && qr[k] != op.aux) ++k;
                if (k == qr.size()) qr.push_back(op.aux);
                if (k >= (size_t)kMaxPairLists) throw std::runtime_error("a mixed window with too many (q, r)");
                op.aux |= (uint32_t)k << 16;
            }
            if (!qr.empty()) {
                pd.pair_lists = (int32_t)qr.size();
                pd.pair_qr = (int32_t)P.pass_qr.size();
                P.pass_qr.insert(P.pass_qr.end(), qr.begin(), qr.end());
            }
        }
        for (const auto &it : items) {
            P.items.push_back(it);
            if (kind != PASS_COLG) P.tile_bytes_moved += 2ull * it.count * line_bytes;
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
        P.ops.push_back(blank(OP_DIAG));
        for (int g = 0; g < ga_count; ++g) push_group_ops(OP_GA, ga_first + g, g, w);
        add_pass(PASS_COL, PF_FROM_PREV, ob, chunk_columns(w), w, ga_count ? ga_first : 0, ga_count);
    }
    // pass 1: beta groups (ROW)
    if (gb_count) {
        int w = lines_for(gb_count);
        if (w < 1) throw std::runtime_error("tile budget too small for the beta excitation groups");
        int ob = (int)P.ops.size();
        for (int g = 0; g < gb_count; ++g) push_group_ops(OP_GB, gb_first + g, g, w);
        add_pass(PASS_ROW, 0, ob, chunk_rows(w), w, gb_first, gb_count);
    }
    if (original) {
        // applyUnitary: the pqqp phase, then every pqqr term in map order; the
        // alpha-pair terms (p_ alpha) precede the beta-pair ones
        int ob = (int)P.ops.size();
        if (!seg[kPQQP].empty()) {
            OpDev d = blank(OP_DIAG);
            d.index = 1;
            P.ops.push_back(d);
        }
        for (const auto &sg : singles)
            if (sg.spin == 0) push_single(sg, W);
        if ((int)P.ops.size() > ob) add_pass(PASS_COL, 0, ob, chunk_columns(W), W, 0, 0);
        ob = (int)P.ops.size();
        for (const auto &sg : singles)
            if (sg.spin == 1) push_single(sg, W);
        if ((int)P.ops.size() > ob) add_pass(PASS_ROW, 0, ob, chunk_rows(W), W, 0, 0);
    }
    // alpha-leading double excitations, p by p: AAAA_p (COL, all columns) then ABBA_p windows
    auto max_orbit = [&](int tsize, int k0, int k1) {
        uint64_t m = 0;
        for (int k = k0; k <= k1; ++k) m = std::max<uint64_t>(m, B(tsize, k));
        return m;
    };
    // Mixed windows need two columns at least.  When the LDS tile holds fewer
    // (complex128 with C above ~5000), their passes rotate in place in global
    // memory (PASS_COLG) with windows of up to kMaxLinesPlan columns.
    const bool gwin = max_orbit(2, std::max(1, e - (n - 2)), std::min(e, 1)) > (uint64_t)W;
    const int WW = gwin ? kMaxLinesPlan : W;
    size_t ri = 0;
    while (ri < rots.size() && rots[ri].cls != OP_BBBB) {
        int p = rots[ri].p;
        std::vector<Rot> aaaa, abba;
        while (ri < rots.size() && rots[ri].cls != OP_BBBB && rots[ri].p == p) {
            (rots[ri].cls == OP_AAAA ? aaaa : abba).push_back(rots[ri]);
            ++ri;
        }
        if (!aaaa.empty()) {
            int ob = (int)P.ops.size();
            push_groups(aaaa, false, W);
            add_pass(PASS_COL, 0, ob, chunk_columns(W), W, 0, 0);
        }
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
                    if (popcnt(nm) > kMaxPairLists + 1) break;   // q and at most kMaxPairLists r
                    int ts = popcnt(nm);
                    int k0 = std::max(1, e - (n - ts)), k1 = std::min(e, ts - 1);
                    if (max_orbit(ts, k0, k1) > (uint64_t)WW) break;
                    tmask = nm;
                }
                ++aj;
            }
            if (aj == ai) throw std::runtime_error("window limit too small for a mixed double excitation");
            std::vector<int> tpos = positions(tmask);
            int ts = (int)tpos.size();
            int k0 = std::max(1, e - (n - ts)), k1 = std::min(e, ts - 1);
            int ob = (int)P.ops.size();
            for (size_t x = ai; x < aj;) {
                size_t y = x;
                while (y < aj && abba[y].r == abba[x].r) ++y;
                push_groups(std::vector<Rot>(abba.begin() + x, abba.begin() + y), true, std::max(1, W / 3));
                x = y;
            }
            // orbits: outside string y (e-k electrons on n-ts positions) x all k-patterns of T
            std::vector<ItemDev> items;
            ItemDev cur{(uint32_t)P.col_list.size(), 0};
            for (int k = k0; k <= k1; ++k) {
                uint32_t ny = (uint32_t)B(n - ts, e - k), npat = (uint32_t)B(ts, k);
                for (uint32_t yi = 0; yi < ny; ++yi) {
                    uint32_t y = expand_outside(unrank_of(yi, e - k, n - ts, B.tab), tpos);
                    if (cur.count + npat > (uint32_t)WW) {
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
            add_pass(gwin ? PASS_COLG : PASS_COL, 0, ob, items, WW, 0, 0);
            ai = aj;
        }
    }
    // pure-beta double excitations (ROW)
    {
        int ob = (int)P.ops.size();
        std::vector<Rot> bbbb(rots.begin() + ri, rots.end());
        push_groups(bbbb, false, W);
        if ((int)P.ops.size() > ob) add_pass(PASS_ROW, 0, ob, chunk_rows(W), W, 0, 0);
    }
    if (order == 2) {
        // mirrored half: passes and the ops in each pass in reverse order; a
        // fused group's schedule rows reversed, so every orbit sees the group's
        // terms backwards (rows of a single term or an excitation group commute)
        const size_t nf = P.passes.size();
        std::map<uint32_t, uint32_t> reversed;   // schedule rows shared by several ops are reversed once
        for (size_t k = nf; k-- > 0;) {
            PassDev pd = P.passes[k];
            pd.flags = 0;
            const int ob = (int)P.ops.size();
            for (int o = P.passes[k].op_end - 1; o >= P.passes[k].op_begin; --o) {
                OpDev op = P.ops[o];
                if ((op.type == OP_FCC || op.type == OP_FAB) && op.g > 1) {
                    auto hit = reversed.find(op.tab);
                    if (hit != reversed.end()) {
                        op.tab = hit->second;
                    } else {
                        const uint32_t ntab = (uint32_t)P.sched.size();
                        for (int i = op.lg - 1; i >= 0; --i)
                            for (int q = 0; q < op.ns; ++q) P.sched.push_back(P.sched[op.tab + (size_t)i * op.ns + q]);
                        reversed.emplace(op.tab, ntab);
                        op.tab = ntab;
                    }
                }
                P.ops.push_back(op);
            }
            pd.op_begin = ob;
            pd.op_end = (int)P.ops.size();
            for (int i = 0; i < pd.item_count; ++i)
                if (pd.kind != PASS_COLG) P.tile_bytes_moved += 2ull * P.items[pd.item_begin + i].count * line_bytes;
            P.passes.push_back(pd);
        }
        P.pair_updates *= 2;
        P.diag_updates *= 2;
    }
    if (P.sched.empty()) P.sched.push_back(pack_sched(kSchedPad, kSchedPad, 0, 0));
    if (P.pass_qr.empty()) P.pass_qr.push_back(0u);
    // the block norm is summed by the last pass; it must see every amplitude
    {
        const PassDev &last = P.passes.back();
        uint64_t covered = 0;
        for (int i = 0; i < last.item_count; ++i) covered += P.items[last.item_begin + i].count;
        if (covered != C || last.kind == PASS_COLG) add_pass(PASS_COL, 0, (int)P.ops.size(), chunk_columns(W), W, 0, 0);
    }
    P.passes.back().flags |= PF_EMIT_NORM;
    return P;
}

Flow build_flow(const Plan &P) {
    Flow F;
    const uint32_t C = P.C;
    if (P.passes.empty() || !(P.passes[0].flags & PF_FROM_PREV))
        throw std::runtime_error("dataflow order: the first pass must read the previous block");
    uint32_t nitems = 0;
    for (const PassDev &ps : P.passes) {
        F.pass_first.push_back(nitems);
        nitems += (uint32_t)ps.item_count;
    }
    F.item_pass.assign(nitems, 0u);
    F.dep_off.assign(nitems + 1, 0u);
    // last writer of each column: a COL item (row_pass = false, id = item number), or
    // every item of ROW pass id (row_pass = true)
    struct Writer {
        bool row_pass = false;
        uint32_t id = 0;
        bool set = false;
    };
    std::vector<Writer> last(C);
    std::vector<uint32_t> mark(nitems, 0xffffffffu);   // dedup: the item whose list holds it
    std::vector<uint32_t> list;
    auto add = [&](uint32_t self, uint32_t d) {
        if (mark[d] != self) {
            mark[d] = self;
            list.push_back(d);
        }
    };
    auto overlaps = [](const ItemDev &a, const ItemDev &b) {
        return a.begin < b.begin + b.count && b.begin < a.begin + a.count;
    };
    for (size_t k = 0; k < P.passes.size(); ++k) {
        const PassDev &ps = P.passes[k];
        if (k > 0 && (ps.flags & PF_FROM_PREV))
            throw std::runtime_error("dataflow order: only the first pass may read the previous block");
        const bool row = ps.kind == PASS_ROW;
        for (int i = 0; i < ps.item_count; ++i) {
            const uint32_t x = F.pass_first[k] + (uint32_t)i;
            const ItemDev &it = P.items[ps.item_begin + i];
            F.item_pass[x] = (uint32_t)k;
            list.clear();
            if (k > 0) {
                auto follow = [&](const Writer &w) {
                    if (!w.set) throw std::runtime_error("dataflow order: a column is read before it is written");
                    if (!w.row_pass) {
                        add(x, w.id);
                        return;
                    }
                    const PassDev &wp = P.passes[w.id];
                    for (int y = 0; y < wp.item_count; ++y)
                        if (!row || overlaps(P.items[wp.item_begin + y], it)) add(x, F.pass_first[w.id] + (uint32_t)y);
                };
                if (row) {
                    // a range of rows of every column; consecutive columns mostly share a writer
                    for (uint32_t c = 0; c < C; ++c)
                        if (c == 0 || last[c].row_pass != last[c - 1].row_pass || last[c].id != last[c - 1].id)
                            follow(last[c]);
                } else {
                    for (uint32_t l = 0; l < it.count; ++l) follow(last[P.col_list[it.begin + l]]);
                }
            }
            std::sort(list.begin(), list.end());
            F.deps.insert(F.deps.end(), list.begin(), list.end());
            F.dep_off[x + 1] = (uint32_t)F.deps.size();
        }
        // the pass's writes become the last ones (its items are disjoint)
        if (row) {
            for (uint32_t c = 0; c < C; ++c) last[c] = Writer{true, (uint32_t)k, true};
        } else {
            for (int i = 0; i < ps.item_count; ++i) {
                const ItemDev &it = P.items[ps.item_begin + i];
                for (uint32_t l = 0; l < it.count; ++l)
                    last[P.col_list[it.begin + l]] = Writer{false, F.pass_first[k] + (uint32_t)i, true};
            }
        }
    }
    for (uint32_t c = 0; c < C; ++c)
        if (!last[c].set) throw std::runtime_error("dataflow order: a column is never written");
    return F;
}

}  // namespace qpe
