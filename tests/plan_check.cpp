// Host-side check of the plan's op schedules (built by tests/test_plan_schedules.py
// with the system C++ compiler against the product's plan.cpp / hamiltonian.cpp).
//
// usage: plan_check <hamiltonian.txt> <elem_bytes> <tile_bytes> <order> <term_order>
// hamiltonian.txt: "n e tau", then "O p q v" and "T i j k l v" lines.
//
// Checks, for every op with a schedule:
//   * 1 <= ns <= kThreadsPlan, 1 <= lg <= kSchedMax, the rows lie inside the table;
//   * the lane/slot split the pass kernel derives from ns without a division
//     (lane = (x * ceil(2^20 / ns)) >> 20, L = kThreadsPlan / ns) equals x / ns;
//   * entries hold ranks below C and a term slot below the group size;
//   * every double-excitation group (the ops sharing its first term) lists each
//     of its terms' determinant pairs exactly once: C(n-2, e-1) alpha pairs per
//     mixed term, C(n-4, e-2) pairs per same-spin term.
// Prints "ok <ops> <entries>" or "FAIL <reason>" (exit status 1).
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "plan.hpp"

using namespace qpe;

static int fail(const std::string &why) {
    std::printf("FAIL %s\n", why.c_str());
    return 1;
}

int main(int argc, char **argv) {
    if (argc < 6) return fail("usage");
    FILE *f = std::fopen(argv[1], "r");
    if (!f) return fail("cannot open hamiltonian");
    int n = 0, e = 0;
    double tau = 0;
    if (std::fscanf(f, "%d %d %lf", &n, &e, &tau) != 3) return fail("bad header");
    Hamiltonian h;
    h.SetNumOrbitals(n);
    h.SetNumOccupied(e);
    char kind[8];
    while (std::fscanf(f, "%7s", kind) == 1) {
        if (kind[0] == 'O') {
            int p, q;
            double v;
            if (std::fscanf(f, "%d %d %lf", &p, &q, &v) != 3) return fail("bad one-body line");
            h.AddOneBodyTerm(p, q, v);
        } else {
            int a, b, c, d;
            double v;
            if (std::fscanf(f, "%d %d %d %d %lf", &a, &b, &c, &d, &v) != 5) return fail("bad two-body line");
            h.AddTwoBodyTerm(a, b, c, d, v);
        }
    }
    std::fclose(f);
    h.IndexTerms();
    h.SetTimestep(tau);
    const int order = std::atoi(argv[4]);
    const Plan P = build_plan(h, (size_t)std::atoi(argv[2]), (size_t)std::atol(argv[3]), true, order,
                              std::atoi(argv[5]));
    if (order == 2 && P.ops.size() % 2) return fail("second-order plan with an odd op count");

    // (type, first term, group size, half of a second-order step) -> entries (ranks,
    // slot) over the ops of the group (a second-order step repeats every group mirrored)
    std::map<std::tuple<int, int, int, int>, std::multiset<std::tuple<uint32_t, uint32_t, uint32_t>>> groups;
    size_t entries = 0;
    for (size_t k = 0; k < P.ops.size(); ++k) {
        const OpDev &o = P.ops[k];
        if (o.type == OP_DIAG) continue;
        const std::string at = "op " + std::to_string(k) + " type " + std::to_string(o.type);
        if (o.ns < 1 || o.ns > kThreadsPlan) return fail(at + ": ns out of range");
        if (o.lg < 1 || o.lg > kSchedMax) return fail(at + ": lg out of range");
        if ((size_t)o.tab + (size_t)o.lg * o.ns > P.sched.size()) return fail(at + ": rows outside the table");
        const uint32_t ns = (uint32_t)o.ns, magic = ((1u << 20) + ns - 1) / ns;
        for (uint32_t x = 0; x < (uint32_t)kThreadsPlan; ++x)
            if ((uint32_t)(((uint64_t)x * magic) >> 20) != x / ns) return fail(at + ": lane split");
        if (o.type != OP_FCC && o.type != OP_FAB && o.type != OP_PQQR) continue;
        for (int i = 0; i < o.lg; ++i)
            for (uint32_t t = 0; t < ns; ++t) {
                const uint32_t en = P.sched[(size_t)o.tab + (size_t)i * ns + t];
                const uint32_t r1 = en & kSchedPad, r2 = (en >> kSchedRankBits) & kSchedPad, slot = (en >> 28) & 7u;
                if (r1 == kSchedPad) continue;
                if (r1 >= P.C || r2 >= P.C) return fail(at + ": rank beyond C");
                if ((int)slot >= o.g) return fail(at + ": term slot beyond the group");
                groups[{o.type, o.index, o.g, order == 2 && 2 * k >= P.ops.size() ? 1 : 0}].insert({r1, r2, slot});
                ++entries;
            }
    }
    for (const auto &kv : groups) {
        const int type = std::get<0>(kv.first), g = std::get<2>(kv.first);
        if (type == OP_PQQR) continue;
        const uint32_t per = type == OP_FAB ? P.ca_pairs : P.ccaa_pairs;
        std::vector<uint32_t> count(g, 0);
        const std::tuple<uint32_t, uint32_t, uint32_t> *prev = nullptr;
        for (const auto &en : kv.second) {
            if (prev && *prev == en) return fail("group at term " + std::to_string(std::get<1>(kv.first)) + ": repeated pair");
            prev = &en;
            ++count[std::get<2>(en)];
        }
        for (int s = 0; s < g; ++s)
            if (count[s] != per)
                return fail("group at term " + std::to_string(std::get<1>(kv.first)) + ": term " + std::to_string(s) +
                            " has " + std::to_string(count[s]) + " pairs, expected " + std::to_string(per));
    }
    std::printf("ok %zu %zu\n", P.ops.size(), entries);
    return 0;
}
