// Host-side check of the dataflow order of a Trotter step (built by
// tests/test_flow_order.py with the system C++ compiler against the product's
// plan.cpp / hamiltonian.cpp).
//
// usage: flow_check <hamiltonian.txt> <elem_bytes> <tile_bytes> <order> <term_order> [drop]
// hamiltonian.txt as for plan_check.  drop = k > 0 removes the k-th dependency
// (counted over the whole table) before checking: the check must then fail.
//
// Checks:
//   * item_pass matches the passes' item ranges; every dependency is an item of
//     an earlier pass; first-pass items have none;
//   * for every amplitude (row, column) of the block, the items that touch it
//     (COL items: every row of their columns; ROW items: their rows of every
//     column), taken in pass order, each list the previous one as a dependency;
//   * a simulated run (workers claim items of two chained steps in index order,
//     start when the dependencies are done -- first-pass items when the previous
//     step is done --, random durations) finishes with fewer workers than items.
// Prints "ok <items> <deps>" or "FAIL <reason>" (exit status 1).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <queue>
#include <random>
#include <string>
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
    const Plan P = build_plan(h, (size_t)std::atoi(argv[2]), (size_t)std::atol(argv[3]), true, std::atoi(argv[4]),
                              std::atoi(argv[5]));
    Flow F = build_flow(P);
    const long drop = argc > 6 ? std::atol(argv[6]) : 0;
    if (drop > 0) {
        if ((size_t)drop > F.deps.size()) return fail("nothing to drop");
        const size_t at = (size_t)drop - 1;
        F.deps.erase(F.deps.begin() + (long)at);
        for (auto &o : F.dep_off)
            if (o > at) --o;
    }
    const uint32_t N = (uint32_t)F.item_pass.size(), C = P.C;
    if (F.dep_off.size() != N + 1 || F.pass_first.size() != P.passes.size()) return fail("table sizes");
    // item number x of pass k is item P.items[item_begin + x - pass_first[k]]
    auto item_of = [&](uint32_t x) -> const ItemDev & {
        const uint32_t k = F.item_pass[x];
        return P.items[P.passes[k].item_begin + (x - F.pass_first[k])];
    };

    // pass membership and backward dependencies
    for (size_t k = 0; k < P.passes.size(); ++k) {
        const PassDev &ps = P.passes[k];
        if (ps.kind == PASS_COLG) return fail("global-memory windows are not dataflow items");
        if (F.pass_first[k] + (uint32_t)ps.item_count > N) return fail("pass_first beyond the items");
        for (int i = 0; i < ps.item_count; ++i) {
            const uint32_t x = F.pass_first[k] + (uint32_t)i;
            if (F.item_pass[x] != k) return fail("item_pass of item " + std::to_string(x));
            if (k == 0 && F.dep_off[x + 1] != F.dep_off[x]) return fail("first-pass item with dependencies");
            for (uint32_t d = F.dep_off[x]; d < F.dep_off[x + 1]; ++d)
                if (F.deps[d] >= N || F.item_pass[F.deps[d]] >= k)
                    return fail("item " + std::to_string(x) + " depends on a later or same-pass item");
        }
    }
    auto has_dep = [&](uint32_t x, uint32_t d) {
        return std::binary_search(F.deps.begin() + F.dep_off[x], F.deps.begin() + F.dep_off[x + 1], d);
    };

    // every amplitude: consecutive touchers are direct dependencies
    std::vector<uint32_t> prev((size_t)C * C, 0xffffffffu);   // last toucher of cell (row a, column b) at a + b*C
    for (size_t k = 0; k < P.passes.size(); ++k) {
        const PassDev &ps = P.passes[k];
        for (int i = 0; i < ps.item_count; ++i) {
            const uint32_t x = F.pass_first[k] + (uint32_t)i;
            const ItemDev &it = item_of(x);
            auto touch = [&](size_t cell) -> bool {
                const uint32_t w = prev[cell];
                if (w != 0xffffffffu && !has_dep(x, w)) return false;
                return true;
            };
            if (ps.kind == PASS_ROW) {
                for (uint32_t b = 0; b < C; ++b)
                    for (uint32_t a = it.begin; a < it.begin + it.count; ++a)
                        if (!touch(a + (size_t)b * C))
                            return fail("ROW item " + std::to_string(x) + " misses the writer of (" + std::to_string(a) +
                                        ", " + std::to_string(b) + ")");
            } else {
                for (uint32_t l = 0; l < it.count; ++l) {
                    const uint32_t b = P.col_list[it.begin + l];
                    for (uint32_t a = 0; a < C; ++a)
                        if (!touch(a + (size_t)b * C))
                            return fail("COL item " + std::to_string(x) + " misses the writer of (" +
                                        std::to_string(a) + ", " + std::to_string(b) + ")");
                }
            }
        }
        for (int i = 0; i < ps.item_count; ++i) {
            const uint32_t x = F.pass_first[k] + (uint32_t)i;
            const ItemDev &it = item_of(x);
            if (ps.kind == PASS_ROW) {
                for (uint32_t b = 0; b < C; ++b)
                    for (uint32_t a = it.begin; a < it.begin + it.count; ++a) prev[a + (size_t)b * C] = x;
            } else {
                for (uint32_t l = 0; l < it.count; ++l)
                    for (uint32_t a = 0; a < C; ++a) prev[a + (size_t)P.col_list[it.begin + l] * C] = x;
            }
        }
    }
    for (uint32_t v : prev)
        if (v == 0xffffffffu) return fail("an amplitude no item writes");

    // simulated run of two chained steps: claim in index order, event-driven
    {
        const int steps = 2;
        const uint64_t total = (uint64_t)steps * N;
        const int workers = std::max<int>(2, std::min<int>(17, (int)N / 3));
        std::mt19937 rng(12345);
        std::uniform_real_distribution<double> dur(1.0, 10.0);
        std::vector<double> finish(total, -1.0);   // completion time of claim t (step t / N)
        // workers: the time each becomes free
        std::priority_queue<double, std::vector<double>, std::greater<double>> freeq;
        for (int w = 0; w < workers; ++w) freeq.push(0.0);
        std::vector<double> step_done(steps, -1.0);
        std::vector<uint32_t> left(steps, N);
        for (uint64_t t = 0; t < total; ++t) {
            const double claim = freeq.top();
            freeq.pop();
            const uint64_t s = t / N;
            const uint32_t x = (uint32_t)(t % N);
            double start = claim;
            if (F.item_pass[x] == 0) {
                if (s > 0) {
                    if (step_done[s - 1] < 0) return fail("simulated run: a step starts before the previous one ends");
                    start = std::max(start, step_done[s - 1]);
                }
            } else {
                for (uint32_t d = F.dep_off[x]; d < F.dep_off[x + 1]; ++d) {
                    const double fd = finish[s * N + F.deps[d]];
                    if (fd < 0) return fail("simulated run: a dependency was never claimed");
                    start = std::max(start, fd);
                }
            }
            finish[t] = start + dur(rng);
            freeq.push(finish[t]);
            if (--left[s] == 0) {
                double m = 0;
                for (uint32_t y = 0; y < N; ++y) m = std::max(m, finish[s * N + y]);
                step_done[s] = m;
            }
        }
        for (int s = 0; s < steps; ++s)
            if (step_done[s] < 0) return fail("simulated run did not finish");
    }
    std::printf("ok %u %zu\n", N, F.deps.size());
    return 0;
}
