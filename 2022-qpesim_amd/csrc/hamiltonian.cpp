// Term compiler; see hamiltonian.hpp.  The canonical forms and the per-key
// accumulation order must equal the reference's (cpp/hamiltonian.cc:325-657),
// because the evolution's term order and coefficients are part of the
// reference's result.
#include "hamiltonian.hpp"

#include <algorithm>
#include <set>
#include <stdexcept>

namespace qpe {

namespace {

using Quad = std::array<SpinOrb, 4>;

// Canonical key + coefficient of one spin-resolved two-body product
// a+_0 a+_1 a_2 a_3.  Returns 0 for products the reference discards.
// PQQP: cpp/hamiltonian.cc:467-497, PQQR: :499-545, PQRS: :547-619.
double canonicalize(Quad &t, double v) {
    std::set<SpinOrb> distinct(t.begin(), t.end());
    SpinOrb &p = t[0], &q = t[1], &r = t[2], &s = t[3];
    switch (distinct.size()) {
    case 2: {  // n_p n_q type: halve, make the creation pair ascending
        if (r == s || p == q) return 0.0;
        double c = v / 2.0;
        if (p == r) { std::swap(r, s); c = -c; }
        if (!(p < q)) { std::swap(p, q); std::swap(r, s); }
        return c;
    }
    case 3: {  // a+_p n_q a_r type: quarter, move the repeated orbital inside
        if (p == q || r == s) return 0.0;
        double c = v / 4.0;
        if (q == r) {
        } else if (q == s) {
            c = -c; std::swap(r, s);
        } else if (p == r) {
            c = -c; std::swap(p, q);
        } else if (p == s) {
            std::swap(p, q); std::swap(r, s);
        } else {
            throw std::runtime_error("unknown pqqr product");
        }
        if (s < p) std::swap(p, s);
        return c;
    }
    case 4: {  // double excitation: quarter, lower pair first, alphas outside
        int alphas = (p.spin == 0) + (q.spin == 0) + (r.spin == 0) + (s.spin == 0);
        double c = v / 4.0;
        if (!(std::min(p, q) < std::min(r, s))) { std::swap(p, r); std::swap(q, s); }
        if (alphas == 2) {
            if (q.spin == 0) { std::swap(p, q); c = -c; }
            if (r.spin == 0) { std::swap(r, s); c = -c; }
        } else {
            if (q.orbital < p.orbital) { std::swap(p, q); c = -c; }
            if (r.orbital < s.orbital) { std::swap(r, s); c = -c; }
        }
        return c;
    }
    default:
        throw std::runtime_error("unknown two-body product");
    }
}

}  // namespace

void Hamiltonian::accumulate(const OneKey &k, double delta) {
    if (delta == 0.0) return;  // cpp/hamiltonian.cc:628-632
    one_body_[k] = one_body_[k] + delta;
}

void Hamiltonian::accumulate(const TwoKey &k, double delta) {
    if (delta == 0.0) return;  // cpp/hamiltonian.cc:646-650
    two_body_[k] = two_body_[k] + delta;
}

void Hamiltonian::AddOneBodyTerm(int p, int q, double coeff) {
    if (coeff == 0.0) return;
    indexed_ = false;
    for (int spin = 0; spin < 2; ++spin) accumulate(OneKey{{p, spin}, {q, spin}}, coeff);
}

void Hamiltonian::AddTwoBodyTerm(int p, int q, int r, int s, double coeff) {
    if (coeff == 0.0) return;
    indexed_ = false;
    if (p == q && q == r && r == s) {  // n_{p,up} n_{p,down}: cpp/hamiltonian.cc:338-350
        accumulate(TwoKey{{p, 0}, {p, 1}, {p, 1}, {p, 0}}, coeff);
        return;
    }
    // the eight real-orbital symmetries of the integral, visited in sorted order
    // without repeats (cpp/hamiltonian.cc:400-417 keeps them in a std::set)
    std::set<std::array<int, 4>> images = {{p, q, r, s}, {s, q, r, p}, {p, r, q, s}, {s, r, q, p},
                                           {q, p, s, r}, {r, p, s, q}, {q, s, p, r}, {r, s, p, q}};
    for (const auto &im : images) {
        // spin assignment: creation/annihilation outer pair shares one spin,
        // inner pair the other (cpp/hamiltonian.cc:440-463, inner spin varies slowest)
        for (int inner = 0; inner < 2; ++inner)
            for (int outer = 0; outer < 2; ++outer) {
                Quad t = {SpinOrb{im[0], outer}, SpinOrb{im[1], inner}, SpinOrb{im[2], inner},
                          SpinOrb{im[3], outer}};
                double c = canonicalize(t, coeff);
                accumulate(TwoKey{t[0], t[1], t[2], t[3]}, c);
            }
    }
}

void Hamiltonian::IndexTerms() {
    const int n = static_cast<int>(num_orbitals_);
    auto half_up = [n](const SpinOrb &s) { return s.spin * n + s.orbital - 1; };
    for (auto &v : segmented_) v.clear();
    // HalfUp is monotone in (spin, orbital), so map order carries over unchanged.
    for (const auto &kv : one_body_) {
        IndexedTerm t;
        t.idx[0] = half_up(kv.first.first);
        t.idx[1] = half_up(kv.first.second);
        t.coeff = kv.second;
        segmented_[t.idx[0] == t.idx[1] ? kPP : kPQ].push_back(t);
    }
    for (const auto &kv : two_body_) {
        IndexedTerm t;
        t.idx[0] = half_up(std::get<0>(kv.first));
        t.idx[1] = half_up(std::get<1>(kv.first));
        t.idx[2] = half_up(std::get<2>(kv.first));
        t.idx[3] = half_up(std::get<3>(kv.first));
        t.coeff = kv.second;
        TermClass c = (t.idx[1] == t.idx[2]) ? (t.idx[0] == t.idx[3] ? kPQQP : kPQQR) : kPQRS;
        segmented_[c].push_back(t);
    }
    indexed_ = true;
}

}  // namespace qpe
