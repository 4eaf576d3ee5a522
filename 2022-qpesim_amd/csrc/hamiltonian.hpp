// Host-side term compiler: the product's own restatement of the reference's
// Hamiltonian data model (christopherkang/2022-qpesim, cpp/hamiltonian.hpp:24-184,
// cpp/hamiltonian.cc:160-783).  It turns spatial-orbital integrals into the
// canonical spin-orbital terms, indexes them HalfUp (alpha 0..n-1, beta n..2n-1)
// and returns them as the five map-ordered lists the evolution consumes.
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <vector>

namespace qpe {

// SpinOrbital, cpp/types.hpp:31-98: 1-based orbital, spin Up (0) < Down (1);
// ordered by spin first, then orbital.
struct SpinOrb {
    int orbital = 0;
    int spin = 0;
    bool operator<(const SpinOrb &o) const {
        if (spin != o.spin) return spin < o.spin;
        return orbital < o.orbital;
    }
    bool operator==(const SpinOrb &o) const { return orbital == o.orbital && spin == o.spin; }
    bool operator!=(const SpinOrb &o) const { return !(*this == o); }
};

enum TermClass { kPP = 0, kPQ = 1, kPQQP = 2, kPQQR = 3, kPQRS = 4 };

struct IndexedTerm {
    int idx[4] = {-1, -1, -1, -1};  // HalfUp indices; one-body terms use idx[0..1]
    double coeff = 0.0;
};

class Hamiltonian {
  public:
    // AddOneBodyTerm, cpp/hamiltonian.cc:301-323 (p >= q, 1-based)
    void AddOneBodyTerm(int p, int q, double coeff);
    // AddTwoBodyTerm, cpp/hamiltonian.cc:325-398 (a+_p a+_q a_r a_s, 1-based)
    void AddTwoBodyTerm(int p, int q, int r, int s, double coeff);
    // IndexTerms(HalfUp), cpp/hamiltonian.cc:160-197
    void IndexTerms();
    void SetTimestep(double t) { t_ = t; }
    void SetNumOrbitals(unsigned n) { num_orbitals_ = n; }
    void SetNumOccupied(unsigned e) { num_occupied_ = e; }

    unsigned num_orbitals() const { return num_orbitals_; }
    unsigned num_occupied() const { return num_occupied_; }
    double timestep() const { return t_; }
    bool indexed() const { return indexed_; }

    // PrintIndexedSegmentedHamiltonianTerms, cpp/hamiltonian.cc:253-299
    const std::array<std::vector<IndexedTerm>, 5> &segmented() const { return segmented_; }

  private:
    using OneKey = std::pair<SpinOrb, SpinOrb>;
    using TwoKey = std::tuple<SpinOrb, SpinOrb, SpinOrb, SpinOrb>;
    void accumulate(const OneKey &k, double delta);
    void accumulate(const TwoKey &k, double delta);

    std::map<OneKey, double> one_body_;
    std::map<TwoKey, double> two_body_;
    std::array<std::vector<IndexedTerm>, 5> segmented_;
    unsigned num_orbitals_ = 0;
    unsigned num_occupied_ = 0;
    double t_ = 0.0;
    bool indexed_ = false;
};

}  // namespace qpe
