// Broombridge ingest: a reader for the YAML subset that Broombridge v0.1/v0.2
// molecular-Hamiltonian files are written in, and the product's restatement of
// Hamiltonian::IngestBroombridge (christopherkang/2022-qpesim,
// cpp/hamiltonian.cc:8-158, ParsePolishNotation :146-158).
//
// YAML subset: block mappings and sequences (indentation, "- " items, nested
// "- - x"), flow mappings and sequences ("{a: 1}", "[1, (1a)+]", spanning
// lines, trailing commas), plain / single- / double-quoted scalars, "#"
// comments, a leading "---".  Anchors, aliases, tags and block scalars ("|",
// ">") are rejected with an error naming the line.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "hamiltonian.hpp"

namespace qpe {

struct FormatError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct YamlNode {
    enum Kind { kNull, kScalar, kSeq, kMap };
    Kind kind = kNull;
    std::string scalar;
    int line = 0;  // 1-based source line
    std::vector<YamlNode> seq;
    std::vector<std::pair<std::string, YamlNode>> map;

    const YamlNode *find(const std::string &key) const;  // first entry with this key
    const YamlNode &at(const std::string &key) const;    // throws FormatError when absent
    const YamlNode &at(size_t i) const;
    double as_double() const;      // finite or not; the caller checks
    unsigned as_unsigned() const;  // decimal digits only
};

YamlNode parse_yaml(const std::string &text);

// One determinant of the suggested initial state: its index in block 0
// (C(n,e) * rank(alpha) + rank(beta)) and its coefficient with the
// Jordan-Wigner sign of the operator string applied.
struct Determinant {
    uint64_t address = 0;
    double coeff = 0.0;
};

struct BroombridgeData {
    double identity = 0.0;  // coulomb_repulsion + energy_offset
    unsigned num_orbitals = 0;
    unsigned num_occupied = 0;
    std::vector<std::pair<std::pair<unsigned, unsigned>, double>> one_body;  // (p >= q), 1-based
    std::vector<std::pair<std::array<unsigned, 4>, double>> two_body;        // (i, j, k, l), 1-based
    std::vector<Determinant> initial;                                         // file order
};

// Parses and checks a whole Broombridge document.  prior_orbitals is the
// orbital count already held by the Hamiltonian the data will be added to:
// like the reference, the count is the running maximum of the integral indices.
BroombridgeData read_broombridge(const std::string &text, unsigned prior_orbitals);

// Adds the integrals to h and sets its orbital / occupied counts.
void apply_broombridge(const BroombridgeData &d, Hamiltonian &h);

}  // namespace qpe
