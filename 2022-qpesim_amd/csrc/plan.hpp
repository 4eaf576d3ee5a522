// Evolution plan: one Trotter step U (the reference's applyUnitary_reordered,
// cpp/time_evolution_over_terms.hpp:374-562) compiled into a short list of
// tile passes over one block [alpha][beta] of the register.
//
// Block layout in HBM: column-major within a block, amplitude (a, b) of block
// j at j*C*C + b*C + a (a = alpha rank, b = beta rank).  A "column" (fixed
// beta string, all alpha strings) is contiguous.
//
// Pass kinds
//   COL  a workgroup stages a set of whole columns in LDS and applies, in
//        reference order, the diagonal phase, the alpha excitation groups, the
//        pure-alpha double excitations and mixed alpha-beta double excitations
//        whose beta orbitals lie in the pass's orbit set T (the columns a
//        workgroup holds are closed under those terms).
//   ROW  a workgroup stages a set of whole rows (fixed alpha string) and
//        applies beta excitation groups or pure-beta double excitations.
//
// Exact fusions (mathematically identical to the reference's sequence):
//   * all pp and pqqp terms commute: one diagonal phase per amplitude;
//   * a pq term and the pqqr terms applied right after it (TEOT:428-513) act
//     on the same determinant pairs with the same Jordan-Wigner sign and commute,
//     so their product is one rotation whose angle is t*(h_pq + sum of the
//     pqqr coefficients whose inner orbital is occupied).
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "hamiltonian.hpp"

namespace qpe {

enum OpType : int32_t {
    OP_DIAG = 0,   // diagonal phase (all pp + pqqp)
    OP_GA = 1,     // alpha excitation group (pq + gated pqqr)
    OP_GB = 2,     // beta excitation group
    OP_AAAA = 3,   // term classes of pqrs (the ops below carry them)
    OP_BBBB = 4,
    OP_ABBA = 5,
    OP_FCC = 6,    // consecutive AAAA (COL pass) or BBBB (ROW pass) terms, one or more
    OP_FAB = 7,    // consecutive ABBA terms sharing (p, q, r), one or more
    OP_PQQR = 8    // one pqqr term of applyUnitary's order: like FCC, on the lines whose
                   // string has a bit of aux (aux = 0: every line)
};

// COLG: like COL, but the lines stay in global memory (windows of mixed double
// excitations when the LDS tile cannot hold two columns); its ops are FAB only.
enum PassKind : int32_t { PASS_COL = 0, PASS_ROW = 1, PASS_COLG = 2 };
constexpr int kMaxLinesPlan = 64;   // lines per workgroup (line tables in LDS)
// Column-pair lists of a mixed-window pass: one per distinct (q, r) of its ops, at most
// kMaxPairLists (a window holds at most that many r values), each of at most
// kMaxLinesPlan / 2 pairs; built in LDS when an item starts.
constexpr int kMaxPairLists = 8;
// LDS bytes of the pass kernel's header (line tables, pair lists, op rings) before the tile
constexpr size_t kLdsHeader = 6208;

enum PassFlags : int32_t {
    PF_FROM_PREV = 1,  // read block j-1 (scaled by its renormalisation factor), write block j
    PF_EMIT_NORM = 2   // write per-workgroup FP64 partial sums of |a|^2
};

// Every op but DIAG carries a schedule: ns <= kThreadsPlan slots of
// lg <= kSchedMax entries, entry i of slot t at sched[tab + i*ns + t].  Thread
// x of the pass kernel serves slot x % ns as lane x / ns of L = kThreadsPlan / ns
// (threads past ns * L idle); it applies its slot's entries in order, each to
// its share (lane, lane + L, ...) of the tile's lines (FCC, GA, GB, PQQR) or
// column pairs (FAB).
struct OpDev {
    int32_t type;
    int32_t index;   // FCC/FAB/PQQR: first term (coefficients coef[index .. index+g)); GA/GB: offset of
                     // the pair angles; DIAG: diagonal table set
    uint32_t tab;    // offset of the schedule
    uint32_t aux;    // GA/GB: staged-angle slot; FAB: q | r << 8 (beta orbitals) | pair list << 16;
                     // PQQR: line gate mask
    int32_t lg;      // schedule rows
    int32_t ns;      // schedule slots (kThreadsPlan)
    int32_t g;       // terms in the group
    uint32_t split;  // set on upload: ceil(2^20 / ns) | (kThreadsPlan / ns) << 21, so that the
                     // kernel splits threads into slots and lanes without a division
};

// Schedule entry: two 14-bit ranks, the term's slot in the group (3 bits) and
// the Jordan-Wigner parity.  Ranks equal to kSchedPad mark padding.
constexpr uint32_t kSchedRankBits = 14;
constexpr uint32_t kSchedPad = (1u << kSchedRankBits) - 1u;
constexpr int kMaxGroup = 8;
constexpr int kSchedMax = 4;        // schedule rows per op (entries per thread)
// threads per workgroup of the pass kernel (a build-time choice: 256, 512 or 1024)
#ifndef QPE_PASS_THREADS
#define QPE_PASS_THREADS 1024
#endif
constexpr int kThreadsPlan = QPE_PASS_THREADS;
static_assert(kThreadsPlan == 256 || kThreadsPlan == 512 || kThreadsPlan == 1024, "pass kernel workgroup size");
constexpr int kFuseSupportCC = 7;   // max spin-orbitals touched by a fused AAAA/BBBB group
constexpr int kFuseSupportAB = 5;   // max alpha orbitals touched by a fused ABBA group
inline uint32_t pack_sched(uint32_t r1, uint32_t r2, uint32_t slot, uint32_t par) {
    return r1 | (r2 << kSchedRankBits) | (slot << 28) | (par << 31);
}

// A pair-table entry packs two ranks and the Jordan-Wigner parity into 32 bits:
// bits 0-14 rank of the first string, 15-29 rank of the second, 31 parity.

constexpr uint32_t kRankBits = 15;
constexpr uint32_t kRankMask = (1u << kRankBits) - 1u;
inline uint32_t pack_pair(uint32_t r1, uint32_t r2, uint32_t par) {
    return r1 | (r2 << kRankBits) | (par << 31);
}

// Rotations are stored as (1 - cos a, sin a).  Applying u' = u - (1-cos a) u
// + ... keeps the small quantity 1 - cos a at full relative precision when the
// state is complex64: rounding cos a (close to 1) to FP32 would scale every
// amplitude a term touches by the same factor on every Trotter step.
inline double versine(double a) {
    const double h = std::sin(0.5 * a);
    return 2.0 * h * h;
}

struct GroupDev {
    int32_t x, y;    // orbitals of the excitation inside its spin string (x > y)
    int32_t spin;    // 0 alpha, 1 beta
    int32_t pad;
};

struct TermDev {
    int32_t p, q, r, s;   // local orbital indices (alpha p,s / beta q,r for ABBA)
    double cs, sn;        // cos(c t), sin(c t)
};

struct PassDev {
    int32_t kind;
    int32_t flags;
    int32_t op_begin, op_end;
    int32_t item_begin, item_count;
    int32_t max_lines;    // columns (COL) or rows (ROW) per workgroup
    int32_t group_first;  // first excitation group whose angle table is staged in LDS
    int32_t group_count;  // number of staged groups (0: none)
    int32_t pair_lists;   // mixed windows: distinct (q, r) of the ops (FAB op aux bits 16-18 index them)
    int32_t pair_qr;      // offset of their q | r << 8 values in Plan::pass_qr
    int32_t pad2;
};

struct ItemDev {
    uint32_t begin;       // COL: offset into the column list; ROW: first row
    uint32_t count;       // number of columns / rows
};

struct Plan {
    int n = 0, e = 0;
    int order = 1;             // Trotter order (1: the reference's sequence; 2: symmetric)
    int term_order = 0;        // 0: applyUnitary_reordered's term order; 1: applyUnitary's
    uint32_t C = 0;
    uint32_t ca_pairs = 0;     // C(n-2, e-1)
    uint32_t ccaa_pairs = 0;   // C(n-4, e-2)
    size_t elem_bytes = 16;
    size_t tile_bytes = 0;     // LDS bytes per workgroup for the amplitude tile

    std::vector<uint32_t> strings;      // [C] the colex-ordered strings of one spin
    std::vector<double> theta_a, theta_b, cross;   // diag, per table set: [C], [C], [n][C]
    std::vector<GroupDev> groups;
    std::vector<double> grp_a;          // GA/GB ops: [op.index + i*ns + t][2] 1-cos, sin of the pair angle part
    std::vector<double> grp_b;          // [G][C][2] 1-cos, sin of the other-string angle part
    std::vector<TermDev> terms;
    std::vector<OpDev> ops;
    std::vector<PassDev> passes;
    std::vector<ItemDev> items;
    std::vector<uint32_t> col_list;     // column ranks of COL items
    std::vector<uint32_t> ca_table;     // host only: [(x*n+y)*ca_pairs + k]: rank(f|x), rank(f|y), parity
    std::vector<uint32_t> sched;        // op schedules: [op.tab + i*op.ns + slot]
    std::vector<uint32_t> pass_qr;      // mixed windows' pair lists: q | r << 8, per pass from pair_qr
    std::vector<double> coef;           // [term][2]: 1-cos(c t), sin(c t) of every double excitation
    size_t lds_fixed = kLdsHeader;      // LDS bytes before the tile

    // work accounting per U
    uint64_t pair_updates = 0;   // determinant-pair rotations of the reference's terms
    uint64_t diag_updates = 0;   // diagonal multiplies of the reference's terms
    uint64_t tile_bytes_moved = 0;   // HBM bytes the passes load + store
    int max_items = 0;
};

// Builds the plan.  tile_bytes: LDS budget for the amplitude tile.  fuse: group
// consecutive double excitations with small joint support into one op.
// order: 1 reproduces the reference's first-order product; 2 builds the
// symmetric second-order step (forward sequence at t/2, then reversed at t/2).
// term_order: 0 is applyUnitary_reordered's sequence (TEOT:374-562: pp, pqqp,
// each pq with the pqqr terms it gates, pqrs), the one time_evolve_over_terms
// runs; 1 is applyUnitary's (TEOT:174-372: pp, pq, pqqp, every pqqr, pqrs),
// the one time_evolve_iqpe runs.
// Throws std::runtime_error with a message on unsupported sizes.
Plan build_plan(const Hamiltonian &h, size_t elem_bytes, size_t tile_bytes, bool fuse = true, int order = 1,
                int term_order = 0);

// Dataflow order of one Trotter step: the items of every pass, numbered in pass
// order (a second-order step's mirrored passes reuse the forward passes' item
// ranges, so an item here is a (pass, item) pair), with, for each, the items of
// earlier passes of the same step whose writes it reads.  A COL item reads and writes whole columns: it follows the
// last writer of each of its columns (a COL item, or every item of a ROW pass).
// A ROW item reads and writes a range of rows of every column: it follows the
// last COL writer of every column and the items of a last ROW pass whose rows
// overlap its own.  Items of the first pass (PF_FROM_PREV) read the previous
// block only and follow the previous step as a whole.  Every dependency points
// to an earlier item, so items claimed in index order can always finish.
struct Flow {
    std::vector<uint32_t> item_pass;   // [items] pass of each item
    std::vector<uint32_t> pass_first;  // [passes] number of the pass's first item
    std::vector<uint32_t> dep_off;     // [items + 1] offsets into deps
    std::vector<uint32_t> deps;        // item numbers
};
Flow build_flow(const Plan &P);

}  // namespace qpe
