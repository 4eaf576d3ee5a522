/*
 * qpe_oracle.h -- CPU restatement of the reference QPE hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path (libqpesim.so) never calls
 * into it and never falls back to it.
 *
 * Everything here restates christopherkang/2022-qpesim in plain C99:
 *   combinatorics   cpp/util.hpp:49-94, cpp/offset_computer.hpp:16-27
 *   term compiler   cpp/hamiltonian.cc:160-197, 253-657 (AddOneBodyTerm,
 *                   AddTwoBodyTerm, _Reorder*, _UpdateDataModel_*, IndexTerms,
 *                   PrintIndexedSegmentedHamiltonianTerms)
 *   run tables      cpp/time_evolution_over_terms.hpp:1160-1411 (discover*_BVs)
 *   evolution       cpp/time_evolution_over_terms.hpp:92-139, 374-1158
 *   QIFT            cpp/state_vector.hpp:115-179 (MKL forward DFT, /sqrt(M))
 *   P(k)            cpp/state_vector.hpp:341-372
 *
 * Deviation (documented): state indices are 64-bit here; the reference uses
 * 32-bit `unsigned` offsets (time_evolution_over_terms.hpp:37,130-131) and
 * overflows once 2^m * C(n,e)^2 >= 2^32.  Below that size results coincide.
 */
#ifndef QPE_ORACLE_H
#define QPE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* term classes, in the order PrintIndexedSegmentedHamiltonianTerms returns them */
enum { ORC_PP = 0, ORC_PQ = 1, ORC_PQQP = 2, ORC_PQQR = 3, ORC_PQRS = 4 };

/* run-table kinds (discover*_BVs) */
enum { ORC_N = 0, ORC_NN = 1, ORC_CA = 2, ORC_CNA = 3, ORC_CCAA = 4 };

typedef struct {
    int32_t cls;        /* ORC_PP..ORC_PQRS */
    int32_t idx[4];     /* HalfUp spin-orbital indices; pp/pq use idx[0..1] */
    double coeff;       /* accumulated coefficient */
} orc_term;

uint64_t orc_nchoosek(int n, int k);
uint64_t orc_rank(uint64_t bv);
uint64_t orc_next_comb(uint64_t x);
unsigned orc_parity(uint64_t x);

/* Term compiler.
 * one_pq[2*i], one_pq[2*i+1]: Broombridge one-electron indices (1-based), in file order.
 * two_ilj k: Broombridge two-electron entries [a,b,c,d] (1-based, Mulliken), in file order;
 *            read as i=a, l=b, j=c, k=d and fed to AddTwoBodyTerm(i,j,k,l).
 * n_orb: number of spatial orbitals (max_number_of_orbitals).
 * Returns the number of indexed terms written to out (<= cap), or -1 on error.
 * counts[5] receives the per-class counts; terms are written class by class in
 * map order (pp, pq, pqqp, pqqr, pqrs). */
int64_t orc_compile_terms(int n_orb,
                          int64_t n_one, const int32_t *one_pq, const double *one_val,
                          int64_t n_two, const int32_t *two_abcd, const double *two_val,
                          orc_term *out, int64_t cap, int64_t counts[5]);

/* Run table of one discover*_BVs call.  args = orbital indices within one spin
 * string (0..n-1), in the order the reference passes them.  Returns the number
 * of parts written (<= cap) or -1. */
int64_t orc_run_table(int kind, int n, int e, const int32_t *args,
                      uint64_t *off01, uint64_t *off10, uint64_t *niter,
                      uint32_t *mid_parity, int64_t cap);

/* time_evolve_over_terms on a full [M][C][C] complex128 state (interleaved
 * re,im).  Applies blocks j = j_begin .. j_end-1 (j_begin >= 1): block j =
 * copy(block j-1), then one applyUnitary_reordered.  n = spatial orbitals,
 * e = electrons per spin.  Terms as produced by orc_compile_terms. */
int orc_time_evolve(int n, int e, int m, const orc_term *terms, int64_t n_terms,
                    double t, double *state, int64_t j_begin, int64_t j_end);

/* The same with original != 0: applyUnitary's order (TEOT:174-372) per block. */
int orc_time_evolve_ordered(int n, int e, int m, const orc_term *terms, int64_t n_terms,
                            double t, double *state, int64_t j_begin, int64_t j_end, int original);

/* Apply exactly one applyUnitary_reordered to one block (C*C complex128). */
int orc_apply_unitary(int n, int e, const orc_term *terms, int64_t n_terms,
                      double t, double *block);

/* Apply exactly one applyUnitary (original order: pp, pq, pqqp, every pqqr,
 * pqrs; TEOT:174-372) to one block. */
int orc_apply_unitary_original(int n, int e, const orc_term *terms, int64_t n_terms,
                               double t, double *block);

/* time_evolve_iqpe, TEOT:141-155: num_of_applications applyUnitary calls in
 * place on one block. */
int orc_time_evolve_iqpe(int n, int e, const orc_term *terms, int64_t n_terms, double t, double *block,
                         int64_t num_of_applications);

/* Count det-term updates of one applyUnitary_reordered (pairs + diagonal amps). */
int orc_count_updates(int n, int e, const orc_term *terms, int64_t n_terms,
                      uint64_t *diag_updates, uint64_t *pair_updates);

/* StateVector::qift on [M][ld] complex128, in place. */
int orc_qift(int m, uint64_t ld, double *state);

/* IdentifyPhaseAmplitudes on [M][ld] complex128. */
int orc_phase_probabilities(int m, uint64_t ld, const double *state, double *probs);

/* threads the OpenMP runtime will use */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
