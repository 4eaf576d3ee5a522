/*
 * qpesim.h -- C ABI of the MI355X-native QPE simulator core (libqpesim.so).
 *
 * The entry points stand in for the reference call sequence of
 * christopherkang/2022-qpesim, cpp/main.cc:99-134:
 *
 *   Hamiltonian h; h.AddOneBodyTerm/AddTwoBodyTerm (via IngestBroombridge);
 *   h.IndexTerms(HalfUp); h.SetTimestep(t);
 *   StateVector sv(m, n_orb, n_occ); sv[i] = ...; sv[0..C^2) *= 1/sqrt(2^m);
 *   time_evolve_over_terms(sv, h, m); sv.qift(); probs = sv.IdentifyPhaseAmplitudes();
 *
 * Plain C types only.  Every call returns QPE_OK (0) or a QPE_ERR_* code; the
 * message of the last failure on the calling thread is qpe_last_error().
 * A handle is used by one host thread at a time; each state handle owns one
 * HIP stream, and calls on it are stream-ordered.  Amplitudes cross the ABI as
 * interleaved (re, im) doubles in the reference's index order
 * j*C^2 + rank_alpha*C + rank_beta (cpp/time_evolution_over_terms.hpp:169-171).
 */
#ifndef QPESIM_H
#define QPESIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum qpe_status {
    QPE_OK = 0,
    QPE_ERR_ARG = 1,      /* invalid handle or argument */
    QPE_ERR_OOM = 2,      /* request larger than the free device memory */
    QPE_ERR_HIP = 3,      /* a HIP runtime call failed */
    QPE_ERR_NUMERIC = 4,  /* non-finite amplitudes detected */
    QPE_ERR_STATE = 5,    /* call out of sequence (e.g. evolve before IndexTerms) */
    QPE_ERR_STALLED = 6   /* a wait of the dataflow evolution exceeded QPE_OPT_WAIT_LIMIT_MS;
                             the launch stopped and the blocks of that call are invalid */
};

enum qpe_dtype {
    QPE_COMPLEX128 = 0,   /* FP64 state and rotations (reference precision) */
    QPE_COMPLEX64 = 1     /* FP32 state and rotations; FP64 angles, norms, P(k) */
};

/* register layouts */
enum qpe_layout {
    QPE_LAYOUT_ALPHA_BETA = 0,    /* the reference's [2^m][C(n,e)][C(n,e)] product space (default) */
    QPE_LAYOUT_SINGLE_STRING = 1  /* [2^m][C(2n,2e)]: one string of 2n spin-orbitals, interleaved
                                     Jordan-Wigner order (no counterpart in the reference) */
};

/* term classes, in the order PrintIndexedSegmentedHamiltonianTerms returns them */
enum qpe_term_class { QPE_PP = 0, QPE_PQ = 1, QPE_PQQP = 2, QPE_PQQR = 3, QPE_PQRS = 4 };

typedef struct qpe_hamiltonian *qpe_hamiltonian_t;
typedef struct qpe_state *qpe_state_t;

/* ---------------- Hamiltonian (term compiler, host) ---------------- */

/* Hamiltonian::Hamiltonian(), cpp/hamiltonian.hpp:27 */
int qpe_hamiltonian_create(qpe_hamiltonian_t *out);
int qpe_hamiltonian_destroy(qpe_hamiltonian_t h);

/* Hamiltonian::AddOneBodyTerm, cpp/hamiltonian.cc:301-323.  p, q are 1-based
 * spatial orbitals with p >= q (IngestBroombridge swaps, cpp/hamiltonian.cc:29-32). */
int qpe_hamiltonian_add_one_body_term(qpe_hamiltonian_t h, int p, int q, double coeff);

/* Hamiltonian::AddTwoBodyTerm, cpp/hamiltonian.cc:325-398: term a+_p a+_q a_r a_s
 * over 1-based spatial orbitals; a Broombridge Mulliken entry [a,b,c,d] is
 * passed as (p,q,r,s) = (a,c,d,b) (cpp/hamiltonian.cc:43-48). */
int qpe_hamiltonian_add_two_body_term(qpe_hamiltonian_t h, int p, int q, int r, int s, double coeff);

/* IngestBroombridge fills these from the file (cpp/hamiltonian.cc:34-35,113-115);
 * Hamiltonian::SetNumOccupied, cpp/hamiltonian.hpp:72. */
int qpe_hamiltonian_set_num_orbitals(qpe_hamiltonian_t h, unsigned num_orbitals);
int qpe_hamiltonian_set_num_occupied(qpe_hamiltonian_t h, unsigned num_occupied);

/* Hamiltonian::IndexTerms(SpinIndex::HalfUp), cpp/hamiltonian.cc:160-197 */
int qpe_hamiltonian_index_terms(qpe_hamiltonian_t h);

/* Hamiltonian::SetTimestep, cpp/hamiltonian.hpp:71 */
int qpe_hamiltonian_set_timestep(qpe_hamiltonian_t h, double t);

/* Hamiltonian::IngestBroombridge, cpp/hamiltonian.cc:8-144 (the Hamiltonian
 * half; qpe_state_ingest_broombridge below is the whole call).  Reads a
 * Broombridge v0.1/v0.2 file: one-body [p,q,v] (swapped to p >= q) and
 * two-body Mulliken [i,l,j,k,v] -> AddTwoBodyTerm(i,j,k,l) integrals of
 * integral_sets[0], 1-based; identity = coulomb_repulsion + energy_offset;
 * num_orbitals = the largest integral index (running maximum with what h held);
 * num_occupied and the initial determinants from
 * initial_state_suggestions[0].state.superposition, each entry
 * [coefficient, ladder operators..., vacuum] applied right to left with the
 * Jordan-Wigner sign (-1)^(occupied modes below), modes alpha 0..n-1 then beta
 * n..2n-1 (cpp/hamiltonian.cc:55-128, ParsePolishNotation :146-158).  The file
 * is checked completely before h changes.  QPE_ERR_ARG (message names the
 * line) for an unreadable file, malformed YAML or integrals, an operator string
 * other than "(<k><a|b>)" / "(<k><a|b>)+", creation on an occupied or
 * annihilation of an empty orbital (the reference asserts), unequal alpha and
 * beta counts or determinants of different electron counts (the reference
 * computes a meaningless address). */
int qpe_hamiltonian_ingest_broombridge(qpe_hamiltonian_t h, const char *path);
/* The same from a document held in memory (length bytes, no NUL needed). */
int qpe_hamiltonian_ingest_broombridge_text(qpe_hamiltonian_t h, const char *text, uint64_t length);

/* The initial state of the last ingest: *count determinants in file order,
 * address[i] = C(n,e)*rank(alpha) + rank(beta) within block 0 and coeff[i]
 * with its sign applied (cpp/hamiltonian.cc:117-128).  With address and coeff
 * both NULL only *count is written; otherwise cap must be >= *count. */
int qpe_hamiltonian_initial_state(qpe_hamiltonian_t h, uint64_t *address, double *coeff, uint64_t cap,
                                  uint64_t *count);

/* Hamiltonian::GetIdentityTerm, cpp/hamiltonian.hpp:56 (0 until an ingest) */
int qpe_hamiltonian_identity_term(qpe_hamiltonian_t h, double *identity);

/* Hamiltonian::num_orbitals() / num_occupied(), cpp/hamiltonian.hpp:68-69 */
int qpe_hamiltonian_get_sizes(qpe_hamiltonian_t h, unsigned *num_orbitals, unsigned *num_occupied);

/* sizes of the five lists PrintIndexedSegmentedHamiltonianTerms returns,
 * cpp/hamiltonian.cc:253-299 */
int qpe_hamiltonian_term_counts(qpe_hamiltonian_t h, uint64_t counts[5]);

/* Copies list `cls` (map order): idx[4*i..4*i+3] HalfUp indices (-1 pads the
 * one-body lists), coeff[i] the accumulated coefficient.  cap = entries. */
int qpe_hamiltonian_get_terms(qpe_hamiltonian_t h, int cls, int32_t *idx, double *coeff, uint64_t cap);

/* ---------------- StateVector (device) ---------------- */

/* StateVector(precision, num_orbitals, num_occupied), cpp/state_vector.hpp:20-42:
 * a zero-filled [2^m][C(n,e)][C(n,e)] register, here resident in HBM. */
int qpe_state_create(qpe_state_t *out, unsigned precision, unsigned num_orbitals,
                     unsigned num_occupied, int dtype);
int qpe_state_destroy(qpe_state_t s);

/* No counterpart in the reference, whose StateVector is the alpha x beta product
 * space: a register in either layout.  QPE_LAYOUT_SINGLE_STRING holds
 * [2^precision][C(2n, 2e)] amplitudes; spin-orbital 2p is alpha orbital p and
 * 2p+1 beta orbital p; determinant x (a bit string of 2n bits with 2e set) is
 * stored at its colex rank; qpe_state_set/get use that index order.  Its amplitude
 * is sign * the alpha x beta amplitude of the same determinant, sign = (-1)^(number
 * of pairs alpha a, beta b with b < a) (the reordering of the creation operators).
 * The evolution works on the (e alpha, e beta) sector: each step converts block j-1
 * to the alpha x beta layout, applies U there and writes block j back (amplitudes
 * of other sectors in block 0 are rejected, QPE_ERR_ARG); QIFT, P(k), gates and
 * scaling act on the single-string register directly.  Up to 17 spatial orbitals;
 * time_evolve_iqpe and the batch need the alpha x beta layout. */
int qpe_state_create_layout(qpe_state_t *out, unsigned precision, unsigned num_orbitals, unsigned num_occupied,
                            int dtype, int layout);

/* number of amplitudes 2^m * C(n,e)^2 (single-string layout: 2^m * C(2n,2e)) */
int qpe_state_size(qpe_state_t s, uint64_t *n_amplitudes);

/* StateVector::operator[] writes/reads, batched: amplitudes [offset, offset+count)
 * in reference index order, interleaved (re, im) doubles.  Host memory is
 * borrowed for the duration of the call. */
int qpe_state_set(qpe_state_t s, uint64_t offset, uint64_t count, const double *amplitudes);
int qpe_state_get(qpe_state_t s, uint64_t offset, uint64_t count, double *amplitudes);

/* sv[j*C^2 + i] *= factor for all i: the implicit Hadamards of cpp/main.cc:111-115 */
int qpe_state_scale_block(qpe_state_t s, uint64_t block, double factor);

/* StateVector::apply_gate(gate, qubit), cpp/state_vector.hpp:51-74: a 2x2 gate
 * on ancilla qubit `qubit` < precision.  gate = {g00, g01, g10, g11}, each an
 * interleaved (re, im) pair (Gate1D, cpp/gate.hpp:7-20, row-major).  Qubit q is
 * bit (precision - q - 1) of the block index j (big-endian); for every pair of
 * blocks j0 (bit 0), j1 (bit 1) and every amplitude a of a block:
 * a_j0' = g00 a_j0 + g01 a_j1, a_j1' = g10 a_j0 + g11 a_j1 (FP64 arithmetic). */
int qpe_state_apply_gate(qpe_state_t s, const double *gate, unsigned qubit);

/* StateVector::apply_controlled_gate(gate, control, qubit),
 * cpp/state_vector.hpp:76-113: the same on the block pairs whose control qubit
 * is 1.  control != qubit, both < precision (QPE_ERR_ARG otherwise, where the
 * reference asserts). */
int qpe_state_apply_controlled_gate(qpe_state_t s, const double *gate, unsigned control, unsigned qubit);

/* StateVector Hamiltonian::IngestBroombridge(path, precision),
 * cpp/hamiltonian.cc:8-144: qpe_hamiltonian_ingest_broombridge, then a
 * [2^precision][C][C] state of this dtype whose block 0 holds the initial
 * determinants (a later duplicate address overwrites an earlier one).  On any
 * error *out is NULL and h is unchanged. */
int qpe_state_ingest_broombridge(qpe_state_t *out, qpe_hamiltonian_t h, const char *path, unsigned precision,
                                 int dtype);

/* ---------------- the hot path ---------------- */

/* time_evolve_over_terms(sv, h, m), cpp/time_evolution_over_terms.hpp:92-139:
 * block j = U * block(j-1) for j = 1..2^m-1, U one first-order Trotter step in
 * the reference's term order.  Limit: C(n, e) <= 16383 (the pass schedules hold
 * 14-bit string ranks; C(17,5) = 6188 is the largest workload here); larger
 * states can be created, set, transformed and read, but evolving them returns
 * QPE_ERR_ARG. */
int qpe_time_evolve_over_terms(qpe_state_t s, qpe_hamiltonian_t h);

/* time_evolve_over_terms restricted to blocks j_begin .. j_end-1 (1 <= j_begin
 * <= j_end <= 2^m): the reference's loop (TEOT:116-137) cut into resumable
 * pieces, for callers that report progress or checkpoint between them.  A call
 * with j_begin > 1 continues the chain of an earlier call with the same
 * Hamiltonian that ended at j_begin (QPE_ERR_STATE otherwise).  Calls covering
 * 1..2^m in order give exactly qpe_time_evolve_over_terms's result. */
int qpe_time_evolve_blocks(qpe_state_t s, qpe_hamiltonian_t h, uint64_t j_begin, uint64_t j_end);

/* time_evolve_iqpe(sv, h, num_of_applications), cpp/time_evolution_over_terms.hpp:
 * 141-155: applyUnitary (TEOT:174-372; term order pp, pq, pqqp, every pqqr term
 * whether or not a pq term gates it, pqrs) applied num_of_applications times in
 * place to block 1 (the reference calls it with exp = 1), whatever
 * QPE_OPT_TERM_ORDER says.  Block 1 is not renormalised (complex64 too); its
 * final squared norm is block_norms()[1].  Blocks other than 1 are untouched;
 * a later qpe_time_evolve_blocks must start at j_begin = 1. */
int qpe_time_evolve_iqpe(qpe_state_t s, qpe_hamiltonian_t h, uint64_t num_of_applications);

/* Batching form (no counterpart in the reference, whose main evolves one
 * Hamiltonian): time_evolve_over_terms for `count` independent instances in
 * one launch sequence.  states[i] is evolved under hams[i] exactly as
 * qpe_time_evolve_over_terms would (same results, bit for bit); all states
 * share the amplitude type and the ancilla count.  When every instance's plan
 * allows the dataflow evolution (QPE_OPT_EXECUTION), the whole chain of every
 * instance is one persistent launch: its workgroups claim the instances' tile
 * items step by step, interleaved, each starting as soon as the items it reads
 * from are done, so instances of unequal size proceed at their own pace.
 * Otherwise, or with QPE_BATCH_PASSES, launch k of a Trotter step runs pass k
 * of every instance.  Either sequence is captured into a hipGraph and launched
 * once (QPE_BATCH_EAGER: plain launches instead).
 * elapsed_ms (may be NULL): device time of the evolution.  Each state's
 * evolve timing is set to the batch time. */
#define QPE_BATCH_EAGER 1
#define QPE_BATCH_PASSES 2
int qpe_batch_time_evolve(const qpe_state_t *states, const qpe_hamiltonian_t *hams, int count, int flags,
                          double *elapsed_ms);

/* Autotuner (no counterpart in the reference): times one Trotter step of h on
 * a two-block scratch state for each candidate tile width (lines per
 * workgroup, which sets how many workgroups and hence CUs the instance spreads
 * over) and term grouping (QPE_OPT_ENGINE 0/1), and sets the best on s.  The
 * cost of a candidate is its step time, scaled up when its widest pass needs
 * more workgroups than share_cus (the CUs the instance may take in a batch).
 * result (may be NULL) receives {tile bytes, engine, microseconds per step}. */
int qpe_autotune(qpe_state_t s, qpe_hamiltonian_t h, int share_cus, double *result);

/* Autotuner of the batching form (no counterpart in the reference): times every candidate
 * of every instance as qpe_autotune does, then times one batched step of two policies on
 * two-block scratch states and sets the faster: (A) every instance at its fewest passes per
 * step (a launch of the batch runs pass k of every member, so this gives the fewest
 * launches); (B) the instance whose fastest step is slowest at that candidate and every
 * other instance at its fastest candidate with no more passes than that one.
 * results (may be NULL): [count][3] as qpe_autotune's result.  Call before
 * qpe_batch_time_evolve. */
int qpe_autotune_batch(const qpe_state_t *states, const qpe_hamiltonian_t *hams, int count, double *results);

/* StateVector::qift(), cpp/state_vector.hpp:115-179: in place, forward DFT
 * (e^{-2 pi i jk/M}) along the ancilla axis, natural order, divided by sqrt(M).
 * For 2^precision > 1024 the default kernel needs a second register's worth of free
 * device memory for the duration of the call (allocated and freed here); without it,
 * 2^precision <= 4096 uses the radix-2 kernel and larger ones return QPE_ERR_OOM. */
int qpe_qift(qpe_state_t s);

/* StateVector::IdentifyPhaseAmplitudes(), cpp/state_vector.hpp:341-372:
 * probs[k] = sum_d |y_k(d)|^2, k < 2^m (FP64, compensated sums). */
int qpe_identify_phase_amplitudes(qpe_state_t s, double *probs);

/* No counterpart in the reference.  For callers that cannot allocate memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): qpe_qift followed by qpe_identify_phase_amplitudes, with the same
 * kernels and results, but no allocation and no host transfer.  probs_device is
 * caller-owned DEVICE memory of 2^precision doubles (QPE_ERR_ARG otherwise); it
 * receives P(k).  The per-workgroup P(k) partial sums live in a static device
 * array of 2^20 doubles (8 MiB per device, held until the process ends); states
 * of 2^precision > 4096 outcomes are rejected (QPE_ERR_ARG: the QIFT kernels above
 * that size need a second register).  Calls that overlap
 * on one device share that array and race: serialise them.  Afterwards the state
 * holds the QIFT output and its renormalisation factors are 1; a later
 * qpe_identify_phase_amplitudes computes P(k) from the register itself. */
int qpe_qift_noalloc(qpe_state_t s, double *probs_device);

/* Phase post-processing of cpp/main.cc:149-167 (host only): phase =
 * (sum_k k*probs[k]) / 2^precision, summed in k order;
 * out[0] = -2*pi*phase + identity, the reference's energy estimate (it does not
 *          divide by the timestep);
 * out[1] = -2*pi*phase / timestep + identity, the same phase read as
 *          e^{-i E t}, i.e. divided by the timestep;
 * out[2] = 2*pi / (timestep * 2^precision), the reference's margin of error. */
int qpe_energy_estimate(const double *probs, unsigned precision, double identity, double timestep, double out[3]);

/* ---------------- options, diagnostics ---------------- */

enum qpe_option {
    QPE_OPT_RENORMALIZE = 0,   /* 1: renormalise every block to block 0's norm (default: 1 for
                                  complex64, 0 for complex128) */
    QPE_OPT_ENGINE = 1,        /* 0: tile passes with fused groups of terms (default);
                                  1: tile passes with one term per op */
    QPE_OPT_MAX_TILE_BYTES = 2, /* LDS bytes a tile may use (0: device maximum) */
    QPE_OPT_TROTTER_ORDER = 3,  /* 1: the reference's first-order product (default);
                                   2: symmetric second order, the first-order sequence at
                                   t/2 followed by the same sequence reversed at t/2 (no
                                   counterpart in the reference) */
    QPE_OPT_TERM_ORDER = 4,     /* term order of one step of time_evolve_over_terms:
                                   0: applyUnitary_reordered (TEOT:374-562; default, the
                                   call TEOT:136 makes); 1: applyUnitary (TEOT:174-372, the
                                   call left commented out at TEOT:135) */
    QPE_OPT_QIFT_FORM = 5,      /* kernel of qpe_qift / qpe_qift_noalloc: 0 chosen by size
                                   (default); 1 in-register DFTs (one launch for
                                   2^precision <= 1024, two up to 2^20); 2 DFT as a matrix
                                   product on the matrix cores (2^precision 16..64);
                                   3 radix-2 stages in LDS (2^precision <= 4096).
                                   All give the same QIFT within rounding. */
    QPE_OPT_EXECUTION = 6,      /* how qpe_time_evolve_over_terms / _blocks run the passes:
                                   1: one launch per pass; 2: one dataflow launch for all
                                   the steps of the call, its items starting as soon as the
                                   items they read from are done (QPE_ERR_ARG unless every
                                   pass is tiled in LDS and blocks are below 2^31 bytes);
                                   0: 2 for complex128 when the plan allows it, else 1
                                   (default).  Results are identical. */
    QPE_OPT_WAIT_LIMIT_MS = 7   /* limit of one wait of a dataflow item for the items it
                                   depends on (default 10000 ms; 0: no waiting, so a run
                                   stops at its first wait); QPE_ERR_STALLED past it */
};
int qpe_state_set_option(qpe_state_t s, int option, int64_t value);

/* timings of the last calls in milliseconds (device events):
 * out[0] evolve, out[1] qift, out[2] P(k); out[3] passes per U; out[4] launches per U */
int qpe_state_timings(qpe_state_t s, double out[5]);

/* per-block squared norms measured during the last evolve (FP64, compensated);
 * norms[j] for j < 2^m, before renormalisation. */
int qpe_state_block_norms(qpe_state_t s, double *norms);

/* Determinant pairs of one term as the device computes them (rank/unrank and
 * Jordan-Wigner parity on the GPU): kind 2 = CA(p, r), kind 4 = CCAA(p,q,r,s),
 * kind 3 = CNA(p,q,r) in one spin string of n orbitals with e electrons,
 * argument order as discoverCA/CNA/CCAA_BVs (cpp/time_evolution_over_terms.hpp:
 * 1160-1324).  Writes idx01[i], idx10[i], parity[i] for every pair; *count
 * receives the number of pairs. */
int qpe_debug_pairs(int kind, int n, int e, const int32_t *args, uint32_t *idx01,
                    uint32_t *idx10, uint32_t *parity, uint64_t cap, uint64_t *count);

/* Host-only description of the tile program one Trotter step compiles to, for
 * a state of this dtype and an LDS tile budget (0: 160 KiB - 4 KiB):
 * out[0] passes per step, out[1] kernel launches per step, out[2] determinant
 * pair rotations per step of the reference's terms, out[3] diagonal updates per
 * step, out[4] HBM bytes the passes load+store per step, out[5] C(n,e),
 * out[6] fused excitation groups, out[7] double-excitation terms. */
int qpe_plan_info(qpe_hamiltonian_t h, int dtype, uint64_t tile_bytes, double out[8]);

/* message of the last failed call on this thread ("" if none) */
const char *qpe_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* QPESIM_H */
