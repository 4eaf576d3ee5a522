// Device engine: owns the device copy of a Plan and launches its passes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "plan.hpp"

namespace qpe {

// throws std::runtime_error carrying the HIP error text
void hip_check(hipError_t e, const char *what);

struct DevicePlan {
    int dtype = 0;           // 0 complex128, 1 complex64
    Plan plan;
    uint32_t *binom = nullptr;
    uint32_t *strings = nullptr;
    double *theta_a = nullptr, *theta_b = nullptr, *cross = nullptr;
    GroupDev *groups = nullptr;
    double *grp_a = nullptr, *grp_b = nullptr;
    OpDev *ops = nullptr;
    ItemDev *items = nullptr;
    uint32_t *col_list = nullptr;
    uint32_t *sched = nullptr;
    uint32_t *pass_qr = nullptr;   // mixed windows' pair lists (Plan::pass_qr)
    double *coef = nullptr;      // [op][kMaxGroup][2]: (1 - cos, sin) of each op's terms, in the
                                 // state's type (float pairs for complex64)
    double *partials = nullptr;   // [max_items] norm partial sums
    size_t lds_bytes = 0;         // largest dynamic LDS of a pass launch
    std::vector<size_t> pass_lds; // dynamic LDS of each pass launch
    std::vector<uint8_t> pass_ops; // op-type set of each pass's kernel (kOpsAll / kOpsFab / kOpsFcc)
    // dataflow order of a step (build_flow), for plans whose passes are all LDS-tiled
    // and whose blocks are below 2^31 bytes
    bool flow_ok = false;
    uint32_t flow_items = 0;       // items per step
    uint32_t flow_norm_items = 0;  // items of the pass that sums the block norm
    uint32_t *flow_item_pass = nullptr, *flow_pass_first = nullptr, *flow_dep_off = nullptr, *flow_deps = nullptr;
    PassDev *flow_passes = nullptr;
    uint32_t *flow_done = nullptr; // [flow_items] completion stamps
    uint32_t *flow_ctr = nullptr;  // [4] claim, norm count, finished step, abort

    void upload(const Plan &p, int dtype, hipStream_t s);
    void release();
    ~DevicePlan() { release(); }
};

// Batched launches: one launch runs pass k of every member that has one.
struct BatchWork {
    int32_t inst;          // member index
    uint32_t block_begin;  // first workgroup of this member's items in the launch
    int32_t pad0, pad1;
    PassDev pass;
};
struct FinalizeArgs {
    const double *partials;
    int count;
    int pad;
    double *norms, *scale;
    int *status;
};

// One member of a batched evolution (all members share dtype and M).
struct BatchMember {
    const DevicePlan *dp;
    void *state;
    double *norms, *scale;
    int renorm;
    int *status;
};

// Device tables and the captured graph of a batched evolution.
struct BatchRun {
    int dtype = 0;
    uint64_t M = 0;
    size_t lds_bytes = 0;
    void *args = nullptr;          // PassArgs[count]
    BatchWork *work = nullptr;     // all launches' work lists, concatenated
    FinalizeArgs *fin = nullptr;   // [count]
    uint64_t *jstep = nullptr;     // device step counter
    unsigned *done = nullptr;      // finalize completion counter
    int count = 0;
    struct Launch {
        size_t work_off;
        int nwork;
        uint32_t blocks;
        size_t lds;                // dynamic LDS: the largest of its members' pass k
    };
    std::vector<Launch> launches;  // one step
    // The chain of M-1 steps as graph launches: `graph` holds graph_steps steps and runs
    // graph_reps times, `tail` the remaining tail_steps; a graph holds at most
    // kMaxGraphLaunches kernel launches
    static constexpr uint64_t kMaxGraphLaunches = 1u << 15;
    hipGraphExec_t graph = nullptr, tail = nullptr;
    uint64_t graph_steps = 0, graph_reps = 0, tail_steps = 0;
    // Dataflow form (every member's plan allows it): the whole chain as one persistent
    // launch whose claims interleave the members' items step by step (flow_order: one
    // step's (member, item) claims); the per-member stamp resets and the launch are
    // captured into flow_graph.
    bool flow = false;
    void *flow_args = nullptr;       // FlowArgs[count]
    void *flow_order = nullptr;      // uint2[flow_per_step]
    uint32_t flow_per_step = 0;
    uint32_t *flow_gctr = nullptr;   // [4] launch counters
    std::vector<const DevicePlan *> flow_plans;
    hipGraphExec_t flow_graph = nullptr;

    // use_flow: the dataflow form when every member's plan allows it
    void build(const std::vector<BatchMember> &members, uint64_t M, hipStream_t s, bool use_flow = true);
    // captures the chain's graphs (once)
    void instantiate();
    // runs steps 1..M-1; with use_graph as graph launches
    void run(hipStream_t s, bool use_graph);
    void release();
    ~BatchRun() { release(); }
};

// Blocks j_begin..j_end-1 of `state` (device, column-major blocks of C*C
// amplitudes): block j = U block j-1.  norms[M] and scale[M] are device arrays;
// norms and scale of blocks below j_begin must be set.  renorm selects whether
// each block is rescaled to block 0's norm when it is read.
// execution: EXEC_AUTO runs the dataflow kernel when the plan allows it,
// EXEC_PASSES one launch per pass, EXEC_DATAFLOW the dataflow kernel (throws
// when the plan does not allow it).  flow_wait_ms bounds each wait of a
// dataflow item; a launch whose wait runs out stops with *status_flag = 2.
enum Execution { EXEC_AUTO = 0, EXEC_PASSES = 1, EXEC_DATAFLOW = 2 };
void evolve(const DevicePlan &dp, void *state, uint64_t j_begin, uint64_t j_end, double *norms, double *scale,
            int renorm, int *status_flag, hipStream_t s, int *launches_per_step, int execution = EXEC_AUTO,
            uint32_t flow_wait_ms = 10000);

// time_evolve_iqpe (cpp/time_evolution_over_terms.hpp:141-155): the plan's
// step applied `count` times in place to block `block`, without
// renormalisation; norms[block] receives the final squared norm.
void evolve_in_place(const DevicePlan &dp, void *state, uint64_t block, uint64_t count, double *norms, double *scale,
                     int *status_flag, hipStream_t s, int *launches);

// squared norm of block j into norms[j] (and scale[j] = 1)
void block_norm(int dtype, uint64_t D, void *state, uint64_t j, double *norms, double *scale,
                double *partials, int nparts, hipStream_t s);

// QIFT kernels: 0 picks per size; the others exist for 2^m in 2..2^20 (register DFTs),
// 16..64 (matrix cores) and 2..4096 (radix-2 stages in LDS)
enum QiftForm { QIFT_AUTO = 0, QIFT_REGISTER = 1, QIFT_MFMA = 2, QIFT_LDS_RADIX2 = 3 };
bool qift_form_available(int form, int m);
// whether qift() of this form and size needs `work` (a second register, 2^m * D amplitudes)
bool qift_needs_work(int form, int m);

// QIFT in place with the renormalisation factors applied, and P(k) partials.
// max_grid caps the rows of 2^m partial sums written to pk_partial (*grid_out: rows written).
void qift(int dtype, void *state, int m, uint64_t D, const double *scale, double *pk_partial,
          int *grid_out, double *twiddles, size_t max_lds, hipStream_t s, int max_grid = 1024,
          int form = QIFT_AUTO, void *work = nullptr);

// The same QIFT and P(k) with no allocation and no host transfer: the partial sums
// in a static device array of kNoAllocScratchDoubles doubles (one per device), P(k)
// written to probs_dev (device memory, 2^m doubles), scale[] reset to 1 by a fill
// kernel.  Calls on one device share the array: overlapping calls race.
constexpr uint64_t kNoAllocScratchDoubles = 1ull << 20;   // 8 MiB
constexpr int kMaxScratchDevices = 64;
void qift_noalloc(int dtype, void *state, int m, uint64_t D, double *scale, double *twiddles, size_t max_lds,
                  double *probs_dev, hipStream_t s, int form = QIFT_AUTO);
void reduce_pk(const double *pk_partial, int grid, int m, double *probs_dev, hipStream_t s);

// reference-order host data <-> device layout, with dtype conversion
void scatter_in(int dtype, void *state, const double *staging, uint64_t offset, uint64_t count,
                uint32_t C, hipStream_t s);
void gather_out(int dtype, const void *state, double *staging, uint64_t offset, uint64_t count,
                uint32_t C, hipStream_t s);
void scale_block(int dtype, void *state, uint64_t j, uint64_t D, double factor, hipStream_t s);

// One step cur = U prev with explicit block pointers (block index j: norms[j], scale[j]).
void evolve_step(const DevicePlan &dp, void *prev, void *cur, uint64_t j, double *norms, double *scale, int renorm,
                 int *status_flag, hipStream_t s);

// Single-string (interleaved Jordan-Wigner) register: host staging copies in its own
// index order; conversion of one block to / from the alpha x beta layout of the plan
// (C x C, column-major) with the blocked-to-interleaved sign; check for amplitudes
// outside the (e alpha, e beta) sector (sets *flag).
void single_string_copy(int dtype, void *state, double *staging, uint64_t offset, uint64_t count, bool to_state,
                        hipStream_t s);
void single_string_convert(const DevicePlan &dp, void *ab, void *k1, bool to_ab, hipStream_t s);
void single_string_outside(const DevicePlan &dp, const void *k1, uint64_t D1, int *flag, hipStream_t s);

// 2x2 gate g (row-major, interleaved re/im) on bit tbit of the block index of a
// [2^m][D] register; cbit >= 0: only on blocks whose bit cbit is 1
void apply_gate(int dtype, void *state, uint64_t D, int m, int tbit, int cbit, const double *g, hipStream_t s);

// device-computed determinant pairs of one term (for bit-exact index tests)
void debug_pairs(int kind, int n, int e, const int32_t *args, uint32_t *idx01, uint32_t *idx10,
                 uint32_t *parity, uint64_t count, hipStream_t s);

int device_max_lds();

}  // namespace qpe
