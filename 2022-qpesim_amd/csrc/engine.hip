// gfx950 kernels of the QPE hot path and their launchers.
//
//   pass_kernel   one tile pass of a Trotter step: <R, 1> COL or ROW tiles in
//                 LDS, <R, 2> mixed windows rotated in global memory (COLG);
//                 batch_pass_kernel runs either kind for a batch of instances
//   norm_finalize FP64 compensated block norm + renormalisation factor
//   qift_kernel   batched radix-2 DFT along the ancilla axis in LDS, with the
//                 1/sqrt(M) normalisation and the P(k) partial sums fused
//
// Amplitudes: complex128 (double2) or complex64 (float2).  Angles, their
// cos/sin, norms and P(k) are always FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "combinadic.hpp"
#include "engine.hpp"

namespace qpe {

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

constexpr int kThreads = kThreadsPlan;   // threads per pass workgroup
constexpr int kWaves = kThreads / 64;
// the wave that stages coefficients (wave 0 stages descriptors)
constexpr int kCoefWave = kWaves >= 8 ? 4 : kWaves / 2;
constexpr int kMaxLines = 64;
constexpr int kBinomWords = kBinomRows * kBinomCols;

template <typename R> struct Vec;
template <> struct Vec<double> { using type = double2; };
template <> struct Vec<float> { using type = float2; };

// Neumaier compensated accumulation
struct KSum {
    double s = 0.0, c = 0.0;
    __device__ void add(double x) {
        double t = s + x;
        if (fabs(s) >= fabs(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    __device__ double value() const { return s + c; }
};

// Amplitude update x + d.  complex128 rounds to nearest.  complex64 rounds
// stochastically: d is dithered by a uniform fraction of one ulp of x, in steps
// of ulp/256, so the stored value is unbiased.  df is that fraction times 2^-23
// (an odd multiple of 2^-32 in [-127.5, 127.5] * 2^-31): x with its mantissa
// cleared is 2^e = 2^23 ulp(x), so df * 2^e is the dither (0 for zero and
// subnormal x).  Round to nearest would drop every update below half an ulp:
// the large amplitudes of a state would never lose the (1 - cos) part of the
// small rotations that feed the small ones, a norm and population bias that
// grows step after step.
template <typename R> __device__ inline R upd(R x, R d, float df);
template <> __device__ inline double upd<double>(double x, double d, float) { return x + d; }
template <> __device__ inline float upd<float>(float x, float d, float df) {
    return x + fmaf(df, __int_as_float(__float_as_int(x) & 0x7f800000), d);
}

__device__ inline int rbyte(uint32_t r, int k) { return (int)__builtin_amdgcn_sbfe(r, 8 * k, 8); }
// dither factor from byte k of r: (signed byte + 1/2) * 2^-31, the scaling upd expects
__device__ inline float dfac(uint32_t r, int k) { return fmaf((float)rbyte(r, k), 0x1p-31f, 0x1p-32f); }

// 32-bit mixer for the dither bits (counter based: deterministic for a given plan)
__device__ inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Dither factors of one op for one thread: four signed random bytes + 1/2 of
// a hashed (step, op, tile item, thread) word, scaled by 2^-31 for upd.  Schedule entry i takes them in
// a signed order of its own (entry_dither) and the k-th line of a chunk in the
// k-th cyclic order of that (turn).  The factor set is symmetric, so every
// ordering is uniform and every update stays unbiased (the dither is
// independent of the data); only rounding errors within one op of one thread
// correlate.
struct Dither {
    float a, b, c, d;
};
__device__ inline Dither make_dither(uint32_t r) {
    return {dfac(r, 0), dfac(r, 1), dfac(r, 2), dfac(r, 3)};
}
__device__ inline Dither entry_dither(const Dither &z, int i) {
    return (i & 3) == 0 ? z : (i & 3) == 1 ? Dither{-z.c, z.d, -z.a, z.b}
                            : (i & 3) == 2 ? Dither{z.b, -z.a, z.d, -z.c} : Dither{-z.d, -z.c, z.b, z.a};
}
// v with its sign flipped when bit 31 of m is set (m: a parity moved to bit 31)
__device__ inline float flip_sign(float v, uint32_t m) { return __int_as_float(__float_as_int(v) ^ (int)(m & 0x80000000u)); }
__device__ inline double flip_sign(double v, uint32_t m) {
    return __hiloint2double(__double2hiint(v) ^ (int)(m & 0x80000000u), __double2loint(v));
}
__device__ inline Dither turn(const Dither &z, int k) {
    return (k & 3) == 0 ? z : (k & 3) == 1 ? Dither{z.b, z.c, z.d, z.a}
                            : (k & 3) == 2 ? Dither{z.c, z.d, z.a, z.b} : Dither{z.d, z.a, z.b, z.c};
}

// 2x2 symmetric rotation {c, i*mu; i*mu, c} on the pair (u, v), given as
// h = 1 - c.  The whole update -h*u + i*mu*v is formed first and added to u in
// one rounding (upd).  z: the complex64 dither factors.
template <typename V, typename R>
__device__ inline void rotate(V &u, V &v, R h, R mu, const Dither &z) {
    V nu, nv;
    nu.x = upd<R>(u.x, -h * u.x - mu * v.y, z.a);
    nu.y = upd<R>(u.y, -h * u.y + mu * v.x, z.b);
    nv.x = upd<R>(v.x, -h * v.x - mu * u.y, z.c);
    nv.y = upd<R>(v.y, -h * v.y + mu * u.x, z.d);
    u = nu;
    v = nv;
}

struct PassArgs {
    const uint32_t *binom;
    const uint32_t *strings;
    const double *theta_a, *theta_b, *cross;
    const GroupDev *groups;
    const double *grp_a, *grp_b;
    const OpDev *ops;
    const ItemDev *items;
    const uint32_t *col_list;
    const uint32_t *sched;
    const uint32_t *pass_qr;   // q | r << 8 of the mixed windows' pair lists
    const double *coef;   // [op][kMaxGroup] (1 - cos, sin), op-major, in the state's type (float pairs for complex64)
    double *partials;
    const double *scale;
    void *state;          // batched launches: the instance's register (column-major blocks)
    uint64_t D;           // amplitudes per block
    int n, e;
    uint32_t C, ca_pairs, ccaa_pairs;
    int renorm;
};

// Barrier for LDS traffic only: outstanding global loads (the prefetched pair
// entries of later ops) stay in flight across it.
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Barrier between two ops: LDS only for tiles in LDS; a full workgroup barrier
// (global memory ordered too) when the ops work on global memory.
__device__ inline void op_barrier(bool global_tile) {
    if (global_tile) __syncthreads();
    else lds_barrier();
}


// Op pipeline.  Ops run in batches of kBatch.  Descriptors and term
// coefficients of batch b+1 are staged in LDS rings while batch b runs, so no
// global or scalar load sits between two ops (scalar and LDS traffic share one
// wait counter).  Each op's schedule entries (kSchedMax loads per thread, a
// static count so the compiler's wait counts stay exact) are loaded while the
// previous op runs, into two alternating register sets.
constexpr int kBatch = 8;
constexpr int kOpsAll = 0, kOpsFab = 1, kOpsFcc = 2;   // op-type sets of a pass kernel instantiation
constexpr int kDescWords = sizeof(OpDev) / 4;   // 8
// lines (column pairs) whose LDS loads are issued together: 2 rather than 4 keeps the kernel at
// 112 VGPRs and 81 KB of code (measured 11% faster per Trotter step; 1 gains nothing more)
constexpr int kChunk = 2;
static_assert(kThreads == kThreadsPlan, "schedules are laid out for the pass kernel's thread count");

// Schedule entries of one op for this thread.  The op's ns slots each get
// L = kThreads / ns threads: thread t serves slot t % ns as lane t / ns and
// takes lines (column pairs) lane, lane + L, ...
struct EntSet {
    uint32_t e[kSchedMax];
    int slot, lane, L;
};

__device__ inline uint32_t s_rank1(uint32_t e) { return e & kSchedPad; }
__device__ inline uint32_t s_rank2(uint32_t e) { return (e >> kSchedRankBits) & kSchedPad; }
__device__ inline uint32_t s_slot(uint32_t e) { return (e >> 28) & 7u; }
__device__ inline uint32_t s_par(uint32_t e) { return e >> 31; }

// Staging of batch b: descriptor words (threads 0..63) and coefficients
// (threads 256..319, from the op-major table coef[op][kMaxGroup], already in
// the state's type) are fetched into registers with independent global loads,
// and committed to the LDS rings later, so their latency overlaps the ops of
// the running batch (no conversion between the load and the commit, which
// would make the fetch wait for its load).
template <typename R>
struct Staged {
    uint32_t w;
    typename Vec<R>::type c;
};
template <typename R>
__device__ inline Staged<R> fetch_batch(const PassArgs &A, int o0, int o1, int b) {
    using V = typename Vec<R>::type;
    const int first = o0 + b * kBatch;
    // wave 0 fetches descriptor words, wave kCoefWave coefficients: a wave-uniform branch, so
    // the two never share a (linearised) path with a pending load
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int lane = (int)threadIdx.x & 63;
    Staged<R> st;   // each wave sets only the field it commits
    if (wave == 0) {
        const int i = lane / kDescWords, word = lane - i * kDescWords;
        uint32_t v = 0u;
        if (first + i < o1) v = reinterpret_cast<const uint32_t *>(A.ops + first + i)[word];
        st.w = v;
    } else if (wave == kCoefWave) {
        const int i = lane / kMaxGroup, slot = lane - i * kMaxGroup;
        V c;
        c.x = 0;
        c.y = 0;
        if (first + i < o1) c = reinterpret_cast<const V *>(A.coef)[(size_t)(first + i) * kMaxGroup + slot];
        st.c = c;
    }
    return st;
}
template <typename R>
__device__ inline void commit_batch(const Staged<R> &st, OpDev *ring, double2 *cring, int b) {
    using V = typename Vec<R>::type;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int lane = (int)threadIdx.x & 63;
    if (wave == 0) reinterpret_cast<uint32_t *>(ring + (b & 1) * kBatch)[lane] = st.w;
    else if (wave == kCoefWave) reinterpret_cast<V *>(cring)[(b & 1) * kBatch * kMaxGroup + lane] = st.c;
}
static_assert(kBatch * kDescWords == 64 && kBatch * kMaxGroup == 64, "one wave stages each ring");

__device__ inline void load_entries(const PassArgs &A, const OpDev &d, bool valid, EntSet &E) {
    uint32_t off = 0, st = 0;
    int count = 0;
    E.slot = 0; E.lane = 1; E.L = 1;   // no work unless the op has a schedule
    if (valid && d.type != OP_DIAG) {
        // lane = tid / ns and L = kThreads / ns from the descriptor's magic (exact for tid, ns <= 1024)
        const int ns = d.ns;
        E.lane = (int)(__umul24((uint32_t)threadIdx.x, d.split & 0x1fffffu) >> 20);
        E.slot = (int)threadIdx.x - (int)__umul24((uint32_t)E.lane, (uint32_t)ns);
        E.L = (int)(d.split >> 21);
        if (E.lane < E.L) {
            off = d.tab + (uint32_t)E.slot;
            st = (uint32_t)ns;
            count = d.lg;
        }
    }
    // kSchedMax loads for every thread (rows past lg repeat row 0): a static count keeps
    // the wait counts exact; 32-bit element offsets from the table base
    uint32_t row = 0;
#pragma unroll
    for (int i = 0; i < kSchedMax; ++i) {
        E.e[i] = A.sched[off + (i < count ? row : 0u)];
        row += st;
    }
}

// Column pairs of (q, r) in the tile: lines l01 (r occupied, q empty) and its
// partner l10, as byte offsets line_base[l] * line_bytes from the tile base
// (line_base: the line's position in LDS, or its column rank when the tile is
// the block in global memory); bit 0 of .x holds the Jordan-Wigner parity of
// the pair (offsets are multiples of 8).  Run by one whole wave.
__device__ inline void build_pair_list(uint32_t qr, const uint32_t *line_str, const uint32_t *line_base,
                                       int2 *pair_list, int *pair_count, int nl, uint32_t line_bytes) {
    const int tid = threadIdx.x & 63;
    const int q = (int)(qr & 0xff), r = (int)((qr >> 8) & 0xff);
    {
        // lane l holds line l's string (no string has all 32 bits set); the partner
        // search reads the other lanes' strings with readlane, not LDS round trips
        const uint32_t s = tid < nl ? line_str[tid] : ~0u;
        const bool cand = tid < nl && ((s >> r) & 1u) && !((s >> q) & 1u);
        const uint32_t ps = s ^ (1u << q) ^ (1u << r);
        int partner = -1;
        for (int l = 0; l < nl; ++l)
            if ((uint32_t)__builtin_amdgcn_readlane((int)s, l) == ps) partner = l;
        const bool is01 = cand && partner >= 0;
        const int parb = popcnt(s & between_mask(q, r)) & 1;
        uint64_t mask = __ballot(is01);
        if (is01)
            pair_list[__popcll(mask & ((1ull << tid) - 1ull))] =
                make_int2((int)(line_base[tid] * line_bytes) | parb, (int)(line_base[partner] * line_bytes));
        if (tid == 0) *pair_count = __popcll(mask);
    }
}

// rotate (line[r1], line[r2]) on lines lane, lane + L, ... of the tile with one (h, mu)
template <typename R, typename V>
__device__ inline void rotate_lines(V *tile, uint32_t line_bytes, int nl, int lane, int L, uint32_t base0, uint32_t r1,
                                    uint32_t r2,
                                    R h, R mu, const Dither &z) {
    char *tb = reinterpret_cast<char *>(tile);
    const uint32_t o1 = r1 * (uint32_t)sizeof(V), o2 = r2 * (uint32_t)sizeof(V);
    const uint32_t step = __umul24((uint32_t)L, line_bytes);
    uint32_t base = base0;   // lane * line_bytes
    for (int l0 = lane; l0 < nl; l0 += kChunk * L, base += kChunk * step) {
        V u[kChunk], v[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; ++k)
            if (l0 + k * L < nl) {
                u[k] = *reinterpret_cast<const V *>(tb + base + k * step + o1);
                v[k] = *reinterpret_cast<const V *>(tb + base + k * step + o2);
            }
#pragma unroll
        for (int k = 0; k < kChunk; ++k)
            if (l0 + k * L < nl) {
                rotate(u[k], v[k], h, mu, turn(z, k));
                *reinterpret_cast<V *>(tb + base + k * step + o1) = u[k];
                *reinterpret_cast<V *>(tb + base + k * step + o2) = v[k];
            }
    }
}

// rotate_lines on the lines whose string has a bit of gate (one pqqr term of
// applyUnitary's order whose inner orbital lies in the other string)
template <typename R, typename V>
__device__ inline void rotate_lines_gated(V *tile, uint32_t line_bytes, int nl, int lane, int L, uint32_t base0,
                                          uint32_t r1,
                                          uint32_t r2, R h, R mu, const uint32_t *line_str, uint32_t gate,
                                          const Dither &z) {
    char *tb = reinterpret_cast<char *>(tile);
    const uint32_t o1 = r1 * (uint32_t)sizeof(V), o2 = r2 * (uint32_t)sizeof(V);
    const uint32_t step = __umul24((uint32_t)L, line_bytes);
    uint32_t base = base0;   // lane * line_bytes
    for (int l = lane, k = 0; l < nl; l += L, ++k, base += step) {
        if (gate && !(line_str[l] & gate)) continue;
        V *pu = reinterpret_cast<V *>(tb + base + o1);
        V *pv = reinterpret_cast<V *>(tb + base + o2);
        V u = *pu, v = *pv;
        rotate(u, v, h, mu, turn(z, k));
        *pu = u;
        *pv = v;
    }
}

// rotate (l01[rs], l10[rp]) on column pairs lane, lane + L, ... of the pair list
template <typename R, typename V>
__device__ inline void rotate_pairs(V *tile, const int2 *pair_list, int ncp, int lane, int L, uint32_t rs,
                                    uint32_t rp, uint32_t par, R h, R sn, const Dither &z) {
    char *tb = reinterpret_cast<char *>(tile);
    const uint32_t os = rs * (uint32_t)sizeof(V), op = rp * (uint32_t)sizeof(V);
    const R mneg = -sn;   // mu = sn when the entry's and the pair's parities differ, else -sn
    const uint32_t pm = par << 31;
    for (int x0 = lane; x0 < ncp; x0 += kChunk * L) {
        int2 pl[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; ++k) pl[k] = pair_list[min(x0 + k * L, ncp - 1)];
        V u[kChunk], v[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; ++k)
            if (x0 + k * L < ncp) {
                u[k] = *reinterpret_cast<const V *>(tb + ((uint32_t)pl[k].x & ~7u) + os);
                v[k] = *reinterpret_cast<const V *>(tb + (uint32_t)pl[k].y + op);
            }
#pragma unroll
        for (int k = 0; k < kChunk; ++k)
            if (x0 + k * L < ncp) {
                const R mu = flip_sign(mneg, pm ^ ((uint32_t)pl[k].x << 31));
                rotate(u[k], v[k], h, mu, turn(z, k));
                *reinterpret_cast<V *>(tb + ((uint32_t)pl[k].x & ~7u) + os) = u[k];
                *reinterpret_cast<V *>(tb + (uint32_t)pl[k].y + op) = v[k];
            }
    }
}

// rotate_pairs for a thread whose column pairs all fit one chunk, with the
// pair-list entries already in registers (pl[k], valid for k < npl): the
// pair list is read once per op instead of once per schedule entry
template <typename R, typename V>
__device__ inline void rotate_pairs_held(V *tile, const uint32_t *px, const uint32_t *py, const uint32_t *pb,
                                         int npl, uint32_t rs, uint32_t rp, uint32_t par, R h, R sn,
                                         const Dither &z) {
    char *tb = reinterpret_cast<char *>(tile);
    const uint32_t os = rs * (uint32_t)sizeof(V), op = rp * (uint32_t)sizeof(V);
    const R mneg = -sn;
    const uint32_t pm = par << 31;
    V u[kChunk], v[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k)
        if (k < npl) {
            u[k] = *reinterpret_cast<const V *>(tb + px[k] + os);
            v[k] = *reinterpret_cast<const V *>(tb + py[k] + op);
        }
#pragma unroll
    for (int k = 0; k < kChunk; ++k)
        if (k < npl) {
            rotate(u[k], v[k], h, flip_sign(mneg, pm ^ pb[k]), turn(z, k));
            *reinterpret_cast<V *>(tb + px[k] + os) = u[k];
            *reinterpret_cast<V *>(tb + py[k] + op) = v[k];
        }
}

// excitation group: the angle of pair (entry, line) is the pair part (ha, sa)
// plus the line's other-string part (gang[l]), combined by angle addition in
// the (1 - cos, sin) form
template <typename R, typename V>
__device__ inline void rotate_group(V *tile, uint32_t line_bytes, int nl, int lane, int L, uint32_t base0, uint32_t r1,
                                    uint32_t r2,
                                    uint32_t par, double ha, double sa, const double2 *ga, const Dither &z) {
    char *tb = reinterpret_cast<char *>(tile);
    const uint32_t o1 = r1 * (uint32_t)sizeof(V), o2 = r2 * (uint32_t)sizeof(V);
    const uint32_t step = __umul24((uint32_t)L, line_bytes);
    uint32_t base = base0;   // lane * line_bytes
    for (int l0 = lane; l0 < nl; l0 += kChunk * L, base += kChunk * step) {
        V u[kChunk], v[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; ++k)
            if (l0 + k * L < nl) {
                u[k] = *reinterpret_cast<const V *>(tb + base + k * step + o1);
                v[k] = *reinterpret_cast<const V *>(tb + base + k * step + o2);
            }
#pragma unroll
        for (int k = 0; k < kChunk; ++k)
            if (l0 + k * L < nl) {
                const double2 b = ga[l0 + k * L];
                const double hh = (ha + b.x) - ha * b.x + sa * b.y;
                const double ss = (sa + b.y) - sa * b.x - ha * b.y;
                rotate(u[k], v[k], (R)hh, (R)(par ? ss : -ss), turn(z, k));
                *reinterpret_cast<V *>(tb + base + k * step + o1) = u[k];
                *reinterpret_cast<V *>(tb + base + k * step + o2) = v[k];
            }
    }
}

// OPS: the op types a kernel carries: kOpsAll, kOpsFab (passes of mixed windows, whose
// ops are all FAB) or kOpsFcc (same-spin passes, all FCC).  A kernel with one op path has
// 40% less code and 20 fewer VGPRs, which measured 3.4% off a complex64 step at 32
// spin-orbitals / 8 electrons for the FAB passes alone.
template <typename R, int OPS = kOpsAll>
__device__ inline void run_op(const PassArgs &A, const OpDev &d, const typename Vec<R>::type *coef, const EntSet &E,
                              typename Vec<R>::type *tile, const double2 *gang, const uint32_t *line_rank,
                              const uint32_t *line_str, const uint32_t *line_base, const int2 *pair_lists,
                              const int *pair_counts, int nl, uint32_t C, uint32_t &cur_list, int &ncp,
                              uint32_t tseed) {
    using V = typename Vec<R>::type;
    const int tid = threadIdx.x;
    const int type = d.type;
    const uint32_t line_bytes = C * (uint32_t)sizeof(V);
    // dither: tseed is the thread's random word for this op, a hash of the thread index
    // and a (scalar) hash of (step, op, tile item).  Xoring a fixed per-thread hash instead
    // would tie the threads' dithers together the same way in every op: measured max|dP|
    // 3.6e-6 against 7.9e-7 at 32 spin-orbitals / 8 electrons / 10 ancillas
    const Dither z = make_dither(tseed);
    const uint32_t lbase = __umul24((uint32_t)E.lane, line_bytes);   // the thread's first line (line-local ops)
    if constexpr (OPS == kOpsFcc) {
        if (E.lane >= E.L) return;
        if (nl <= E.L) {
            // at most one line per thread (a tile has fewer lines than a slot has threads):
            // no line loop, and the coefficients of all rows read together (rows past lg
            // repeat row 0, so every slot index is in range).  A third fewer instructions per
            // row; the same-spin passes are bound by VALU issue on their busiest SIMD
            if (E.lane >= nl) return;
            V cf[kSchedMax];
#pragma unroll
            for (int i = 0; i < kSchedMax; ++i) cf[i] = coef[s_slot(E.e[i])];
            char *tb = reinterpret_cast<char *>(tile) + lbase;
#pragma unroll
            for (int i = 0; i < kSchedMax; ++i) {
                if (i >= d.lg) break;
                const uint32_t e = E.e[i];
                if (s_rank1(e) == kSchedPad) continue;
                V *pu = reinterpret_cast<V *>(tb + s_rank1(e) * (uint32_t)sizeof(V));
                V *pv = reinterpret_cast<V *>(tb + s_rank2(e) * (uint32_t)sizeof(V));
                V u = *pu, v = *pv;
                rotate(u, v, cf[i].x, flip_sign(cf[i].y, ~s_par(e) << 31), entry_dither(z, i));
                *pu = u;
                *pv = v;
            }
            return;
        }
#pragma unroll
        for (int i = 0; i < kSchedMax; ++i) {
            if (i >= d.lg) break;
            const uint32_t e = E.e[i];
            if (s_rank1(e) == kSchedPad) continue;
            const V cf = coef[s_slot(e)];
            rotate_lines<R>(tile, line_bytes, nl, E.lane, E.L, lbase, s_rank1(e), s_rank2(e), cf.x,
                            flip_sign(cf.y, ~s_par(e) << 31), entry_dither(z, i));
        }
        return;
    }
    if (type == OP_FAB) {
        // the pass's pair lists were built when the item started; aux bits 16-18 pick this op's
        const uint32_t li = (d.aux >> 16) & (uint32_t)(kMaxPairLists - 1);
        if (li != cur_list) {
            cur_list = li;
            ncp = pair_counts[li];
        }
        const int2 *pair_list = pair_lists + li * (kMaxLines / 2);
        if (E.lane >= E.L) return;
        if (ncp <= kChunk * E.L) {
            // the thread's column pairs fit one chunk: read them once for all entries
            // (byte offsets of both columns and the pair's parity in bit 31)
            uint32_t px[kChunk], py[kChunk], pb[kChunk];
            int npl = 0;
#pragma unroll
            for (int k = 0; k < kChunk; ++k) {
                const int2 pl = pair_list[min(E.lane + k * E.L, max(ncp - 1, 0))];
                px[k] = (uint32_t)pl.x & ~7u;
                py[k] = (uint32_t)pl.y;
                pb[k] = (uint32_t)pl.x << 31;
                npl += E.lane + k * E.L < ncp ? 1 : 0;
            }
#pragma unroll
            for (int i = 0; i < kSchedMax; ++i) {
                if (i >= d.lg) break;
                const uint32_t e = E.e[i];
                if (s_rank1(e) == kSchedPad) continue;
                const V cf = coef[s_slot(e)];
                rotate_pairs_held<R>(tile, px, py, pb, npl, s_rank1(e), s_rank2(e), s_par(e), cf.x, cf.y,
                                     entry_dither(z, i));
            }
            return;
        }
#pragma unroll
        for (int i = 0; i < kSchedMax; ++i) {
            if (i >= d.lg) break;
            const uint32_t e = E.e[i];
            if (s_rank1(e) == kSchedPad) continue;
            const V cf = coef[s_slot(e)];
            rotate_pairs<R>(tile, pair_list, ncp, E.lane, E.L, s_rank1(e), s_rank2(e), s_par(e), cf.x, cf.y,
                            entry_dither(z, i));
        }
    } else if constexpr (OPS == kOpsAll) {
      if (type == OP_FCC) {
        if (E.lane >= E.L) return;
#pragma unroll
        for (int i = 0; i < kSchedMax; ++i) {
            if (i >= d.lg) break;
            const uint32_t e = E.e[i];
            if (s_rank1(e) == kSchedPad) continue;
            const V cf = coef[s_slot(e)];
            rotate_lines<R>(tile, line_bytes, nl, E.lane, E.L, lbase, s_rank1(e), s_rank2(e), cf.x,
                            flip_sign(cf.y, ~s_par(e) << 31), entry_dither(z, i));
        }
    } else if (type == OP_PQQR) {
        if (E.lane >= E.L) return;
#pragma unroll
        for (int i = 0; i < kSchedMax; ++i) {
            if (i >= d.lg) break;
            const uint32_t e = E.e[i];
            if (s_rank1(e) == kSchedPad) continue;
            const V cf = coef[s_slot(e)];
            rotate_lines_gated<R>(tile, line_bytes, nl, E.lane, E.L, lbase, s_rank1(e), s_rank2(e), cf.x,
                                  flip_sign(cf.y, ~s_par(e) << 31), line_str, d.aux, entry_dither(z, i));
        }
    } else if (type == OP_GA || type == OP_GB) {
        if (E.lane >= E.L) return;
        const double2 *ga = gang + (size_t)d.aux * nl;
#pragma unroll
        for (int i = 0; i < kSchedMax; ++i) {
            if (i >= d.lg) break;
            const uint32_t e = E.e[i];
            if (s_rank1(e) == kSchedPad) continue;
            const double *gp = A.grp_a + 2ull * ((size_t)d.index + (size_t)i * d.ns + E.slot);
            rotate_group<R>(tile, line_bytes, nl, E.lane, E.L, lbase, s_rank1(e), s_rank2(e), s_par(e), gp[0], gp[1], ga,
                            entry_dither(z, i));
        }
    } else if (type == OP_DIAG) {
        const uint32_t total = (uint32_t)nl * C;
        for (uint32_t idx = tid; idx < total; idx += kThreads) {
            uint32_t l = idx / C, a = idx - l * C;
            uint32_t b = line_rank[l];
            uint32_t sa = A.strings[a];
            const uint64_t set = (uint64_t)d.index;   // diagonal table set
            double th = A.theta_a[set * C + a] + A.theta_b[set * C + b];
            const double *cross = A.cross + set * (uint64_t)A.n * C;
            while (sa) {
                int p = __builtin_ctz(sa);
                th += cross[(uint64_t)p * C + b];
                sa &= sa - 1;
            }
            // e^{-i th} = 1 - h - i s with h = 2 sin^2(th/2), s = sin th
            double s2, c2;
            sincos(0.5 * th, &s2, &c2);
            V v = tile[idx], w;
            const R h = (R)(2.0 * s2 * s2), s = (R)(2.0 * s2 * c2);
            const uint32_t r = mix32(tseed + idx * 0x9E3779B9u);
            w.x = upd<R>(v.x, s * v.y - h * v.x, dfac(r, 0));
            w.y = upd<R>(v.y, -s * v.x - h * v.y, dfac(r, 1));
            tile[idx] = w;
        }
      }
    }
}

// Tile copies between HBM and LDS in 16-byte units, copy_unroll<V>() loads in
// flight per thread before their LDS writes.  COL tiles: whole lines (columns)
// of C amplitudes; the line and its LDS image are 16-byte aligned when
// C * sizeof(V) is a multiple of 16, else the copy goes element by element.
// complex64 tiles: 12 (a ten-column tile at C = 1820 is 9100 units, one round
// for 1024 threads); complex128: 8 (register budget)
template <typename V> constexpr int copy_unroll() { return sizeof(V) == 8 ? 12 : 8; }
// 16-byte copy unit: a native vector type, loaded and stored as one value (a
// struct type would be copied with memcpy, which can keep the staging array
// out of registers)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Block traffic of the dataflow evolution (FLOW = true): the workgroup that
// writes a column and the one that reads it next may sit on different XCDs,
// whose L2 caches are not coherent with one another, so tiles are loaded and
// stored with device-scope coherence (sc1, the policy the compiler gives
// agent-scope relaxed atomics) through a buffer descriptor over the block.
// Offsets are bytes from the block base (blocks below 2^31 bytes); the
// descriptor's record count bounds every access to the block.
constexpr int kCohPolicy = 16;   // sc1
struct Coh {
    __amdgpu_buffer_rsrc_t r;
    __device__ Coh(const void *base, uint32_t bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000)) {}
    template <typename U> __device__ U load(uint32_t off) const {
        if constexpr (sizeof(U) == 16)
            return __builtin_bit_cast(U, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, kCohPolicy));
        else
            return __builtin_bit_cast(U, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, kCohPolicy));
    }
    template <typename U> __device__ void store(uint32_t off, const U &v) const {
        if constexpr (sizeof(U) == 16)
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, (int)off, 0, kCohPolicy);
        else {
            typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)off, 0, kCohPolicy);
        }
    }
};
// coherent FP64 load / store of one value written or read by other workgroups of the launch
__device__ inline double coh_load_f64(const double *p) {
    return __builtin_bit_cast(double, __hip_atomic_load(reinterpret_cast<const uint64_t *>(p), __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT));
}
__device__ inline void coh_store_f64(double *p, double v) {
    __hip_atomic_store(reinterpret_cast<uint64_t *>(p), __builtin_bit_cast(uint64_t, v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Unrolled copies are branch-free: an index past the end is clamped to the
// last element, whose value such a thread then loads and writes again (the
// same value the owning thread writes).  Guarded writes would let the compiler
// sink each load next to its write and serialise the copy.  blk_bytes: the
// block's size (FLOW copies only).
template <typename V, typename U, bool FLOW>
__device__ __forceinline__ void copy_cols_in(const V *src, V *tile, const uint32_t *line_rank, int nl, uint32_t C, double sc,
                                             uint32_t blk_bytes) {
    constexpr int kU = copy_unroll<V>();
    const uint32_t upl = (uint32_t)(C * sizeof(V) / sizeof(U));   // units per line
    const uint32_t total = (uint32_t)nl * upl;
    const U *s = reinterpret_cast<const U *>(src);
    U *t = reinterpret_cast<U *>(tile);
    const Coh coh(src, blk_bytes);
    for (uint32_t i0 = threadIdx.x; i0 < total; i0 += kThreads * kU) {
        U r[kU];
        uint32_t ii[kU];
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            ii[k] = min(i0 + (uint32_t)k * kThreads, total - 1);
            const uint32_t l = ii[k] / upl, a = ii[k] - l * upl;
            if constexpr (FLOW) r[k] = coh.load<U>((line_rank[l] * upl + a) * (uint32_t)sizeof(U));
            else r[k] = s[(uint64_t)line_rank[l] * upl + a];
        }
#pragma unroll
        for (int k = 0; k < kU; ++k) t[ii[k]] = r[k];
    }
    if (sc != 1.0) {
        __syncthreads();
        const uint32_t n = (uint32_t)nl * C;
        for (uint32_t i = threadIdx.x; i < n; i += kThreads) {
            V v = tile[i];
            v.x *= (decltype(v.x))sc;
            v.y *= (decltype(v.y))sc;
            tile[i] = v;
        }
    }
}

// ROW tiles: rows r0 .. r0+nl-1 (alpha strings) of all C columns, transposed
// into lines of C amplitudes; global runs of nl consecutive amplitudes.
template <typename V, bool FLOW>
__device__ __forceinline__ void copy_rows_in(const V *src, V *tile, uint32_t r0, int nl, uint32_t C, double sc,
                                             uint32_t blk_bytes) {
    constexpr int kU = copy_unroll<V>();
    const uint32_t total = (uint32_t)nl * C;
    const auto f = (decltype(V().x))sc;
    const Coh coh(src, blk_bytes);
    for (uint32_t i0 = threadIdx.x; i0 < total; i0 += kThreads * kU) {
        V r[kU];
        uint32_t tb[kU];
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            const uint32_t i = min(i0 + (uint32_t)k * kThreads, total - 1);
            const uint32_t b = i / (uint32_t)nl, l = i - b * (uint32_t)nl;
            if constexpr (FLOW) r[k] = coh.load<V>((b * C + r0 + l) * (uint32_t)sizeof(V));
            else r[k] = src[(uint64_t)b * C + r0 + l];
            tb[k] = l * C + b;
        }
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            V v = r[k];
            v.x *= f;
            v.y *= f;
            tile[tb[k]] = v;
        }
    }
}

template <typename V, bool FLOW>
__device__ __forceinline__ void copy_rows_out(V *dst, const V *tile, uint32_t r0, int nl, uint32_t C, uint32_t blk_bytes) {
    constexpr int kU = copy_unroll<V>();
    const uint32_t total = (uint32_t)nl * C;
    const Coh coh(dst, blk_bytes);
    for (uint32_t i0 = threadIdx.x; i0 < total; i0 += kThreads * kU) {
        V r[kU];
        uint64_t gi[kU];
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            const uint32_t i = min(i0 + (uint32_t)k * kThreads, total - 1);
            const uint32_t b = i / (uint32_t)nl, l = i - b * (uint32_t)nl;
            r[k] = tile[l * C + b];
            gi[k] = (uint64_t)b * C + r0 + l;
        }
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            if constexpr (FLOW) coh.store<V>((uint32_t)gi[k] * (uint32_t)sizeof(V), r[k]);
            else dst[gi[k]] = r[k];
        }
    }
}

template <typename V, typename U, bool FLOW>
__device__ __forceinline__ void copy_cols_out(V *dst, const V *tile, const uint32_t *line_rank, int nl, uint32_t C,
                                              uint32_t blk_bytes) {
    constexpr int kU = copy_unroll<V>();
    const uint32_t upl = (uint32_t)(C * sizeof(V) / sizeof(U));
    const uint32_t total = (uint32_t)nl * upl;
    U *d = reinterpret_cast<U *>(dst);
    const U *t = reinterpret_cast<const U *>(tile);
    const Coh coh(dst, blk_bytes);
    for (uint32_t i0 = threadIdx.x; i0 < total; i0 += kThreads * kU) {
        U r[kU];
        uint64_t gi[kU];
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            const uint32_t i = min(i0 + (uint32_t)k * kThreads, total - 1);
            const uint32_t l = i / upl, a = i - l * upl;
            r[k] = t[i];
            gi[k] = (uint64_t)line_rank[l] * upl + a;
        }
#pragma unroll
        for (int k = 0; k < kU; ++k) {
            if constexpr (FLOW) coh.store<U>((uint32_t)gi[k] * (uint32_t)sizeof(U), r[k]);
            else d[gi[k]] = r[k];
        }
    }
}

// The op loop of one tile item.  Ops run in batches of kBatch whose
// descriptors and coefficients were staged in the LDS rings; each op's
// schedule entries load one op ahead into the other register set.  GLOBAL: the
// lines are columns of the block in global memory (gtile), else the LDS tile.
template <typename R, bool GLOBAL, int OPS = kOpsAll>
__device__ __forceinline__ void op_loop(const PassArgs &A, const PassDev &pass, uint32_t item, uint64_t j,
                                        typename Vec<R>::type *gtile, OpDev *ring, double2 *cring, EntSet &E0,
                                        EntSet &E1, const double2 *gang, const uint32_t *line_rank,
                                        const uint32_t *line_str, const uint32_t *line_base, const int2 *pair_lists,
                                        const int *pair_counts, int nl, uint32_t C) {
    using V = typename Vec<R>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    V *tile = GLOBAL ? gtile : reinterpret_cast<V *>(smem + kLdsHeader);
    const int o0 = pass.op_begin, o1 = pass.op_end;
    uint32_t cur_list = 0xffffffffu;   // pair list of the last mixed-window op, and its count
    int ncp = 0;
    const int nbatch = (o1 - o0 + kBatch - 1) / kBatch;
    for (int b = 0; b < nbatch; ++b) {
        // batch b+1 is fetched now and committed before the barrier of op kBatch/2: every
        // wave has then left batch b-1, whose ring half it reuses, and the entries of its
        // first op are read (at op kBatch-2) two barriers later
        const bool more = b + 1 < nbatch;
        Staged<R> st;
        if (more) st = fetch_batch<R>(A, o0, o1, b + 1);
        // two ops per iteration: the even op runs on E0 while E1 loads, the odd op the other way round
        for (int i = 0; i < kBatch; i += 2) {
            const int o = o0 + b * kBatch + i;
            if (o >= o1) break;
            if (i == kBatch / 2 && more) commit_batch<R>(st, ring, cring, b + 1);
            const OpDev *rb = ring + (b & 1) * kBatch;
            const V *cb = reinterpret_cast<const V *>(cring) + (size_t)((b & 1) * kBatch) * kMaxGroup;
            load_entries(A, rb[i + 1], o + 1 < o1, E1);
            {
                // descriptors are read before the barrier (their ring half was committed
                // at least one barrier earlier)
                const OpDev d = rb[i];
                op_barrier(GLOBAL);
                run_op<R, OPS>(A, d, cb + i * kMaxGroup, E0, tile, gang, line_rank, line_str, line_base, pair_lists,
                          pair_counts, nl, C, cur_list, ncp,
                          mix32(mix32((uint32_t)j * 0x27D4EB2Fu ^ (uint32_t)o * 0x165667B1u ^ item * 0x9E3779B1u) ^ threadIdx.x));
            }
            if (o + 1 >= o1) break;
            // entries of op o+2 (the next batch's first op after the last pair)
            const OpDev &dn = (i + 2 < kBatch) ? rb[i + 2] : ring[((b + 1) & 1) * kBatch];
            load_entries(A, dn, o + 2 < o1, E0);
            {
                const OpDev d = rb[i + 1];
                op_barrier(GLOBAL);
                run_op<R, OPS>(A, d, cb + (i + 1) * kMaxGroup, E1, tile, gang, line_rank, line_str, line_base, pair_lists,
                          pair_counts, nl, C, cur_list, ncp,
                          mix32(mix32((uint32_t)j * 0x27D4EB2Fu ^ (uint32_t)(o + 1) * 0x165667B1u ^ item * 0x9E3779B1u) ^
                              threadIdx.x));
            }
        }
    }
}

struct NoGate {
    __device__ bool operator()() const { return true; }
};

// One tile item of one pass: load the item's lines, run the pass's ops, write back.
// FLOW: an item of the dataflow evolution (block traffic, the renormalisation
// factor and the norm partial coherent at device scope; LDS-tiled passes only).
// gate() runs, in every thread, after the item's tables and first op are staged
// and before its tile is read; false abandons the item (returns false).
template <typename R, int MODE, bool FLOW = false, class Gate = NoGate, int OPS = kOpsAll>
__device__ __forceinline__ bool pass_body(const PassArgs &A, const PassDev &pass, uint32_t item,
                                          typename Vec<R>::type *prev, typename Vec<R>::type *cur, uint64_t j,
                                          Gate gate = Gate()) {
    static_assert(!FLOW || MODE == 1, "dataflow items are LDS-tiled");
    using V = typename Vec<R>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    // LDS header (bytes): line ranks [0,256), line strings [256,512), norm reduction
    // [512,768), LDS line index [1040,1296), descriptor ring [1296,1808), coefficient ring
    // [1808,3856), dataflow broadcast words [3856,3868), the mixed windows' pair lists
    // [4096,6144) and their counts [6144,6176); the tile starts at kLdsHeader
    uint32_t *line_rank = reinterpret_cast<uint32_t *>(smem);            // [kMaxLines]
    uint32_t *line_str = line_rank + kMaxLines;                         // [kMaxLines]
    double *red = reinterpret_cast<double *>(smem + 512);               // [2 * kWaves], after the ops
    int2 *pair_lists = reinterpret_cast<int2 *>(smem + 4096);           // [kMaxPairLists][kMaxLines / 2]
    int *pair_counts = reinterpret_cast<int *>(smem + 6144);            // [kMaxPairLists]
    static_assert(4096 + kMaxPairLists * (kMaxLines / 2) * 8 == 6144 && 6144 + 4 * kMaxPairLists <= (int)kLdsHeader,
                  "pair lists overlap the tile");
    OpDev *ring = reinterpret_cast<OpDev *>(smem + 1296);              // [2][kBatch]
    double2 *cring = reinterpret_cast<double2 *>(smem + 1808);          // [2][kBatch][kMaxGroup]
    static_assert(16 * kWaves <= 256, "norm reduction does not fit its LDS region");
    static_assert(1808 + 2 * kBatch * kMaxGroup * 16 + 16 <= 4096, "LDS header overlaps the pair lists");
    // lines stay in global memory (MODE 1: never, 2: always, 0: as the pass says)
    const bool gmode = MODE == 2 ? true : MODE == 1 ? false : pass.kind == PASS_COLG;
    V *tile = reinterpret_cast<V *>(smem + kLdsHeader);
    uint32_t *line_pos = reinterpret_cast<uint32_t *>(smem + 1040);     // [kMaxLines] LDS line index (0..63)

    const int tid = threadIdx.x;
    const ItemDev it = A.items[pass.item_begin + item];
    const int nl = (int)it.count;
    const uint32_t C = A.C;
    // a mixed-window item of the dataflow kernel (FLOW, OPS == kOpsFab) is a COL pass in the
    // middle of a step with no staged group angles: its body carries none of the other
    // passes' paths (the dataflow kernel holds three bodies; measured 3% off the batch)
    constexpr bool kMid = FLOW && OPS == kOpsFab;
    const bool col = kMid || pass.kind != PASS_ROW;   // COL and COLG lines are columns from the column list
    double2 *gang = reinterpret_cast<double2 *>(smem + kLdsHeader + (size_t)pass.max_lines * C * sizeof(V));
    const int o0 = pass.op_begin, o1 = pass.op_end;

    if (tid < nl) {
        uint32_t r = col ? A.col_list[it.begin + tid] : it.begin + tid;
        line_rank[tid] = r;
        line_pos[tid] = (uint32_t)tid;
        line_str[tid] = A.strings[r];
    }
    commit_batch<R>(fetch_batch<R>(A, o0, o1, 0), ring, cring, 0);
    __syncthreads();
    // mixed windows: wave w builds the pair list of the pass's w-th (q, r); the first op's
    // barrier publishes them
    {
        const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
        if (w < pass.pair_lists)
            build_pair_list(A.pass_qr[pass.pair_qr + w], line_str, gmode ? line_rank : line_pos,
                            pair_lists + w * (kMaxLines / 2), pair_counts + w, nl,
                            C * (uint32_t)sizeof(V));
    }
    for (int i = tid; !kMid && i < pass.group_count * nl; i += kThreads) {
        int g = i / nl, l = i - g * nl;
        const double *gb = A.grp_b + 2ull * (pass.group_first + g) * C + 2ull * line_rank[l];
        gang[i] = make_double2(gb[0], gb[1]);
    }
    EntSet E0, E1;
    load_entries(A, ring[0], o0 < o1, E0);   // overlaps the tile load below
    if (!gate()) return false;

    // ---- load the tile ---------------------------------------------------
    const bool from_prev = !kMid && (pass.flags & PF_FROM_PREV);
    const V *src = from_prev ? prev : cur;
    const double sc = (from_prev && A.renorm) ? (FLOW ? coh_load_f64(A.scale + j - 1) : A.scale[j - 1]) : 1.0;
    const uint32_t total = (uint32_t)nl * C;
    const bool vec16 = (C * sizeof(V)) % 16 == 0;
    const uint32_t blk_bytes = FLOW ? (uint32_t)(A.D * sizeof(V)) : 0u;
    if (gmode) {
        // in place in global memory: nothing to stage
    } else if (col) {
        if (vec16) copy_cols_in<V, u32x4, FLOW>(src, tile, line_rank, nl, C, sc, blk_bytes);
        else copy_cols_in<V, V, FLOW>(src, tile, line_rank, nl, C, sc, blk_bytes);
    } else {
        copy_rows_in<V, FLOW>(src, tile, it.begin, nl, C, sc, blk_bytes);
    }

    // ---- ops ------------------------------------------------------------------
    // two instantiations, so that each sees a single address space for the tile
    // (LDS accesses stay ds_read/ds_write, global ones global_load/store)
    if (gmode)
        op_loop<R, true, OPS>(A, pass, item, j, cur, ring, cring, E0, E1, gang, line_rank, line_str, line_rank, pair_lists,
                         pair_counts, nl, C);
    else
        op_loop<R, false, OPS>(A, pass, item, j, nullptr, ring, cring, E0, E1, gang, line_rank, line_str, line_pos,
                          pair_lists, pair_counts, nl, C);
    __syncthreads();

    // ---- FP64 norm of the tile (last pass of a step) ----------------------
    if (!kMid && (pass.flags & PF_EMIT_NORM)) {
        KSum acc;
        for (uint32_t idx = tid; idx < total; idx += kThreads) {
            V v = tile[idx];
            acc.add((double)v.x * (double)v.x + (double)v.y * (double)v.y);
        }
        double s = acc.s, c = acc.c;
        for (int off = 32; off > 0; off >>= 1) {
            s += __shfl_down(s, off);
            c += __shfl_down(c, off);
        }
        if ((tid & 63) == 0) {
            red[2 * (tid >> 6)] = s;
            red[2 * (tid >> 6) + 1] = c;
        }
        __syncthreads();
        if (tid == 0) {
            KSum tot;
            for (int w = 0; w < kWaves; ++w) {
                tot.add(red[2 * w]);
                tot.add(red[2 * w + 1]);
            }
            if (FLOW) coh_store_f64(A.partials + item, tot.value());
            else A.partials[item] = tot.value();
        }
    }

    // ---- store the tile -----------------------------------------------------
    if (gmode) {
        // the ops wrote global memory directly
    } else if (col) {
        if (vec16) copy_cols_out<V, u32x4, FLOW>(cur, tile, line_rank, nl, C, blk_bytes);
        else copy_cols_out<V, V, FLOW>(cur, tile, line_rank, nl, C, blk_bytes);
    } else {
        copy_rows_out<V, FLOW>(cur, tile, it.begin, nl, C, blk_bytes);
    }
    return true;
}

// MODE 1: LDS-tiled passes only, MODE 2: global-memory windows only (separate
// kernels keep the common one small)
template <typename R, int MODE, int OPS = kOpsAll>
__global__ void __launch_bounds__(kThreads)
pass_kernel(PassArgs A, PassDev pass, typename Vec<R>::type *prev, typename Vec<R>::type *cur, uint64_t j) {
    pass_body<R, MODE, false, NoGate, OPS>(A, pass, blockIdx.x, prev, cur, j);
}

// Pointers loaded from a table carry no address space, so the compiler would
// emit flat accesses (which also count against the LDS wait counter).  The
// device reads the table through a mirror whose pointers are declared in the
// global address space; converting them to the generic PassArgs keeps that
// knowledge.
#define QPE_GPTR(T) __attribute__((address_space(1))) T *
struct PassArgsG {
    QPE_GPTR(const uint32_t) binom;
    QPE_GPTR(const uint32_t) strings;
    QPE_GPTR(const double) theta_a;
    QPE_GPTR(const double) theta_b;
    QPE_GPTR(const double) cross;
    QPE_GPTR(const GroupDev) groups;
    QPE_GPTR(const double) grp_a;
    QPE_GPTR(const double) grp_b;
    QPE_GPTR(const OpDev) ops;
    QPE_GPTR(const ItemDev) items;
    QPE_GPTR(const uint32_t) col_list;
    QPE_GPTR(const uint32_t) sched;
    QPE_GPTR(const uint32_t) pass_qr;
    QPE_GPTR(const double) coef;
    QPE_GPTR(double) partials;
    QPE_GPTR(const double) scale;
    QPE_GPTR(char) state;
    uint64_t D;
    int n, e;
    uint32_t C, ca_pairs, ccaa_pairs;
    int renorm;
};
static_assert(sizeof(PassArgsG) == sizeof(PassArgs), "PassArgsG mirrors PassArgs");

__device__ inline PassArgs from_global_args(const PassArgsG &g) {
    PassArgs A;
    A.binom = (const uint32_t *)g.binom; A.strings = (const uint32_t *)g.strings;
    A.theta_a = (const double *)g.theta_a; A.theta_b = (const double *)g.theta_b; A.cross = (const double *)g.cross;
    A.groups = (const GroupDev *)g.groups; A.grp_a = (const double *)g.grp_a; A.grp_b = (const double *)g.grp_b;
    A.ops = (const OpDev *)g.ops; A.items = (const ItemDev *)g.items; A.col_list = (const uint32_t *)g.col_list;
    A.sched = (const uint32_t *)g.sched; A.pass_qr = (const uint32_t *)g.pass_qr;
    A.coef = (const double *)g.coef; A.partials = (double *)g.partials;
    A.scale = (const double *)g.scale; A.state = (void *)g.state;
    A.D = g.D; A.n = g.n; A.e = g.e; A.C = g.C; A.ca_pairs = g.ca_pairs; A.ccaa_pairs = g.ccaa_pairs;
    A.renorm = g.renorm;
    return A;
}

// Batched form: one launch runs one pass of several instances.  Workgroups
// [work[w].block_begin, work[w+1].block_begin) are the items of work[w]; the
// step index is read from device memory (*jp) so that one captured graph of
// a step serves every step.
template <typename R>
__global__ void __launch_bounds__(kThreads)
batch_pass_kernel(const PassArgs *__restrict__ args, const BatchWork *__restrict__ work, int nwork,
                  const uint64_t *__restrict__ jp) {
    using V = typename Vec<R>::type;
    int lo = 0, hi = nwork - 1;   // last w with block_begin <= blockIdx.x
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (work[mid].block_begin <= blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const BatchWork &bw = work[lo];
    const PassArgs A = from_global_args(reinterpret_cast<const PassArgsG *>(args)[bw.inst]);
    const uint64_t j = *jp;
    V *base = reinterpret_cast<V *>(A.state);
    pass_body<R, 0>(A, bw.pass, blockIdx.x - bw.block_begin, base + (j - 1) * A.D, base + j * A.D, j);
}

// sum the per-workgroup partials of block j; set norms[j] and scale[j]
__device__ inline void finalize_body(const double *partials, int count, uint64_t j, double *norms, double *scale,
                                     int *status) {
    __shared__ double red[2 * 16];
    KSum acc;
    for (int i = threadIdx.x; i < count; i += blockDim.x) acc.add(partials[i]);
    double s = acc.s, c = acc.c;
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_down(s, off);
        c += __shfl_down(c, off);
    }
    if ((threadIdx.x & 63) == 0) {
        red[2 * (threadIdx.x >> 6)] = s;
        red[2 * (threadIdx.x >> 6) + 1] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        KSum tot;
        for (int w = 0; w < (int)(blockDim.x / 64); ++w) {
            tot.add(red[2 * w]);
            tot.add(red[2 * w + 1]);
        }
        double v = tot.value();
        norms[j] = v;
        double f = 1.0;
        if (j > 0 && v > 0.0 && isfinite(v)) f = sqrt(norms[0] / v);
        scale[j] = f;
        if (!isfinite(v) && status) *status = 1;
    }
}

__global__ void norm_finalize(const double *partials, int count, uint64_t j, double *norms, double *scale,
                              int *status) {
    finalize_body(partials, count, j, norms, scale, status);
}

// resets the batched step counter (j = 1) and the finalize counter
__global__ void batch_reset(uint64_t *jp, unsigned *done) {
    if (threadIdx.x == 0) {
        *jp = 1;
        *done = 0u;
    }
}

// batched form: workgroup i finalizes instance i; the last workgroup to finish
// advances the step counter (*jp) for the next step's launches
__global__ void batch_finalize(const FinalizeArgs *fa, int count, uint64_t *jp, unsigned *done) {
    const FinalizeArgs f = fa[blockIdx.x];
    const uint64_t j = *jp;
    finalize_body(f.partials, f.count, j, f.norms, f.scale, f.status);
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(done, 1u) == (unsigned)count - 1u) {
            *done = 0u;
            *jp = j + 1;
            __threadfence();
        }
    }
}

// ---- dataflow evolution ------------------------------------------------------
// One launch runs Trotter steps j_begin..j_end-1 of one register.  Its
// workgroups (one per CU) claim items from a device counter in order -- step
// by step, pass by pass, item by item --, wait until the items whose writes
// their tile reads are done (the lists of build_flow; first-pass items wait
// for the previous step), run the item and stamp it done with its step
// number.  The workgroup that completes a step's last norm item sums the block
// norm and publishes norms[j], scale[j] before the step counts as done.  Every
// dependency is an earlier claim, so the lowest unfinished claim can always
// proceed, whether or not all workgroups are resident; every wait is bounded
// by a wall-clock limit, after which the launch aborts with status 2.  Items
// run in the same order per column as the pass-by-pass launches, with the
// same rounding dither, so results are identical to theirs.
struct FlowArgs {
    const uint32_t *item_pass;   // [ipS] pass of each item
    const uint32_t *pass_first;  // [passes] number of the pass's first item
    const uint32_t *dep_off;     // [ipS + 1]
    const uint32_t *deps;        // item numbers
    const PassDev *passes;
    uint32_t *done;              // [ipS] step number of the item's last completion (0: none in this launch)
    uint32_t *ctr;               // this register's counters: [1] norm items done, [2] last finished step
    uint32_t *gctr;              // the launch's counters: [0] next claim, [3] abort (one register: ctr)
    double *norms, *scale;
    int *status;
    uint32_t j_begin, j_end;
    uint32_t ipS;                // items per step
    uint32_t norm_items;         // items of the pass that sums the block norm
    uint64_t wait_ticks;         // limit of one wait in s_memrealtime ticks (100 MHz)
};

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left alone): every vector memory access of
// this wave has completed, stores included
constexpr unsigned kVmcntZero = 0x0F70;

__device__ inline uint32_t coh_load_u32(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void coh_store_u32(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave 0 waits until item x of step j may start: its dependencies stamped j or
// later, and for first-pass items the previous step finished.  false: abort.
__device__ bool flow_wait(const FlowArgs &F, uint32_t x, uint32_t j, bool from_prev) {
    const int lane = threadIdx.x & 63;
    const uint32_t d1 = F.dep_off[x + 1];
    uint32_t i = F.dep_off[x] + (uint32_t)lane;
    bool step_wait = from_prev && lane == 0;
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    for (;;) {
        if (step_wait && coh_load_u32(F.ctr + 2) + 1u >= j) step_wait = false;
        while (i < d1 && coh_load_u32(F.done + F.deps[i]) >= j) i += 64;
        if (__ballot(i < d1 || step_wait) == 0ull) return true;
        if (coh_load_u32(F.gctr + 3)) return false;
        if (__builtin_amdgcn_s_memrealtime() - t0 > F.wait_ticks) {
            coh_store_u32(F.gctr + 3, 1u);
            __hip_atomic_store(F.status, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// block norm of step j from the norm items' partial sums; then step j is done
__device__ void flow_finalize(const PassArgs &A, const FlowArgs &F, uint32_t j, double *red) {
    const int tid = threadIdx.x;
    KSum acc;
    for (uint32_t i = (uint32_t)tid; i < F.norm_items; i += kThreads) acc.add(coh_load_f64(A.partials + i));
    double s = acc.s, c = acc.c;
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_down(s, off);
        c += __shfl_down(c, off);
    }
    if ((tid & 63) == 0) {
        red[2 * (tid >> 6)] = s;
        red[2 * (tid >> 6) + 1] = c;
    }
    __syncthreads();
    if (tid == 0) {
        KSum tot;
        for (int w = 0; w < kWaves; ++w) {
            tot.add(red[2 * w]);
            tot.add(red[2 * w + 1]);
        }
        const double v = tot.value();
        double f = 1.0;
        if (v > 0.0 && isfinite(v)) f = sqrt(coh_load_f64(F.norms) / v);
        coh_store_f64(F.norms + j, v);
        coh_store_f64(F.scale + j, f);
        if (!isfinite(v)) __hip_atomic_store(F.status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(kVmcntZero);
        coh_store_u32(F.ctr + 2, j);
    }
    __syncthreads();
}

// Item x of step j of one register: wait for the items it reads from (after its
// read-only set-up, so their loads overlap the wait), run it, stamp it done, and
// finish the step's norm when it completes the step's last norm item.  false:
// the launch aborted.  bc: LDS broadcast words beside the pair count in the header
// ([1] go, [2] finalize).
template <typename R>
__device__ __forceinline__ bool flow_item(const PassArgs &A, const FlowArgs &F, uint32_t x, uint32_t j,
                                          volatile uint32_t *bc, unsigned char *smem) {
    using V = typename Vec<R>::type;
    const int tid = threadIdx.x;
    V *base = reinterpret_cast<V *>(A.state);
    const uint32_t k = __builtin_amdgcn_readfirstlane(F.item_pass[x]);
    const PassDev ps = F.passes[k];
    const bool from_prev = (ps.flags & PF_FROM_PREV) != 0;
    auto gate = [&]() -> bool {
        if (tid < 64) {
            const bool ok = flow_wait(F, x, j, from_prev);
            if (tid == 0) bc[1] = ok ? 1u : 0u;
        }
        __syncthreads();
        return bc[1] != 0u;
    };
    // the item runs the body compiled for its pass's op types (pad2): one op path in the
    // hot loop, as in the pass launches
    const uint32_t item = x - F.pass_first[k];
    V *prev = base + (uint64_t)(j - 1) * A.D, *cur = base + (uint64_t)j * A.D;
    bool ok;
    if (ps.pad2 == kOpsFab) ok = pass_body<R, 1, true, decltype(gate), kOpsFab>(A, ps, item, prev, cur, j, gate);
    else if (ps.pad2 == kOpsFcc) ok = pass_body<R, 1, true, decltype(gate), kOpsFcc>(A, ps, item, prev, cur, j, gate);
    else ok = pass_body<R, 1, true, decltype(gate)>(A, ps, item, prev, cur, j, gate);
    if (!ok) return false;
    // every thread's block stores and the norm partial have completed before the stamp
    __builtin_amdgcn_s_waitcnt(kVmcntZero);
    __syncthreads();
    if (tid == 0) {
        coh_store_u32(F.done + x, j);
        uint32_t fin = 0u;
        if (ps.flags & PF_EMIT_NORM) {
            const uint32_t c = __hip_atomic_fetch_add(F.ctr + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
            fin = c == (j - F.j_begin + 1u) * F.norm_items ? 1u : 0u;
        }
        bc[2] = fin;
    }
    __syncthreads();
    if (bc[2]) flow_finalize(A, F, j, reinterpret_cast<double *>(smem + 512));
    return true;
}

// the next claim of the launch (wave-uniform; ~0u once the launch aborted), through
// the LDS broadcast word bc[0]
__device__ inline uint32_t flow_claim(uint32_t *gctr, volatile uint32_t *bc) {
    if (threadIdx.x == 0) {
        uint32_t t = 0xffffffffu;
        if (!coh_load_u32(gctr + 3)) t = __hip_atomic_fetch_add(gctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bc[0] = t;
    }
    __syncthreads();
    // wave-uniform: read into a scalar register, so that the item's pass descriptor and
    // everything derived from it stay scalar as in pass_kernel
    return __builtin_amdgcn_readfirstlane(bc[0]);
}

template <typename R>
__global__ void __launch_bounds__(kThreads) flow_kernel(PassArgs A, FlowArgs F) {
    extern __shared__ __align__(16) unsigned char smem[];
    volatile uint32_t *bc = reinterpret_cast<volatile uint32_t *>(smem + 1028);
    const uint32_t total = (F.j_end - F.j_begin) * F.ipS;
    for (;;) {
        const uint32_t t = flow_claim(F.gctr, bc);
        if (t >= total) break;
        if (!flow_item<R>(A, F, t % F.ipS, F.j_begin + t / F.ipS, bc, smem)) break;
    }
}

// Batched form: the steps 1..steps of several registers in one launch.  Claim t
// is step 1 + t / per_step, member and item order[t % per_step]: one step of every
// member, the members' items interleaved (each member's in its own order, which is
// all that deadlock freedom needs: a member's dependencies are earlier claims).
template <typename R>
__global__ void __launch_bounds__(kThreads) flow_batch_kernel(const PassArgs *__restrict__ args,
                                                              const FlowArgs *__restrict__ flows,
                                                              const uint2 *__restrict__ order, uint32_t per_step,
                                                              uint32_t steps, uint32_t *gctr) {
    extern __shared__ __align__(16) unsigned char smem[];
    volatile uint32_t *bc = reinterpret_cast<volatile uint32_t *>(smem + 1028);
    const uint32_t total = steps * per_step;
    for (;;) {
        const uint32_t t = flow_claim(gctr, bc);
        if (t >= total) break;
        const uint2 mi = order[t % per_step];
        const uint32_t inst = __builtin_amdgcn_readfirstlane(mi.x), x = __builtin_amdgcn_readfirstlane(mi.y);
        const PassArgs A = from_global_args(reinterpret_cast<const PassArgsG *>(args)[inst]);
        const FlowArgs F = flows[inst];
        if (!flow_item<R>(A, F, x, 1u + t / per_step, bc, smem)) break;
    }
}

// before a dataflow launch: no item done, no claim, steps below j_begin finished
// (ctr: the register's counters; the launch's, gctr, too when it is the same array)
__global__ void flow_reset(uint32_t *done, uint32_t n, uint32_t *ctr, uint32_t finished) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) done[i] = 0u;
    if (i == 0) {
        ctr[0] = 0u;
        ctr[1] = 0u;
        ctr[2] = finished;
        ctr[3] = 0u;
    }
}

// squared norm of one block, partial sums per workgroup
template <typename R>
__global__ void __launch_bounds__(256) block_norm_kernel(const typename Vec<R>::type *blk, uint64_t D, double *partials) {
    __shared__ double red[2 * 4];
    KSum acc;
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < D; i += (uint64_t)gridDim.x * 256ull) {
        auto v = blk[i];
        acc.add((double)v.x * (double)v.x + (double)v.y * (double)v.y);
    }
    double s = acc.s, c = acc.c;
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_down(s, off);
        c += __shfl_down(c, off);
    }
    if ((threadIdx.x & 63) == 0) {
        red[2 * (threadIdx.x >> 6)] = s;
        red[2 * (threadIdx.x >> 6) + 1] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        KSum tot;
        for (int w = 0; w < 4; ++w) {
            tot.add(red[2 * w]);
            tot.add(red[2 * w + 1]);
        }
        partials[blockIdx.x] = tot.value();
    }
}

// ---- QIFT ------------------------------------------------------------------
constexpr int kQThreads = 512;
constexpr int kQAcc = 32;   // P(k) accumulators per thread: M*w <= kQThreads*kQAcc

__device__ inline uint32_t bitrev(uint32_t x, int bits) { return __brev(x) >> (32 - bits); }

template <typename R>
__global__ void __launch_bounds__(kQThreads)
qift_kernel(typename Vec<R>::type *state, uint64_t D, int m, int w, const double *scale, const double2 *twiddles,
            double *pk_partial) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t M = 1u << m;
    double2 *tw = reinterpret_cast<double2 *>(smem);          // [M/2]
    double2 *tile = tw + (M / 2 > 0 ? M / 2 : 1);              // [M][w]
    const int tid = threadIdx.x;
    const uint32_t elems = M * (uint32_t)w;
    const int nacc = (int)((elems + kQThreads - 1) / kQThreads);
    const double inv = 1.0 / sqrt((double)M);

    for (uint32_t i = tid; i < M / 2; i += kQThreads) tw[i] = twiddles[i];
    double acc[kQAcc], cmp[kQAcc];
#pragma unroll
    for (int i = 0; i < kQAcc; ++i) { acc[i] = 0.0; cmp[i] = 0.0; }

    const uint64_t ntiles = (D + w - 1) / w;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t d0 = t * (uint64_t)w;
        const int cols = (int)((D - d0) < (uint64_t)w ? (D - d0) : (uint64_t)w);
        __syncthreads();
        for (uint32_t idx = tid; idx < elems; idx += kQThreads) {
            uint32_t jj = idx / w, c = idx - jj * w;
            double2 v = make_double2(0.0, 0.0);
            if ((int)c < cols) {
                auto x = state[(uint64_t)jj * D + d0 + c];
                double f = scale[jj];
                v.x = (double)x.x * f;
                v.y = (double)x.y * f;
            }
            tile[bitrev(jj, m) * w + c] = v;
        }
        for (uint32_t len = 2; len <= M; len <<= 1) {
            __syncthreads();
            const uint32_t half = len >> 1, stride = M / len;
            for (uint32_t idx = tid; idx < (M / 2) * w; idx += kQThreads) {
                uint32_t c = idx % w, b = idx / w;
                uint32_t grp = b / half, kk = b - grp * half;
                uint32_t i0 = grp * len + kk, i1 = i0 + half;
                double2 u = tile[i0 * w + c], v = tile[i1 * w + c], z = tw[kk * stride];
                double2 vz = make_double2(v.x * z.x - v.y * z.y, v.x * z.y + v.y * z.x);
                tile[i0 * w + c] = make_double2(u.x + vz.x, u.y + vz.y);
                tile[i1 * w + c] = make_double2(u.x - vz.x, u.y - vz.y);
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kQAcc; ++i) {
            uint32_t idx = tid + (uint32_t)i * kQThreads;
            if (i < nacc && idx < elems) {
                uint32_t kk = idx / w, c = idx - kk * w;
                if ((int)c < cols) {
                    double2 y = tile[idx];
                    y.x *= inv;
                    y.y *= inv;
                    auto &o = state[(uint64_t)kk * D + d0 + c];
                    o.x = (R)y.x;
                    o.y = (R)y.y;
                    double p = y.x * y.x + y.y * y.y;
                    double s = acc[i] + p;
                    if (fabs(acc[i]) >= p) cmp[i] += (acc[i] - s) + p;
                    else cmp[i] += (p - s) + acc[i];
                    acc[i] = s;
                }
            }
        }
    }
    // per-workgroup P(k): combine the w lanes of each k in a fixed order
    __syncthreads();
    double *red = reinterpret_cast<double *>(tile);
#pragma unroll
    for (int i = 0; i < kQAcc; ++i) {
        uint32_t idx = tid + (uint32_t)i * kQThreads;
        if (i < nacc && idx < elems) red[idx] = acc[i] + cmp[i];
    }
    __syncthreads();
    for (uint32_t k = tid; k < M; k += kQThreads) {
        double s = 0.0, c = 0.0;
        for (int q = 0; q < w; ++q) {
            double x = red[k * w + q];
            double t2 = s + x;
            if (fabs(s) >= fabs(x)) c += (s - t2) + x;
            else c += (x - t2) + s;
            s = t2;
        }
        pk_partial[(uint64_t)blockIdx.x * M + k] = s + c;
    }
}

// ---- QIFT, register form (2^m <= 1024) ------------------------------------------
// Y[k] = 2^{-m/2} sum_j scale[j] X[j] e^{-2 pi i jk/M} per column, as two rounds of
// in-register DFTs (the "four-step" split M = R1 * R2, j = j1 + R1 j2, k = k2 + R2 k1):
//   A  thread (j1, column) loads X[j1 + R1 j2] for all j2 straight from HBM, takes the
//      R2-point DFT over j2 in registers and multiplies by e^{-2 pi i j1 k2 / M};
//   C  thread (k2, column) reads the R1 values Z[k2][j1] through LDS, takes the R1-point
//      DFT over j1 and writes Y[k2 + R2 k1] back to HBM.
// LDS holds one exchange of the tile instead of log2(M) radix-2 stages.  A tile is W
// columns: W * sizeof(amplitude) = 128 bytes or more per row.  The DFTs run in the
// state's precision (FP32 for complex64, like the rotations); |Y|^2 and the P(k) sums
// are FP64 with compensation.
template <typename R, int LOGM> struct QiftShape {
    static constexpr int M = 1 << LOGM;
    static constexpr int LR1 = LOGM / 2, LR2 = LOGM - LR1;
    static constexpr int R1 = 1 << LR1, R2 = 1 << LR2;
    static constexpr int S = (int)sizeof(typename Vec<R>::type);
    static constexpr int W0 = 256 / R2 > 128 / S ? 256 / R2 : 128 / S;   // rows of 128 bytes or more
    static constexpr int W = W0 > 64 ? 64 : W0;    // the lanes of one column group share a wave
    static constexpr int T = R2 * W;                // threads (step C: all; step A: R1 * W)
    static constexpr int PAD = 128 / S;             // row pad: the next k2 row starts 128 bytes of banks later
    static constexpr int PITCH = R1 * W + PAD;      // elements per k2 row of the exchange tile
    static constexpr int OWN = R1 / W > 1 ? R1 / W : 1;   // P(k) values each lane keeps after the reduction
    // loads one tile ahead: measured faster for complex128 (0.49 against 0.42 of the HBM
    // peak at 2^m = 1024); complex64 at 2^m = 1024 would spill (0.32 against 0.47)
    static constexpr bool PREFETCH = S == 16;
    static constexpr int LW = W >= 64 ? 6 : W >= 32 ? 5 : W >= 16 ? 4 : W >= 8 ? 3 : W >= 4 ? 2 : W >= 2 ? 1 : 0;
    static constexpr size_t lds_bytes() {
        return (size_t)M * sizeof(double) + ((size_t)(M / 2 > 0 ? M / 2 : 1) + (size_t)R2 * PITCH) * S;
    }
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C2> __device__ __forceinline__ C2 cadd(C2 a, C2 b) { C2 r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename C2> __device__ __forceinline__ C2 csub(C2 a, C2 b) { C2 r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }

// f(integral_constant<int, I>) for I = 0 .. N-1: array indices that are constants from
// the start, so the arrays below live in registers (a runtime loop index, even one the
// compiler unrolls later, can leave an array in memory behind selects of addresses)
template <int I, int N, class F> __device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

template <int L> constexpr int rev_bits(int x) {
    int r = 0;
    for (int i = 0; i < L; ++i) r |= ((x >> i) & 1) << (L - 1 - i);
    return r;
}

// one radix-2 stage of span LEN (DIT, bit-reversed input order)
template <int N, int LEN, typename C2>
__device__ __forceinline__ void dft_stage(C2 (&a)[N], const C2 *tw, int M) {
    constexpr int H = LEN / 2;
    static_for<0, H>([&](auto kk) {
        // kk = 0: the twiddle is 1; kk = LEN/4: it is -i (exact, no multiply)
        constexpr bool unit = kk.value == 0, minus_i = 4 * kk.value == LEN;
        const C2 z = tw[kk.value * (M / LEN)];
        static_for<0, N / LEN>([&](auto g) {
            constexpr int i0 = g.value * LEN + kk.value, i1 = i0 + H;
            const C2 u = a[i0];
            C2 t = a[i1];
            if constexpr (minus_i) {
                const auto x = t.x;
                t.x = t.y;
                t.y = -x;
            } else if constexpr (!unit) {
                t = cmul(t, z);
            }
            a[i0] = cadd(u, t);
            a[i1] = csub(u, t);
        });
        // one twiddle in flight at a time: the scheduler would otherwise load every
        // stage's twiddles up front and run out of registers
        __builtin_amdgcn_sched_barrier(0);
    });
}

// forward DFT of 2^L points in registers, natural order in and out;
// tw[k] = e^{-2 pi i k / M} for k < M / 2
template <int L, typename C2>
__device__ __forceinline__ void reg_dft(C2 (&v)[1 << L], const C2 *tw, int M) {
    constexpr int N = 1 << L;
    if constexpr (L > 0) {
        C2 a[N];
        static_for<0, N>([&](auto i) { a[rev_bits<L>(i.value)] = v[i.value]; });
        static_for<1, L + 1>([&](auto st) { dft_stage<N, (1 << st.value)>(a, tw, M); });
        static_for<0, N>([&](auto i) { v[i.value] = a[i.value]; });
    }
}

// Where one launch of the register-form kernel reads and writes.  Group g (blockIdx.y)
// takes the 2^LOGM-point DFT of the rows g*gi + j*ji of src (j < 2^LOGM) and writes its
// output k to row g*go + k*ko of dst; rows are D amplitudes apart.  The inputs are scaled
// by scale[row] (use_scale), output k of group g is multiplied by e^{-2 pi i g k / Mt}
// (group_twiddle) and by f, and |output|^2 is summed into P(g*gk + k*kk) (emit_pk).  The
// sub-DFT's twiddles e^{-2 pi i k / 2^LOGM} are tw_full[k * tw_step], tw_full holding
// e^{-2 pi i k / Mt} for k < Mt / 2.  One launch with one group is the whole QIFT for
// 2^m <= 1024; two launches with 2^m = R1 * R2 groups split a larger one (four-step).
struct QiftMap {
    uint64_t gi, ji, go, ko, gk, kk;
    uint32_t Mt, tw_step;
    int group_twiddle, use_scale, emit_pk, pad;
    double f;
};

// SPLIT: one of the two launches of a 2^m > 1024 QIFT (groups, strides, group twiddles);
// otherwise the whole QIFT of one launch (the map's defaults, fixed at compile time)
template <typename R, int LOGM, bool SPLIT>
__global__ void __launch_bounds__((QiftShape<R, LOGM>::T))
qift_reg_kernel(const typename Vec<R>::type *src, typename Vec<R>::type *dst, uint64_t D, QiftMap mp,
                const double *scale, const double2 *tw_full, double *pk_partial) {
    using Q = QiftShape<R, LOGM>;
    using V = typename Vec<R>::type;
    constexpr int M = Q::M, R1 = Q::R1, R2 = Q::R2, W = Q::W;
    if constexpr (!SPLIT) {
        mp.gi = 0; mp.ji = 1; mp.go = 0; mp.ko = 1; mp.gk = 0; mp.kk = 1;
        mp.tw_step = 1;
        mp.group_twiddle = 0;
        mp.use_scale = 1;
        mp.emit_pk = 1;
    }
    extern __shared__ __align__(16) unsigned char smem[];
    double *sc = reinterpret_cast<double *>(smem);                     // [M]
    V *tw = reinterpret_cast<V *>(sc + M);                             // [max(1, M/2)]
    V *Z = tw + (M / 2 > 0 ? M / 2 : 1);                               // [R2][PITCH]
    const uint64_t g = blockIdx.y;
    for (int i = threadIdx.x; i < M / 2; i += Q::T) {
        const double2 z2 = tw_full[(uint64_t)i * mp.tw_step];
        V z;
        z.x = (R)z2.x;
        z.y = (R)z2.y;
        tw[i] = z;
    }
    for (int i = threadIdx.x; i < M; i += Q::T) sc[i] = mp.use_scale ? scale[g * mp.gi + (uint64_t)i * mp.ji] : 1.0;
    const R fo = (R)mp.f;
    double acc[Q::OWN], cmp[Q::OWN];
    static_for<0, Q::OWN>([&](auto i) { acc[i.value] = 0.0; cmp[i.value] = 0.0; });
    // tiles in an order that gives workgroups on one XCD (blockIdx mod 8) neighbouring columns
    const uint64_t ntiles = (D + W - 1) / W;
    const uint32_t G = gridDim.x;
    const bool by_xcd = G % 8 == 0;   // the host sizes the grid to a multiple of 8 when it can
    const uint32_t rank = by_xcd ? (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8 : blockIdx.x;
    // step A's raw row values; with PREFETCH they are loaded one tile ahead, so that the
    // loads of tile t + G are in flight while tile t finishes
    V x[R2];
    auto load_rows = [&](uint64_t tt, uint64_t Dt, int c, int j1) {
        const uint64_t d0 = tt * (uint64_t)W;
        const bool live = (uint64_t)c < Dt - d0;
        const V *s0 = src + (g * mp.gi + (uint64_t)j1 * mp.ji) * Dt + d0 + (live ? c : 0);
        const uint64_t step = (uint64_t)R1 * mp.ji * Dt;
        static_for<0, R2>([&](auto q) { x[q.value] = s0[(uint64_t)q.value * step]; });
    };
    if (Q::PREFETCH && rank < ntiles && (int)threadIdx.x / W < R1) load_rows(rank, D, threadIdx.x % W, threadIdx.x / W);
    __syncthreads();
    for (uint64_t t = rank; t < ntiles; t += G) {
        // opaque copies of D and of the thread index per tile keep the compiler from
        // hoisting every row's offset, scale factor and twiddle address out of the loop
        // (they would hold hundreds of registers)
        uint64_t Dt = D;
        int tid = threadIdx.x;
        asm volatile("" : "+s"(Dt), "+v"(tid)::"memory");
        const int c = tid % W;          // column inside the tile (both steps)
        const int row = tid / W;        // step A: j1 (active when < R1); step C: k2
        const uint64_t d0 = t * (uint64_t)W;
        const int cols = (int)((Dt - d0) < (uint64_t)W ? (Dt - d0) : (uint64_t)W);
        const bool live = c < cols;
        if (row < R1) {
            const int j1 = row;
            if (!Q::PREFETCH) load_rows(t, Dt, c, j1);
            V v[R2];
            static_for<0, R2>([&](auto q) {
                const R f = live ? (R)sc[j1 + R1 * q.value] : (R)0;
                v[q.value].x = x[q.value].x * f;
                v[q.value].y = x[q.value].y * f;
            });
            reg_dft<Q::LR2>(v, tw, M);
            static_for<1, R2>([&](auto q) {
                const int e = j1 * q.value;   // < M
                V z = tw[e & (M / 2 - 1)];
                if (e >= M / 2) {
                    z.x = -z.x;
                    z.y = -z.y;
                }
                v[q.value] = cmul(v[q.value], z);
                __builtin_amdgcn_sched_barrier(0);
            });
            static_for<0, R2>([&](auto q) { Z[q.value * Q::PITCH + j1 * W + c] = v[q.value]; });
            if (Q::PREFETCH && t + G < ntiles) load_rows(t + G, Dt, c, j1);
        }
        __syncthreads();
        const int k2 = row;
        V u[R1];
        static_for<0, R1>([&](auto q) { u[q.value] = Z[k2 * Q::PITCH + q.value * W + c]; });
        reg_dft<Q::LR1>(u, tw, M);
        double p[R1];
        V *d1 = dst + (g * mp.go + (uint64_t)k2 * mp.ko) * Dt + d0 + c;
        const uint64_t ostep = (uint64_t)R2 * mp.ko * Dt;
        static_for<0, R1>([&](auto q) {
            V y;
            y.x = u[q.value].x * fo;
            y.y = u[q.value].y * fo;
            if (SPLIT && mp.group_twiddle) {
                // e^{-2 pi i g k / Mt} for output k = k2 + R2 q of this group (g k < Mt)
                const uint64_t e = g * (uint64_t)(k2 + R2 * q.value);
                double2 z2 = tw_full[e & (mp.Mt / 2 - 1)];
                if (e >= mp.Mt / 2) {
                    z2.x = -z2.x;
                    z2.y = -z2.y;
                }
                V z;
                z.x = (R)z2.x;
                z.y = (R)z2.y;
                y = cmul(y, z);
            }
            p[q.value] = live ? (double)y.x * (double)y.x + (double)y.y * (double)y.y : 0.0;
            if (live) d1[(uint64_t)q.value * ostep] = y;
        });
        if (mp.emit_pk) {
            // sum over the W columns of the tile: a halving exchange across the lanes of the
            // column group leaves each lane OWN of the R1 sums (every lane all of them when R1 < W)
            static_for<0, Q::LW>([&](auto st) {
                constexpr int mask = 1 << st.value;
                constexpr int n = R1 >> st.value;      // values still held before this step
                const bool hi = (c & mask) != 0;
                if constexpr (n > 1) {
                    constexpr int h = n / 2;
                    static_for<0, h>([&](auto i) {
                        const double lo_v = p[i.value], hi_v = p[h + i.value];
                        const double send = hi ? lo_v : hi_v;
                        const double keep = hi ? hi_v : lo_v;
                        p[i.value] = keep + __shfl_xor(send, mask, W);
                    });
                } else {
                    p[0] += __shfl_xor(p[0], mask, W);
                }
            });
            static_for<0, Q::OWN>([&](auto i) {
                const double x = p[i.value], a = acc[i.value];
                const double s2 = a + x;
                if (fabs(a) >= fabs(x)) cmp[i.value] += (a - s2) + x;
                else cmp[i.value] += (x - s2) + a;
                acc[i.value] = s2;
            });
        }
        __syncthreads();
    }
    if (!mp.emit_pk) return;
    // the k1 values a lane holds: bit b of its column picks half b of what remained
    const int c = threadIdx.x % W, k2 = threadIdx.x / W;
    int base = 0, n = R1;
    bool writer = true;   // one lane of each group holding the same sums writes them
    for (int mask = 1; mask < W; mask <<= 1) {
        if (n > 1) {
            n >>= 1;
            if (c & mask) base += n;
        } else if (c & mask) {
            writer = false;
        }
    }
    if (writer) {
        static_for<0, Q::OWN>([&](auto i) {
            const uint64_t k = (uint64_t)(k2 + R2 * (base + i.value));
            pk_partial[(uint64_t)blockIdx.x * mp.Mt + g * mp.gk + k * mp.kk] = acc[i.value] + cmp[i.value];
        });
    }
}

// ---- QIFT as a matrix product on the matrix cores (2^m in 16..64) ------------------
// Y = F X per 16-column block, F[k][j] = e^{-2 pi i jk/M} scale[j] / sqrt(M), in real
// form: rows 2k, 2k+1 of Y~ are Re, Im of Y[k]; F~ has 2x2 blocks [[Fr, -Fi], [Fi, Fr]].
// One wave owns a block: v_mfma_*_16x16x4 with A = a 16x4 slice of F~ (lane i + 16 kk
// holds row i, column kk), B = 4 rows of X~ (lane n + 16 kk holds row kk, column n) and
// the 16x16 result in 4 registers per lane n + 16 g: register r holds row 4 g + r (FP32)
// or row g + 4 r (FP64) of column n.  8 M^2 flops per column against
// 2 * sizeof(amplitude) * M bytes.
template <typename R> struct MfmaAcc;
template <> struct MfmaAcc<float> {
    typedef float type __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ type fma(float a, float b, type c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <> struct MfmaAcc<double> {
    typedef double type __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ type fma(double a, double b, type c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

// complex row of Y that lane group g stores as its h-th value (h = 0, 1) of 16-row block rb
template <typename R> __device__ __forceinline__ int mfma_out_row(int rb, int h, int g) {
    if constexpr (sizeof(R) == 4) return 8 * rb + 2 * g + h;                       // registers (2h, 2h+1)
    else return 8 * rb + (g >> 1) + 4 * (g & 1) + 2 * h;   // Re and Im from lanes g and g ^ 1
}

template <int LOGM> struct MfmaShape {
    static constexpr int M = 1 << LOGM;
    static constexpr int RB = M / 8;   // 16-row blocks of Y~ (2M rows)
    static constexpr int KS = M / 2;   // 4-row steps of X~ (2M rows)
    static constexpr int T = 256;      // four waves, one 16-column block each at a time
    template <typename R> static constexpr size_t lds_bytes() { return (size_t)RB * KS * 64 * sizeof(R); }
};

template <typename R, int LOGM>
__global__ void __launch_bounds__(256)
qift_mfma_kernel(typename Vec<R>::type *state, uint64_t D, const double *scale, const double2 *twiddles,
                 double *pk_partial) {
    using S = MfmaShape<LOGM>;
    using V = typename Vec<R>::type;
    using A4 = typename MfmaAcc<R>::type;
    constexpr int M = S::M, RB = S::RB, KS = S::KS;
    extern __shared__ __align__(16) unsigned char smem[];
    R *Fl = reinterpret_cast<R *>(smem);   // [RB][KS][64]: each lane's A operand, in order
    const double inv = 1.0 / sqrt((double)M);
    for (int e = threadIdx.x; e < RB * KS * 64; e += S::T) {
        const int ln = e & 63, rk = e >> 6, ks = rk % KS, rb = rk / KS;
        const int row = 16 * rb + (ln & 15), col = 4 * ks + (ln >> 4);
        const int k = row >> 1, a = row & 1, j = col >> 1, b = col & 1;
        const int ex = (j * k) & (M - 1);
        double2 z = twiddles[ex & (M / 2 - 1)];
        if (ex >= M / 2) {
            z.x = -z.x;
            z.y = -z.y;
        }
        const double f = scale[j] * inv;
        const double fr = z.x * f, fi = z.y * f;
        Fl[e] = (R)(a == b ? fr : (a == 0 ? -fi : fi));
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    double acc[2 * RB], cmp[2 * RB];
    static_for<0, 2 * RB>([&](auto i) { acc[i.value] = 0.0; cmp[i.value] = 0.0; });
    const uint64_t nblk = (D + 15) / 16;
    const uint64_t gw = (uint64_t)blockIdx.x * 4 + wave, nw = (uint64_t)gridDim.x * 4;
    for (uint64_t blk = gw; blk < nblk; blk += nw) {
        // opaque per block: keeps the row offsets from being hoisted into registers
        uint64_t Dt = D;
        int lane = threadIdx.x & 63;
        asm volatile("" : "+s"(Dt), "+v"(lane)::"memory");
        const int n = lane & 15, g = lane >> 4;
        const uint64_t d = blk * 16 + n;
        const bool live = d < Dt;
        const V *src = state + (live ? d : 0) + (uint64_t)(g >> 1) * Dt;   // complex row 2 ks + g/2
        V xb[KS];
        static_for<0, KS>([&](auto ks) { xb[ks.value] = src[(uint64_t)(2 * ks.value) * Dt]; });
        A4 C[RB];
        static_for<0, RB>([&](auto rb) { C[rb.value] = A4{0, 0, 0, 0}; });
        static_for<0, KS>([&](auto ks) {
            const R bv = live ? ((g & 1) ? xb[ks.value].y : xb[ks.value].x) : (R)0;
            static_for<0, RB>([&](auto rb) {
                const R av = Fl[(rb.value * KS + ks.value) * 64 + lane];
                C[rb.value] = MfmaAcc<R>::fma(av, bv, C[rb.value]);
            });
        });
        V *dst = state + d;
        static_for<0, RB>([&](auto rb) {
            V y0, y1;
            if constexpr (sizeof(R) == 4) {
                y0.x = C[rb.value][0];
                y0.y = C[rb.value][1];
                y1.x = C[rb.value][2];
                y1.y = C[rb.value][3];
            } else {
                // even g holds the real parts of its rows, odd g the imaginary parts:
                // swap halves with the partner lane (g ^ 1, lane ^ 16)
                const bool odd = (g & 1) != 0;
                const R r0 = __shfl_xor(odd ? C[rb.value][0] : C[rb.value][2], 16);
                const R r1 = __shfl_xor(odd ? C[rb.value][1] : C[rb.value][3], 16);
                y0.x = odd ? r0 : C[rb.value][0];
                y0.y = odd ? C[rb.value][2] : r0;
                y1.x = odd ? r1 : C[rb.value][1];
                y1.y = odd ? C[rb.value][3] : r1;
            }
            if (live) {
                dst[(uint64_t)mfma_out_row<R>(rb.value, 0, g) * Dt] = y0;
                dst[(uint64_t)mfma_out_row<R>(rb.value, 1, g) * Dt] = y1;
            }
            const double p0 = live ? (double)y0.x * (double)y0.x + (double)y0.y * (double)y0.y : 0.0;
            const double p1 = live ? (double)y1.x * (double)y1.x + (double)y1.y * (double)y1.y : 0.0;
            static_for<0, 2>([&](auto h) {
                constexpr int i = 2 * rb.value + h.value;
                const double x = h.value ? p1 : p0, a0 = acc[i], s2 = a0 + x;
                if (fabs(a0) >= fabs(x)) cmp[i] += (a0 - s2) + x;
                else cmp[i] += (x - s2) + a0;
                acc[i] = s2;
            });
        });
    }
    // sum over the 16 columns (lanes n) of each row group; one row of partial sums per wave
    const int lane = threadIdx.x & 63, n = lane & 15, g = lane >> 4;
    static_for<0, 2 * RB>([&](auto i) {
        double v = acc[i.value] + cmp[i.value];
        v += __shfl_xor(v, 1, 16);
        v += __shfl_xor(v, 2, 16);
        v += __shfl_xor(v, 4, 16);
        v += __shfl_xor(v, 8, 16);
        if (n == 0) pk_partial[gw * M + mfma_out_row<R>(i.value >> 1, i.value & 1, g)] = v;
    });
}

__global__ void pk_reduce(const double *pk_partial, int grid, uint32_t M, double *probs) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M) return;
    double s = 0.0, c = 0.0;
    for (int g = 0; g < grid; ++g) {
        double x = pk_partial[(uint64_t)g * M + k];
        double t = s + x;
        if (fabs(s) >= fabs(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    probs[k] = s + c;
}

// ---- host <-> device layout ---------------------------------------------
template <typename R>
__global__ void scatter_kernel(typename Vec<R>::type *state, const double2 *staging, uint64_t offset, uint64_t count,
                               uint32_t C) {
    const uint64_t D = (uint64_t)C * C;
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256ull) {
        uint64_t ref = offset + i, jj = ref / D, rem = ref - jj * D;
        uint64_t a = rem / C, b = rem - a * C;
        double2 v = staging[i];
        auto &o = state[jj * D + b * C + a];
        o.x = (R)v.x;
        o.y = (R)v.y;
    }
}

template <typename R>
__global__ void gather_kernel(const typename Vec<R>::type *state, double2 *staging, uint64_t offset, uint64_t count,
                              uint32_t C) {
    const uint64_t D = (uint64_t)C * C;
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256ull) {
        uint64_t ref = offset + i, jj = ref / D, rem = ref - jj * D;
        uint64_t a = rem / C, b = rem - a * C;
        auto v = state[jj * D + b * C + a];
        staging[i] = make_double2((double)v.x, (double)v.y);
    }
}

// Single-string register <-> alpha x beta block.  dir 0: ab[b*C + a] = sign *
// k1[rank(interleave(a, b))]; dir 1: k1[...] = sign * ab[b*C + a].  One thread per
// alpha x beta amplitude (column-major, alpha fastest): the reads (dir 0) or writes
// (dir 1) on the single-string side are a permutation of the Sz = 0 entries.
template <typename R>
__global__ void __launch_bounds__(256) layout_kernel(typename Vec<R>::type *ab, typename Vec<R>::type *k1,
                                                     const uint32_t *strings, const uint32_t *tab_g, uint32_t C, int n,
                                                     int dir) {
    __shared__ uint32_t tab[kBinomWords];
    for (int i = threadIdx.x; i < kBinomWords; i += blockDim.x) tab[i] = tab_g[i];
    __syncthreads();
    const uint64_t D = (uint64_t)C * C;
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < D; i += (uint64_t)gridDim.x * 256ull) {
        const uint32_t b = (uint32_t)(i / C), a = (uint32_t)(i - (uint64_t)b * C);
        const uint32_t sa = strings[a], sb = strings[b];
        const uint32_t r = interleaved_rank(sa, sb, n, tab);
        const bool neg = interleave_parity(sa, sb) != 0u;
        auto v = dir == 0 ? k1[r] : ab[i];
        if (neg) {
            v.x = -v.x;
            v.y = -v.y;
        }
        if (dir == 0) ab[i] = v;
        else k1[r] = v;
    }
}

// flags (sets *flag) a non-zero single-string amplitude outside the (e alpha, e beta) sector
template <typename R>
__global__ void __launch_bounds__(256) outside_sector_kernel(const typename Vec<R>::type *k1, uint64_t D1,
                                                             const uint32_t *tab_g, int n, int e, int *flag) {
    __shared__ uint32_t tab[kBinomWords];
    for (int i = threadIdx.x; i < kBinomWords; i += blockDim.x) tab[i] = tab_g[i];
    __syncthreads();
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < D1; i += (uint64_t)gridDim.x * 256ull) {
        const auto v = k1[i];
        if (v.x == 0 && v.y == 0) continue;
        uint32_t sa = 0, sb = 0;
        if (!interleaved_unrank((uint32_t)i, 2 * e, n, tab, &sa, &sb) || popcnt(sa) != e) atomicOr(flag, 1);
    }
}

// single-string register <-> host staging (interleaved re, im doubles), with dtype conversion
template <typename R>
__global__ void copy_in_kernel(typename Vec<R>::type *state, const double2 *staging, uint64_t offset, uint64_t count) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256ull) {
        const double2 v = staging[i];
        auto &o = state[offset + i];
        o.x = (R)v.x;
        o.y = (R)v.y;
    }
}
template <typename R>
__global__ void copy_out_kernel(const typename Vec<R>::type *state, double2 *staging, uint64_t offset, uint64_t count) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256ull) {
        const auto v = state[offset + i];
        staging[i] = make_double2((double)v.x, (double)v.y);
    }
}

template <typename R>
__global__ void scale_kernel(typename Vec<R>::type *blk, uint64_t D, double f) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < D; i += (uint64_t)gridDim.x * 256ull) {
        auto v = blk[i];
        v.x = (R)((double)v.x * f);
        v.y = (R)((double)v.y * f);
        blk[i] = v;
    }
}

// 2x2 gate on ancilla bit tbit of the block index (blocks j0 with the bit 0
// and j1 = j0 | bit), every amplitude of the two blocks; cbit >= 0 restricts
// it to the blocks whose bit cbit is 1.  Threads stride over the amplitudes of
// a block (coalesced); gridDim.y strides over the block pairs.  FP64 arithmetic.
__device__ inline uint64_t insert_zero_bit(uint64_t x, int b) {
    return ((x >> b) << (b + 1)) | (x & ((1ull << b) - 1ull));
}

template <typename R>
__global__ void __launch_bounds__(256) gate_kernel(typename Vec<R>::type *amp, uint64_t D, uint64_t npairs, int tbit,
                                                   int cbit, double2 g00, double2 g01, double2 g10, double2 g11) {
    for (uint64_t pi = blockIdx.y; pi < npairs; pi += gridDim.y) {
        uint64_t j0;
        if (cbit < 0) {
            j0 = insert_zero_bit(pi, tbit);
        } else {
            const int lo = tbit < cbit ? tbit : cbit, hi = tbit < cbit ? cbit : tbit;
            j0 = insert_zero_bit(insert_zero_bit(pi, lo), hi) | (1ull << cbit);
        }
        typename Vec<R>::type *b0 = amp + j0 * D, *b1 = amp + (j0 | (1ull << tbit)) * D;
        for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < D; i += (uint64_t)gridDim.x * 256ull) {
            const auto u = b0[i], v = b1[i];
            const double ur = u.x, ui = u.y, vr = v.x, vi = v.y;
            typename Vec<R>::type nu, nv;
            nu.x = (R)(g00.x * ur - g00.y * ui + g01.x * vr - g01.y * vi);
            nu.y = (R)(g00.x * ui + g00.y * ur + g01.x * vi + g01.y * vr);
            nv.x = (R)(g10.x * ur - g10.y * ui + g11.x * vr - g11.y * vi);
            nv.y = (R)(g10.x * ui + g10.y * ur + g11.x * vi + g11.y * vr);
            b0[i] = nu;
            b1[i] = nv;
        }
    }
}

__global__ void debug_pairs_kernel(int kind, int n, int e, int a0, int a1, int a2, int a3, uint32_t *o1,
                                   uint32_t *o2, uint32_t *op, uint64_t count) {
    __shared__ uint32_t tab[kBinomWords];
    if (threadIdx.x == 0) fill_binomials(tab);
    __syncthreads();
    uint64_t k = blockIdx.x * 256ull + threadIdx.x;
    if (k >= count) return;
    uint32_t r1 = 0, r2 = 0, par = 0;
    if (kind == 2) ca_pair((uint32_t)k, a1, a0, n, e, tab, &r1, &r2, &par);
    else if (kind == 3) cna_pair((uint32_t)k, a0, a1, a2, n, e, tab, &r1, &r2, &par);
    else ccaa_pair((uint32_t)k, a0, a1, a2, a3, n, e, tab, &r1, &r2, &par);
    o1[k] = r1;
    o2[k] = r2;
    op[k] = par;
}

unsigned grid_for(uint64_t n) {
    uint64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    return (unsigned)(g ? g : 1);
}

template <typename T>
void to_device(T **dst, const std::vector<T> &v, hipStream_t s) {
    size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
    hip_check(hipMalloc(reinterpret_cast<void **>(dst), bytes), "hipMalloc(plan)");
    if (!v.empty()) hip_check(hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "plan upload");
}

}  // namespace

int device_max_lds() {
    int dev = 0, v = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "max LDS");
    return v;
}

void DevicePlan::upload(const Plan &p, int dt, hipStream_t s) {
    release();
    plan = p;
    dtype = dt;
    std::vector<uint32_t> tab(kBinomWords);
    fill_binomials(tab.data());
    to_device(&binom, tab, s);
    to_device(&strings, p.strings, s);
    to_device(&theta_a, p.theta_a, s);
    to_device(&theta_b, p.theta_b, s);
    to_device(&cross, p.cross, s);
    to_device(&groups, p.groups, s);
    to_device(&grp_a, p.grp_a, s);
    to_device(&grp_b, p.grp_b, s);
    std::vector<OpDev> ops_dev = p.ops;
    for (OpDev &o : ops_dev) {
        const uint32_t ns = (uint32_t)std::max(1, std::min(o.ns, kThreadsPlan));
        o.split = ((1u << 20) + ns - 1) / ns | (uint32_t)(kThreadsPlan / ns) << 21;
    }
    to_device(&ops, ops_dev, s);
    to_device(&items, p.items, s);
    to_device(&col_list, p.col_list, s);
    to_device(&sched, p.sched, s);
    to_device(&pass_qr, p.pass_qr, s);
    // coefficients op-major, kMaxGroup slots per op (slot k of an op: term index + k),
    // so that staging an op's coefficients needs no descriptor first; stored in the
    // state's type (complex64: rounded to float here, once)
    std::vector<double> opc(p.ops.size() * kMaxGroup * 2, 0.0);
    for (size_t o = 0; o < p.ops.size(); ++o) {
        const OpDev &od = p.ops[o];
        if (od.type < OP_AAAA) continue;
        for (int k = 0; k < od.g && k < kMaxGroup; ++k) {
            opc[(o * kMaxGroup + k) * 2] = p.coef[2 * ((size_t)od.index + k)];
            opc[(o * kMaxGroup + k) * 2 + 1] = p.coef[2 * ((size_t)od.index + k) + 1];
        }
    }
    if (dt == 1) {
        std::vector<float> opf(opc.size());
        for (size_t k = 0; k < opc.size(); ++k) opf[k] = (float)opc[k];
        float *f = nullptr;
        to_device(&f, opf, s);
        coef = reinterpret_cast<double *>(f);
    } else {
        to_device(&coef, opc, s);
    }
    hip_check(hipMalloc(reinterpret_cast<void **>(&partials), std::max(1, p.max_items) * sizeof(double)), "hipMalloc(partials)");
    lds_bytes = 0;
    pass_lds.clear();
    pass_ops.clear();
    for (const auto &ps : p.passes) {
        // mixed windows: COL passes of FAB ops only, never the first or last of a step, no group angles
        bool fab = ps.kind == PASS_COL && ps.op_end > ps.op_begin && ps.group_count == 0 &&
                   !(ps.flags & (PF_FROM_PREV | PF_EMIT_NORM));
        bool fcc = ps.kind != PASS_COLG && ps.op_end > ps.op_begin;
        for (int o = ps.op_begin; o < ps.op_end; ++o) {
            fab = fab && p.ops[o].type == OP_FAB;
            fcc = fcc && p.ops[o].type == OP_FCC;
        }
        pass_ops.push_back(fab ? kOpsFab : fcc ? kOpsFcc : kOpsAll);
        if (ps.max_lines > kMaxLines) throw std::runtime_error("plan exceeds the per-tile line limit");
        const size_t tile = ps.kind == PASS_COLG ? 0 : (size_t)ps.max_lines * p.C * p.elem_bytes;
        size_t b = kLdsHeader + tile + (size_t)ps.group_count * ps.max_lines * 16;
        pass_lds.push_back(b);
        lds_bytes = std::max(lds_bytes, b);
    }
    // dataflow tables: every pass LDS-tiled, block offsets below 2^31 bytes
    bool eligible = !p.passes.empty() && (uint64_t)p.C * p.C * p.elem_bytes < (1ull << 31);
    for (const auto &ps : p.passes) eligible = eligible && ps.kind != PASS_COLG;
    if (eligible) {
        const Flow fl = build_flow(p);
        flow_items = (uint32_t)fl.item_pass.size();
        flow_norm_items = (uint32_t)p.passes.back().item_count;
        to_device(&flow_item_pass, fl.item_pass, s);
        to_device(&flow_pass_first, fl.pass_first, s);
        to_device(&flow_dep_off, fl.dep_off, s);
        std::vector<uint32_t> deps = fl.deps;
        if (deps.empty()) deps.push_back(0u);
        to_device(&flow_deps, deps, s);
        // the dataflow kernel picks each item's op-set body from pad2
        std::vector<PassDev> fp = p.passes;
        for (size_t k = 0; k < fp.size(); ++k) fp[k].pad2 = pass_ops[k];
        to_device(&flow_passes, fp, s);
        hip_check(hipMalloc(reinterpret_cast<void **>(&flow_done), flow_items * sizeof(uint32_t)), "hipMalloc(flow stamps)");
        hip_check(hipMalloc(reinterpret_cast<void **>(&flow_ctr), 4 * sizeof(uint32_t)), "hipMalloc(flow counters)");
        flow_ok = true;
    }
    hip_check(hipStreamSynchronize(s), "plan upload sync");
}

void DevicePlan::release() {
    void *ptrs[] = {binom, strings, theta_a, theta_b, cross, groups, grp_a, grp_b, ops, items, col_list,
                    partials, sched, pass_qr, coef, flow_item_pass, flow_pass_first, flow_dep_off, flow_deps,
                    flow_passes, flow_done, flow_ctr};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    flow_ok = false;
    flow_items = flow_norm_items = 0;
    flow_item_pass = flow_pass_first = flow_dep_off = flow_deps = nullptr;
    flow_passes = nullptr;
    flow_done = flow_ctr = nullptr;
    binom = nullptr; strings = nullptr; theta_a = theta_b = cross = nullptr; groups = nullptr;
    grp_a = grp_b = nullptr; ops = nullptr; items = nullptr; col_list = nullptr; partials = nullptr;
    sched = nullptr;
    pass_qr = nullptr;
    coef = nullptr;
}

static PassArgs make_args(const DevicePlan &dp, void *state, double *scale, int renorm) {
    const Plan &p = dp.plan;
    PassArgs A;
    A.binom = dp.binom; A.strings = dp.strings;
    A.theta_a = dp.theta_a; A.theta_b = dp.theta_b; A.cross = dp.cross;
    A.groups = dp.groups; A.grp_a = dp.grp_a; A.grp_b = dp.grp_b;
    A.ops = dp.ops; A.items = dp.items; A.col_list = dp.col_list;
    A.sched = dp.sched; A.pass_qr = dp.pass_qr; A.coef = dp.coef;
    A.partials = dp.partials; A.scale = scale;
    A.state = state;
    A.D = (uint64_t)p.C * p.C;
    A.n = p.n; A.e = p.e; A.C = p.C; A.ca_pairs = p.ca_pairs; A.ccaa_pairs = p.ccaa_pairs;
    A.renorm = renorm;
    return A;
}

static int norm_items_of(const Plan &p) {
    int k = 0;
    for (const auto &ps : p.passes)
        if (ps.flags & PF_EMIT_NORM) k = ps.item_count;
    return k;
}

// one step with explicit block pointers: cur = U prev, block index j (norms[j], scale[j];
// scale[j-1] renormalises prev; j also seeds the rounding dither)
template <typename R>
static void evolve_pointers_t(const DevicePlan &dp, void *prev_v, void *cur_v, uint64_t j, double *norms, double *scale,
                              int renorm, int *status_flag, hipStream_t s) {
    using V = typename Vec<R>::type;
    const Plan &p = dp.plan;
    const PassArgs A = make_args(dp, cur_v, scale, renorm);
    V *prev = reinterpret_cast<V *>(prev_v), *cur = reinterpret_cast<V *>(cur_v);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&pass_kernel<R, 1>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dp.lds_bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&pass_kernel<R, 2>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dp.lds_bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&pass_kernel<R, 2, kOpsFab>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dp.lds_bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&pass_kernel<R, 1, kOpsFab>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dp.lds_bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&pass_kernel<R, 1, kOpsFcc>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dp.lds_bytes);
    for (size_t k = 0; k < p.passes.size(); ++k) {
        const PassDev &ps = p.passes[k];
        if (ps.kind == PASS_COLG)   // windows in global memory: FAB ops only
            hipLaunchKernelGGL((pass_kernel<R, 2, kOpsFab>), dim3(ps.item_count), dim3(kThreads), dp.pass_lds[k], s, A,
                               ps, prev, cur, j);
        else if (dp.pass_ops[k] == kOpsFab)
            hipLaunchKernelGGL((pass_kernel<R, 1, kOpsFab>), dim3(ps.item_count), dim3(kThreads), dp.pass_lds[k], s, A,
                               ps, prev, cur, j);
        else if (dp.pass_ops[k] == kOpsFcc)
            hipLaunchKernelGGL((pass_kernel<R, 1, kOpsFcc>), dim3(ps.item_count), dim3(kThreads), dp.pass_lds[k], s, A,
                               ps, prev, cur, j);
        else
            hipLaunchKernelGGL((pass_kernel<R, 1>), dim3(ps.item_count), dim3(kThreads), dp.pass_lds[k], s, A, ps, prev,
                               cur, j);
    }
    hipLaunchKernelGGL(norm_finalize, dim3(1), dim3(1024), 0, s, (const double *)dp.partials, norm_items_of(p), j, norms,
                       scale, status_flag);
    hip_check(hipGetLastError(), "pass launch");
}

template <typename R>
static void evolve_t(const DevicePlan &dp, void *state, uint64_t j_begin, uint64_t j_end, double *norms,
                     double *scale, int renorm, int *status_flag, hipStream_t s, int *launches) {
    using V = typename Vec<R>::type;
    const Plan &p = dp.plan;
    const uint64_t D = (uint64_t)p.C * p.C;
    V *base = reinterpret_cast<V *>(state);
    for (uint64_t j = j_begin; j < j_end; ++j)
        evolve_pointers_t<R>(dp, base + (j - 1) * D, base + j * D, j, norms, scale, renorm, status_flag, s);
    if (launches) *launches = (int)p.passes.size() + 1;
}

// Steps j_begin..j_end-1 as dataflow launches (one launch unless the claims of all
// steps exceed 2^31), preceded each by a reset of the stamps and counters.
template <typename R>
static void evolve_flow_t(const DevicePlan &dp, void *state, uint64_t j_begin, uint64_t j_end, double *norms,
                          double *scale, int renorm, int *status_flag, hipStream_t s, uint32_t wait_ms,
                          int *launches) {
    const PassArgs A = make_args(dp, state, scale, renorm);
    const void *fn = reinterpret_cast<const void *>(&flow_kernel<R>);
    hip_check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dp.lds_bytes), "flow LDS");
    int dev = 0, cus = 0, per_cu = 0;
    hip_check(hipGetDevice(&dev), "device");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kThreads, dp.lds_bytes), "flow occupancy");
    if (per_cu < 1) throw std::runtime_error("dataflow kernel does not fit a compute unit");
    FlowArgs F;
    F.item_pass = dp.flow_item_pass;
    F.pass_first = dp.flow_pass_first;
    F.dep_off = dp.flow_dep_off;
    F.deps = dp.flow_deps;
    F.passes = dp.flow_passes;
    F.done = dp.flow_done;
    F.ctr = dp.flow_ctr;
    F.gctr = dp.flow_ctr;
    F.norms = norms;
    F.scale = scale;
    F.status = status_flag;
    F.ipS = dp.flow_items;
    F.norm_items = dp.flow_norm_items;
    F.wait_ticks = (uint64_t)wait_ms * 100000ull;
    const uint64_t max_steps = std::max<uint64_t>(1, (1ull << 31) / dp.flow_items);
    for (uint64_t j0 = j_begin; j0 < j_end; j0 += max_steps) {
        const uint64_t j1 = std::min(j_end, j0 + max_steps);
        F.j_begin = (uint32_t)j0;
        F.j_end = (uint32_t)j1;
        const uint64_t claims = (j1 - j0) * dp.flow_items;
        const uint32_t grid = (uint32_t)std::min<uint64_t>(claims, (uint64_t)cus * per_cu);
        hipLaunchKernelGGL(flow_reset, dim3((dp.flow_items + 255) / 256), dim3(256), 0, s, dp.flow_done,
                           dp.flow_items, dp.flow_ctr, (uint32_t)(j0 - 1));
        hipLaunchKernelGGL((flow_kernel<R>), dim3(grid), dim3(kThreads), dp.lds_bytes, s, A, F);
    }
    hip_check(hipGetLastError(), "dataflow launch");
    if (launches) *launches = 2;
}

template <typename R>
static void evolve_in_place_t(const DevicePlan &dp, void *state, uint64_t block, uint64_t count, double *norms,
                              double *scale, int *status_flag, hipStream_t s, int *launches) {
    using V = typename Vec<R>::type;
    const Plan &p = dp.plan;
    const PassArgs A = make_args(dp, state, scale, 0);
    V *cur = reinterpret_cast<V *>(state) + block * A.D;
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&pass_kernel<R, 1>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dp.lds_bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&pass_kernel<R, 2>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dp.lds_bytes);
    const int norm_items = norm_items_of(p);
    for (uint64_t k = 0; k < count; ++k) {
        // the step index only seeds the rounding dither (no renormalisation reads scale[])
        for (size_t q = 0; q < p.passes.size(); ++q)
            if (p.passes[q].kind == PASS_COLG)
                hipLaunchKernelGGL((pass_kernel<R, 2>), dim3(p.passes[q].item_count), dim3(kThreads), dp.pass_lds[q],
                                   s, A, p.passes[q], cur, cur, k + 1);
            else
                hipLaunchKernelGGL((pass_kernel<R, 1>), dim3(p.passes[q].item_count), dim3(kThreads), dp.pass_lds[q],
                                   s, A, p.passes[q], cur, cur, k + 1);
        hipLaunchKernelGGL(norm_finalize, dim3(1), dim3(1024), 0, s, (const double *)dp.partials, norm_items, block,
                           norms, scale, status_flag);
    }
    hip_check(hipGetLastError(), "pass launch");
    if (launches) *launches = (int)p.passes.size() + 1;
}

// ---- batched evolution ---------------------------------------------------------

void BatchRun::build(const std::vector<BatchMember> &members, uint64_t M_, hipStream_t s, bool use_flow) {
    release();
    if (members.empty()) throw std::runtime_error("empty batch");
    count = (int)members.size();
    dtype = members[0].dp->dtype;
    M = M_;
    std::vector<PassArgs> args;
    std::vector<FinalizeArgs> fin;
    size_t maxp = 0;
    lds_bytes = 0;
    for (const auto &m : members) {
        if (m.dp->dtype != dtype) throw std::runtime_error("batch members must share the amplitude type");
        args.push_back(make_args(*m.dp, m.state, m.scale, m.renorm));
        FinalizeArgs f;
        f.partials = m.dp->partials;
        f.count = norm_items_of(m.dp->plan);
        f.pad = 0;
        f.norms = m.norms;
        f.scale = m.scale;
        f.status = m.status;
        fin.push_back(f);
        maxp = std::max(maxp, m.dp->plan.passes.size());
        lds_bytes = std::max(lds_bytes, m.dp->lds_bytes);
    }
    // launch k of a step: pass k of every member that has one, larger grids first
    std::vector<BatchWork> work;
    launches.clear();
    for (size_t k = 0; k < maxp; ++k) {
        std::vector<int> order;
        for (int i = 0; i < count; ++i)
            if (k < members[i].dp->plan.passes.size()) order.push_back(i);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return members[a].dp->plan.passes[k].item_count > members[b].dp->plan.passes[k].item_count;
        });
        Launch L;
        L.work_off = work.size();
        L.nwork = (int)order.size();
        L.blocks = 0;
        L.lds = 0;
        for (int i : order) {
            L.lds = std::max(L.lds, members[i].dp->pass_lds[k]);
            BatchWork w;
            w.inst = i;
            w.block_begin = L.blocks;
            w.pad0 = w.pad1 = 0;
            w.pass = members[i].dp->plan.passes[k];
            L.blocks += (uint32_t)w.pass.item_count;
            work.push_back(w);
        }
        launches.push_back(L);
    }
    PassArgs *a_dev = nullptr;
    to_device(&a_dev, args, s);
    this->args = a_dev;
    to_device(&this->work, work, s);
    to_device(&this->fin, fin, s);
    hip_check(hipMalloc(reinterpret_cast<void **>(&jstep), sizeof(uint64_t)), "hipMalloc(step counter)");
    hip_check(hipMalloc(reinterpret_cast<void **>(&done), sizeof(unsigned)), "hipMalloc(done counter)");
    // dataflow form: every member's plan has its dependency tables
    uint64_t per_step = 0;
    bool eligible = use_flow;
    for (const auto &m : members) {
        eligible = eligible && m.dp->flow_ok;
        per_step += m.dp->flow_items;
    }
    if (eligible && per_step > 0 && (M - 1) * per_step < (1ull << 31)) {
        // one step's claims: member i's items at positions (x + 1/2) / items_i, merged in
        // position order (ties by member), so every member advances through its step at
        // the same relative pace
        std::vector<std::pair<double, uint64_t>> pos;
        pos.reserve(per_step);
        for (int i = 0; i < count; ++i) {
            const uint32_t n = members[i].dp->flow_items;
            for (uint32_t x = 0; x < n; ++x) pos.push_back({(x + 0.5) / n, ((uint64_t)i << 32) | x});
        }
        std::stable_sort(pos.begin(), pos.end(),
                         [](const std::pair<double, uint64_t> &a, const std::pair<double, uint64_t> &b) {
                             return a.first < b.first;
                         });
        std::vector<uint32_t> ord;
        ord.reserve(2 * per_step);
        for (const auto &p : pos) {
            ord.push_back((uint32_t)(p.second >> 32));
            ord.push_back((uint32_t)p.second);
        }
        std::vector<FlowArgs> fa;
        for (int i = 0; i < count; ++i) {
            const DevicePlan &dp = *members[i].dp;
            FlowArgs F;
            F.item_pass = dp.flow_item_pass;
            F.pass_first = dp.flow_pass_first;
            F.dep_off = dp.flow_dep_off;
            F.deps = dp.flow_deps;
            F.passes = dp.flow_passes;
            F.done = dp.flow_done;
            F.ctr = dp.flow_ctr;
            F.gctr = nullptr;   // set below
            F.norms = members[i].norms;
            F.scale = members[i].scale;
            F.status = members[i].status;
            F.j_begin = 1;
            F.j_end = (uint32_t)M;
            F.ipS = dp.flow_items;
            F.norm_items = dp.flow_norm_items;
            F.wait_ticks = 10000ull * 100000ull;   // 10 s
            fa.push_back(F);
            flow_plans.push_back(&dp);
        }
        hip_check(hipMalloc(reinterpret_cast<void **>(&flow_gctr), 4 * sizeof(uint32_t)), "hipMalloc(batch counters)");
        for (auto &F : fa) F.gctr = flow_gctr;
        FlowArgs *fd = nullptr;
        to_device(&fd, fa, s);
        flow_args = fd;
        uint32_t *od = nullptr;
        to_device(&od, ord, s);
        flow_order = od;
        flow_per_step = (uint32_t)per_step;
        flow = true;
    }
    hip_check(hipStreamSynchronize(s), "batch upload sync");
}

// the dataflow form's launch sequence: every member's stamps and counters reset, the
// launch counters reset, one persistent launch for steps 1..M-1
template <typename R>
static void batch_flow_launch(const BatchRun &b, hipStream_t s) {
    for (const DevicePlan *dp : b.flow_plans)
        hipLaunchKernelGGL(flow_reset, dim3((dp->flow_items + 255) / 256), dim3(256), 0, s, dp->flow_done,
                           dp->flow_items, dp->flow_ctr, 0u);
    hipLaunchKernelGGL(flow_reset, dim3(1), dim3(64), 0, s, (uint32_t *)nullptr, 0u, b.flow_gctr, 0u);
    const void *fn = reinterpret_cast<const void *>(&flow_batch_kernel<R>);
    int dev = 0, cus = 0, per_cu = 0;
    hip_check(hipGetDevice(&dev), "device");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kThreads, b.lds_bytes), "flow occupancy");
    if (per_cu < 1) throw std::runtime_error("dataflow kernel does not fit a compute unit");
    const uint64_t claims = (b.M - 1) * b.flow_per_step;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(claims, (uint64_t)cus * per_cu);
    hipLaunchKernelGGL((flow_batch_kernel<R>), dim3(grid), dim3(kThreads), b.lds_bytes, s,
                       reinterpret_cast<const PassArgs *>(b.args), reinterpret_cast<const FlowArgs *>(b.flow_args),
                       reinterpret_cast<const uint2 *>(b.flow_order), b.flow_per_step, (uint32_t)(b.M - 1), b.flow_gctr);
}

template <typename R>
static void batch_step(const BatchRun &b, hipStream_t s) {
    for (const auto &L : b.launches) {
        if (L.blocks == 0) continue;
        hipLaunchKernelGGL(batch_pass_kernel<R>, dim3(L.blocks), dim3(kThreads), L.lds, s,
                           reinterpret_cast<const PassArgs *>(b.args), (const BatchWork *)(b.work + L.work_off),
                           L.nwork, (const uint64_t *)b.jstep);
    }
    hipLaunchKernelGGL(batch_finalize, dim3(b.count), dim3(1024), 0, s, (const FinalizeArgs *)b.fin, b.count, b.jstep,
                       b.done);
}

void BatchRun::run(hipStream_t s, bool use_graph) {
    if (flow) {
        const void *fn = dtype == 0 ? reinterpret_cast<const void *>(&flow_batch_kernel<double>)
                                    : reinterpret_cast<const void *>(&flow_batch_kernel<float>);
        hip_check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes), "flow LDS");
        if (!use_graph) {
            if (dtype == 0) batch_flow_launch<double>(*this, s);
            else batch_flow_launch<float>(*this, s);
            hip_check(hipGetLastError(), "batch dataflow launch");
            return;
        }
        instantiate();
        hip_check(hipGraphLaunch(flow_graph, s), "graph launch");
        return;
    }
    hipLaunchKernelGGL(batch_reset, dim3(1), dim3(64), 0, s, jstep, done);
    if (dtype == 0)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&batch_pass_kernel<double>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    else
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&batch_pass_kernel<float>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (!use_graph) {
        for (uint64_t j = 1; j < M; ++j) {
            if (dtype == 0) batch_step<double>(*this, s);
            else batch_step<float>(*this, s);
        }
        hip_check(hipGetLastError(), "batch launch");
        return;
    }
    instantiate();
    for (uint64_t r = 0; r < graph_reps; ++r) hip_check(hipGraphLaunch(graph, s), "graph launch");
    if (tail) hip_check(hipGraphLaunch(tail, s), "graph launch");
}

// steps kernels of the chain captured into one executable graph
template <typename R>
static hipGraphExec_t capture_steps(const BatchRun &b, uint64_t steps) {
    hipStream_t cs;
    hip_check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "capture stream");
    hip_check(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture");
    for (uint64_t j = 0; j < steps; ++j) batch_step<R>(b, cs);
    hipGraph_t g = nullptr;
    hipError_t e1 = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    hip_check(e1, "end capture");
    hipGraphExec_t exec = nullptr;
    hipError_t e2 = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    hip_check(e2, "graph instantiate");
    return exec;
}

void BatchRun::instantiate() {
    if (flow) {
        if (flow_graph) return;
        const void *fn = dtype == 0 ? reinterpret_cast<const void *>(&flow_batch_kernel<double>)
                                    : reinterpret_cast<const void *>(&flow_batch_kernel<float>);
        hip_check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes), "flow LDS");
        hipStream_t cs;
        hip_check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "capture stream");
        hip_check(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture");
        if (dtype == 0) batch_flow_launch<double>(*this, cs);
        else batch_flow_launch<float>(*this, cs);
        hipGraph_t g = nullptr;
        hipError_t e1 = hipStreamEndCapture(cs, &g);
        (void)hipStreamDestroy(cs);
        hip_check(e1, "end capture");
        hipError_t e2 = hipGraphInstantiate(&flow_graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(e2, "graph instantiate");
        return;
    }
    if (graph || tail) return;
    if (dtype == 0)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&batch_pass_kernel<double>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    else
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&batch_pass_kernel<float>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    // each step's kernels read the step from *jstep, so one graph of S steps replayed
    // (M-1)/S times and a graph of the remaining steps run the whole chain
    uint64_t per_step = 1;   // the finalize
    for (const auto &L : launches)
        if (L.blocks) ++per_step;
    const uint64_t total = M - 1;
    uint64_t S = std::max<uint64_t>(1, kMaxGraphLaunches / per_step);
    if (S > total) S = total;
    graph_steps = S;
    graph_reps = S ? total / S : 0;
    tail_steps = S ? total - graph_reps * S : 0;
    if (graph_reps)
        graph = dtype == 0 ? capture_steps<double>(*this, S) : capture_steps<float>(*this, S);
    if (tail_steps)
        tail = dtype == 0 ? capture_steps<double>(*this, tail_steps) : capture_steps<float>(*this, tail_steps);
}

void BatchRun::release() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (tail) (void)hipGraphExecDestroy(tail);
    graph = tail = nullptr;
    graph_steps = graph_reps = tail_steps = 0;
    if (flow_graph) (void)hipGraphExecDestroy(flow_graph);
    flow_graph = nullptr;
    void *ptrs[] = {args, work, fin, jstep, done, flow_args, flow_order, flow_gctr};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    args = nullptr; work = nullptr; fin = nullptr; jstep = nullptr; done = nullptr;
    flow_args = nullptr; flow_order = nullptr; flow_gctr = nullptr;
    flow = false;
    flow_per_step = 0;
    flow_plans.clear();
    launches.clear();
}
void evolve_in_place(const DevicePlan &dp, void *state, uint64_t block, uint64_t count, double *norms, double *scale,
                     int *status_flag, hipStream_t s, int *launches) {
    if (dp.dtype == 0) evolve_in_place_t<double>(dp, state, block, count, norms, scale, status_flag, s, launches);
    else evolve_in_place_t<float>(dp, state, block, count, norms, scale, status_flag, s, launches);
}

void evolve(const DevicePlan &dp, void *state, uint64_t j_begin, uint64_t j_end, double *norms, double *scale,
            int renorm, int *status_flag, hipStream_t s, int *launches, int execution, uint32_t flow_wait_ms) {
    if (execution == EXEC_DATAFLOW && !dp.flow_ok)
        throw std::runtime_error("the dataflow evolution needs LDS-tiled passes and blocks below 2^31 bytes");
    // auto: the dataflow kernel for complex128, whose tiles are half as wide (twice the
    // items per pass, so more passes overlap); complex64 measured faster pass by pass
    // (32 spin-orbitals / 8 electrons: 53.4 against 62.9 ms per complex128 step, 32.9
    // against 32.6 ms per complex64 step; 28 spin-orbitals / 10 electrons: complex128 43.0
    // against 45.9 ms, complex64 28.1 against 25.8 ms)
    const bool flow = execution == EXEC_DATAFLOW || (execution == EXEC_AUTO && dp.dtype == 0);
    if (flow && dp.flow_ok) {
        if (dp.dtype == 0)
            evolve_flow_t<double>(dp, state, j_begin, j_end, norms, scale, renorm, status_flag, s, flow_wait_ms, launches);
        else
            evolve_flow_t<float>(dp, state, j_begin, j_end, norms, scale, renorm, status_flag, s, flow_wait_ms, launches);
        return;
    }
    if (dp.dtype == 0) evolve_t<double>(dp, state, j_begin, j_end, norms, scale, renorm, status_flag, s, launches);
    else evolve_t<float>(dp, state, j_begin, j_end, norms, scale, renorm, status_flag, s, launches);
}

void block_norm(int dtype, uint64_t D, void *state, uint64_t j, double *norms, double *scale, double *partials,
                int nparts, hipStream_t s) {
    if (dtype == 0)
        hipLaunchKernelGGL(block_norm_kernel<double>, dim3(nparts), dim3(256), 0, s,
                           reinterpret_cast<const double2 *>(state) + j * D, D, partials);
    else
        hipLaunchKernelGGL(block_norm_kernel<float>, dim3(nparts), dim3(256), 0, s,
                           reinterpret_cast<const float2 *>(state) + j * D, D, partials);
    hipLaunchKernelGGL(norm_finalize, dim3(1), dim3(1024), 0, s, (const double *)partials, nparts, j, norms, scale,
                       (int *)nullptr);
    hip_check(hipGetLastError(), "block norm launch");
}

// one launch of the register-form kernel with `groups` groups; max_rows caps the rows of
// P(k) partial sums when the map emits them (*rows_out: rows written)
template <typename R, int LOGM, bool SPLIT>
static bool launch_qift_reg(const void *src, void *dst, uint64_t D, const QiftMap &mp, uint32_t groups,
                            const double *scale, const double *twiddles, double *pk_partial, int max_rows,
                            size_t max_lds, int *rows_out, hipStream_t s) {
    using Q = QiftShape<R, LOGM>;
    const size_t lds = Q::lds_bytes();
    if (lds > max_lds) return false;
    const uint64_t ntiles = (D + Q::W - 1) / Q::W;
    // about 1024 workgroups in all, at most max_rows along x when P(k) rows are written
    uint64_t gx = std::max<uint64_t>(1, 1024 / groups);
    if (mp.emit_pk) gx = std::min<uint64_t>(gx, (uint64_t)max_rows);
    gx = std::min<uint64_t>(gx, ntiles);
    if (gx >= 8) gx &= ~7ull;
    if (gx < 1) throw std::runtime_error("no workgroup fits the P(k) partial sums");
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&qift_reg_kernel<R, LOGM, SPLIT>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((qift_reg_kernel<R, LOGM, SPLIT>), dim3((unsigned)gx, groups), dim3(Q::T), lds, s,
                       reinterpret_cast<const typename Vec<R>::type *>(src),
                       reinterpret_cast<typename Vec<R>::type *>(dst), D, mp, scale,
                       reinterpret_cast<const double2 *>(twiddles), pk_partial);
    hip_check(hipGetLastError(), "qift launch");
    if (rows_out) *rows_out = (int)gx;
    return true;
}

template <typename R, bool SPLIT>
static bool qift_reg_launch(int logr, const void *src, void *dst, uint64_t D, const QiftMap &mp, uint32_t groups,
                            const double *scale, const double *twiddles, double *pk_partial, int max_rows,
                            size_t max_lds, int *rows_out, hipStream_t s) {
#define QPE_QIFT_REG(L)                                                                                         \
    case L:                                                                                                     \
        return launch_qift_reg<R, L, SPLIT>(src, dst, D, mp, groups, scale, twiddles, pk_partial, max_rows, max_lds, \
                                            rows_out, s);
    if constexpr (SPLIT) {
        // the halves of 2^11 .. 2^20
        switch (logr) {
            QPE_QIFT_REG(5) QPE_QIFT_REG(6) QPE_QIFT_REG(7) QPE_QIFT_REG(8) QPE_QIFT_REG(9) QPE_QIFT_REG(10)
        default: return false;
        }
    } else {
        switch (logr) {
            QPE_QIFT_REG(1) QPE_QIFT_REG(2) QPE_QIFT_REG(3) QPE_QIFT_REG(4) QPE_QIFT_REG(5)
            QPE_QIFT_REG(6) QPE_QIFT_REG(7) QPE_QIFT_REG(8) QPE_QIFT_REG(9) QPE_QIFT_REG(10)
        default: return false;
        }
    }
#undef QPE_QIFT_REG
}

// The register-form QIFT: one launch for 2^m <= 1024; for 2^m up to 2^20 the four-step
// split 2^m = R1 * R2 (R1 = 2^floor(m/2)) as two launches: A, the R2-point DFTs of the rows
// j1 + R1 j2 in place, times e^{-2 pi i j1 k2 / 2^m}; B, the R1-point DFTs of the rows
// j1 + R1 k2 into `work` (a second register) in natural order, which is then copied back.
template <typename R>
static bool qift_reg(int m, void *state, void *work, uint64_t D, const double *scale, const double *twiddles,
                     double *pk_partial, int max_rows, size_t max_lds, int *rows_out, hipStream_t s) {
    const uint64_t M = 1ull << m;
    QiftMap mp{};
    mp.Mt = (uint32_t)M;
    mp.f = 1.0;
    if (m <= 10) {
        mp.gi = 0; mp.ji = 1; mp.go = 0; mp.ko = 1; mp.gk = 0; mp.kk = 1;
        mp.tw_step = 1;
        mp.use_scale = 1;
        mp.emit_pk = 1;
        mp.f = 1.0 / std::sqrt((double)M);
        return qift_reg_launch<R, false>(m, state, state, D, mp, 1, scale, twiddles, pk_partial, max_rows, max_lds,
                                         rows_out, s);
    }
    if (m > 20) return false;
    if (!work) throw std::runtime_error("the QIFT of 2^m > 1024 needs a second register's worth of device memory");
    const int mb = m / 2, ma = m - mb;   // R1 = 2^mb groups of R2 = 2^ma points, then R2 groups of R1 points
    const uint64_t R1 = 1ull << mb, R2 = 1ull << ma;
    QiftMap a{};
    a.gi = 1; a.ji = R1; a.go = 1; a.ko = R1;
    a.Mt = (uint32_t)M;
    a.tw_step = (uint32_t)(M / R2);
    a.group_twiddle = 1;
    a.use_scale = 1;
    a.f = 1.0;
    if (!qift_reg_launch<R, true>(ma, state, state, D, a, (uint32_t)R1, scale, twiddles, nullptr, max_rows, max_lds,
                                  nullptr, s))
        return false;
    QiftMap b{};
    b.gi = R1; b.ji = 1; b.go = 1; b.ko = R2; b.gk = 1; b.kk = R2;
    b.Mt = (uint32_t)M;
    b.tw_step = (uint32_t)(M / R1);
    b.emit_pk = 1;
    b.f = 1.0 / std::sqrt((double)M);
    if (!qift_reg_launch<R, true>(mb, state, work, D, b, (uint32_t)R2, scale, twiddles, pk_partial, max_rows, max_lds,
                                  rows_out, s))
        throw std::runtime_error("QIFT second step did not launch");
    hip_check(hipMemcpyAsync(state, work, M * D * sizeof(typename Vec<R>::type), hipMemcpyDeviceToDevice, s),
              "QIFT copy back");
    return true;
}

template <typename R, int LOGM>
static bool launch_qift_mfma(void *state, uint64_t D, const double *scale, const double *twiddles, double *pk_partial,
                             int max_grid, size_t max_lds, int *grid_out, hipStream_t s) {
    using S = MfmaShape<LOGM>;
    const size_t lds = S::template lds_bytes<R>();
    if (lds > max_lds) return false;
    const uint64_t nblk = (D + 15) / 16;
    // one row of partial sums per wave (4 per workgroup)
    int grid = (int)std::min<uint64_t>((nblk + 3) / 4, (uint64_t)std::min(1024, max_grid) / 4);
    if (grid < 1) throw std::runtime_error("no workgroup fits the P(k) partial sums");
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&qift_mfma_kernel<R, LOGM>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((qift_mfma_kernel<R, LOGM>), dim3(grid), dim3(S::T), lds, s,
                       reinterpret_cast<typename Vec<R>::type *>(state), D, scale,
                       reinterpret_cast<const double2 *>(twiddles), pk_partial);
    hip_check(hipGetLastError(), "qift launch");
    *grid_out = grid * 4;
    return true;
}

template <typename R>
static bool qift_mfma(int m, void *state, uint64_t D, const double *scale, const double *twiddles, double *pk_partial,
                      int max_grid, size_t max_lds, int *grid_out, hipStream_t s) {
    switch (m) {
    case 4: return launch_qift_mfma<R, 4>(state, D, scale, twiddles, pk_partial, max_grid, max_lds, grid_out, s);
    case 5: return launch_qift_mfma<R, 5>(state, D, scale, twiddles, pk_partial, max_grid, max_lds, grid_out, s);
    case 6: return launch_qift_mfma<R, 6>(state, D, scale, twiddles, pk_partial, max_grid, max_lds, grid_out, s);
    default: return false;
    }
}

bool qift_form_available(int form, int m) {
    if (form == QIFT_MFMA) return m >= 4 && m <= 6;
    if (form == QIFT_REGISTER) return m >= 1 && m <= 20;
    if (form == QIFT_LDS_RADIX2) return m >= 1 && m <= 12;
    return true;
}

bool qift_needs_work(int form, int m) { return form != QIFT_MFMA && form != QIFT_LDS_RADIX2 && m > 10; }

void qift(int dtype, void *state, int m, uint64_t D, const double *scale, double *pk_partial, int *grid_out,
          double *twiddles, size_t max_lds, hipStream_t s, int max_grid, int form, void *work) {
    if (form == QIFT_MFMA &&
        (dtype == 0 ? qift_mfma<double>(m, state, D, scale, twiddles, pk_partial, max_grid, max_lds, grid_out, s)
                    : qift_mfma<float>(m, state, D, scale, twiddles, pk_partial, max_grid, max_lds, grid_out, s)))
        return;
    // the register form (2^m <= 2^20); radix-2 stages in LDS when asked for
    if (form != QIFT_LDS_RADIX2 &&
        (dtype == 0
             ? qift_reg<double>(m, state, work, D, scale, twiddles, pk_partial, max_grid, max_lds, grid_out, s)
             : qift_reg<float>(m, state, work, D, scale, twiddles, pk_partial, max_grid, max_lds, grid_out, s)))
        return;
    const uint32_t M = 1u << m;
    const size_t tw_bytes = (size_t)std::max<uint32_t>(1, M / 2) * sizeof(double2);
    int w = 1;
    // widest column tile that fits LDS and the accumulator budget, capped at 32
    while (w * 2 <= 32 && (size_t)M * (w * 2) <= (size_t)kQThreads * kQAcc &&
           tw_bytes + (size_t)M * (w * 2) * sizeof(double2) <= max_lds)
        w *= 2;
    if (tw_bytes + (size_t)M * w * sizeof(double2) > max_lds || (size_t)M * w > (size_t)kQThreads * kQAcc)
        throw std::runtime_error("2^m too large for the QIFT tile");
    const uint64_t ntiles = (D + w - 1) / w;
    int grid = (int)std::min<uint64_t>(ntiles, (uint64_t)std::min(1024, max_grid));
    if (grid < 1) throw std::runtime_error("no workgroup fits the P(k) partial sums");
    size_t lds = tw_bytes + (size_t)M * w * sizeof(double2);
    *grid_out = grid;
    if (dtype == 0) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&qift_kernel<double>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(qift_kernel<double>, dim3(grid), dim3(kQThreads), lds, s, reinterpret_cast<double2 *>(state),
                           D, m, w, scale, reinterpret_cast<const double2 *>(twiddles), pk_partial);
    } else {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&qift_kernel<float>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(qift_kernel<float>, dim3(grid), dim3(kQThreads), lds, s, reinterpret_cast<float2 *>(state),
                           D, m, w, scale, reinterpret_cast<const double2 *>(twiddles), pk_partial);
    }
    hip_check(hipGetLastError(), "qift launch");
}

// ---- allocation-free QIFT + P(k) ---------------------------------------------
// The per-workgroup P(k) partial sums of qift_noalloc live in this file-scope
// array (one copy per device, held until the process ends), so that entry point
// allocates nothing.  Host code never passes the array's name to a kernel: it
// uses the device address hipGetSymbolAddress returns, looked up once per device.
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
__global__ void fill_kernel(double *p, uint64_t n, double v) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256ull) p[i] = v;
}

std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxScratchDevices] = {};

// device address of `scratch` on the current device
double *scratch_address() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxScratchDevices) throw std::runtime_error("device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    return g_scratch_ptr[dev];
}
}  // namespace

void qift_noalloc(int dtype, void *state, int m, uint64_t D, double *scale, double *twiddles, size_t max_lds,
                  double *probs_dev, hipStream_t s, int form) {
    const uint64_t M = 1ull << m;
    if (M > kNoAllocScratchDoubles)
        throw std::runtime_error("2^m above the allocation-free QIFT's scratch (" +
                                 std::to_string(kNoAllocScratchDoubles) + " P(k) partial sums)");
    if (qift_needs_work(form, m))
        throw std::runtime_error("the allocation-free QIFT covers 2^m <= 1024 (larger ones need a second register)");
    double *part = scratch_address();
    int grid = 0;
    // every workgroup writes all 2^m of its partial sums, so the scratch needs no clearing
    qift(dtype, state, m, D, scale, part, &grid, twiddles, max_lds, s, (int)(kNoAllocScratchDoubles / M), form);
    reduce_pk(part, grid, m, probs_dev, s);
    // the QIFT applied the renormalisation factors: reset them to 1 (a fill kernel, no memset)
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)std::min<uint64_t>((M + 255) / 256, 1024)), dim3(256), 0, s, scale,
                       M, 1.0);
    hip_check(hipGetLastError(), "scale reset launch");
}

void reduce_pk(const double *pk_partial, int grid, int m, double *probs_dev, hipStream_t s) {
    const uint32_t M = 1u << m;
    hipLaunchKernelGGL(pk_reduce, dim3((M + 255) / 256), dim3(256), 0, s, pk_partial, grid, M, probs_dev);
    hip_check(hipGetLastError(), "pk reduce launch");
}

void scatter_in(int dtype, void *state, const double *staging, uint64_t offset, uint64_t count, uint32_t C,
                hipStream_t s) {
    if (dtype == 0)
        hipLaunchKernelGGL(scatter_kernel<double>, dim3(grid_for(count)), dim3(256), 0, s,
                           reinterpret_cast<double2 *>(state), reinterpret_cast<const double2 *>(staging), offset, count, C);
    else
        hipLaunchKernelGGL(scatter_kernel<float>, dim3(grid_for(count)), dim3(256), 0, s,
                           reinterpret_cast<float2 *>(state), reinterpret_cast<const double2 *>(staging), offset, count, C);
    hip_check(hipGetLastError(), "scatter launch");
}

void gather_out(int dtype, const void *state, double *staging, uint64_t offset, uint64_t count, uint32_t C,
                hipStream_t s) {
    if (dtype == 0)
        hipLaunchKernelGGL(gather_kernel<double>, dim3(grid_for(count)), dim3(256), 0, s,
                           reinterpret_cast<const double2 *>(state), reinterpret_cast<double2 *>(staging), offset, count, C);
    else
        hipLaunchKernelGGL(gather_kernel<float>, dim3(grid_for(count)), dim3(256), 0, s,
                           reinterpret_cast<const float2 *>(state), reinterpret_cast<double2 *>(staging), offset, count, C);
    hip_check(hipGetLastError(), "gather launch");
}

void single_string_copy(int dtype, void *state, double *staging, uint64_t offset, uint64_t count, bool to_state,
                        hipStream_t s) {
    if (dtype == 0) {
        if (to_state)
            hipLaunchKernelGGL(copy_in_kernel<double>, dim3(grid_for(count)), dim3(256), 0, s,
                               reinterpret_cast<double2 *>(state), reinterpret_cast<const double2 *>(staging), offset, count);
        else
            hipLaunchKernelGGL(copy_out_kernel<double>, dim3(grid_for(count)), dim3(256), 0, s,
                               reinterpret_cast<const double2 *>(state), reinterpret_cast<double2 *>(staging), offset, count);
    } else {
        if (to_state)
            hipLaunchKernelGGL(copy_in_kernel<float>, dim3(grid_for(count)), dim3(256), 0, s,
                               reinterpret_cast<float2 *>(state), reinterpret_cast<const double2 *>(staging), offset, count);
        else
            hipLaunchKernelGGL(copy_out_kernel<float>, dim3(grid_for(count)), dim3(256), 0, s,
                               reinterpret_cast<const float2 *>(state), reinterpret_cast<double2 *>(staging), offset, count);
    }
    hip_check(hipGetLastError(), "single-string copy launch");
}

void single_string_convert(const DevicePlan &dp, void *ab, void *k1, bool to_ab, hipStream_t s) {
    const Plan &p = dp.plan;
    const uint64_t D = (uint64_t)p.C * p.C;
    if (dp.dtype == 0)
        hipLaunchKernelGGL(layout_kernel<double>, dim3(grid_for(D)), dim3(256), 0, s, reinterpret_cast<double2 *>(ab),
                           reinterpret_cast<double2 *>(k1), (const uint32_t *)dp.strings, (const uint32_t *)dp.binom, p.C,
                           p.n, to_ab ? 0 : 1);
    else
        hipLaunchKernelGGL(layout_kernel<float>, dim3(grid_for(D)), dim3(256), 0, s, reinterpret_cast<float2 *>(ab),
                           reinterpret_cast<float2 *>(k1), (const uint32_t *)dp.strings, (const uint32_t *)dp.binom, p.C,
                           p.n, to_ab ? 0 : 1);
    hip_check(hipGetLastError(), "layout launch");
}

void single_string_outside(const DevicePlan &dp, const void *k1, uint64_t D1, int *flag, hipStream_t s) {
    const Plan &p = dp.plan;
    if (dp.dtype == 0)
        hipLaunchKernelGGL(outside_sector_kernel<double>, dim3(grid_for(D1)), dim3(256), 0, s,
                           reinterpret_cast<const double2 *>(k1), D1, (const uint32_t *)dp.binom, p.n, p.e, flag);
    else
        hipLaunchKernelGGL(outside_sector_kernel<float>, dim3(grid_for(D1)), dim3(256), 0, s,
                           reinterpret_cast<const float2 *>(k1), D1, (const uint32_t *)dp.binom, p.n, p.e, flag);
    hip_check(hipGetLastError(), "sector check launch");
}

void evolve_step(const DevicePlan &dp, void *prev, void *cur, uint64_t j, double *norms, double *scale, int renorm,
                 int *status_flag, hipStream_t s) {
    if (dp.dtype == 0) evolve_pointers_t<double>(dp, prev, cur, j, norms, scale, renorm, status_flag, s);
    else evolve_pointers_t<float>(dp, prev, cur, j, norms, scale, renorm, status_flag, s);
}

void scale_block(int dtype, void *state, uint64_t j, uint64_t D, double factor, hipStream_t s) {
    if (dtype == 0)
        hipLaunchKernelGGL(scale_kernel<double>, dim3(grid_for(D)), dim3(256), 0, s,
                           reinterpret_cast<double2 *>(state) + j * D, D, factor);
    else
        hipLaunchKernelGGL(scale_kernel<float>, dim3(grid_for(D)), dim3(256), 0, s,
                           reinterpret_cast<float2 *>(state) + j * D, D, factor);
    hip_check(hipGetLastError(), "scale launch");
}

void apply_gate(int dtype, void *state, uint64_t D, int m, int tbit, int cbit, const double *g, hipStream_t s) {
    const uint64_t npairs = (1ull << m) >> (cbit < 0 ? 1 : 2);
    const unsigned gx = (unsigned)std::min<uint64_t>((D + 255) / 256, 1024);
    const unsigned gy = (unsigned)std::min<uint64_t>(npairs, std::max<uint64_t>(1, 4096 / gx));
    const double2 g00 = make_double2(g[0], g[1]), g01 = make_double2(g[2], g[3]);
    const double2 g10 = make_double2(g[4], g[5]), g11 = make_double2(g[6], g[7]);
    if (dtype == 0)
        hipLaunchKernelGGL(gate_kernel<double>, dim3(gx, gy), dim3(256), 0, s, reinterpret_cast<double2 *>(state), D,
                           npairs, tbit, cbit, g00, g01, g10, g11);
    else
        hipLaunchKernelGGL(gate_kernel<float>, dim3(gx, gy), dim3(256), 0, s, reinterpret_cast<float2 *>(state), D,
                           npairs, tbit, cbit, g00, g01, g10, g11);
    hip_check(hipGetLastError(), "gate launch");
}

void debug_pairs(int kind, int n, int e, const int32_t *args, uint32_t *o1, uint32_t *o2, uint32_t *op,
                 uint64_t count, hipStream_t s) {
    hipLaunchKernelGGL(debug_pairs_kernel, dim3(grid_for(count)), dim3(256), 0, s, kind, n, e, args[0], args[1],
                       args[2], args[3], o1, o2, op, count);
    hip_check(hipGetLastError(), "debug pairs launch");
}

}  // namespace qpe
