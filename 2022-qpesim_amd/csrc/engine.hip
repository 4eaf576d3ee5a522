// gfx950 kernels of the QPE hot path and their launchers.
//
//   pass_kernel   one tile pass of a Trotter step (COL or ROW tiles in LDS)
//   norm_finalize FP64 compensated block norm + renormalisation factor
//   qift_kernel   batched radix-2 DFT along the ancilla axis in LDS, with the
//                 1/sqrt(M) normalisation and the P(k) partial sums fused
//
// Amplitudes: complex128 (double2) or complex64 (float2).  Angles, their
// cos/sin, norms and P(k) are always FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "combinadic.hpp"
#include "engine.hpp"

namespace qpe {

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;
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

// 2x2 symmetric rotation {c, i*mu; i*mu, c} on the pair (u, v)
template <typename V, typename R>
__device__ inline void rotate(V &u, V &v, R c, R mu) {
    V nu, nv;
    nu.x = c * u.x - mu * v.y;
    nu.y = c * u.y + mu * v.x;
    nv.x = c * v.x - mu * u.y;
    nv.y = c * v.y + mu * u.x;
    u = nu;
    v = nv;
}

struct PassArgs {
    const uint32_t *binom;
    const uint32_t *strings;
    const double *theta_a, *theta_b, *cross;
    const GroupDev *groups;
    const double *grp_a, *grp_b;
    const TermDev *terms;
    const OpDev *ops;
    const ItemDev *items;
    const uint32_t *col_list;
    const uint32_t *ca_table, *ccaa_table;
    double *partials;
    const double *scale;
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

__device__ inline uint32_t rank1(uint32_t e) { return e & kRankMask; }
__device__ inline uint32_t rank2(uint32_t e) { return (e >> kRankBits) & kRankMask; }
__device__ inline uint32_t par_of(uint32_t e) { return e >> 31; }

// Work split of one op over the workgroup: thread t owns the determinant pairs
// k = kk + i*np... of its class, i.e. k0 = t mod np (and k0 + kThreads,
// k0 + 2*kThreads when np > kThreads) and the lines lane, lane + L, ...
// with L = max(1, kThreads / np) threads per pair.  Computed once per kernel
// for the two pair classes (single excitations CA, double excitations CCAA).
constexpr int kMaxEnt = 3;   // np <= 3 * kThreads

struct Split {
    int k0, lane, L, nent;   // nent: pairs this thread owns (0..kMaxEnt)
};

__device__ inline Split make_split(int np) {
    Split sp;
    const int t = threadIdx.x;
    if (np <= 0) {
        sp.k0 = 0; sp.lane = 0; sp.L = 1; sp.nent = 0;
    } else if (np <= kThreads) {
        sp.L = kThreads / np;
        sp.k0 = t % np;
        sp.lane = t / np;
        sp.nent = sp.lane < sp.L ? 1 : 0;
    } else {
        sp.L = 1; sp.lane = 0; sp.k0 = t;
        sp.nent = (np - t + kThreads - 1) / kThreads;
        if (sp.nent > kMaxEnt) sp.nent = kMaxEnt;
    }
    return sp;
}

// Op pipeline.  Ops run in batches of kBatch.  Descriptors of batch b+1 sit
// in an LDS ring (written one batch ahead, so reading them needs no global
// load); each thread's pair entries for batch b+1 are loaded into registers
// while batch b runs and handed over at the batch boundary, so their latency
// overlaps a whole batch.  No scalar loads happen inside the loop: scalar and
// LDS traffic share one wait counter.
constexpr int kBatch = 8;
constexpr int kWaitVmcnt0 = 0x0F70;   // s_waitcnt vmcnt(0) expcnt(7) lgkmcnt(15) (gfx9 encoding)
constexpr int kDescWords = sizeof(OpDev) / 4;   // 8

struct BatchEnt {
    uint32_t e[kBatch][kMaxEnt];
};

__device__ inline const OpDev &desc_at(const OpDev *ring, int slot, int i) { return ring[slot * kBatch + i]; }

// stage the descriptors of batch b (ops o0 + b*kBatch ...) into ring slot b % 2
__device__ inline void stage_descs(const PassArgs &A, OpDev *ring, int o0, int o1, int b) {
    const int first = o0 + b * kBatch;
    const int w = threadIdx.x;
    if (w < kBatch * kDescWords) {
        const int i = w / kDescWords, word = w - i * kDescWords;
        uint32_t v = 0;
        if (first + i < o1) v = reinterpret_cast<const uint32_t *>(A.ops + first + i)[word];
        reinterpret_cast<uint32_t *>(ring + (b & 1) * kBatch)[w] = v;
    }
}

__device__ inline void load_batch(const PassArgs &A, const OpDev *ring, int o0, int o1, int b, const Split &ca,
                                  const Split &cc, BatchEnt &E) {
#pragma unroll
    for (int i = 0; i < kBatch; ++i) {
        const int o = o0 + b * kBatch + i;
        const OpDev &d = desc_at(ring, b & 1, i);
        const int type = o < o1 ? d.type : -1;
        const bool dbl = type == OP_AAAA || type == OP_BBBB;
        const Split &sp = dbl ? cc : ca;
        const uint32_t *tab = (dbl ? A.ccaa_table : A.ca_table) + d.tab;
#pragma unroll
        for (int x = 0; x < kMaxEnt; ++x)
            E.e[i][x] = (type > OP_DIAG && x < sp.nent) ? tab[sp.k0 + x * kThreads] : 0u;
    }
}

template <typename R>
__device__ inline void run_op(const PassArgs &A, const OpDev &d, const uint32_t *ent, const Split &sca,
                              const Split &scc, typename Vec<R>::type *tile, const double2 *gang,
                              const uint32_t *line_rank, const uint32_t *line_str, int *pair_list, int *pair_count,
                              int nl, uint32_t C, uint32_t &cur_qr, int &ncp) {
    using V = typename Vec<R>::type;
    const int tid = threadIdx.x;
    const int type = d.type;
    if (type == OP_DIAG) {
        const uint32_t total = (uint32_t)nl * C;
        for (uint32_t idx = tid; idx < total; idx += kThreads) {
            uint32_t l = idx / C, a = idx - l * C;
            uint32_t b = line_rank[l];
            uint32_t sa = A.strings[a];
            double th = A.theta_a[a] + A.theta_b[b];
            while (sa) {
                int p = __builtin_ctz(sa);
                th += A.cross[(uint64_t)p * C + b];
                sa &= sa - 1;
            }
            double sn, cs;
            sincos(th, &sn, &cs);
            V v = tile[idx], w;
            R c = (R)cs, s = (R)(-sn);
            w.x = c * v.x - s * v.y;
            w.y = c * v.y + s * v.x;
            tile[idx] = w;
        }
    } else if (type == OP_ABBA) {
        if (d.aux != cur_qr) {
            cur_qr = d.aux;
            const int q = (int)(cur_qr & 0xff), r = (int)(cur_qr >> 8);
            if (tid < 64) {
                bool is01 = false;
                int partner = -1, parb = 0;
                if (tid < nl) {
                    uint32_t s = line_str[tid];
                    if (((s >> r) & 1u) && !((s >> q) & 1u)) {
                        uint32_t ps = s ^ (1u << q) ^ (1u << r);
                        for (int l = 0; l < nl; ++l)
                            if (line_str[l] == ps) partner = l;
                        is01 = partner >= 0;
                        parb = popcnt(s & between_mask(q, r)) & 1;
                    }
                }
                uint64_t mask = __ballot(is01);
                if (is01) pair_list[__popcll(mask & ((1ull << tid) - 1ull))] = tid | (partner << 8) | (parb << 16);
                if (tid == 0) *pair_count = __popcll(mask);
            }
            lds_barrier();
            ncp = *pair_count;
        }
        const R c = (R)d.cs;
#pragma unroll
        for (int i = 0; i < kMaxEnt; ++i) {
            if (i >= sca.nent) break;
            const uint32_t e = ent[i];
            const uint32_t rs_ = rank1(e), rp_ = rank2(e), para = par_of(e);
            for (int x = sca.lane; x < ncp; x += sca.L) {
                const int pl = pair_list[x];
                V *l01 = tile + (size_t)(pl & 0xff) * C;
                V *l10 = tile + (size_t)((pl >> 8) & 0xff) * C;
                const R mu = (R)(((para ^ (uint32_t)(pl >> 16)) & 1u) ? d.sn : -d.sn);
                rotate(l01[rs_], l10[rp_], c, mu);
            }
        }
    } else if (type == OP_GA || type == OP_GB) {
        const double2 *ga = gang + (size_t)d.aux * nl;
        const double *gp = A.grp_a + 2ull * d.index * A.ca_pairs;
#pragma unroll
        for (int i = 0; i < kMaxEnt; ++i) {
            if (i >= sca.nent) break;
            const uint32_t e = ent[i];
            const uint32_t rx = rank1(e), ry = rank2(e), par = par_of(e);
            const int k = sca.k0 + i * kThreads;
            const double ca = gp[2 * k], sa = gp[2 * k + 1];
            for (int l = sca.lane; l < nl; l += sca.L) {
                const double2 b = ga[l];
                const double c = ca * b.x - sa * b.y;
                const double s = sa * b.x + ca * b.y;
                V *line = tile + (size_t)l * C;
                rotate(line[rx], line[ry], (R)c, (R)(par ? s : -s));
            }
        }
    } else if (type >= 0) {   // AAAA / BBBB
        const R c = (R)d.cs;
#pragma unroll
        for (int i = 0; i < kMaxEnt; ++i) {
            if (i >= scc.nent) break;
            const uint32_t e = ent[i];
            const uint32_t r01 = rank1(e), r10 = rank2(e);
            const R mu = (R)(par_of(e) ? d.sn : -d.sn);
            for (int l = scc.lane; l < nl; l += scc.L) {
                V *line = tile + (size_t)l * C;
                rotate(line[r01], line[r10], c, mu);
            }
        }
    }
}

template <typename R>
__global__ void __launch_bounds__(kThreads)
pass_kernel(PassArgs A, PassDev pass, typename Vec<R>::type *prev, typename Vec<R>::type *cur, uint64_t j) {
    using V = typename Vec<R>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t *line_rank = reinterpret_cast<uint32_t *>(smem);            // [kMaxLines]
    uint32_t *line_str = line_rank + kMaxLines;                         // [kMaxLines]
    int *pair_list = reinterpret_cast<int *>(line_str + kMaxLines);     // [kMaxLines]
    int *pair_count = pair_list + kMaxLines;                            // [4]
    double *red = reinterpret_cast<double *>(pair_count + 4);           // [2 * kWaves]
    OpDev *ring = reinterpret_cast<OpDev *>(smem + 2048);              // [2][kBatch]
    V *tile = reinterpret_cast<V *>(smem + 4096);

    const int tid = threadIdx.x;
    const ItemDev it = A.items[pass.item_begin + blockIdx.x];
    const int nl = (int)it.count;
    const uint32_t C = A.C;
    const bool col = pass.kind == PASS_COL;
    double2 *gang = reinterpret_cast<double2 *>(smem + 4096 + (size_t)pass.max_lines * C * sizeof(V));
    const int o0 = pass.op_begin, o1 = pass.op_end;

    if (tid < nl) {
        uint32_t r = col ? A.col_list[it.begin + tid] : it.begin + tid;
        line_rank[tid] = r;
        line_str[tid] = A.strings[r];
    }
    stage_descs(A, ring, o0, o1, 0);
    __syncthreads();
    for (int i = tid; i < pass.group_count * nl; i += kThreads) {
        int g = i / nl, l = i - g * nl;
        const double *gb = A.grp_b + 2ull * (pass.group_first + g) * C + 2ull * line_rank[l];
        gang[i] = make_double2(gb[0], gb[1]);
    }
    const Split sca = make_split((int)A.ca_pairs), scc = make_split((int)A.ccaa_pairs);
    BatchEnt Ecur, Enext;
    load_batch(A, ring, o0, o1, 0, sca, scc, Ecur);   // overlaps the tile load below

    // ---- load the tile ---------------------------------------------------
    const bool from_prev = pass.flags & PF_FROM_PREV;
    const V *src = from_prev ? prev : cur;
    R sc = (R)1;
    if (from_prev && A.renorm) sc = (R)A.scale[j - 1];
    const uint32_t total = (uint32_t)nl * C;
    if (col) {
        for (int l = 0; l < nl; ++l) {
            const V *s_line = src + (uint64_t)line_rank[l] * C;
            V *t_line = tile + (size_t)l * C;
            for (uint32_t a = tid; a < C; a += kThreads) {
                V v = s_line[a];
                v.x *= sc; v.y *= sc;
                t_line[a] = v;
            }
        }
    } else {
        const uint32_t r0 = it.begin;
        for (uint32_t idx = tid; idx < total; idx += kThreads) {
            uint32_t b = idx / nl, l = idx - b * nl;
            V v = src[(uint64_t)b * C + r0 + l];
            v.x *= sc; v.y *= sc;
            tile[l * C + b] = v;
        }
    }

    // ---- ops ------------------------------------------------------------------
    uint32_t cur_qr = 0xffffffffu;
    int ncp = 0;
    const int nbatch = (o1 - o0 + kBatch - 1) / kBatch;
    for (int b = 0; b < nbatch; ++b) {
        // this batch's entries (loaded one batch ago) must land before the next
        // batch's loads are issued: the vector-memory counter retires in order
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        // descriptors of batch b+1 into the other ring slot, then its entries
        __syncthreads();
        if (b + 1 < nbatch) stage_descs(A, ring, o0, o1, b + 1);
        __syncthreads();
        if (b + 1 < nbatch) load_batch(A, ring, o0, o1, b + 1, sca, scc, Enext);
#pragma unroll
        for (int i = 0; i < kBatch; ++i) {
            if (o0 + b * kBatch + i >= o1) break;
            const OpDev d = desc_at(ring, b & 1, i);
            lds_barrier();
            run_op<R>(A, d, Ecur.e[i], sca, scc, tile, gang, line_rank, line_str, pair_list, pair_count, nl, C,
                      cur_qr, ncp);
        }
        Ecur = Enext;
    }
    __syncthreads();

    // ---- FP64 norm of the tile (last pass of a step) ----------------------
    if (pass.flags & PF_EMIT_NORM) {
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
            A.partials[blockIdx.x] = tot.value();
        }
    }

    // ---- store the tile -----------------------------------------------------
    if (col) {
        for (int l = 0; l < nl; ++l) {
            V *d_line = cur + (uint64_t)line_rank[l] * C;
            const V *t_line = tile + (size_t)l * C;
            for (uint32_t a = tid; a < C; a += kThreads) d_line[a] = t_line[a];
        }
    } else {
        const uint32_t r0 = it.begin;
        for (uint32_t idx = tid; idx < total; idx += kThreads) {
            uint32_t b = idx / nl, l = idx - b * nl;
            cur[(uint64_t)b * C + r0 + l] = tile[l * C + b];
        }
    }
}

// sum the per-workgroup partials of block j; set norms[j] and scale[j]
__global__ void norm_finalize(const double *partials, int count, uint64_t j, double *norms, double *scale,
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

template <typename R>
__global__ void scale_kernel(typename Vec<R>::type *blk, uint64_t D, double f) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < D; i += (uint64_t)gridDim.x * 256ull) {
        auto v = blk[i];
        v.x = (R)((double)v.x * f);
        v.y = (R)((double)v.y * f);
        blk[i] = v;
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
    to_device(&terms, p.terms, s);
    to_device(&ops, p.ops, s);
    to_device(&items, p.items, s);
    to_device(&col_list, p.col_list, s);
    to_device(&ca_table, p.ca_table, s);
    to_device(&ccaa_table, p.ccaa_table, s);
    hip_check(hipMalloc(reinterpret_cast<void **>(&partials), std::max(1, p.max_items) * sizeof(double)), "hipMalloc(partials)");
    lds_bytes = 0;
    for (const auto &ps : p.passes) {
        if (ps.max_lines > kMaxLines) throw std::runtime_error("plan exceeds the per-tile line limit");
        size_t b = 4096 + (size_t)ps.max_lines * p.C * p.elem_bytes + (size_t)ps.group_count * ps.max_lines * 16;
        lds_bytes = std::max(lds_bytes, b);
    }
    hip_check(hipStreamSynchronize(s), "plan upload sync");
}

void DevicePlan::release() {
    void *ptrs[] = {binom, strings, theta_a, theta_b, cross, groups, grp_a, grp_b, terms, ops, items, col_list,
                    partials, ca_table, ccaa_table};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    binom = nullptr; strings = nullptr; theta_a = theta_b = cross = nullptr; groups = nullptr;
    grp_a = grp_b = nullptr; terms = nullptr; ops = nullptr; items = nullptr; col_list = nullptr; partials = nullptr;
    ca_table = ccaa_table = nullptr;
}

template <typename R>
static void evolve_t(const DevicePlan &dp, void *state, uint64_t M, double *norms, double *scale, int renorm,
                     int *status_flag, hipStream_t s, int *launches) {
    using V = typename Vec<R>::type;
    const Plan &p = dp.plan;
    PassArgs A;
    A.binom = dp.binom; A.strings = dp.strings;
    A.theta_a = dp.theta_a; A.theta_b = dp.theta_b; A.cross = dp.cross;
    A.groups = dp.groups; A.grp_a = dp.grp_a; A.grp_b = dp.grp_b;
    A.terms = dp.terms; A.ops = dp.ops; A.items = dp.items; A.col_list = dp.col_list;
    A.ca_table = dp.ca_table; A.ccaa_table = dp.ccaa_table;
    A.partials = dp.partials; A.scale = scale;
    A.n = p.n; A.e = p.e; A.C = p.C; A.ca_pairs = p.ca_pairs; A.ccaa_pairs = p.ccaa_pairs;
    A.renorm = renorm;
    const uint64_t D = (uint64_t)p.C * p.C;
    V *base = reinterpret_cast<V *>(state);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&pass_kernel<R>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dp.lds_bytes);
    int norm_items = 0;
    for (const auto &ps : p.passes)
        if (ps.flags & PF_EMIT_NORM) norm_items = ps.item_count;
    for (uint64_t j = 1; j < M; ++j) {
        V *prev = base + (j - 1) * D, *cur = base + j * D;
        for (const auto &ps : p.passes) {
            hipLaunchKernelGGL(pass_kernel<R>, dim3(ps.item_count), dim3(kThreads), dp.lds_bytes, s, A, ps, prev, cur, j);
        }
        hipLaunchKernelGGL(norm_finalize, dim3(1), dim3(1024), 0, s, (const double *)dp.partials, norm_items, j, norms,
                           scale, status_flag);
    }
    hip_check(hipGetLastError(), "pass launch");
    if (launches) *launches = (int)p.passes.size() + 1;
}

void evolve(const DevicePlan &dp, void *state, uint64_t M, double *norms, double *scale, int renorm,
            int *status_flag, hipStream_t s, int *launches) {
    if (dp.dtype == 0) evolve_t<double>(dp, state, M, norms, scale, renorm, status_flag, s, launches);
    else evolve_t<float>(dp, state, M, norms, scale, renorm, status_flag, s, launches);
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

void qift(int dtype, void *state, int m, uint64_t D, const double *scale, double *pk_partial, int *grid_out,
          double *twiddles, size_t max_lds, hipStream_t s) {
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
    int grid = (int)std::min<uint64_t>(ntiles, 1024);
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

void scale_block(int dtype, void *state, uint64_t j, uint64_t D, double factor, hipStream_t s) {
    if (dtype == 0)
        hipLaunchKernelGGL(scale_kernel<double>, dim3(grid_for(D)), dim3(256), 0, s,
                           reinterpret_cast<double2 *>(state) + j * D, D, factor);
    else
        hipLaunchKernelGGL(scale_kernel<float>, dim3(grid_for(D)), dim3(256), 0, s,
                           reinterpret_cast<float2 *>(state) + j * D, D, factor);
    hip_check(hipGetLastError(), "scale launch");
}

void debug_pairs(int kind, int n, int e, const int32_t *args, uint32_t *o1, uint32_t *o2, uint32_t *op,
                 uint64_t count, hipStream_t s) {
    hipLaunchKernelGGL(debug_pairs_kernel, dim3(grid_for(count)), dim3(256), 0, s, kind, n, e, args[0], args[1],
                       args[2], args[3], o1, o2, op, count);
    hip_check(hipGetLastError(), "debug pairs launch");
}

}  // namespace qpe
