// C ABI of libqpesim.so (include/qpesim.h): handles, argument checks, error
// mapping and the device-side StateVector.  No exceptions cross the ABI.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qpesim.h"
#include "broombridge.hpp"
#include "combinadic.hpp"
#include "engine.hpp"
#include "hamiltonian.hpp"
#include "plan.hpp"

namespace {

thread_local std::string g_last_error;

struct ApiError : std::runtime_error {
    int code;
    ApiError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

template <typename F>
int guarded(F &&f) {
    try {
        f();
        g_last_error.clear();
        return QPE_OK;
    } catch (const ApiError &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return QPE_ERR_OOM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return QPE_ERR_HIP;
    }
}

void require(bool ok, int code, const char *msg) {
    if (!ok) throw ApiError(code, msg);
}

uint64_t choose(int n, int k) {
    uint32_t tab[qpe::kBinomRows * qpe::kBinomCols];
    qpe::fill_binomials(tab);
    return qpe::binom(tab, n, k);
}

// Process-wide generation numbers: every Hamiltonian mutation, state and device
// plan gets a value no other object ever had, so caches keyed on them cannot
// mistake a new object for a freed one that lived at the same address.
std::atomic<uint64_t> g_generation{1};
uint64_t next_generation() { return g_generation.fetch_add(1, std::memory_order_relaxed); }

}  // namespace

struct qpe_hamiltonian {
    qpe::Hamiltonian h;
    uint64_t version = next_generation();   // changes (to a fresh generation) on every mutation
    double identity = 0.0;                    // GetIdentityTerm
    std::vector<qpe::Determinant> initial;    // initial state of the last Broombridge ingest
};

struct qpe_state {
    const uint64_t id = next_generation();
    unsigned m = 0, n = 0, e = 0;
    int dtype = QPE_COMPLEX128;
    int layout = QPE_LAYOUT_ALPHA_BETA;
    uint32_t C = 0;            // strings per spin, C(n, e)
    uint64_t D = 0, M = 0;     // amplitudes per block (C^2, or C(2n, 2e) single-string), blocks
    void *ws = nullptr;        // single-string layout: two alpha x beta blocks the passes work on
    size_t elem = 16;
    void *amp = nullptr;
    double *norms = nullptr, *scale = nullptr, *pk_partial = nullptr, *twiddles = nullptr, *probs = nullptr;
    double *norm_partials = nullptr;
    int *status = nullptr;
    hipStream_t stream = nullptr;
    int renorm = 0;
    int engine = 0;
    int64_t max_tile_bytes = 0;
    int order = 1;   // Trotter order
    int term_order = 0;        // QPE_OPT_TERM_ORDER of the chain
    int qift_form = 0;         // QPE_OPT_QIFT_FORM
    int execution = 0;         // QPE_OPT_EXECUTION
    uint32_t wait_ms = 10000;  // QPE_OPT_WAIT_LIMIT_MS
    int pk_rows = 0;           // rows of pk_partial
    int plan_term_order = -1;  // term order of the cached plan
    uint64_t evolved_to = 0;   // blocks < evolved_to hold U^j psi (chained qpe_time_evolve_blocks)
    bool scale_pending = false;
    bool pk_valid = false;
    int pk_grid = 0;
    double t_evolve = 0, t_qift = 0, t_pk = 0;
    int passes = 0, launches = 0;
    std::unique_ptr<qpe::DevicePlan> plan;
    uint64_t plan_id = 0;                  // generation of the device plan (0: none)
    const qpe_hamiltonian *plan_h = nullptr;
    uint64_t plan_version = ~0ull;         // the Hamiltonian generation the plan was built from
    int max_lds = 0;

    ~qpe_state() {
        plan.reset();
        void *ptrs[] = {amp, norms, scale, pk_partial, twiddles, probs, norm_partials, status, ws};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

constexpr int kNormParts = 512;
constexpr int kPkGridMax = 1024;

void check_state(qpe_state_t s) { require(s != nullptr, QPE_ERR_ARG, "null state handle"); }
void check_ham(qpe_hamiltonian_t h) { require(h != nullptr, QPE_ERR_ARG, "null hamiltonian handle"); }

std::string read_text_file(const char *path) {
    require(path != nullptr, QPE_ERR_ARG, "null path");
    std::ifstream f(path, std::ios::binary);
    if (!f) throw ApiError(QPE_ERR_ARG, std::string("cannot open ") + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

qpe::BroombridgeData parse_broombridge(const qpe_hamiltonian *h, const std::string &text) {
    try {
        return qpe::read_broombridge(text, h->h.num_orbitals());
    } catch (const qpe::FormatError &e) {
        throw ApiError(QPE_ERR_ARG, std::string("Broombridge: ") + e.what());
    }
}

void adopt_broombridge(qpe_hamiltonian *h, qpe::BroombridgeData &&d) {
    qpe::apply_broombridge(d, h->h);
    h->identity = d.identity;
    h->initial = std::move(d.initial);
    h->version = next_generation();
}

void set_ones(double *dev, uint64_t count, hipStream_t st) {
    std::vector<double> ones(count, 1.0);
    qpe::hip_check(hipMemcpyAsync(dev, ones.data(), count * sizeof(double), hipMemcpyHostToDevice, st), "copy ones");
    qpe::hip_check(hipStreamSynchronize(st), "sync");
}

// fold pending renormalisation factors into the stored amplitudes
void materialize_scale(qpe_state_t s) {
    if (!s->scale_pending) return;
    std::vector<double> sc(s->M);
    qpe::hip_check(hipMemcpyAsync(sc.data(), s->scale, s->M * sizeof(double), hipMemcpyDeviceToHost, s->stream), "scale copy");
    qpe::hip_check(hipStreamSynchronize(s->stream), "sync");
    for (uint64_t j = 0; j < s->M; ++j)
        if (sc[j] != 1.0) qpe::scale_block(s->dtype, s->amp, j, s->D, sc[j], s->stream);
    set_ones(s->scale, s->M, s->stream);
    s->scale_pending = false;
}

// The last batch's device tables and captured graph, reused while the members,
// their plans and the flags are unchanged.  The key holds generation numbers
// (state ids, plan ids), never addresses; a member's destruction or new plan
// drops the cached run at once, so no graph outlives the buffers it points at.
struct BatchCache {
    std::vector<uint64_t> key;             // (state id, plan id, renorm) per member
    std::unique_ptr<qpe::BatchRun> run;
};
BatchCache g_batch;
std::mutex g_batch_mu;

void drop_batch_with(uint64_t state_id) {
    std::lock_guard<std::mutex> lock(g_batch_mu);
    if (!g_batch.run) return;
    for (size_t i = 0; i < g_batch.key.size(); i += 3)
        if (g_batch.key[i] == state_id) {
            g_batch.run.reset();
            g_batch.key.clear();
            return;
        }
}

}  // namespace

extern "C" {

const char *qpe_last_error(void) { return g_last_error.c_str(); }

int qpe_hamiltonian_create(qpe_hamiltonian_t *out) {
    return guarded([&] {
        require(out != nullptr, QPE_ERR_ARG, "null output pointer");
        *out = new qpe_hamiltonian();
    });
}

int qpe_hamiltonian_destroy(qpe_hamiltonian_t h) {
    return guarded([&] { delete h; });
}

int qpe_hamiltonian_add_one_body_term(qpe_hamiltonian_t h, int p, int q, double coeff) {
    return guarded([&] {
        check_ham(h);
        require(p >= 1 && q >= 1 && p >= q, QPE_ERR_ARG, "one-body indices must be 1-based with p >= q");
        require(std::isfinite(coeff), QPE_ERR_ARG, "non-finite coefficient");
        h->h.AddOneBodyTerm(p, q, coeff);
        h->version = next_generation();
    });
}

int qpe_hamiltonian_add_two_body_term(qpe_hamiltonian_t h, int p, int q, int r, int s, double coeff) {
    return guarded([&] {
        check_ham(h);
        require(p >= 1 && q >= 1 && r >= 1 && s >= 1, QPE_ERR_ARG, "two-body indices must be 1-based");
        require(std::isfinite(coeff), QPE_ERR_ARG, "non-finite coefficient");
        h->h.AddTwoBodyTerm(p, q, r, s, coeff);
        h->version = next_generation();
    });
}

int qpe_hamiltonian_set_num_orbitals(qpe_hamiltonian_t h, unsigned n) {
    return guarded([&] {
        check_ham(h);
        require(n >= 1 && n <= 32, QPE_ERR_ARG, "num_orbitals must be in [1, 32]");
        h->h.SetNumOrbitals(n);
        h->version = next_generation();
    });
}

int qpe_hamiltonian_set_num_occupied(qpe_hamiltonian_t h, unsigned e) {
    return guarded([&] {
        check_ham(h);
        h->h.SetNumOccupied(e);
        h->version = next_generation();
    });
}

int qpe_hamiltonian_index_terms(qpe_hamiltonian_t h) {
    return guarded([&] {
        check_ham(h);
        require(h->h.num_orbitals() >= 1, QPE_ERR_STATE, "set num_orbitals before IndexTerms");
        h->h.IndexTerms();
        h->version = next_generation();
    });
}

int qpe_hamiltonian_set_timestep(qpe_hamiltonian_t h, double t) {
    return guarded([&] {
        check_ham(h);
        require(std::isfinite(t), QPE_ERR_ARG, "non-finite timestep");
        h->h.SetTimestep(t);
        h->version = next_generation();
    });
}

int qpe_hamiltonian_term_counts(qpe_hamiltonian_t h, uint64_t counts[5]) {
    return guarded([&] {
        check_ham(h);
        require(counts != nullptr, QPE_ERR_ARG, "null counts");
        require(h->h.indexed(), QPE_ERR_STATE, "IndexTerms has not been called");
        for (int c = 0; c < 5; ++c) counts[c] = h->h.segmented()[c].size();
    });
}

int qpe_hamiltonian_get_terms(qpe_hamiltonian_t h, int cls, int32_t *idx, double *coeff, uint64_t cap) {
    return guarded([&] {
        check_ham(h);
        require(cls >= 0 && cls < 5, QPE_ERR_ARG, "term class out of range");
        require(h->h.indexed(), QPE_ERR_STATE, "IndexTerms has not been called");
        const auto &v = h->h.segmented()[cls];
        require(cap >= v.size(), QPE_ERR_ARG, "capacity too small");
        for (size_t i = 0; i < v.size(); ++i) {
            for (int k = 0; k < 4; ++k)
                if (idx) idx[4 * i + k] = v[i].idx[k];
            if (coeff) coeff[i] = v[i].coeff;
        }
    });
}

int qpe_hamiltonian_ingest_broombridge(qpe_hamiltonian_t h, const char *path) {
    return guarded([&] {
        check_ham(h);
        adopt_broombridge(h, parse_broombridge(h, read_text_file(path)));
    });
}

int qpe_hamiltonian_ingest_broombridge_text(qpe_hamiltonian_t h, const char *text, uint64_t length) {
    return guarded([&] {
        check_ham(h);
        require(text != nullptr || length == 0, QPE_ERR_ARG, "null text");
        adopt_broombridge(h, parse_broombridge(h, std::string(text ? text : "", (size_t)length)));
    });
}

int qpe_hamiltonian_initial_state(qpe_hamiltonian_t h, uint64_t *address, double *coeff, uint64_t cap,
                                  uint64_t *count) {
    return guarded([&] {
        check_ham(h);
        require(count != nullptr, QPE_ERR_ARG, "null count");
        *count = h->initial.size();
        if (!address && !coeff) return;
        require(cap >= h->initial.size(), QPE_ERR_ARG, "capacity too small");
        for (size_t i = 0; i < h->initial.size(); ++i) {
            if (address) address[i] = h->initial[i].address;
            if (coeff) coeff[i] = h->initial[i].coeff;
        }
    });
}

int qpe_hamiltonian_identity_term(qpe_hamiltonian_t h, double *identity) {
    return guarded([&] {
        check_ham(h);
        require(identity != nullptr, QPE_ERR_ARG, "null output");
        *identity = h->identity;
    });
}

int qpe_hamiltonian_get_sizes(qpe_hamiltonian_t h, unsigned *num_orbitals, unsigned *num_occupied) {
    return guarded([&] {
        check_ham(h);
        if (num_orbitals) *num_orbitals = h->h.num_orbitals();
        if (num_occupied) *num_occupied = h->h.num_occupied();
    });
}

int qpe_state_create(qpe_state_t *out, unsigned precision, unsigned num_orbitals, unsigned num_occupied, int dtype) {
    return qpe_state_create_layout(out, precision, num_orbitals, num_occupied, dtype, QPE_LAYOUT_ALPHA_BETA);
}

int qpe_state_create_layout(qpe_state_t *out, unsigned precision, unsigned num_orbitals, unsigned num_occupied,
                            int dtype, int layout) {
    return guarded([&] {
        require(out != nullptr, QPE_ERR_ARG, "null output pointer");
        *out = nullptr;
        require(precision >= 1 && precision <= 20, QPE_ERR_ARG, "precision must be in [1, 20]");
        require(num_orbitals >= 2 && num_orbitals <= 32, QPE_ERR_ARG, "num_orbitals must be in [2, 32]");
        require(num_occupied >= 1 && num_occupied < num_orbitals, QPE_ERR_ARG, "need 1 <= num_occupied < num_orbitals");
        require(dtype == QPE_COMPLEX128 || dtype == QPE_COMPLEX64, QPE_ERR_ARG, "unknown dtype");
        require(layout == QPE_LAYOUT_ALPHA_BETA || layout == QPE_LAYOUT_SINGLE_STRING, QPE_ERR_ARG, "unknown layout");
        uint64_t C = choose((int)num_orbitals, (int)num_occupied);
        require(C >= 1 && C <= 65535, QPE_ERR_ARG, "C(num_orbitals, num_occupied) exceeds 65535");
        uint64_t D = C * C;
        if (layout == QPE_LAYOUT_SINGLE_STRING) {
            // 2n spin-orbitals: ranks of strings of up to 34 bits (the binomial table's rows)
            require(num_orbitals <= 17, QPE_ERR_ARG, "the single-string layout holds up to 17 spatial orbitals");
            D = choose(2 * (int)num_orbitals, 2 * (int)num_occupied);
        }
        std::unique_ptr<qpe_state> s(new qpe_state());
        s->m = precision; s->n = num_orbitals; s->e = num_occupied; s->dtype = dtype; s->layout = layout;
        s->C = (uint32_t)C; s->D = D; s->M = 1ull << precision;
        s->elem = dtype == QPE_COMPLEX128 ? 16 : 8;
        s->renorm = dtype == QPE_COMPLEX64 ? 1 : 0;
        const uint64_t bytes = s->M * s->D * s->elem;
        int ndev = 0;
        require(hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0, QPE_ERR_HIP, "no HIP device");
        size_t free_b = 0, total_b = 0;
        qpe::hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        require(bytes + (64ull << 20) < free_b, QPE_ERR_OOM, "state does not fit in free device memory");
        qpe::hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
        qpe::hip_check(hipMalloc(&s->amp, bytes), "hipMalloc(state)");
        qpe::hip_check(hipMemsetAsync(s->amp, 0, bytes, s->stream), "zero state");
        qpe::hip_check(hipMalloc(reinterpret_cast<void **>(&s->norms), s->M * sizeof(double)), "hipMalloc(norms)");
        qpe::hip_check(hipMalloc(reinterpret_cast<void **>(&s->scale), s->M * sizeof(double)), "hipMalloc(scale)");
        qpe::hip_check(hipMalloc(reinterpret_cast<void **>(&s->probs), s->M * sizeof(double)), "hipMalloc(probs)");
        // rows of 2^m P(k) partial sums: up to kPkGridMax, at most 64 MB in all
        s->pk_rows = (int)std::min<uint64_t>(kPkGridMax, std::max<uint64_t>(8, (8ull << 20) / s->M));
        qpe::hip_check(hipMalloc(reinterpret_cast<void **>(&s->pk_partial), (size_t)s->pk_rows * s->M * sizeof(double)),
                       "hipMalloc(pk)");
        qpe::hip_check(hipMalloc(reinterpret_cast<void **>(&s->norm_partials), kNormParts * sizeof(double)), "hipMalloc");
        qpe::hip_check(hipMalloc(reinterpret_cast<void **>(&s->status), sizeof(int)), "hipMalloc(status)");
        qpe::hip_check(hipMemsetAsync(s->status, 0, sizeof(int), s->stream), "zero status");
        qpe::hip_check(hipMemsetAsync(s->norms, 0, s->M * sizeof(double), s->stream), "zero norms");
        // forward DFT twiddles e^{-2 pi i k / M}
        const uint64_t half = std::max<uint64_t>(1, s->M / 2);
        std::vector<double> tw(2 * half);
        for (uint64_t k = 0; k < half; ++k) {
            long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)s->M;
            tw[2 * k] = (double)std::cos(a);
            tw[2 * k + 1] = (double)-std::sin(a);
        }
        qpe::hip_check(hipMalloc(reinterpret_cast<void **>(&s->twiddles), tw.size() * sizeof(double)), "hipMalloc(tw)");
        qpe::hip_check(hipMemcpyAsync(s->twiddles, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice, s->stream),
                       "twiddles");
        set_ones(s->scale, s->M, s->stream);
        s->max_lds = qpe::device_max_lds();
        *out = s.release();
    });
}

int qpe_state_destroy(qpe_state_t s) {
    return guarded([&] {
        if (!s) return;
        (void)hipStreamSynchronize(s->stream);
        drop_batch_with(s->id);   // a cached batch graph must not outlive this state's buffers
        delete s;
    });
}

int qpe_state_size(qpe_state_t s, uint64_t *n) {
    return guarded([&] {
        check_state(s);
        require(n != nullptr, QPE_ERR_ARG, "null output");
        *n = s->M * s->D;
    });
}

int qpe_state_set(qpe_state_t s, uint64_t offset, uint64_t count, const double *amps) {
    return guarded([&] {
        check_state(s);
        require(amps != nullptr || count == 0, QPE_ERR_ARG, "null amplitudes");
        require(offset <= s->M * s->D && count <= s->M * s->D - offset, QPE_ERR_ARG, "range outside the state");
        if (count == 0) return;
        materialize_scale(s);
        const uint64_t chunk = 1ull << 26;
        double *staging = nullptr;
        qpe::hip_check(hipMalloc(reinterpret_cast<void **>(&staging), std::min(count, chunk) * 2 * sizeof(double)), "staging");
        try {
            for (uint64_t done = 0; done < count; done += chunk) {
                uint64_t c = std::min(chunk, count - done);
                qpe::hip_check(hipMemcpyAsync(staging, amps + 2 * done, c * 2 * sizeof(double), hipMemcpyHostToDevice,
                                              s->stream), "upload");
                if (s->layout == QPE_LAYOUT_SINGLE_STRING)
                    qpe::single_string_copy(s->dtype, s->amp, staging, offset + done, c, true, s->stream);
                else
                    qpe::scatter_in(s->dtype, s->amp, staging, offset + done, c, s->C, s->stream);
            }
            qpe::hip_check(hipStreamSynchronize(s->stream), "sync");
        } catch (...) {
            (void)hipFree(staging);
            throw;
        }
        (void)hipFree(staging);
        s->pk_valid = false;
    });
}

int qpe_state_get(qpe_state_t s, uint64_t offset, uint64_t count, double *amps) {
    return guarded([&] {
        check_state(s);
        require(amps != nullptr || count == 0, QPE_ERR_ARG, "null amplitudes");
        require(offset <= s->M * s->D && count <= s->M * s->D - offset, QPE_ERR_ARG, "range outside the state");
        if (count == 0) return;
        materialize_scale(s);
        const uint64_t chunk = 1ull << 26;
        double *staging = nullptr;
        qpe::hip_check(hipMalloc(reinterpret_cast<void **>(&staging), std::min(count, chunk) * 2 * sizeof(double)), "staging");
        try {
            for (uint64_t done = 0; done < count; done += chunk) {
                uint64_t c = std::min(chunk, count - done);
                if (s->layout == QPE_LAYOUT_SINGLE_STRING)
                    qpe::single_string_copy(s->dtype, s->amp, staging, offset + done, c, false, s->stream);
                else
                    qpe::gather_out(s->dtype, s->amp, staging, offset + done, c, s->C, s->stream);
                qpe::hip_check(hipMemcpyAsync(amps + 2 * done, staging, c * 2 * sizeof(double), hipMemcpyDeviceToHost,
                                              s->stream), "download");
                qpe::hip_check(hipStreamSynchronize(s->stream), "sync");
            }
        } catch (...) {
            (void)hipFree(staging);
            throw;
        }
        (void)hipFree(staging);
    });
}

int qpe_state_scale_block(qpe_state_t s, uint64_t block, double factor) {
    return guarded([&] {
        check_state(s);
        require(block < s->M, QPE_ERR_ARG, "block out of range");
        require(std::isfinite(factor), QPE_ERR_ARG, "non-finite factor");
        materialize_scale(s);
        qpe::scale_block(s->dtype, s->amp, block, s->D, factor, s->stream);
        qpe::hip_check(hipStreamSynchronize(s->stream), "sync");
        s->pk_valid = false;
    });
}

static void apply_ancilla_gate(qpe_state_t s, const double *gate, int control, unsigned qubit) {
    check_state(s);
    require(gate != nullptr, QPE_ERR_ARG, "null gate");
    for (int k = 0; k < 8; ++k) require(std::isfinite(gate[k]), QPE_ERR_ARG, "non-finite gate entry");
    require(qubit < s->m, QPE_ERR_ARG, "qubit must be below the precision");
    if (control >= 0) {
        require((unsigned)control < s->m, QPE_ERR_ARG, "control must be below the precision");
        require((unsigned)control != qubit, QPE_ERR_ARG, "control and target must differ");
    }
    materialize_scale(s);
    // qubit q is bit m - q - 1 of the block index (big-endian, cpp/state_vector.hpp:56, 84-85)
    const int tbit = (int)(s->m - qubit - 1), cbit = control >= 0 ? (int)(s->m - (unsigned)control - 1) : -1;
    qpe::apply_gate(s->dtype, s->amp, s->D, (int)s->m, tbit, cbit, gate, s->stream);
    qpe::hip_check(hipStreamSynchronize(s->stream), "sync");
    s->pk_valid = false;
    s->evolved_to = 0;
}

int qpe_state_apply_gate(qpe_state_t s, const double *gate, unsigned qubit) {
    return guarded([&] { apply_ancilla_gate(s, gate, -1, qubit); });
}

int qpe_state_apply_controlled_gate(qpe_state_t s, const double *gate, unsigned control, unsigned qubit) {
    return guarded([&] {
        check_state(s);
        require(control < 64, QPE_ERR_ARG, "control must be below the precision");
        apply_ancilla_gate(s, gate, (int)control, qubit);
    });
}

int qpe_state_ingest_broombridge(qpe_state_t *out, qpe_hamiltonian_t h, const char *path, unsigned precision,
                                 int dtype) {
    return guarded([&] {
        require(out != nullptr, QPE_ERR_ARG, "null output pointer");
        *out = nullptr;
        check_ham(h);
        qpe::BroombridgeData d = parse_broombridge(h, read_text_file(path));
        qpe_state_t s = nullptr;
        auto check = [&](int rc) {
            if (rc == QPE_OK) return;
            std::string msg = g_last_error;
            if (s) qpe_state_destroy(s);
            throw ApiError(rc, msg);
        };
        check(qpe_state_create(&s, precision, d.num_orbitals, d.num_occupied, dtype));
        // block 0: the determinants in file order, a repeated address keeping its last coefficient
        std::map<uint64_t, double> amps;
        for (const auto &det : d.initial) amps[det.address] = det.coeff;
        for (const auto &kv : amps) {
            const double a[2] = {kv.second, 0.0};
            check(qpe_state_set(s, kv.first, 1, a));
        }
        adopt_broombridge(h, std::move(d));
        *out = s;
    });
}

int qpe_energy_estimate(const double *probs, unsigned precision, double identity, double timestep, double out[3]) {
    return guarded([&] {
        require(probs != nullptr && out != nullptr, QPE_ERR_ARG, "null pointer");
        require(precision >= 1 && precision <= 30, QPE_ERR_ARG, "precision must be in [1, 30]");
        require(std::isfinite(timestep) && timestep != 0.0, QPE_ERR_ARG, "timestep must be finite and non-zero");
        const uint64_t M = 1ull << precision;
        double total = 0.0;
        for (uint64_t k = 0; k < M; ++k) total += (double)k * probs[k];
        const double phase = total / (double)M;
        const double two_pi = 2 * 2 * std::acos(0.0);
        out[0] = -1.0 * two_pi * phase + identity;
        out[1] = -1.0 * two_pi * phase / timestep + identity;
        out[2] = two_pi / (timestep * (double)M);
    });
}

int qpe_state_set_option(qpe_state_t s, int option, int64_t value) {
    return guarded([&] {
        check_state(s);
        switch (option) {
        case QPE_OPT_RENORMALIZE: s->renorm = value ? 1 : 0; break;
        case QPE_OPT_ENGINE:
            require(value == 0 || value == 1, QPE_ERR_ARG, "engine must be 0 (fused groups) or 1 (one term per op)");
            s->engine = (int)value;
            s->plan_version = ~0ull;
            break;
        case QPE_OPT_TROTTER_ORDER:
            require(value == 1 || value == 2, QPE_ERR_ARG, "Trotter order must be 1 or 2");
            s->order = (int)value;
            s->plan_version = ~0ull;
            break;
        case QPE_OPT_TERM_ORDER:
            require(value == 0 || value == 1, QPE_ERR_ARG, "term order must be 0 (reordered) or 1 (applyUnitary)");
            s->term_order = (int)value;
            s->plan_version = ~0ull;
            break;
        case QPE_OPT_QIFT_FORM:
            require(value >= 0 && value <= 3, QPE_ERR_ARG, "QIFT form must be 0 (auto), 1, 2 or 3");
            require(qpe::qift_form_available((int)value, (int)s->m), QPE_ERR_ARG,
                    "this QIFT form does not cover 2^precision (register DFTs: 2..1024; matrix cores: 16..64)");
            s->qift_form = (int)value;
            break;
        case QPE_OPT_EXECUTION:
            require(value >= 0 && value <= 2, QPE_ERR_ARG, "execution must be 0 (auto), 1 (passes) or 2 (dataflow)");
            s->execution = (int)value;
            break;
        case QPE_OPT_WAIT_LIMIT_MS:
            require(value >= 0 && value <= 3600000, QPE_ERR_ARG, "wait limit must be 0..3600000 ms");
            s->wait_ms = (uint32_t)value;
            break;
        case QPE_OPT_MAX_TILE_BYTES:
            require(value >= 0, QPE_ERR_ARG, "negative tile budget");
            s->max_tile_bytes = value;
            s->plan_version = ~0ull;
            break;
        default: throw ApiError(QPE_ERR_ARG, "unknown option");
        }
    });
}

// Everything before the timed evolution: argument checks, the plan (cached per
// Hamiltonian version and options), the status flag, and block 0's norm.
static void prepare_evolve(qpe_state_t s, qpe_hamiltonian_t h, int term_order) {
    check_state(s);
    check_ham(h);
    require(h->h.indexed(), QPE_ERR_STATE, "IndexTerms has not been called");
    require(h->h.num_orbitals() == s->n && h->h.num_occupied() == s->e, QPE_ERR_ARG,
            "hamiltonian (num_orbitals, num_occupied) does not match the state");
    materialize_scale(s);
    if (!s->plan || s->plan_h != h || s->plan_version != h->version || s->plan_term_order != term_order) {
        size_t budget = (size_t)s->max_lds > qpe::kLdsHeader ? (size_t)s->max_lds - qpe::kLdsHeader : 0;
        if (s->max_tile_bytes > 0) budget = std::min(budget, (size_t)s->max_tile_bytes);
        qpe::Plan p;
        try {
            p = qpe::build_plan(h->h, s->elem, budget, s->engine == 0, s->order, term_order);
        } catch (const std::runtime_error &e) {
            throw ApiError(QPE_ERR_ARG, e.what());
        }
        std::unique_ptr<qpe::DevicePlan> dp(new qpe::DevicePlan());
        dp->upload(p, s->dtype == QPE_COMPLEX128 ? 0 : 1, s->stream);
        drop_batch_with(s->id);   // the cached batch graph points at the plan being replaced
        s->plan = std::move(dp);
        s->plan_id = next_generation();
        s->plan_h = h;
        s->plan_version = h->version;
        s->plan_term_order = term_order;
    }
    qpe::hip_check(hipMemsetAsync(s->status, 0, sizeof(int), s->stream), "zero status");
    set_ones(s->scale, s->M, s->stream);
    qpe::block_norm(s->dtype == QPE_COMPLEX128 ? 0 : 1, s->D, s->amp, 0, s->norms, s->scale, s->norm_partials, kNormParts, s->stream);
}

static void finish_evolve(qpe_state_t s, float ms) {
    s->t_evolve = ms;
    s->passes = (int)s->plan->plan.passes.size();
    s->scale_pending = s->renorm != 0;
    s->pk_valid = false;
    int status = 0;
    qpe::hip_check(hipMemcpy(&status, s->status, sizeof(int), hipMemcpyDeviceToHost), "status copy");
    require(status != 2, QPE_ERR_STALLED,
            "evolution stalled: a dataflow item waited longer than the wait limit for the items it depends on");
    require(status == 0, QPE_ERR_NUMERIC, "non-finite amplitudes during evolution");
}

int qpe_time_evolve_over_terms(qpe_state_t s, qpe_hamiltonian_t h) {
    uint64_t M = 0;
    const int rc = guarded([&] {
        check_state(s);
        M = s->M;
    });
    if (rc != QPE_OK) return rc;
    return qpe_time_evolve_blocks(s, h, 1, M);
}

int qpe_time_evolve_blocks(qpe_state_t s, qpe_hamiltonian_t h, uint64_t j_begin, uint64_t j_end) {
    return guarded([&] {
        check_state(s);
        require(j_begin >= 1 && j_begin <= j_end && j_end <= s->M, QPE_ERR_ARG, "block range outside 1..2^m");
        if (j_begin == 1) {
            prepare_evolve(s, h, s->term_order);
        } else {
            // continuing a chain: block j_begin-1 and its norm come from an earlier call with the same plan
            check_ham(h);
            require(s->plan && s->plan_h == h && s->plan_version == h->version && s->evolved_to == j_begin &&
                        s->plan_term_order == s->term_order,
                    QPE_ERR_STATE, "blocks before j_begin have not been evolved with this Hamiltonian");
        }
        const bool k1 = s->layout == QPE_LAYOUT_SINGLE_STRING;
        if (k1) {
            if (!s->ws)
                qpe::hip_check(hipMalloc(&s->ws, 2 * (size_t)s->C * s->C * s->elem), "hipMalloc(alpha x beta blocks)");
            if (j_begin == 1) {
                // the passes evolve the (e alpha, e beta) sector only
                qpe::hip_check(hipMemsetAsync(s->status, 0, sizeof(int), s->stream), "zero status");
                qpe::single_string_outside(*s->plan, s->amp, s->D, s->status, s->stream);
                int outside = 0;
                qpe::hip_check(hipMemcpy(&outside, s->status, sizeof(int), hipMemcpyDeviceToHost), "status copy");
                require(outside == 0, QPE_ERR_ARG,
                        "block 0 has amplitudes outside the sector of num_occupied alpha and beta electrons");
            }
        }
        hipEvent_t e0, e1;
        qpe::hip_check(hipEventCreate(&e0), "event");
        qpe::hip_check(hipEventCreate(&e1), "event");
        qpe::hip_check(hipEventRecord(e0, s->stream), "event record");
        if (k1) {
            // block j-1 in alpha x beta order, one step, block j back in single-string order;
            // the two alpha x beta blocks swap roles, so each block is gathered once
            char *wsb = static_cast<char *>(s->ws);
            const size_t bb = (size_t)s->C * s->C * s->elem;
            char *amp = static_cast<char *>(s->amp);
            int cur = 0;
            qpe::single_string_convert(*s->plan, wsb, amp + (j_begin - 1) * s->D * s->elem, true, s->stream);
            for (uint64_t j = j_begin; j < j_end; ++j) {
                qpe::evolve_step(*s->plan, wsb + cur * bb, wsb + (1 - cur) * bb, j, s->norms, s->scale, s->renorm,
                                 s->status, s->stream);
                qpe::single_string_convert(*s->plan, wsb + (1 - cur) * bb, amp + j * s->D * s->elem, false, s->stream);
                cur = 1 - cur;
            }
            s->launches = (int)s->plan->plan.passes.size() + 2;
        } else {
            require(s->execution != qpe::EXEC_DATAFLOW || s->plan->flow_ok, QPE_ERR_ARG,
                    "the dataflow evolution needs every pass tiled in LDS and blocks below 2^31 bytes");
            qpe::evolve(*s->plan, s->amp, j_begin, j_end, s->norms, s->scale, s->renorm, s->status, s->stream,
                        &s->launches, s->execution, s->wait_ms);
        }
        qpe::hip_check(hipEventRecord(e1, s->stream), "event record");
        qpe::hip_check(hipEventSynchronize(e1), "evolve sync");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        finish_evolve(s, j_begin == 1 ? ms : (float)s->t_evolve + ms);
        s->evolved_to = j_end;
    });
}

int qpe_time_evolve_iqpe(qpe_state_t s, qpe_hamiltonian_t h, uint64_t num_of_applications) {
    return guarded([&] {
        check_state(s);
        require(s->layout == QPE_LAYOUT_ALPHA_BETA, QPE_ERR_ARG, "time_evolve_iqpe needs the alpha x beta layout");
        prepare_evolve(s, h, 1);
        hipEvent_t e0, e1;
        qpe::hip_check(hipEventCreate(&e0), "event");
        qpe::hip_check(hipEventCreate(&e1), "event");
        qpe::hip_check(hipEventRecord(e0, s->stream), "event record");
        qpe::evolve_in_place(*s->plan, s->amp, 1, num_of_applications, s->norms, s->scale, s->status, s->stream,
                             &s->launches);
        qpe::hip_check(hipEventRecord(e1, s->stream), "event record");
        qpe::hip_check(hipEventSynchronize(e1), "evolve sync");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const int renorm = s->renorm;
        s->renorm = 0;   // no renormalisation in place: the scale factors stay 1
        finish_evolve(s, ms);
        s->renorm = renorm;
        s->evolved_to = 0;   // block 1 no longer continues a chain
    });
}

// ---- batched evolution --------------------------------------------------------

int qpe_batch_time_evolve(const qpe_state_t *states, const qpe_hamiltonian_t *hams, int count, int flags,
                          double *elapsed_ms) {
    return guarded([&] {
        require(states != nullptr && hams != nullptr && count >= 1, QPE_ERR_ARG, "empty batch");
        require((flags & ~(QPE_BATCH_EAGER | QPE_BATCH_PASSES)) == 0, QPE_ERR_ARG, "unknown batch flags");
        for (int i = 0; i < count; ++i) {
            check_state(states[i]);
            for (int k = 0; k < i; ++k) require(states[k] != states[i], QPE_ERR_ARG, "a state appears twice in the batch");
            require(states[i]->dtype == states[0]->dtype, QPE_ERR_ARG, "batch members must share the amplitude type");
            require(states[i]->M == states[0]->M, QPE_ERR_ARG, "batch members must share the ancilla count");
            require(states[i]->layout == QPE_LAYOUT_ALPHA_BETA, QPE_ERR_ARG, "batch members need the alpha x beta layout");
        }
        for (int i = 0; i < count; ++i) prepare_evolve(states[i], hams[i], states[i]->term_order);
        for (int i = 0; i < count; ++i) qpe::hip_check(hipStreamSynchronize(states[i]->stream), "batch prepare sync");
        std::lock_guard<std::mutex> lock(g_batch_mu);
        std::vector<uint64_t> key;
        for (int i = 0; i < count; ++i) {
            key.push_back(states[i]->id);
            key.push_back(states[i]->plan_id);
            key.push_back((uint64_t)states[i]->renorm);
        }
        const bool use_flow = (flags & QPE_BATCH_PASSES) == 0;
        key.push_back(use_flow ? 1u : 0u);
        if (!g_batch.run || g_batch.key != key) {
            std::vector<qpe::BatchMember> mem;
            for (int i = 0; i < count; ++i) {
                qpe::BatchMember b;
                b.dp = states[i]->plan.get();
                b.state = states[i]->amp;
                b.norms = states[i]->norms;
                b.scale = states[i]->scale;
                b.renorm = states[i]->renorm;
                b.status = states[i]->status;
                mem.push_back(b);
            }
            g_batch.run.reset(new qpe::BatchRun());
            g_batch.key.clear();
            g_batch.run->build(mem, states[0]->M, states[0]->stream, use_flow);
            g_batch.key = key;
        }
        hipStream_t st = states[0]->stream;
        const bool graph = (flags & QPE_BATCH_EAGER) == 0;
        if (graph) g_batch.run->instantiate();   // capture and instantiation stay outside the timed region
        hipEvent_t e0, e1;
        qpe::hip_check(hipEventCreate(&e0), "event");
        qpe::hip_check(hipEventCreate(&e1), "event");
        qpe::hip_check(hipEventRecord(e0, st), "event record");
        g_batch.run->run(st, graph);
        qpe::hip_check(hipEventRecord(e1, st), "event record");
        qpe::hip_check(hipEventSynchronize(e1), "batch sync");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (elapsed_ms) *elapsed_ms = ms;
        for (int i = 0; i < count; ++i) {
            states[i]->launches = g_batch.run->flow ? 1 : (int)g_batch.run->launches.size() + 1;
            finish_evolve(states[i], ms);
            states[i]->evolved_to = states[i]->M;
        }
    });
}

// ---- autotuner ------------------------------------------------------------------

}  // extern "C"

namespace {
// one timed configuration of an instance
struct TuneCandidate {
    int64_t budget;   // QPE_OPT_MAX_TILE_BYTES
    int engine;       // QPE_OPT_ENGINE
    double us;        // one Trotter step
    int wgs;          // workgroups of the widest pass
    int passes;       // passes per step
};

// Times one Trotter step of every candidate tile width (lines per workgroup, from the widest
// that fits down to 1) and grouping of instance (s, h) on a two-block scratch state, with
// the given execution (QPE_OPT_EXECUTION; a candidate the dataflow kernel cannot run is
// timed pass by pass).
std::vector<TuneCandidate> tune_candidates(qpe_state_t s, qpe_hamiltonian_t h, int execution) {
    check_state(s);
    check_ham(h);
    require(h->h.indexed(), QPE_ERR_STATE, "IndexTerms has not been called");
    require(s->layout == QPE_LAYOUT_ALPHA_BETA, QPE_ERR_ARG, "the autotuner needs the alpha x beta layout");
    const size_t full = (size_t)s->max_lds > qpe::kLdsHeader ? (size_t)s->max_lds - qpe::kLdsHeader : 0;
    const size_t line = (size_t)s->C * s->elem;
    std::vector<size_t> budgets;
    for (size_t w = std::min<size_t>(full / line, 64); w >= 1; w /= 2) budgets.push_back(std::min(full, w * line + 64));
    qpe_state_t t = nullptr;
    int rc = qpe_state_create(&t, 1, s->n, s->e, s->dtype);
    if (rc != QPE_OK) throw ApiError(rc, "autotune scratch state: " + std::string(qpe_last_error()));
    std::unique_ptr<qpe_state, int (*)(qpe_state_t)> guard(t, qpe_state_destroy);
    const double one[2] = {1.0, 0.0};
    std::vector<TuneCandidate> out;
    for (size_t b : budgets) {
        for (int engine = 0; engine < 2; ++engine) {
            t->max_tile_bytes = (int64_t)b;
            t->engine = engine;
            t->order = s->order;
            t->term_order = s->term_order;
            t->execution = execution;
            t->plan_version = ~0ull;
            t->plan.reset();
            double us = 0.0;
            bool ok = true;
            for (int rep = 0; rep < 3 && ok; ++rep) {   // first run builds the plan
                if (rep == 0 && t->execution == qpe::EXEC_DATAFLOW && qpe_state_set(t, 0, 1, one) == QPE_OK &&
                    qpe_time_evolve_over_terms(t, h) == QPE_ERR_ARG)
                    t->execution = qpe::EXEC_PASSES;   // global-memory windows: no dataflow form
                if (qpe_state_set(t, 0, 1, one) != QPE_OK || qpe_time_evolve_over_terms(t, h) != QPE_OK) ok = false;
                else if (rep > 0) us += t->t_evolve * 1e3 / 2.0;
            }
            if (!ok) continue;   // e.g. a budget too small for a mixed double excitation
            TuneCandidate c;
            c.budget = (int64_t)b;
            c.engine = engine;
            c.us = us;
            c.wgs = 0;
            for (const auto &ps : t->plan->plan.passes) c.wgs = std::max(c.wgs, ps.item_count);
            c.passes = (int)t->plan->plan.passes.size();
            out.push_back(c);
        }
    }
    require(!out.empty(), QPE_ERR_ARG, "no tile configuration fits this instance");
    return out;
}

void apply_choice(qpe_state_t s, const TuneCandidate &c, double *result) {
    s->max_tile_bytes = c.budget;
    s->engine = c.engine;
    s->plan_version = ~0ull;
    if (result) {
        result[0] = (double)c.budget;
        result[1] = (double)c.engine;
        result[2] = c.us;
    }
}
}  // namespace

extern "C" {

int qpe_autotune(qpe_state_t s, qpe_hamiltonian_t h, int share_cus, double *result) {
    return guarded([&] {
        require(share_cus >= 1, QPE_ERR_ARG, "share_cus must be positive");
        const std::vector<TuneCandidate> cands = tune_candidates(s, h, s->execution);
        size_t best = 0;
        double best_cost = 1e300;
        for (size_t i = 0; i < cands.size(); ++i) {
            const double cost = cands[i].us * std::max(1.0, (double)cands[i].wgs / (double)share_cus);
            if (cost < best_cost) {
                best_cost = cost;
                best = i;
            }
        }
        apply_choice(s, cands[best], result);
    });
}

int qpe_autotune_batch(const qpe_state_t *states, const qpe_hamiltonian_t *hams, int count, double *results) {
    return guarded([&] {
        require(states != nullptr && hams != nullptr && count >= 1, QPE_ERR_ARG, "empty batch");
        std::vector<std::vector<TuneCandidate>> cands;
        // an instance of a dataflow batch runs its own chain: its candidates are timed that way
        for (int i = 0; i < count; ++i) cands.push_back(tune_candidates(states[i], hams[i], qpe::EXEC_DATAFLOW));
        // fastest candidate of instance i with at most max_passes passes per step (none: the fewest passes)
        auto fastest = [&](int i, int max_passes) {
            size_t best = 0;
            bool found = false;
            for (size_t k = 0; k < cands[i].size(); ++k) {
                const TuneCandidate &c = cands[i][k];
                if (c.passes > max_passes) continue;
                if (!found || c.us < cands[i][best].us) best = k;
                found = true;
            }
            if (!found)
                for (size_t k = 0; k < cands[i].size(); ++k)
                    if (cands[i][k].passes < cands[i][best].passes) best = k;
            return best;
        };
        // Four policies, each timed as one batched step on two-block scratch states:
        //  A  every instance at its fewest passes (the fastest of those): the fewest launches;
        //  B  the instance whose fastest step is slowest at that fastest candidate, every other
        //     instance at its fastest candidate with no more passes than that one;
        //  C  every instance at its fastest candidate (the dataflow batch runs each instance's
        //     chain at its own pace);
        //  D  every instance at half its widest tile (fused groups): twice the items per pass,
        //     so the instances of a dataflow batch share the CUs in smaller pieces (measured at
        //     48 instances, 2^8 outcomes: 3.24 s complex64 and 2.97 s complex128 against 3.65 s
        //     and 3.39 s with the widest tiles).
        std::vector<size_t> pol_a(count), pol_b(count), pol_c(count), pol_d(count);
        for (int i = 0; i < count; ++i) {
            pol_c[i] = fastest(i, INT32_MAX);
            int64_t widest = 0;
            for (const auto &c : cands[i]) widest = std::max(widest, c.budget);
            size_t pick = pol_c[i];
            bool found = false;
            for (size_t k = 0; k < cands[i].size(); ++k) {
                const TuneCandidate &c = cands[i][k];
                if (c.engine != 0 || c.budget > widest / 2 + 64) continue;
                if (!found || c.budget > cands[i][pick].budget) pick = k;
                found = true;
            }
            pol_d[i] = pick;
        }
        for (int i = 0; i < count; ++i) {
            int minp = INT32_MAX;
            for (const auto &c : cands[i]) minp = std::min(minp, c.passes);
            pol_a[i] = fastest(i, minp);
        }
        int crit = 0;
        double crit_us = -1.0;
        for (int i = 0; i < count; ++i) {
            const double us = cands[i][fastest(i, INT32_MAX)].us;
            if (us > crit_us) {
                crit_us = us;
                crit = i;
            }
        }
        const size_t cbest = fastest(crit, INT32_MAX);
        for (int i = 0; i < count; ++i) pol_b[i] = i == crit ? cbest : fastest(i, cands[crit][cbest].passes);
        auto time_policy = [&](const std::vector<size_t> &pol) {
            std::vector<qpe_state_t> t(count, nullptr);
            auto cleanup = [&] {
                for (qpe_state_t x : t)
                    if (x) qpe_state_destroy(x);
            };
            try {
                const double one[2] = {1.0, 0.0};
                for (int i = 0; i < count; ++i) {
                    // three steps per timed batch, so that steps overlap as in a long chain
                    int rc = qpe_state_create(&t[i], 2, states[i]->n, states[i]->e, states[i]->dtype);
                    if (rc != QPE_OK) throw ApiError(rc, "autotune scratch state: " + std::string(qpe_last_error()));
                    t[i]->max_tile_bytes = cands[i][pol[i]].budget;
                    t[i]->engine = cands[i][pol[i]].engine;
                    t[i]->order = states[i]->order;
                    t[i]->term_order = states[i]->term_order;
                    t[i]->renorm = states[i]->renorm;
                }
                double best_ms = 1e300;
                for (int rep = 0; rep < 3; ++rep) {   // the first run builds plans and the graph
                    for (int i = 0; i < count; ++i)
                        if (qpe_state_set(t[i], 0, 1, one) != QPE_OK)
                            throw ApiError(QPE_ERR_STATE, "autotune scratch state: " + std::string(qpe_last_error()));
                    double ms = 0.0;
                    int rc = qpe_batch_time_evolve(t.data(), hams, count, 0, &ms);
                    if (rc != QPE_OK) throw ApiError(rc, "autotune batch: " + std::string(qpe_last_error()));
                    if (rep > 0) best_ms = std::min(best_ms, ms);
                }
                cleanup();
                return best_ms;
            } catch (...) {
                cleanup();
                throw;
            }
        };
        const double ms_a = time_policy(pol_a);
        const double ms_b = pol_b == pol_a ? ms_a : time_policy(pol_b);
        const double ms_c = pol_c == pol_a ? ms_a : pol_c == pol_b ? ms_b : time_policy(pol_c);
        const double ms_d = time_policy(pol_d);
        const std::vector<size_t> *pol = &pol_a;
        double best = ms_a;
        if (ms_b < best) { best = ms_b; pol = &pol_b; }
        if (ms_c < best) { best = ms_c; pol = &pol_c; }
        if (ms_d < best) { best = ms_d; pol = &pol_d; }
        for (int i = 0; i < count; ++i) apply_choice(states[i], cands[i][(*pol)[i]], results ? results + 3 * i : nullptr);
    });
}

int qpe_qift(qpe_state_t s) {
    return guarded([&] {
        check_state(s);
        // 2^m > 1024: the QIFT's second step writes a second register (freed again below);
        // without room for it, 2^m <= 4096 falls back to the radix-2 stages in LDS
        int form = s->qift_form;
        void *work = nullptr;
        if (qpe::qift_needs_work(form, (int)s->m)) {
            const size_t bytes = s->M * s->D * s->elem;
            size_t free_b = 0, total_b = 0;
            qpe::hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
            if (bytes + (64ull << 20) < free_b && hipMalloc(&work, bytes) == hipSuccess) {
                // room for the second register
            } else {
                (void)hipGetLastError();
                work = nullptr;
                require(form == qpe::QIFT_AUTO && s->m <= 12, QPE_ERR_OOM,
                        "the QIFT of 2^precision > 1024 needs a second register's worth of free device memory");
                form = qpe::QIFT_LDS_RADIX2;
            }
        }
        std::unique_ptr<void, hipError_t (*)(void *)> work_guard(work, hipFree);
        hipEvent_t e0, e1;
        qpe::hip_check(hipEventCreate(&e0), "event");
        qpe::hip_check(hipEventCreate(&e1), "event");
        qpe::hip_check(hipEventRecord(e0, s->stream), "event record");
        qpe::qift(s->dtype, s->amp, (int)s->m, s->D, s->scale, s->pk_partial, &s->pk_grid, s->twiddles,
                  (size_t)s->max_lds, s->stream, s->pk_rows, form, work);
        qpe::hip_check(hipEventRecord(e1, s->stream), "event record");
        qpe::hip_check(hipEventSynchronize(e1), "qift sync");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        s->t_qift = ms;
        set_ones(s->scale, s->M, s->stream);  // the QIFT applied them
        s->scale_pending = false;
        s->pk_valid = true;
    });
}

int qpe_qift_noalloc(qpe_state_t s, double *probs_device) {
    return guarded([&] {
        check_state(s);
        require(probs_device != nullptr, QPE_ERR_ARG, "null probs");
        require(s->M <= qpe::kNoAllocScratchDoubles, QPE_ERR_ARG,
                "2^precision exceeds the allocation-free QIFT's limit of 2^20 outcomes");
        int form = s->qift_form;
        if (qpe::qift_needs_work(form, (int)s->m)) {
            // the register form above 1024 outcomes needs a second register: radix-2 stages instead
            require(form == qpe::QIFT_AUTO && s->m <= 12, QPE_ERR_ARG,
                    "the allocation-free QIFT covers 2^precision <= 4096");
            form = qpe::QIFT_LDS_RADIX2;
        }
        hipPointerAttribute_t at;
        std::memset(&at, 0, sizeof(at));
        const hipError_t pe = hipPointerGetAttributes(&at, probs_device);
        if (pe != hipSuccess) (void)hipGetLastError();   // the failed query must not surface in a later check
        require(pe == hipSuccess && (at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged), QPE_ERR_ARG,
                "probs must point to device memory (2^precision doubles)");
        qpe::qift_noalloc(s->dtype, s->amp, (int)s->m, s->D, s->scale, s->twiddles, (size_t)s->max_lds, probs_device,
                          s->stream, form);
        qpe::hip_check(hipStreamSynchronize(s->stream), "qift sync");
        s->scale_pending = false;   // the QIFT applied the factors and reset them to 1
        s->pk_valid = false;        // no partial sums of this QIFT are kept with the state
    });
}

int qpe_identify_phase_amplitudes(qpe_state_t s, double *probs) {
    return guarded([&] {
        check_state(s);
        require(probs != nullptr, QPE_ERR_ARG, "null probs");
        hipEvent_t e0, e1;
        qpe::hip_check(hipEventCreate(&e0), "event");
        qpe::hip_check(hipEventCreate(&e1), "event");
        qpe::hip_check(hipEventRecord(e0, s->stream), "event record");
        if (s->pk_valid) {
            qpe::reduce_pk(s->pk_partial, s->pk_grid, (int)s->m, s->probs, s->stream);
        } else {
            // P(k) = squared norm of block k of the current register
            materialize_scale(s);
            for (uint64_t j = 0; j < s->M; ++j)
                qpe::block_norm(s->dtype == QPE_COMPLEX128 ? 0 : 1, s->D, s->amp, j, s->probs, s->scale, s->norm_partials, kNormParts, s->stream);
            set_ones(s->scale, s->M, s->stream);
        }
        qpe::hip_check(hipEventRecord(e1, s->stream), "event record");
        qpe::hip_check(hipMemcpyAsync(probs, s->probs, s->M * sizeof(double), hipMemcpyDeviceToHost, s->stream), "probs");
        qpe::hip_check(hipStreamSynchronize(s->stream), "sync");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        s->t_pk = ms;
    });
}

int qpe_state_timings(qpe_state_t s, double out[5]) {
    return guarded([&] {
        check_state(s);
        require(out != nullptr, QPE_ERR_ARG, "null output");
        out[0] = s->t_evolve;
        out[1] = s->t_qift;
        out[2] = s->t_pk;
        out[3] = s->passes;
        out[4] = s->launches;
    });
}

int qpe_state_block_norms(qpe_state_t s, double *norms) {
    return guarded([&] {
        check_state(s);
        require(norms != nullptr, QPE_ERR_ARG, "null output");
        qpe::hip_check(hipMemcpyAsync(norms, s->norms, s->M * sizeof(double), hipMemcpyDeviceToHost, s->stream), "norms");
        qpe::hip_check(hipStreamSynchronize(s->stream), "sync");
    });
}

int qpe_plan_info(qpe_hamiltonian_t h, int dtype, uint64_t tile_bytes, double out[8]) {
    return guarded([&] {
        check_ham(h);
        require(out != nullptr, QPE_ERR_ARG, "null output");
        require(dtype == QPE_COMPLEX128 || dtype == QPE_COMPLEX64, QPE_ERR_ARG, "unknown dtype");
        require(h->h.indexed(), QPE_ERR_STATE, "IndexTerms has not been called");
        size_t budget = tile_bytes ? (size_t)tile_bytes : (size_t)(160 * 1024 - qpe::kLdsHeader);
        qpe::Plan p;
        try {
            p = qpe::build_plan(h->h, dtype == QPE_COMPLEX128 ? 16 : 8, budget);
        } catch (const std::runtime_error &e) {
            throw ApiError(QPE_ERR_ARG, e.what());
        }
        out[0] = (double)p.passes.size();
        out[1] = (double)p.passes.size() + 1;
        out[2] = (double)p.pair_updates;
        out[3] = (double)p.diag_updates;
        out[4] = (double)p.tile_bytes_moved;
        out[5] = (double)p.C;
        out[6] = (double)p.groups.size();
        out[7] = (double)p.terms.size();
    });
}

int qpe_debug_pairs(int kind, int n, int e, const int32_t *args, uint32_t *idx01, uint32_t *idx10, uint32_t *parity,
                    uint64_t cap, uint64_t *count) {
    return guarded([&] {
        require(args && idx01 && idx10 && parity && count, QPE_ERR_ARG, "null argument");
        require(n >= 2 && n <= 32 && e >= 1 && e < n, QPE_ERR_ARG, "bad (n, e)");
        uint64_t c;
        if (kind == 2) c = choose(n - 2, e - 1);
        else if (kind == 3) c = e >= 2 ? choose(n - 3, e - 2) : 0;
        else if (kind == 4) c = e >= 2 ? choose(n - 4, e - 2) : 0;
        else throw ApiError(QPE_ERR_ARG, "kind must be 2 (CA), 3 (CNA) or 4 (CCAA)");
        int na = kind == 2 ? 2 : kind == 3 ? 3 : 4;
        for (int i = 0; i < na; ++i) require(args[i] >= 0 && args[i] < n, QPE_ERR_ARG, "orbital out of range");
        *count = c;
        require(cap >= c, QPE_ERR_ARG, "capacity too small");
        if (c == 0) return;
        int32_t a4[4] = {args[0], args[1], na > 2 ? args[2] : 0, na > 3 ? args[3] : 0};
        uint32_t *d = nullptr;
        qpe::hip_check(hipMalloc(reinterpret_cast<void **>(&d), 3 * c * sizeof(uint32_t)), "hipMalloc");
        try {
            qpe::debug_pairs(kind, n, e, a4, d, d + c, d + 2 * c, c, nullptr);
            qpe::hip_check(hipMemcpy(idx01, d, c * sizeof(uint32_t), hipMemcpyDeviceToHost), "copy");
            qpe::hip_check(hipMemcpy(idx10, d + c, c * sizeof(uint32_t), hipMemcpyDeviceToHost), "copy");
            qpe::hip_check(hipMemcpy(parity, d + 2 * c, c * sizeof(uint32_t), hipMemcpyDeviceToHost), "copy");
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        (void)hipFree(d);
    });
}

}  // extern "C"
