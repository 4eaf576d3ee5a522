"""Python binding of libqpesim.so (ctypes over include/qpesim.h) and the
seeded synthetic-Hamiltonian generator used by the tests and the benchmark.

The binding mirrors the reference call sequence (christopherkang/2022-qpesim,
cpp/main.cc:99-134):

    h = Hamiltonian(); <add integrals>; h.index_terms(); h.set_timestep(tau)
    sv = StateVector(m, n_orb, n_occ, dtype); sv.set(0, psi0); sv.scale_block(0, 1/sqrt(2^m))
    time_evolve_over_terms(sv, h); sv.qift(); probs = sv.identify_phase_amplitudes()

There is no CPU fallback: if the library is missing this module raises.
"""
import ctypes
import itertools
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libqpesim.so")

COMPLEX128, COMPLEX64 = 0, 1
LAYOUT_ALPHA_BETA, LAYOUT_SINGLE_STRING = 0, 1
PP, PQ, PQQP, PQQR, PQRS = range(5)
OPT_RENORMALIZE, OPT_ENGINE, OPT_MAX_TILE_BYTES, OPT_TROTTER_ORDER, OPT_TERM_ORDER, OPT_QIFT_FORM = 0, 1, 2, 3, 4, 5
OPT_EXECUTION, OPT_WAIT_LIMIT_MS = 6, 7
EXEC_AUTO, EXEC_PASSES, EXEC_DATAFLOW = 0, 1, 2
QIFT_AUTO, QIFT_REGISTER, QIFT_MFMA, QIFT_LDS_RADIX2 = 0, 1, 2, 3
ERR_ARG, ERR_OOM, ERR_HIP, ERR_NUMERIC, ERR_STATE, ERR_STALLED = 1, 2, 3, 4, 5, 6

# every entry point include/qpesim.h declares
EXPORTS = [
    "qpe_hamiltonian_create", "qpe_hamiltonian_destroy", "qpe_hamiltonian_add_one_body_term",
    "qpe_hamiltonian_add_two_body_term", "qpe_hamiltonian_set_num_orbitals", "qpe_hamiltonian_set_num_occupied",
    "qpe_hamiltonian_index_terms", "qpe_hamiltonian_set_timestep", "qpe_hamiltonian_term_counts",
    "qpe_hamiltonian_get_terms", "qpe_state_create", "qpe_state_destroy", "qpe_state_size", "qpe_state_set",
    "qpe_state_get", "qpe_state_scale_block", "qpe_time_evolve_over_terms", "qpe_qift",
    "qpe_identify_phase_amplitudes", "qpe_state_set_option", "qpe_state_timings", "qpe_state_block_norms",
    "qpe_debug_pairs", "qpe_plan_info", "qpe_batch_time_evolve", "qpe_autotune", "qpe_time_evolve_blocks", "qpe_last_error",
    "qpe_hamiltonian_ingest_broombridge", "qpe_hamiltonian_ingest_broombridge_text", "qpe_hamiltonian_initial_state",
    "qpe_hamiltonian_identity_term", "qpe_hamiltonian_get_sizes", "qpe_state_ingest_broombridge", "qpe_energy_estimate",
    "qpe_time_evolve_iqpe", "qpe_state_apply_gate", "qpe_state_apply_controlled_gate", "qpe_qift_noalloc",
    "qpe_state_create_layout", "qpe_autotune_batch",
]
BATCH_EAGER, BATCH_PASSES = 1, 2

_lib = None


class QpeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qpesim error {code}: {msg}")
        self.code = code


def load():
    """Load libqpesim.so (built by `make -C 2022-qpesim_amd`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {_HERE}` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    vp, u64, i64, dbl, ci = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64, ctypes.c_double, ctypes.c_int
    P = ctypes.POINTER
    sig = {
        "qpe_hamiltonian_create": [P(vp)], "qpe_hamiltonian_destroy": [vp],
        "qpe_hamiltonian_add_one_body_term": [vp, ci, ci, dbl],
        "qpe_hamiltonian_add_two_body_term": [vp, ci, ci, ci, ci, dbl],
        "qpe_hamiltonian_set_num_orbitals": [vp, ctypes.c_uint], "qpe_hamiltonian_set_num_occupied": [vp, ctypes.c_uint],
        "qpe_hamiltonian_index_terms": [vp], "qpe_hamiltonian_set_timestep": [vp, dbl],
        "qpe_hamiltonian_term_counts": [vp, P(u64)],
        "qpe_hamiltonian_get_terms": [vp, ci, P(ctypes.c_int32), P(dbl), u64],
        "qpe_state_create": [P(vp), ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ci],
        "qpe_state_destroy": [vp], "qpe_state_size": [vp, P(u64)],
        "qpe_state_set": [vp, u64, u64, P(dbl)], "qpe_state_get": [vp, u64, u64, P(dbl)],
        "qpe_state_scale_block": [vp, u64, dbl], "qpe_time_evolve_over_terms": [vp, vp], "qpe_qift": [vp],
        "qpe_identify_phase_amplitudes": [vp, P(dbl)], "qpe_state_set_option": [vp, ci, i64],
        "qpe_state_timings": [vp, P(dbl)], "qpe_state_block_norms": [vp, P(dbl)],
        "qpe_debug_pairs": [ci, ci, ci, P(ctypes.c_int32), P(ctypes.c_uint32), P(ctypes.c_uint32),
                            P(ctypes.c_uint32), u64, P(u64)],
        "qpe_plan_info": [vp, ci, u64, P(dbl)],
        "qpe_batch_time_evolve": [P(vp), P(vp), ci, ci, P(dbl)],
        "qpe_autotune": [vp, vp, ci, P(dbl)],
        "qpe_time_evolve_blocks": [vp, vp, u64, u64],
        "qpe_hamiltonian_ingest_broombridge": [vp, ctypes.c_char_p],
        "qpe_hamiltonian_ingest_broombridge_text": [vp, ctypes.c_char_p, u64],
        "qpe_hamiltonian_initial_state": [vp, P(u64), P(dbl), u64, P(u64)],
        "qpe_hamiltonian_identity_term": [vp, P(dbl)],
        "qpe_hamiltonian_get_sizes": [vp, P(ctypes.c_uint), P(ctypes.c_uint)],
        "qpe_state_ingest_broombridge": [P(vp), vp, ctypes.c_char_p, ctypes.c_uint, ci],
        "qpe_energy_estimate": [P(dbl), ctypes.c_uint, dbl, dbl, P(dbl)],
        "qpe_time_evolve_iqpe": [vp, vp, u64],
        "qpe_state_apply_gate": [vp, P(dbl), ctypes.c_uint],
        "qpe_state_apply_controlled_gate": [vp, P(dbl), ctypes.c_uint, ctypes.c_uint],
        "qpe_qift_noalloc": [vp, vp],
        "qpe_state_create_layout": [P(vp), ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ci, ci],
        "qpe_autotune_batch": [P(vp), P(vp), ci, P(dbl)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ci
    L.qpe_last_error.argtypes = []
    L.qpe_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _call(name, *args):
    L = load()
    rc = getattr(L, name)(*args)
    if rc != 0:
        raise QpeError(rc, L.qpe_last_error().decode())


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _gate_words(gate):
    g = np.asarray(gate, dtype=np.complex128).reshape(2, 2)
    return np.ascontiguousarray(np.stack([g.real, g.imag], axis=-1).ravel())


def hadamard():
    """Hadamard(), cpp/gate.hpp:22-25."""
    r = 1.0 / math.sqrt(2.0)
    return np.array([[r, r], [r, -r]], np.complex128)


def pauli_x():
    """Gate1D(0, 1, 1, 0), the X of the reference's gate tests (cpp/state_vector.hpp:390)."""
    return np.array([[0, 1], [1, 0]], np.complex128)


def rotate(m):
    """Rotate(m), cpp/gate.hpp:27-32: diag(1, exp(2 pi i / 2^m))."""
    return np.array([[1, 0], [0, np.exp(2j * math.pi / 2.0 ** m)]], np.complex128)


def inv_rotate(m):
    """InvRotate(m), cpp/gate.hpp:34-39: diag(1, exp(-2 pi i / 2^m))."""
    return np.array([[1, 0], [0, np.exp(-2j * math.pi / 2.0 ** m)]], np.complex128)


class Hamiltonian:
    """qpe_hamiltonian_t: the reference's Hamiltonian (cpp/hamiltonian.hpp:24-184)."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        _call("qpe_hamiltonian_create", ctypes.byref(self._h))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None and self._h.value:
            _lib.qpe_hamiltonian_destroy(self._h)
            self._h = ctypes.c_void_p()

    def add_one_body_term(self, p, q, coeff):
        _call("qpe_hamiltonian_add_one_body_term", self._h, int(p), int(q), float(coeff))

    def add_two_body_term(self, p, q, r, s, coeff):
        _call("qpe_hamiltonian_add_two_body_term", self._h, int(p), int(q), int(r), int(s), float(coeff))

    def set_num_orbitals(self, n):
        _call("qpe_hamiltonian_set_num_orbitals", self._h, int(n))

    def set_num_occupied(self, e):
        _call("qpe_hamiltonian_set_num_occupied", self._h, int(e))

    def index_terms(self):
        _call("qpe_hamiltonian_index_terms", self._h)

    def set_timestep(self, t):
        _call("qpe_hamiltonian_set_timestep", self._h, float(t))

    def term_counts(self):
        c = (ctypes.c_uint64 * 5)()
        _call("qpe_hamiltonian_term_counts", self._h, c)
        return [int(x) for x in c]

    def terms(self, cls):
        n = self.term_counts()[cls]
        idx = np.zeros((max(n, 1), 4), np.int32)
        co = np.zeros(max(n, 1), np.float64)
        _call("qpe_hamiltonian_get_terms", self._h, int(cls), idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
              _dptr(co), max(n, 1))
        return idx[:n], co[:n]

    def sum_abs_coeff(self):
        return float(sum(np.abs(self.terms(c)[1]).sum() for c in range(5)))

    def ingest_broombridge(self, path, precision=5, dtype=COMPLEX128):
        """Hamiltonian::IngestBroombridge(path, precision), cpp/hamiltonian.cc:8-144:
        adds the file's integrals and returns the StateVector of 2^precision blocks
        whose block 0 holds the file's initial state."""
        s = ctypes.c_void_p()
        _call("qpe_state_ingest_broombridge", ctypes.byref(s), self._h, os.fsencode(path), int(precision), int(dtype))
        n, e = self.sizes()
        return StateVector._adopt(s, precision, n, e, dtype)

    def load_broombridge(self, path=None, text=None):
        """The Hamiltonian half of IngestBroombridge (no device state): integrals,
        orbital and electron counts, identity term and initial determinants."""
        if (path is None) == (text is None):
            raise ValueError("give exactly one of path and text")
        if path is not None:
            _call("qpe_hamiltonian_ingest_broombridge", self._h, os.fsencode(path))
        else:
            b = text.encode() if isinstance(text, str) else bytes(text)
            _call("qpe_hamiltonian_ingest_broombridge_text", self._h, b, len(b))

    def initial_state(self):
        """(addresses in block 0, signed coefficients) of the last ingest, file order."""
        cnt = ctypes.c_uint64()
        _call("qpe_hamiltonian_initial_state", self._h, None, None, 0, ctypes.byref(cnt))
        n = int(cnt.value)
        addr = np.zeros(max(n, 1), np.uint64)
        co = np.zeros(max(n, 1), np.float64)
        _call("qpe_hamiltonian_initial_state", self._h, addr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), _dptr(co),
              max(n, 1), ctypes.byref(cnt))
        return addr[:n], co[:n]

    def identity_term(self):
        """Hamiltonian::GetIdentityTerm, cpp/hamiltonian.hpp:56."""
        v = ctypes.c_double()
        _call("qpe_hamiltonian_identity_term", self._h, ctypes.byref(v))
        return v.value

    def sizes(self):
        """(num_orbitals(), num_occupied()), cpp/hamiltonian.hpp:68-69."""
        n, e = ctypes.c_uint(), ctypes.c_uint()
        _call("qpe_hamiltonian_get_sizes", self._h, ctypes.byref(n), ctypes.byref(e))
        return int(n.value), int(e.value)

    @property
    def handle(self):
        return self._h


class StateVector:
    """qpe_state_t: the reference's StateVector (cpp/state_vector.hpp:16-42) resident in HBM."""

    def __init__(self, precision, num_orbitals, num_occupied, dtype=COMPLEX128, layout=LAYOUT_ALPHA_BETA):
        self._s = ctypes.c_void_p()
        _call("qpe_state_create_layout", ctypes.byref(self._s), int(precision), int(num_orbitals), int(num_occupied),
              int(dtype), int(layout))
        self.m = int(precision)
        self.M = 1 << self.m
        self.n = int(num_orbitals)
        self.e = int(num_occupied)
        self.C = math.comb(self.n, self.e)
        self.layout = layout
        self.D = self.C * self.C if layout == LAYOUT_ALPHA_BETA else math.comb(2 * self.n, 2 * self.e)
        self.dtype = dtype

    @classmethod
    def _adopt(cls, handle, precision, num_orbitals, num_occupied, dtype):
        sv = cls.__new__(cls)
        sv._s = handle
        sv.m = int(precision)
        sv.M = 1 << sv.m
        sv.n = int(num_orbitals)
        sv.e = int(num_occupied)
        sv.C = math.comb(sv.n, sv.e)
        sv.layout = LAYOUT_ALPHA_BETA
        sv.D = sv.C * sv.C
        sv.dtype = dtype
        return sv

    def close(self):
        if getattr(self, "_s", None) and _lib is not None and self._s.value:
            _lib.qpe_state_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __del__(self):
        self.close()

    def size(self):
        n = ctypes.c_uint64()
        _call("qpe_state_size", self._s, ctypes.byref(n))
        return int(n.value)

    def set(self, offset, amps):
        a = np.ascontiguousarray(amps, dtype=np.complex128).ravel()
        _call("qpe_state_set", self._s, int(offset), a.size, _dptr(a))

    def get(self, offset=0, count=None):
        if count is None:
            count = self.size() - offset
        out = np.empty(int(count), np.complex128)
        _call("qpe_state_get", self._s, int(offset), int(count), _dptr(out))
        return out

    def scale_block(self, block, factor):
        _call("qpe_state_scale_block", self._s, int(block), float(factor))

    def set_option(self, opt, value):
        _call("qpe_state_set_option", self._s, int(opt), int(value))

    def apply_gate(self, gate, qubit):
        """StateVector::apply_gate, cpp/state_vector.hpp:51-74 (gate: 2x2 complex)."""
        g = _gate_words(gate)
        _call("qpe_state_apply_gate", self._s, _dptr(g), int(qubit))

    def apply_controlled_gate(self, gate, control, qubit):
        """StateVector::apply_controlled_gate, cpp/state_vector.hpp:76-113."""
        g = _gate_words(gate)
        _call("qpe_state_apply_controlled_gate", self._s, _dptr(g), int(control), int(qubit))

    def qift(self):
        _call("qpe_qift", self._s)

    def qift_noalloc(self, probs_device_ptr):
        """qpe_qift_noalloc: QIFT + P(k) into caller-owned device memory
        (probs_device_ptr: address of 2^m doubles on the device), no allocation."""
        _call("qpe_qift_noalloc", self._s, ctypes.c_void_p(int(probs_device_ptr)))

    def identify_phase_amplitudes(self):
        out = np.zeros(self.M, np.float64)
        _call("qpe_identify_phase_amplitudes", self._s, _dptr(out))
        return out

    def timings(self):
        out = np.zeros(5, np.float64)
        _call("qpe_state_timings", self._s, _dptr(out))
        return {"evolve_ms": out[0], "qift_ms": out[1], "pk_ms": out[2], "passes_per_step": int(out[3]),
                "launches_per_step": int(out[4])}

    def block_norms(self):
        out = np.zeros(self.M, np.float64)
        _call("qpe_state_block_norms", self._s, _dptr(out))
        return out

    @property
    def handle(self):
        return self._s


def time_evolve_over_terms(sv, h):
    """time_evolve_over_terms(sv, h, m), cpp/time_evolution_over_terms.hpp:92-139."""
    _call("qpe_time_evolve_over_terms", sv.handle, h.handle)


def time_evolve_blocks(sv, h, j_begin, j_end):
    """Blocks j_begin..j_end-1 of the chain (resumable pieces of time_evolve_over_terms)."""
    _call("qpe_time_evolve_blocks", sv.handle, h.handle, int(j_begin), int(j_end))


def time_evolve_iqpe(sv, h, num_of_applications):
    """time_evolve_iqpe(sv, h, k), cpp/time_evolution_over_terms.hpp:141-155:
    applyUnitary (original term order, every pqqr term) k times in place on block 1."""
    _call("qpe_time_evolve_iqpe", sv.handle, h.handle, int(num_of_applications))


def batch_time_evolve(states, hams, eager=False, passes=False):
    """Batching form: evolve states[i] under hams[i] in one launch sequence (one
    hipGraph launch unless eager): one persistent dataflow launch when every
    member's plan allows it, unless passes (then launch k runs pass k of every
    member).  Returns the device time in ms."""
    n = len(states)
    if n != len(hams) or n == 0:
        raise ValueError("states and hams must be non-empty and of equal length")
    sa = (ctypes.c_void_p * n)(*[sv.handle.value for sv in states])
    ha = (ctypes.c_void_p * n)(*[h.handle.value for h in hams])
    ms = ctypes.c_double(0.0)
    flags = (BATCH_EAGER if eager else 0) | (BATCH_PASSES if passes else 0)
    _call("qpe_batch_time_evolve", sa, ha, n, flags, ctypes.byref(ms))
    return ms.value


def autotune(sv, h, share_cus=256):
    """Pick the tile width and term grouping of sv for h by timing one step of
    each candidate; returns {tile_bytes, engine, us_per_step}."""
    out = np.zeros(3, np.float64)
    _call("qpe_autotune", sv.handle, h.handle, int(share_cus), _dptr(out))
    return {"tile_bytes": int(out[0]), "engine": int(out[1]), "us_per_step": float(out[2])}


def autotune_batch(states, hams):
    """Batching form's autotuner: times every candidate of every instance, then gives
    the instance with the slowest fastest step its fastest candidate and every other
    instance its fastest candidate with no more passes per step.  Returns one
    {tile_bytes, engine, us_per_step} per instance."""
    n = len(states)
    if n != len(hams) or n == 0:
        raise ValueError("states and hams must be non-empty and of equal length")
    sa = (ctypes.c_void_p * n)(*[sv.handle.value for sv in states])
    ha = (ctypes.c_void_p * n)(*[h.handle.value for h in hams])
    out = np.zeros(3 * n, np.float64)
    _call("qpe_autotune_batch", sa, ha, n, _dptr(out))
    return [{"tile_bytes": int(out[3 * i]), "engine": int(out[3 * i + 1]), "us_per_step": float(out[3 * i + 2])}
            for i in range(n)]


def energy_estimate(probs, precision, identity, timestep):
    """Phase post-processing of cpp/main.cc:149-167: phase = sum_k k P(k) / 2^m;
    `energy` is the reference's estimate -2 pi phase + identity (no division by
    the timestep), `energy_per_time` = -2 pi phase / t + identity, `moe` the
    reference's 2 pi / (t 2^m)."""
    p = np.ascontiguousarray(probs, dtype=np.float64)
    if p.size != 1 << int(precision):
        raise ValueError("probs must hold 2^precision values")
    out = np.zeros(3, np.float64)
    _call("qpe_energy_estimate", _dptr(p), int(precision), float(identity), float(timestep), _dptr(out))
    return {"phase": float(np.dot(np.arange(p.size, dtype=np.float64), p) / p.size), "energy": float(out[0]),
            "energy_per_time": float(out[1]), "moe": float(out[2])}


def run_main(path, precision, steps, dtype=COMPLEX128):
    """The reference program's call sequence (cpp/main.cc:99-167) on one
    Broombridge file: ingest, IndexTerms(HalfUp), SetTimestep(1/steps), block 0
    scaled by 1/sqrt(2^m), time_evolve_over_terms, qift, IdentifyPhaseAmplitudes,
    energy estimate."""
    h = Hamiltonian()
    sv = h.ingest_broombridge(path, precision, dtype)
    try:
        h.index_terms()
        t = 1.0 / float(steps)
        h.set_timestep(t)
        sv.scale_block(0, 1.0 / math.sqrt(1 << int(precision)))
        time_evolve_over_terms(sv, h)
        sv.qift()
        probs = sv.identify_phase_amplitudes()
        est = energy_estimate(probs, precision, h.identity_term(), t)
        est.update(probs=probs, timings=sv.timings(), num_orbitals=sv.n, num_occupied=sv.e)
        return est
    finally:
        sv.close()


def plan_info(h, dtype=COMPLEX128, tile_bytes=0):
    """Host-only summary of the tile program of one Trotter step."""
    out = np.zeros(8, np.float64)
    _call("qpe_plan_info", h.handle, int(dtype), int(tile_bytes), _dptr(out))
    keys = ["passes", "launches", "pair_updates", "diag_updates", "pass_bytes", "C", "groups", "double_terms"]
    return {k: (int(v) if k != "pass_bytes" else float(v)) for k, v in zip(keys, out)}


def debug_pairs(kind, n, e, args):
    """Device-computed (idx01, idx10, parity) of every pair of one term."""
    a = np.zeros(4, np.int32)
    a[:len(args)] = args
    cap = math.comb(n, max(0, e)) + 1
    o1 = np.zeros(cap, np.uint32)
    o2 = np.zeros(cap, np.uint32)
    op = np.zeros(cap, np.uint32)
    cnt = ctypes.c_uint64()
    u32p = ctypes.POINTER(ctypes.c_uint32)
    _call("qpe_debug_pairs", int(kind), int(n), int(e), a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
          o1.ctypes.data_as(u32p), o2.ctypes.data_as(u32p), op.ctypes.data_as(u32p), cap, ctypes.byref(cnt))
    k = int(cnt.value)
    return o1[:k], o2[:k], op[:k]


# --------------------------------------------------------------------------
# seeded synthetic Hamiltonians
# --------------------------------------------------------------------------
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, start=0):
    """Outputs start..start+count-1 of the SplitMix64 stream seeded with `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def normals(seed, count, start=0):
    """Standard normals by Box-Muller over pairs of SplitMix64 uniforms;
    normal 2i uses the cosine branch and 2i+1 the sine branch of pair i."""
    off = start % 2
    npairs = (off + count + 1) // 2
    z = splitmix64(seed, 2 * npairs, 2 * (start // 2))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * npairs)
    out[0::2] = r * np.cos(2.0 * np.pi * u2)
    out[1::2] = r * np.sin(2.0 * np.pi * u2)
    return out[off:off + count]


CUTOFF = 1e-8


def synthetic_integrals(n_spatial, seed):
    """Broombridge-style integral lists for the synthetic family:

    one-body  h_pp = -2.0 + 0.25 p,  h_pq ~ N(0, 0.05^2) (p > q, symmetric)
    two-body  (pq|rs) = N(0, 0.1^2) * exp(-0.3 (|p-q| + |r-s|)) on the 8-fold-unique
              set p >= q, r >= s, pq >= rs (compound index), plus 0.5/(1+|p-r|)
              on (pp|rr);  |value| < 1e-8 dropped.
    Orbitals are 0-based in the formulas; entries are 1-based like Broombridge.
    Normals are drawn in this order: one-body off-diagonals (p ascending, q < p
    ascending), then two-body values in (pq, rs) compound order.
    Returns (one_body [(p, q, v)], two_body [(a, b, c, d, v)] Mulliken (ab|cd))."""
    n = n_spatial
    pairs = [(p, q) for p in range(n) for q in range(p + 1)]
    n_off = n * (n - 1) // 2
    n_two = len(pairs) * (len(pairs) + 1) // 2
    g = normals(seed, n_off + n_two)
    one, k = [], 0
    for p in range(n):
        for q in range(p + 1):
            if p == q:
                v = -2.0 + 0.25 * p
            else:
                v = 0.05 * g[k]
                k += 1
            if abs(v) >= CUTOFF:
                one.append((p + 1, q + 1, float(v)))
    two = []
    for i, (p, q) in enumerate(pairs):
        for j in range(i + 1):
            r, s = pairs[j]
            v = 0.1 * g[k] * math.exp(-0.3 * (abs(p - q) + abs(r - s)))
            k += 1
            if p == q and r == s:
                v += 0.5 / (1.0 + abs(p - r))
            if abs(v) >= CUTOFF:
                two.append((p + 1, q + 1, r + 1, s + 1, float(v)))
    return one, two


def build_hamiltonian(n_spatial, n_occupied, one_body, two_body):
    """Feeds Broombridge entries as IngestBroombridge does (cpp/hamiltonian.cc:22-52):
    one-body [p,q,v] swapped to p >= q; two-body [a,b,c,d,v] read as i=a, l=b,
    j=c, k=d and passed as AddTwoBodyTerm(i, j, k, l).  Then IndexTerms and
    tau = 0.5 / sum |coefficient| over the indexed terms."""
    h = Hamiltonian()
    h.set_num_orbitals(n_spatial)
    h.set_num_occupied(n_occupied)
    for p, q, v in one_body:
        if q > p:
            p, q = q, p
        h.add_one_body_term(p, q, v)
    for a, b, c, d, v in two_body:
        h.add_two_body_term(a, c, d, b, v)
    h.index_terms()
    tau = 0.5 / h.sum_abs_coeff()
    h.set_timestep(tau)
    return h, tau


def initial_state(n_spatial, n_occupied, seed=None, hf_weight=1.0):
    """Block-0 amplitudes: Hartree-Fock (rank 0 in both strings, amplitude +1) or
    normalize(hf_weight*HF + (1-hf_weight)*r), r a unit complex Gaussian vector
    over all C(n,e)^2 determinants drawn from the SplitMix64 stream seed+7919."""
    C = math.comb(n_spatial, n_occupied)
    psi = np.zeros(C * C, np.complex128)
    psi[0] = 1.0
    if hf_weight < 1.0:
        g = normals(seed + 7919, 2 * C * C)
        r = g[0::2] + 1j * g[1::2]
        r /= np.linalg.norm(r)
        psi = hf_weight * psi + (1.0 - hf_weight) * r
        psi /= np.linalg.norm(psi)
    return psi


def single_string_map(n_spatial, n_occupied):
    """Where each alpha x beta amplitude lives in the single-string register.

    Returns (index, sign), two arrays over the C(n,e)^2 determinants in the
    reference's order (alpha rank * C + beta rank): the colex rank of the
    interleaved string (bit 2p = alpha orbital p, bit 2p+1 = beta orbital p) among
    the C(2n, 2e) strings, and the sign (-1)^(pairs alpha a, beta b with b < a)
    that reorders the creation operators from alpha-then-beta to interleaved.
    amp_single[index] = sign * amp_alpha_beta."""
    n, e = int(n_spatial), int(n_occupied)
    strings = np.array(sorted(sum(1 << b for b in c) for c in itertools.combinations(range(n), e)), np.int64)
    C = len(strings)
    sa = np.repeat(strings, C)
    sb = np.tile(strings, C)
    index = np.zeros(C * C, np.int64)
    k = np.zeros(C * C, np.int64)    # electrons placed so far (the 1-based order of the next one)
    par = np.zeros(C * C, np.int64)
    binom = np.array([[math.comb(i, j) for j in range(2 * n + 1)] for i in range(2 * n)], np.int64)
    beta_below = np.zeros(C * C, np.int64)   # beta electrons on orbitals below p
    for p in range(n):
        for spin, s in ((0, sa), (1, sb)):
            occ = (s >> p) & 1
            k += occ
            index += np.where(occ == 1, binom[2 * p + spin, k], 0)
        par += ((sa >> p) & 1) * beta_below   # alpha p moves past every beta below it
        beta_below += (sb >> p) & 1
    return index, np.where(par & 1, -1.0, 1.0)


# the workloads, by name: spatial orbitals, electrons per spin, ancillas, seed, HF weight
CONFIGS = {
    "cfg0": dict(n=6, e=3, m=6, seed=0, hf=1.0),        # n=12 spin-orbitals, N=6
    "cfg1": dict(n=14, e=5, m=8, seed=1, hf=0.9),       # n=28, N=10
    "cfg2": dict(n=16, e=4, m=10, seed=2, hf=1.0),      # n=32, N=8
    "cfg4": dict(n=17, e=5, m=7, seed=4, hf=0.95, order=2),   # n=34, N=10, second-order Trotter
}


def batch_instances(count=48, seed=3):
    """48 instances: spin-orbital count n uniform on {12,14,...,24}, electron count
    N uniform on {4,6,8} with N <= n/2 (redrawn otherwise); instance i uses seed 1000+i."""
    z = splitmix64(seed, 4 * count)
    out, k = [], 0
    while len(out) < count:
        n = 12 + 2 * int(z[k % len(z)] % np.uint64(7))
        N = 4 + 2 * int(z[(k + 1) % len(z)] % np.uint64(3))
        k += 2
        if k >= len(z):
            z = splitmix64(seed + k, 4 * count)
            k = 0
        if N <= n // 2:
            out.append(dict(n=n // 2, e=N // 2, m=8, seed=1000 + len(out), hf=1.0))
    return out
