"""ctypes loader for the CPU oracle (oracle/_build/libqpe_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product (libqpesim.so and its Python binding) never imports this module.

The C code restates christopherkang/2022-qpesim's hot path; see the citations
in qpe_oracle.c.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libqpe_oracle.so")
_lib = None

PP, PQ, PQQP, PQQR, PQRS = range(5)
KIND_N, KIND_NN, KIND_CA, KIND_CNA, KIND_CCAA = range(5)


class OrcTerm(ctypes.Structure):
    _fields_ = [("cls", ctypes.c_int32), ("idx", ctypes.c_int32 * 4), ("coeff", ctypes.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u64, i64, i32, dbl = ctypes.c_uint64, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
        P = ctypes.POINTER
        L.orc_nchoosek.restype = u64
        L.orc_nchoosek.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_rank.restype = u64
        L.orc_rank.argtypes = [u64]
        L.orc_next_comb.restype = u64
        L.orc_next_comb.argtypes = [u64]
        L.orc_parity.restype = ctypes.c_uint
        L.orc_parity.argtypes = [u64]
        L.orc_compile_terms.restype = i64
        L.orc_compile_terms.argtypes = [ctypes.c_int, i64, P(i32), P(dbl), i64, P(i32), P(dbl),
                                        P(OrcTerm), i64, P(i64)]
        L.orc_run_table.restype = i64
        L.orc_run_table.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(i32),
                                    P(u64), P(u64), P(u64), P(ctypes.c_uint32), i64]
        L.orc_time_evolve.restype = ctypes.c_int
        L.orc_time_evolve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(OrcTerm), i64,
                                      dbl, P(dbl), i64, i64]
        L.orc_apply_unitary.restype = ctypes.c_int
        L.orc_apply_unitary.argtypes = [ctypes.c_int, ctypes.c_int, P(OrcTerm), i64, dbl, P(dbl)]
        L.orc_apply_unitary_original.restype = ctypes.c_int
        L.orc_apply_unitary_original.argtypes = [ctypes.c_int, ctypes.c_int, P(OrcTerm), i64, dbl, P(dbl)]
        L.orc_time_evolve_ordered.restype = ctypes.c_int
        L.orc_time_evolve_ordered.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(OrcTerm), i64,
                                              dbl, P(dbl), i64, i64, ctypes.c_int]
        L.orc_time_evolve_iqpe.restype = ctypes.c_int
        L.orc_time_evolve_iqpe.argtypes = [ctypes.c_int, ctypes.c_int, P(OrcTerm), i64, dbl, P(dbl), i64]
        L.orc_count_updates.restype = ctypes.c_int
        L.orc_count_updates.argtypes = [ctypes.c_int, ctypes.c_int, P(OrcTerm), i64, P(u64), P(u64)]
        L.orc_qift.restype = ctypes.c_int
        L.orc_qift.argtypes = [ctypes.c_int, u64, P(dbl)]
        L.orc_phase_probabilities.restype = ctypes.c_int
        L.orc_phase_probabilities.argtypes = [ctypes.c_int, u64, P(dbl), P(dbl)]
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def nchoosek(n, k):
    return int(lib().orc_nchoosek(n, k))


def rank(bv):
    return int(lib().orc_rank(bv))


def parity(x):
    return int(lib().orc_parity(x))


def next_comb(x):
    return int(lib().orc_next_comb(x))


class Terms:
    """Indexed terms in the reference's map order, class by class."""

    def __init__(self, arr, counts):
        self.arr = arr            # ctypes array of OrcTerm
        self.counts = counts      # [pp, pq, pqqp, pqqr, pqrs]

    def __len__(self):
        return len(self.arr)

    def as_lists(self):
        out = [[] for _ in range(5)]
        for t in self.arr:
            c = int(t.cls)
            if c in (PP, PQ):
                key = (int(t.idx[0]), int(t.idx[1]))
            else:
                key = tuple(int(v) for v in t.idx)
            out[c].append((key, float(t.coeff)))
        return out

    def sum_abs(self):
        return float(sum(abs(t.coeff) for t in self.arr))


def compile_terms(n_orb, one_body, two_body):
    """one_body: list of (p, q, v) 1-based Broombridge one-electron entries;
    two_body: list of (a, b, c, d, v) 1-based Mulliken entries, in file order."""
    one_pq = np.array([[e[0], e[1]] for e in one_body], dtype=np.int32).reshape(-1, 2)
    one_v = np.array([e[2] for e in one_body], dtype=np.float64)
    two_i = np.array([[e[0], e[1], e[2], e[3]] for e in two_body], dtype=np.int32).reshape(-1, 4)
    two_v = np.array([e[4] for e in two_body], dtype=np.float64)
    cap = 8 * 4 * max(1, len(two_body)) + 2 * max(1, len(one_body)) + 16
    out = (OrcTerm * cap)()
    counts = np.zeros(5, dtype=np.int64)
    n = lib().orc_compile_terms(n_orb, len(one_body), _ptr(one_pq, ctypes.c_int32), _ptr(one_v, ctypes.c_double),
                                len(two_body), _ptr(two_i, ctypes.c_int32), _ptr(two_v, ctypes.c_double),
                                out, cap, _ptr(counts, ctypes.c_int64))
    if n < 0 or n > cap:
        raise RuntimeError("orc_compile_terms failed")
    arr = (OrcTerm * n)()
    ctypes.memmove(arr, out, n * ctypes.sizeof(OrcTerm))
    return Terms(arr, [int(c) for c in counts])


def run_table(kind, n, e, args):
    size = {KIND_N: nchoosek(n - 1, e - 1), KIND_NN: nchoosek(n - 2, e - 2), KIND_CA: nchoosek(n - 2, e - 1),
            KIND_CNA: nchoosek(n - 3, e - 2), KIND_CCAA: nchoosek(n - 4, e - 2)}[kind]
    cap = max(1, size)
    o1 = np.zeros(cap, np.uint64)
    o2 = np.zeros(cap, np.uint64)
    ni = np.zeros(cap, np.uint64)
    mp = np.zeros(cap, np.uint32)
    a = np.array(list(args) + [0] * (4 - len(args)), np.int32)
    k = lib().orc_run_table(kind, n, e, _ptr(a, ctypes.c_int32), _ptr(o1, ctypes.c_uint64),
                            _ptr(o2, ctypes.c_uint64), _ptr(ni, ctypes.c_uint64), _ptr(mp, ctypes.c_uint32), cap)
    if k < 0:
        raise RuntimeError("orc_run_table failed")
    return o1[:k], o2[:k], ni[:k], mp[:k]


def expand_pairs(kind, n, e, args):
    """Pairs (idx01, idx10, parity&1) of one term, expanded from the run table."""
    o1, o2, ni, mp = run_table(kind, n, e, args)
    out = []
    for a, b, c, p in zip(o1, o2, ni, mp):
        for i in range(int(c)):
            out.append((int(a) + i, int(b) + i, int(p) & 1))
    return out


def time_evolve(n, e, m, terms, t, state, j_begin=1, j_end=None, original=False):
    """state: complex128 array of shape (2^m, C, C) or flat; modified in place.
    original: each block applies applyUnitary's term order instead of
    applyUnitary_reordered's."""
    assert state.dtype == np.complex128 and state.flags.c_contiguous
    if j_end is None:
        j_end = 1 << m
    rc = lib().orc_time_evolve_ordered(n, e, m, terms.arr, len(terms.arr), t,
                                       state.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), j_begin, j_end,
                                       1 if original else 0)
    if rc != 0:
        raise RuntimeError("orc_time_evolve failed")
    return state


def apply_unitary(n, e, terms, t, block):
    assert block.dtype == np.complex128 and block.flags.c_contiguous
    rc = lib().orc_apply_unitary(n, e, terms.arr, len(terms.arr), t,
                                 block.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if rc != 0:
        raise RuntimeError("orc_apply_unitary failed")
    return block


def apply_unitary_original(n, e, terms, t, block):
    """One applyUnitary (original order, every pqqr) on one block, in place."""
    assert block.dtype == np.complex128 and block.flags.c_contiguous
    rc = lib().orc_apply_unitary_original(n, e, terms.arr, len(terms.arr), t,
                                          block.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if rc != 0:
        raise RuntimeError("orc_apply_unitary_original failed")
    return block


def time_evolve_iqpe(n, e, terms, t, block, num_of_applications):
    """time_evolve_iqpe: applyUnitary num_of_applications times on one block, in place."""
    assert block.dtype == np.complex128 and block.flags.c_contiguous
    rc = lib().orc_time_evolve_iqpe(n, e, terms.arr, len(terms.arr), t,
                                    block.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(num_of_applications))
    if rc != 0:
        raise RuntimeError("orc_time_evolve_iqpe failed")
    return block


def count_updates(n, e, terms):
    d = ctypes.c_uint64()
    p = ctypes.c_uint64()
    lib().orc_count_updates(n, e, terms.arr, len(terms.arr), ctypes.byref(d), ctypes.byref(p))
    return int(d.value), int(p.value)


def qift(m, state):
    assert state.dtype == np.complex128 and state.flags.c_contiguous
    ld = state.size >> m
    if lib().orc_qift(m, ld, state.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) != 0:
        raise RuntimeError("orc_qift failed")
    return state


def phase_probabilities(m, state):
    ld = state.size >> m
    probs = np.zeros(1 << m, np.float64)
    lib().orc_phase_probabilities(m, ld, state.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  _ptr(probs, ctypes.c_double))
    return probs


def num_threads():
    return int(lib().orc_num_threads())
