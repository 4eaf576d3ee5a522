"""Broombridge ingest on the CPU: the oracle's restatement of IngestBroombridge
(PyYAML + cpp/hamiltonian.cc:8-144) pinned against the reference's fixture
values, dense Jordan-Wigner matrices and the FCI energies the files record;
then the product's C++ reader (through the C ABI, no GPU call) against the
oracle, bit for bit, on every fixture and on YAML re-serialisations of them."""
import glob
import json
import math
import os

import numpy as np
import pytest
import yaml

import broombridge_oracle as bo
import jw
import oracle
import qpesim

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = os.path.join(HERE, "golden", "broombridge")
FILES = sorted(os.path.basename(p) for p in glob.glob(os.path.join(FIXTURES, "*.y*ml")))
REFERENCE_DATA = "/root/reference/qsharp-verify/broombridge"


def fixture(name):
    return os.path.join(FIXTURES, name)


def product_ingest(path=None, text=None):
    h = qpesim.Hamiltonian()
    h.load_broombridge(path=path, text=text)
    return h


def assert_product_equals_oracle(h, r):
    assert h.sizes() == (r["n"], r["e"])
    assert h.identity_term() == r["identity"]
    addr, co = h.initial_state()
    assert [int(a) for a in addr] == [d[0] for d in r["initial"]]
    assert [float(c) for c in co] == [d[1] for d in r["initial"]]
    h.index_terms()
    T = oracle.compile_terms(r["n"], r["one"], r["two"])
    assert h.term_counts() == T.counts
    for cls, ref in enumerate(T.as_lists()):
        idx, c = h.terms(cls)
        assert len(ref) == len(c)
        for i, (key, v) in enumerate(ref):
            assert tuple(int(x) for x in idx[i][: len(key)]) == key
            assert c[i] == v


# ------------------------------------------------------------------ pins of the oracle
def test_oracle_reads_the_kat_fixture_values():
    """The five one-integral files of the reference's term KATs
    (cpp/hamiltonian.hpp:192-313) hold the integrals restated in kat_integrals.json."""
    kat = json.load(open(os.path.join(HERE, "golden", "kat_integrals.json")))
    for name in ("PP", "PQ", "PQQP", "PQQR", "PQRS"):
        r = bo.ingest_file(fixture(name + ".yml"))
        want = kat[name]
        assert r["n"] == want["n_orb"]
        one = [tuple(e) for e in want["one"]]
        one = [(max(p, q), min(p, q), v) for p, q, v in one]
        assert [tuple(x) for x in r["one"]] == one
        assert [tuple(x) for x in r["two"]] == [tuple(e) for e in want["two"]]


def _dense_ladder(mode, raising, nmodes):
    """Textbook JW matrix of a+_mode / a_mode on 2^nmodes amplitudes; basis index
    = bit string, mode j = bit j, Z on every mode below j."""
    Z = np.diag([1.0, -1.0])
    I = np.eye(2)
    lower = np.array([[0.0, 1.0], [0.0, 0.0]])   # |0><1| : annihilation on one mode
    op = lower.T if raising else lower
    M = np.eye(1)
    for j in reversed(range(nmodes)):                # most significant bit first
        M = np.kron(M, op if j == mode else (Z if j < mode else I))
    return M


def test_oracle_sign_rule_equals_dense_jordan_wigner():
    n = 3
    nm = 2 * n
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(300):
        k = int(rng.integers(1, 7))
        ops = [(int(rng.integers(0, n)), bool(rng.integers(0, 2)), bool(rng.integers(0, 3) > 0)) for _ in range(k)]
        vac = np.zeros(1 << nm)
        vac[0] = 1.0
        psi = vac
        for orb, up, raising in reversed(ops):        # the rightmost operator acts first
            psi = _dense_ladder(orb if up else orb + n, raising, nm) @ psi
        nz = np.flatnonzero(psi)
        try:
            bv, sign = bo.apply_ladder(ops, n)
        except ValueError:
            assert nz.size == 0                        # the string annihilates the vacuum
            continue
        assert nz.tolist() == [bv] and psi[bv] == sign
        checked += 1
    assert checked > 50


# FCI energies recorded in the files; the integrals are rounded to 10 digits, and
# H2_2_sto-6g's integrals were edited by hand (comments in the file).
FCI_CASES = [("h2_ccpvtz_ccsd_1_4008au.yaml", 1e-8), ("h2o_oh0.9_sto3g.nw.out.yaml", 1e-8),
             ("Be-ccpVDZ-5orb.yaml", 1e-6), ("H2_2_sto-6g.yml", 1e-4)]


@pytest.mark.parametrize("name,tol", FCI_CASES)
def test_oracle_integrals_reproduce_the_recorded_fci_energy(name, tol):
    """Lowest eigenvalue of the textbook second-quantised H built from the read
    integrals (Mulliken (pq|rs) with 8-fold symmetry) + identity = the file's
    fci_energy: pins the index order and the identity term."""
    r = bo.ingest_file(fixture(name))
    doc = yaml.safe_load(open(fixture(name)))
    fci = doc["integral_sets"][0]["fci_energy"]["value"]
    H = jw.textbook_hamiltonian(r["n"], r["e"], r["one"], r["two"])
    e0 = np.linalg.eigvalsh(H)[0] + r["identity"]
    assert abs(e0 - fci) < tol


@pytest.mark.parametrize("name", ["Be-ccpVDZ-5orb.yaml", "H2_2_sto-6g.yml"])
def test_oracle_initial_signs_make_the_suggested_ground_state(name):
    """Two-determinant suggestions (HF + a double excitation): with the Jordan-
    Wigner signs the oracle applies, the state overlaps the exact ground state
    more and has a lower energy than with the second sign flipped."""
    r = bo.ingest_file(fixture(name))
    C = math.comb(r["n"], r["e"])
    H = jw.textbook_hamiltonian(r["n"], r["e"], r["one"], r["two"]).real
    w, V = np.linalg.eigh(H)
    psi = np.zeros(C * C)
    for a, c in r["initial"]:
        psi[a] = c
    flip = psi.copy()
    a1, c1 = r["initial"][1]
    flip[a1] = -c1
    ov = lambda x: abs(V[:, 0] @ x) / np.linalg.norm(x)
    en = lambda x: x @ H @ x / (x @ x)
    assert ov(psi) > 0.97 and ov(psi) > ov(flip) + 0.01
    assert en(psi) < en(flip)


def test_oracle_addresses_and_hf_sign():
    r = bo.ingest_file(fixture("H2_2_sto-6g.yml"))
    # (1a)+ (2a)+ (1b)+ (2b)+ |vacuum> = rank 0 in both strings, sign +;
    # the second entry is alpha = beta = orbitals {1, 3}: colex rank 1, address 6*1 + 1
    assert [d[0] for d in r["initial"]] == [0, 7]
    assert r["initial"][0][1] == 0.9950000486809176
    assert r["strings"] == [(0b0011, 0b0011), (0b0101, 0b0101)]


# ------------------------------------------------------------------ product vs oracle
@pytest.mark.parametrize("name", FILES)
def test_product_reads_fixture_like_oracle(name):
    assert_product_equals_oracle(product_ingest(path=fixture(name)), bo.ingest_file(fixture(name)))


@pytest.mark.parametrize("style", ["block", "flow", "mixed", "wrapped"])
@pytest.mark.parametrize("name", ["H2_2_sto-6g.yml", "Be-ccpVDZ-5orb.yaml", "h4_sto6g_0.100.nw.out.yaml"])
def test_product_reads_any_yaml_layout(name, style):
    """The same document re-serialised by PyYAML in block, flow and mixed layouts
    (and with 30-column line wrapping, which splits flow collections and plain
    scalars over lines) must read exactly as the original."""
    doc = yaml.safe_load(open(fixture(name)))
    kw = {"block": dict(default_flow_style=False), "flow": dict(default_flow_style=True),
          "mixed": dict(default_flow_style=None), "wrapped": dict(default_flow_style=None, width=30)}[style]
    text = "--- # re-serialised\n" + yaml.safe_dump(doc, **kw)
    r = bo.ingest(text)
    ref = bo.ingest_file(fixture(name))
    assert r["initial"] == ref["initial"] and r["one"] == ref["one"] and r["two"] == ref["two"]
    assert_product_equals_oracle(product_ingest(text=text), r)


def test_text_and_path_agree():
    text = open(fixture("h2o_oh0.9_sto3g.nw.out.yaml")).read()
    a, b = product_ingest(text=text), product_ingest(path=fixture("h2o_oh0.9_sto3g.nw.out.yaml"))
    assert a.sizes() == b.sizes() and a.identity_term() == b.identity_term()
    a.index_terms()
    b.index_terms()
    for cls in range(5):
        assert np.array_equal(a.terms(cls)[1], b.terms(cls)[1])


@pytest.mark.skipif(not os.path.isdir(REFERENCE_DATA), reason="the reference's molecule files are not present")
def test_product_reads_every_reference_molecule_like_oracle():
    """All Broombridge files of qsharp-verify/: files without an initial state
    and determinants with unequal alpha/beta counts are errors in the product
    (the reference throws / computes a meaningless address); every other file
    reads bit-exactly like the oracle."""
    files = sorted(glob.glob(os.path.join(REFERENCE_DATA, "**", "*.y*ml"), recursive=True))
    assert len(files) > 70
    read = 0
    for f in files:
        doc = bo.load(open(f).read())
        if "initial_state_suggestions" not in doc["integral_sets"][0]:
            with pytest.raises(qpesim.QpeError, match="initial_state_suggestions"):
                product_ingest(path=f)
            continue
        r = bo.ingest_doc(doc)
        if any(bin(a).count("1") != bin(b).count("1") for a, b in r["strings"]):
            with pytest.raises(qpesim.QpeError, match="alpha"):
                product_ingest(path=f)
            continue
        h = product_ingest(path=f)
        assert h.sizes() == (r["n"], r["e"]), f
        assert h.identity_term() == r["identity"], f
        addr, co = h.initial_state()
        assert [int(a) for a in addr] == [d[0] for d in r["initial"]], f
        assert [float(c) for c in co] == [d[1] for d in r["initial"]], f
        read += 1
    assert read > 60


# ------------------------------------------------------------------ errors
BASE = open(os.path.join(FIXTURES, "PQRS.yml")).read()

BAD_DOCUMENTS = [
    ("missing key", BASE.replace("coulomb_repulsion", "coulomb"), "coulomb_repulsion"),
    ("bad operator", BASE.replace("(1a)+", "(1c)+"), "ladder operator"),
    ("double plus", BASE.replace("(1a)+", "(1a)++"), "ladder operator"),
    ("zero orbital", BASE.replace("(1a)+", "(0a)+"), "1-based"),
    ("double creation", BASE.replace("(1b)+", "(1a)+"), "occupied"),
    ("empty annihilation", BASE.replace("(1b)+", "(3b)"), "empty"),
    ("unequal spins", BASE.replace("(1b)+", "(3a)+"), "alpha"),
    ("orbital past integrals", BASE.replace("(1b)+", "(9b)+"), "beyond"),
    ("index zero", BASE.replace("[1, 4, 2, 3, 1.0]", "[0, 4, 2, 3, 1.0]"), "1-based"),
    ("index not integer", BASE.replace("[1, 4, 2, 3, 1.0]", "[1.5, 4, 2, 3, 1.0]"), "unsigned"),
    ("value not number", BASE.replace("[1, 4, 2, 3, 1.0]", "[1, 4, 2, 3, abc]"), "number"),
    ("non-finite value", BASE.replace("[1, 4, 2, 3, 1.0]", "[1, 4, 2, 3, .nan]"), "non-finite"),
    ("unterminated flow", BASE.replace("[1, 4, 2, 3, 1.0]", "[1, 4, 2, 3, 1.0"), "flow"),
    ("anchor", BASE.replace("format: { version", "format: &f { version"), "anchors"),
    ("tab indent", BASE.replace("\n    hamiltonian:", "\n\thamiltonian:"), "tab"),
    ("short entry", BASE.replace("[1, 4, 2, 3, 1.0]", "[1, 4, 2]"), "entries"),
]


@pytest.mark.parametrize("what,text,msg", BAD_DOCUMENTS, ids=[b[0] for b in BAD_DOCUMENTS])
def test_malformed_documents_are_rejected_and_leave_h_unchanged(what, text, msg):
    assert text != BASE
    h = qpesim.Hamiltonian()
    with pytest.raises(qpesim.QpeError) as ei:
        h.load_broombridge(text=text)
    assert ei.value.code == qpesim.ERR_ARG
    assert msg in str(ei.value)
    assert h.sizes() == (0, 0) and h.identity_term() == 0.0
    assert len(h.initial_state()[0]) == 0


def test_missing_file_is_an_argument_error():
    h = qpesim.Hamiltonian()
    with pytest.raises(qpesim.QpeError) as ei:
        h.load_broombridge(path=os.path.join(FIXTURES, "no_such_file.yml"))
    assert ei.value.code == qpesim.ERR_ARG and "cannot open" in str(ei.value)


def test_error_messages_name_the_line():
    bad = BASE.replace("[1, 4, 2, 3, 1.0]", "[1, 4, 2, 3, abc]")
    line = 1 + bad[: bad.index("abc")].count("\n")
    with pytest.raises(qpesim.QpeError, match=f"line {line}:"):
        qpesim.Hamiltonian().load_broombridge(text=bad)


DUPLICATE = open(os.path.join(FIXTURES, "H2_2_sto-6g.yml")).read().replace(
    '- [0.9950000486809176, (1a)+, (2a)+, (1b)+, (2b)+, "|vacuum>"]',
    '- [0.9950000486809176, (1a)+, (2a)+, (1b)+, (2b)+, "|vacuum>"]\n'
    '            - [0.5, (2a)+, (1a)+, (1b)+, (2b)+, "|vacuum>"]')


def test_repeated_determinants_are_kept_in_file_order():
    """Two entries for the same determinant (the second with two operators
    swapped, so with the opposite sign): both are listed in file order; the
    state keeps the last one (cpp/hamiltonian.cc:134-139 assigns in order)."""
    assert DUPLICATE.count("[0.5, (2a)+") == 1
    h = product_ingest(text=DUPLICATE)
    addr, co = h.initial_state()
    r = bo.ingest(DUPLICATE)
    assert [int(a) for a in addr] == [d[0] for d in r["initial"]] == [0, 0, 7]
    assert [float(c) for c in co] == [d[1] for d in r["initial"]]
    assert co[1] == -0.5                       # (2a)+ (1a)+ = -(1a)+ (2a)+


def test_ingest_accumulates_like_repeated_add_calls():
    """A second ingest adds its integrals to what the Hamiltonian holds (the
    reference's AddOneBodyTerm / AddTwoBodyTerm accumulate per key)."""
    twice = qpesim.Hamiltonian()
    twice.load_broombridge(path=fixture("PQRS.yml"))
    twice.load_broombridge(path=fixture("PQRS.yml"))
    twice.index_terms()
    once = product_ingest(path=fixture("PQRS.yml"))
    once.index_terms()
    assert np.array_equal(twice.terms(qpesim.PQRS)[1], 2 * once.terms(qpesim.PQRS)[1])


# ------------------------------------------------------------------ energy estimate
def test_energy_estimate_closed_forms():
    m, ident, t = 6, 1.25, 0.1
    M = 1 << m
    for k0 in (0, 1, 17, M - 1):
        P = np.zeros(M)
        P[k0] = 1.0
        est = qpesim.energy_estimate(P, m, ident, t)
        assert est["phase"] == k0 / M
        assert est["energy"] == pytest.approx(-2 * math.pi * k0 / M + ident, abs=1e-15)
        assert est["energy_per_time"] == pytest.approx(-2 * math.pi * k0 / M / t + ident, abs=1e-13)
        assert est["moe"] == pytest.approx(2 * math.pi / (t * M), rel=1e-15)
    P = np.full(M, 1.0 / M)
    est = qpesim.energy_estimate(P, m, 0.0, 1.0)
    assert est["energy"] == pytest.approx(-2 * math.pi * (M - 1) / 2 / M, rel=1e-14)


def test_energy_estimate_rejects_bad_arguments():
    with pytest.raises(ValueError):
        qpesim.energy_estimate(np.zeros(5), 3, 0.0, 1.0)
    with pytest.raises(qpesim.QpeError):
        qpesim.energy_estimate(np.zeros(8), 3, 0.0, 0.0)
