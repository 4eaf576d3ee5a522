"""CPU restatement of Hamiltonian::IngestBroombridge (christopherkang/2022-qpesim,
cpp/hamiltonian.cc:8-144; ParsePolishNotation :146-158).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the product's
C++ Broombridge reader.  The product never imports it.

Parsing is done by PyYAML's safe loader (on libyaml when present), a YAML
implementation independent of the product's reader, so the two agree only if
both read the file the same way.  The ingest logic follows the reference line by line:

  identity      = coulomb_repulsion.value + energy_offset.value        (:15)
  one-body      [p, q, v], swapped so that p >= q                       (:22-37)
  two-body      [i, l, j, k, v] -> AddTwoBodyTerm(i, j, k, l)           (:39-52)
  num_orbitals  = largest 1-based index of any integral                 (:34-35, :50-51)
  initial state = initial_state_suggestions[0].state.superposition;
                  entry [c, op_1, ..., op_r, vacuum]: the last entry is
                  skipped, op_r .. op_1 act on the empty string in that
                  order, each multiplying c by (-1)^(occupied modes below
                  its mode) after it acts; mode = orbital (alpha) or
                  orbital + n (beta)                                    (:55-112)
  address       = C(n, e) * rank(alpha bits) + rank(beta bits),
                  e = electrons / 2                                     (:113-128)
"""
import math
import re

import yaml

_LADDER = re.compile(r"\((\d+)([ab])\)(\+?)")


def colex_rank(bv):
    """OffsetComputer::rank, cpp/offset_computer.hpp:16-27: sum_k C(pos_k, k)."""
    r, k, pos = 0, 1, 0
    while bv:
        if bv & 1:
            r += math.comb(pos, k)
            k += 1
        bv >>= 1
        pos += 1
    return r


def parse_ladder(text):
    """ParsePolishNotation: '(3b)+' -> (orbital 2 (0-based), spin up False, raising True)."""
    m = _LADDER.fullmatch(str(text))
    if m is None:
        raise ValueError(f"not a ladder operator: {text!r}")
    return int(m.group(1)) - 1, m.group(2) == "a", m.group(3) == "+"


def apply_ladder(ops, n):
    """The operator string ops (written left to right) applied to the empty
    string: returns (bit string over 2n modes, sign)."""
    bv, sign = 0, 1
    for orb, up, raising in reversed(ops):
        mode = orb if up else orb + n
        occupied = (bv >> mode) & 1
        if raising == bool(occupied):
            raise ValueError("creation on an occupied or annihilation of an empty spin orbital")
        bv ^= 1 << mode
        if bin(bv & ((1 << mode) - 1)).count("1") & 1:
            sign = -sign
    return bv, sign


_Loader = getattr(yaml, "CSafeLoader", yaml.SafeLoader)


def load(text):
    return yaml.load(text, Loader=_Loader)


def ingest(text):
    """Returns dict(n, e, identity, one=[(p, q, v)], two=[(i, l, j, k, v)] in file
    (Mulliken) order, initial=[(address, coefficient)], strings=[(alpha bits,
    beta bits)] of each determinant)."""
    return ingest_doc(load(text))


def ingest_doc(doc):
    """ingest() of a document PyYAML has already parsed."""
    integrals = doc["integral_sets"][0]
    ham = integrals["hamiltonian"]
    identity = float(integrals["coulomb_repulsion"]["value"]) + float(integrals["energy_offset"]["value"])
    n = 0
    one = []
    for t in ham["one_electron_integrals"]["values"] or []:
        p, q, v = int(t[0]), int(t[1]), float(t[2])
        if q > p:
            p, q = q, p
        one.append((p, q, v))
        n = max(n, p)
    two = []
    for t in ham["two_electron_integrals"]["values"] or []:
        i, l, j, k, v = int(t[0]), int(t[1]), int(t[2]), int(t[3]), float(t[4])
        two.append((i, l, j, k, v))
        n = max(n, i, j, k, l)
    sup = integrals["initial_state_suggestions"][0]["state"]["superposition"]
    initial, strings, e = [], [], None
    for basis in sup:
        coeff = float(basis[0])
        bv, sign = apply_ladder([parse_ladder(x) for x in basis[1:-1]], n)
        alpha, beta = bv & ((1 << n) - 1), bv >> n
        e = bin(bv).count("1") // 2
        initial.append((math.comb(n, e) * colex_rank(alpha) + colex_rank(beta), sign * coeff))
        strings.append((alpha, beta))
    return {"n": n, "e": e, "identity": identity, "one": one, "two": two, "initial": initial, "strings": strings}


def ingest_file(path):
    with open(path) as f:
        return ingest(f.read())
