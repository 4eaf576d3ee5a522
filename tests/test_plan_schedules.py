"""The plan's op schedules, checked on the host: tests/plan_check.cpp is built
with the system C++ compiler against the product's plan builder and verifies,
for synthetic Hamiltonians of several sizes, tile budgets, Trotter orders and
term orders, that every op's slot count, rows and division-free lane split are
valid and that every double-excitation group lists each of its terms'
determinant pairs exactly once (C(n-2, e-1) per mixed term, C(n-4, e-2) per
same-spin term) -- the closed forms the reference's run tables also give
(cpp/time_evolution_over_terms.hpp:1160-1411)."""
import math
import os
import shutil
import subprocess
import sys

import pytest

import qpesim

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "2022-qpesim_amd", "csrc")
FULL = 160 * 1024 - 6208   # the pass kernel's LDS tile budget (160 KB less its header)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    out = str(tmp_path_factory.mktemp("plan_check") / "plan_check")
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "plan_check.cpp"),
                    os.path.join(CSRC, "plan.cpp"), os.path.join(CSRC, "hamiltonian.cpp"), "-o", out], check=True)
    return out


def write_hamiltonian(path, n, e, seed):
    one, two = qpesim.synthetic_integrals(n, seed)
    with open(path, "w") as f:
        f.write(f"{n} {e} 0.01\n")
        for p, q, v in one:
            if q > p:
                p, q = q, p
            f.write(f"O {p} {q} {v!r}\n")
        for a, b, c, d, v in two:
            f.write(f"T {a} {c} {d} {b} {v!r}\n")   # as build_hamiltonian feeds them


CASES = [
    # n, e, element bytes, tile budget, Trotter order, term order
    (8, 3, 16, FULL, 1, 0),
    (10, 4, 8, FULL, 1, 0),
    (12, 4, 8, FULL, 1, 1),
    (10, 3, 8, FULL, 2, 0),
    (12, 5, 16, 3 * math.comb(12, 5) * 16 + 64, 1, 0),
    (16, 4, 8, FULL, 1, 0),      # the 32 spin-orbital / 8 electron workload's block
    (16, 4, 16, FULL, 1, 0),
]


@pytest.mark.parametrize("n,e,elem,tile,order,term_order", CASES)
def test_plan_schedules_cover_every_pair_once(checker, tmp_path, n, e, elem, tile, order, term_order):
    path = str(tmp_path / "h.txt")
    write_hamiltonian(path, n, e, seed=7 + n)
    r = subprocess.run([checker, path, str(elem), str(tile), str(order), str(term_order)],
                       capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
    assert int(r.stdout.split()[2]) > 0
