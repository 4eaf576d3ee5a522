"""The dataflow order of a Trotter step, checked on the host: tests/flow_check.cpp
is built with the system C++ compiler against the product's plan builder and
verifies, for synthetic Hamiltonians of several sizes, tile budgets, Trotter
orders and term orders, that every dependency points to an earlier pass, that
for every amplitude of the block each item touching it lists the previous
toucher as a dependency (so no item can read a column or row before its last
writer is done, and none can overwrite it while an earlier reader still runs),
and that a simulated run claiming items in index order finishes.  Dropping any
single dependency must make the check fail."""
import math
import os
import shutil
import subprocess
import sys

import pytest

from test_plan_schedules import FULL, write_hamiltonian

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "2022-qpesim_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    out = str(tmp_path_factory.mktemp("flow_check") / "flow_check")
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "flow_check.cpp"),
                    os.path.join(CSRC, "plan.cpp"), os.path.join(CSRC, "hamiltonian.cpp"), "-o", out], check=True)
    return out


CASES = [
    # n, e, element bytes, tile budget, Trotter order, term order
    (6, 3, 16, FULL, 1, 0),
    (8, 3, 16, FULL, 1, 0),
    (8, 3, 8, 6 * math.comb(8, 3) * 8 + 64, 1, 0),     # narrow tiles: many items and windows
    (10, 4, 8, FULL, 1, 0),
    (10, 4, 16, 4 * math.comb(10, 4) * 16 + 64, 2, 0),
    (10, 3, 8, FULL, 1, 1),
    (12, 4, 8, 12 * math.comb(12, 4) * 8 + 64, 1, 0),
]


def run(checker, path, args, drop=0):
    cmd = [checker, path] + [str(a) for a in args] + ([str(drop)] if drop else [])
    return subprocess.run(cmd, capture_output=True, text=True, timeout=300)


@pytest.mark.parametrize("n,e,elem,tile,order,term_order", CASES)
def test_dataflow_order_covers_every_conflict(checker, tmp_path, n, e, elem, tile, order, term_order):
    path = str(tmp_path / "h.txt")
    write_hamiltonian(path, n, e, seed=11 + n)
    args = (elem, tile, order, term_order)
    r = run(checker, path, args)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
    items, deps = (int(x) for x in r.stdout.split()[1:3])
    assert items > 0 and deps > 0
    # a planted mistake: any dropped dependency is caught
    for drop in sorted({1, deps // 3 + 1, deps // 2 + 1, deps}):
        rb = run(checker, path, args, drop)
        assert rb.returncode == 1 and rb.stdout.startswith("FAIL"), (drop, rb.stdout)
