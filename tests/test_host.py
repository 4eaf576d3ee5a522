"""Host logic of the product on the CPU: the C ABI library loads and exports
every declared symbol, the product's term compiler agrees bit-exactly with the
oracle, the tile program accounts for exactly the reference's work, and the
argument checks fail cleanly.  No GPU call is made here."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import oracle
import qpesim

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "qpesim.h")).read()
    declared = set(re.findall(r"\b(qpe_[a-z0-9_]+)\s*\(", header))
    assert declared == set(qpesim.EXPORTS)
    L = qpesim.load()
    for name in declared:
        assert hasattr(L, name), name


def _both(n, e, seed):
    one, two = qpesim.synthetic_integrals(n, seed)
    h, tau = qpesim.build_hamiltonian(n, e, one, two)
    T = oracle.compile_terms(n, one, two)
    return h, tau, T


@pytest.mark.parametrize("n,e,seed", [(2, 1, 1), (4, 2, 2), (6, 3, 0), (8, 3, 5), (11, 4, 9), (14, 5, 1)])
def test_term_compiler_matches_oracle_bit_exactly(n, e, seed):
    h, tau, T = _both(n, e, seed)
    assert h.term_counts() == T.counts
    L = T.as_lists()
    for cls in range(5):
        idx, co = h.terms(cls)
        ref = L[cls]
        assert len(ref) == len(co)
        for i, (key, c) in enumerate(ref):
            assert tuple(int(v) for v in idx[i][: len(key)]) == key
            assert co[i] == c          # bit-exact accumulation order
    assert tau == pytest.approx(0.5 / T.sum_abs(), rel=1e-14)


def test_kat_fixture_through_product():
    # the PQRS fixture (qsharp-verify/broombridge/PQRS.yml) through the ABI
    h = qpesim.Hamiltonian()
    h.set_num_orbitals(4)
    h.set_num_occupied(2)
    h.add_two_body_term(1, 2, 3, 4, 1.0)   # [1,4,2,3] read as i=1,l=4,j=2,k=3
    h.index_terms()
    idx, co = h.terms(qpesim.PQRS)
    d = {tuple(int(v) for v in k): c for k, c in zip(idx, co)}
    assert d[(0, 1, 3, 2)] == -1.0 and d[(4, 6, 7, 5)] == -1.0 and d[(0, 5, 6, 3)] == 1.0


@pytest.mark.parametrize("n,e,seed", [(4, 2, 2), (6, 3, 0), (7, 2, 3), (8, 4, 5)])
@pytest.mark.parametrize("dtype", [qpesim.COMPLEX128, qpesim.COMPLEX64])
def test_plan_accounts_for_reference_work(n, e, seed, dtype):
    h, tau, T = _both(n, e, seed)
    info = qpesim.plan_info(h, dtype)
    dg, pr = oracle.count_updates(n, e, T)
    assert info["pair_updates"] == pr
    assert info["diag_updates"] == dg
    assert info["C"] == math.comb(n, e)


def test_plan_small_tiles_still_valid():
    h, tau, T = _both(8, 3, 4)
    C = math.comb(8, 3)
    big = qpesim.plan_info(h, qpesim.COMPLEX128, 0)
    small = qpesim.plan_info(h, qpesim.COMPLEX128, 3 * C * 16)
    assert small["passes"] > big["passes"]
    with pytest.raises(qpesim.QpeError) as ei:
        qpesim.plan_info(h, qpesim.COMPLEX128, C * 16)   # one column per tile: no mixed window fits
    assert ei.value.code == qpesim.ERR_ARG


def test_string_count_limit_of_the_pass_schedules():
    """Schedule entries hold 14-bit ranks: C(18,6) = 18564 cannot be planned."""
    one, two = qpesim.synthetic_integrals(18, 0)
    h, _ = qpesim.build_hamiltonian(18, 6, one, two)
    with pytest.raises(qpesim.QpeError) as ei:
        qpesim.plan_info(h, qpesim.COMPLEX64, 0)
    assert ei.value.code == qpesim.ERR_ARG and "16383" in str(ei.value)


def test_argument_errors_without_gpu():
    h = qpesim.Hamiltonian()
    with pytest.raises(qpesim.QpeError) as ei:
        h.add_one_body_term(1, 2, 1.0)      # p < q
    assert ei.value.code == qpesim.ERR_ARG
    with pytest.raises(qpesim.QpeError) as ei:
        h.add_two_body_term(0, 1, 1, 1, 1.0)
    assert ei.value.code == qpesim.ERR_ARG
    with pytest.raises(qpesim.QpeError) as ei:
        h.term_counts()                     # before IndexTerms
    assert ei.value.code == qpesim.ERR_STATE
    with pytest.raises(qpesim.QpeError) as ei:
        h.add_one_body_term(1, 1, float("nan"))
    assert ei.value.code == qpesim.ERR_ARG
    for args in [(0, 4, 2), (4, 4, 4), (4, 40, 2), (21, 4, 2)]:
        with pytest.raises(qpesim.QpeError) as ei:
            qpesim.StateVector(*args)
        assert ei.value.code == qpesim.ERR_ARG


def test_synthetic_generator_is_deterministic_and_shaped():
    one, two = qpesim.synthetic_integrals(6, 0)
    one2, two2 = qpesim.synthetic_integrals(6, 0)
    assert one == one2 and two == two2
    assert len(one) == 21 and len(two) == 21 * 22 // 2
    assert one[0] == (1, 1, -2.0) and one[-1] == (6, 6, -2.0 + 0.25 * 5)
    diag = {(a, c): v for a, b, c, d, v in two if a == b and c == d}
    assert abs(diag[(3, 3)] - (0.5 + 0.0)) < 0.5   # Coulomb diagonal dominates noise
    z = qpesim.normals(123, 100000)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1.0) < 0.02
    assert np.array_equal(qpesim.normals(7, 10, start=3), qpesim.normals(7, 13)[3:])


def test_batch_instances_follow_the_spec():
    inst = qpesim.batch_instances()
    assert len(inst) == 48
    for i, x in enumerate(inst):
        assert 6 <= x["n"] <= 12 and x["e"] in (2, 3, 4) and 2 * x["e"] <= x["n"]
        assert x["seed"] == 1000 + i and x["m"] == 8


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N with N > 1 exits with a message (one GPU per box; no multi-GPU path)."""
    import json
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
