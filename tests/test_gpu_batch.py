"""Batching form on the GPU: one launch sequence (a captured hipGraph) for
several instances of uneven size must give, bit for bit, what evolving each
instance on its own gives (the same ops run on the same tile items; complex64
dither bits depend only on the step, op, item, thread, entry and line)."""
import math

import numpy as np
import pytest

import qpesim
from test_gpu_parity import oracle_run, say

pytestmark = pytest.mark.gpu

SHAPES = [(6, 3, 11), (8, 2, 12), (7, 3, 13), (9, 4, 14)]   # (n_spatial, n_occupied, seed)


def _instances(m, dtype, shapes=SHAPES):
    out = []
    for n, e, seed in shapes:
        one, two = qpesim.synthetic_integrals(n, seed)
        h, _ = qpesim.build_hamiltonian(n, e, one, two)
        psi = qpesim.initial_state(n, e, seed, 0.9)
        sv = qpesim.StateVector(m, n, e, dtype)
        sv.set(0, psi / math.sqrt(1 << m))
        out.append((sv, h, psi))
    return out


def _reset(inst, m):
    for sv, _, psi in inst:
        sv.set(0, psi / math.sqrt(1 << m))


def _all(inst):
    return [sv.get() for sv, _, _ in inst]


@pytest.mark.parametrize("dtype", [qpesim.COMPLEX128, qpesim.COMPLEX64])
def test_batch_matches_single_instances(dtype):
    m = 4
    say(f"batch vs single dtype={dtype} m={m}")
    inst = _instances(m, dtype)
    for sv, h, _ in inst:
        qpesim.time_evolve_over_terms(sv, h)
    single = _all(inst)
    _reset(inst, m)
    ms = qpesim.batch_time_evolve([i[0] for i in inst], [i[1] for i in inst])
    assert ms > 0.0
    batch = _all(inst)
    for a, b in zip(single, batch):
        assert np.array_equal(a, b)
    for sv, _, _ in inst:
        sv.close()


def test_batch_graph_replays_and_eager_agree():
    m = 3
    say("batch graph replays vs eager")
    inst = _instances(m, qpesim.COMPLEX64)
    svs, hs = [i[0] for i in inst], [i[1] for i in inst]
    qpesim.batch_time_evolve(svs, hs, eager=True)
    ref = _all(inst)
    for rep in range(3):
        _reset(inst, m)
        _ = [sv.block_norms() for sv in svs]          # eager work between graph replays
        qpesim.batch_time_evolve(svs, hs)
        for a, b in zip(ref, _all(inst)):
            assert np.array_equal(a, b), f"replay {rep}"
    for sv in svs:
        sv.close()


def test_batch_qpe_matches_oracle_probabilities():
    """Each batched instance, taken through the QIFT and P(k), matches the FP64
    oracle as a single evolution does."""
    m = 4
    say("batched QPE vs oracle P(k)")
    inst = _instances(m, qpesim.COMPLEX128, SHAPES[:3])
    qpesim.batch_time_evolve([i[0] for i in inst], [i[1] for i in inst])
    for (sv, h, psi), (n, e, seed) in zip(inst, SHAPES):
        sv.qift()
        P = sv.identify_phase_amplitudes()
        _, _, rP = oracle_run(n, e, m, seed, psi0=psi)
        assert np.max(np.abs(P - rP)) < 1e-12
        sv.close()


def test_autotune_keeps_results_and_reports_a_choice():
    m = 3
    say("autotune")
    inst = _instances(m, qpesim.COMPLEX128, SHAPES[:2])
    for sv, h, _ in inst:
        qpesim.time_evolve_over_terms(sv, h)
    ref = _all(inst)
    _reset(inst, m)
    for sv, h, _ in inst:
        r = qpesim.autotune(sv, h, share_cus=64)
        assert r["tile_bytes"] > 0 and r["engine"] in (0, 1) and r["us_per_step"] > 0.0
    qpesim.batch_time_evolve([i[0] for i in inst], [i[1] for i in inst])
    # FP64: every amplitude sees the same rotations in the same order whatever the tile width
    for a, b in zip(ref, _all(inst)):
        assert np.array_equal(a, b)
    for sv, _, _ in inst:
        sv.close()


def test_batch_rejects_mixed_types_and_duplicates():
    a = _instances(2, qpesim.COMPLEX64, SHAPES[:1])[0]
    b = _instances(2, qpesim.COMPLEX128, SHAPES[1:2])[0]
    with pytest.raises(qpesim.QpeError) as ei:
        qpesim.batch_time_evolve([a[0], b[0]], [a[1], b[1]])
    assert ei.value.code == qpesim.ERR_ARG
    with pytest.raises(qpesim.QpeError) as ei:
        qpesim.batch_time_evolve([a[0], a[0]], [a[1], a[1]])
    assert ei.value.code == qpesim.ERR_ARG
    a[0].close()
    b[0].close()


def test_batch_autotuner_choices():
    """qpe_autotune_batch sets a valid configuration on every instance, and the tuned
    batch still equals the tuned instances run one by one."""
    m = 3
    say("batch autotuner")
    inst = _instances(m, qpesim.COMPLEX64)
    svs, hams = [i[0] for i in inst], [i[1] for i in inst]
    res = qpesim.autotune_batch(svs, hams)
    assert len(res) == len(inst)
    passes = []
    for (sv, h, _), r in zip(inst, res):
        assert r["tile_bytes"] > 0 and r["engine"] in (0, 1) and r["us_per_step"] > 0.0
        qpesim.time_evolve_over_terms(sv, h)
        passes.append(sv.timings()["passes_per_step"])
    say(f"  choices {res}, passes per step {passes}")
    single = _all(inst)
    _reset(inst, m)
    qpesim.batch_time_evolve(svs, hams)
    for a, b in zip(single, _all(inst)):
        assert np.array_equal(a, b)
    for sv in svs:
        sv.close()


@pytest.mark.parametrize("dtype", [qpesim.COMPLEX128, qpesim.COMPLEX64])
def test_batch_dataflow_equals_pass_chain(dtype):
    """The persistent dataflow launch (default) and the chain of pass launches give
    the same registers bit for bit; so do the dataflow launch's graph replays and an
    eager launch, with eager work between the replays."""
    m = 4
    say(f"batch dataflow vs pass chain dtype={dtype}")
    inst = _instances(m, dtype)
    svs, hs = [i[0] for i in inst], [i[1] for i in inst]
    qpesim.batch_time_evolve(svs, hs, passes=True)
    ref = _all(inst)
    assert all(sv.timings()["launches_per_step"] > 1 for sv in svs)
    _reset(inst, m)
    qpesim.batch_time_evolve(svs, hs, eager=True)
    assert all(sv.timings()["launches_per_step"] == 1 for sv in svs)
    for a, b in zip(ref, _all(inst)):
        assert np.array_equal(a, b)
    for rep in range(3):
        _reset(inst, m)
        _ = [sv.block_norms() for sv in svs]
        qpesim.batch_time_evolve(svs, hs)
        for a, b in zip(ref, _all(inst)):
            assert np.array_equal(a, b), f"replay {rep}"
    for sv in svs:
        sv.close()
