"""bench.py's run structure on the CPU, with the device calls replaced by fakes:
a step is one Trotter step, exactly --steps of them are timed after --warmup,
the full QPE runs once per dtype whatever --steps says, and the chained block
ranges handed to qpe_time_evolve_blocks are always valid continuations."""
import importlib
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


@pytest.fixture
def bench():
    return importlib.import_module("bench")


class FakeSV:
    def __init__(self, m, n, e, dtype, layout=0):
        self.m, self.M, self.dtype = m, 1 << m, dtype
        self.ms = 0.0

    def set_option(self, *a):
        pass

    def set(self, off, amps):
        pass

    def size(self):
        return self.M * 4

    def get(self, off=0, count=None):
        return np.ones(count or 4, np.complex128)

    def timings(self):
        return {"evolve_ms": self.ms, "qift_ms": 0.0, "pk_ms": 0.0}

    def close(self):
        pass


def run_main(bench, monkeypatch, capsys, argv):
    calls = {"steps": 0, "qpe": [], "ranges": []}
    state = {"next": None}

    def fake_blocks(sv, h, jb, je):
        assert 1 <= jb <= je <= sv.M
        # a call with jb > 1 must continue the previous call's chain
        assert jb == 1 or state["next"] == jb
        state["next"] = je
        calls["ranges"].append((jb, je))
        sv.ms = (sv.ms if jb > 1 else 0.0) + 1.5 * (je - jb)

    def fake_full_qpe(sv, h, psi0, M):
        calls["qpe"].append(sv.dtype)
        return 1.0, sv.timings() | {"evolve_ms": 10.0}, np.full(M, 1.0 / M)

    monkeypatch.setattr(bench.qpesim, "StateVector", FakeSV)
    monkeypatch.setattr(bench.qpesim, "time_evolve_blocks", fake_blocks)
    monkeypatch.setattr(bench, "full_qpe", fake_full_qpe)
    bench.main(argv)
    line = capsys.readouterr().out.strip().splitlines()[-1]
    steps = sum(je - jb for jb, je in calls["ranges"])
    return json.loads(line), calls, steps


@pytest.mark.parametrize("steps,warmup", [(1, 0), (20, 5), (200, 7)])
def test_full_qpe_count_does_not_grow_with_steps(bench, monkeypatch, capsys, steps, warmup):
    out, calls, nsteps = run_main(bench, monkeypatch, capsys,
                                  ["--config", "cfg0", "--m", "4", "--steps", str(steps), "--warmup", str(warmup),
                                   "--no-cpu"])
    import qpesim
    assert calls["qpe"] == [qpesim.COMPLEX64, qpesim.COMPLEX128]   # one full QPE per dtype
    assert nsteps == steps + warmup                                # chain steps (m=4 wraps past 15)
    assert out["steps"] == steps and out["warmup"] == warmup
    # device time of the timed steps only: 1.5 ms per step in the fake
    assert out["device_ms_per_step"] == pytest.approx(1.5)
    assert out["value"] == pytest.approx(out["updates_per_step"] / (out["ms_per_step"] * 1e-3))
    r = out["roofline"]
    assert r["frac"] == pytest.approx(r["achieved"] / r["peak"])
    assert r["avg_launch_us"] == pytest.approx(1.5e3 / r["launches_per_step"])


def test_runs_planned_flags(bench):
    a = bench.parse(["--steps", "50", "--no-fp64"])
    assert bench.runs_planned(a) == {"warmup_steps": 5, "timed_steps": 50, "full_qpe_fp32": 1, "full_qpe_fp64": 0}
    a = bench.parse(["--no-qpe"])
    assert bench.runs_planned(a)["full_qpe_fp32"] == 0 and bench.runs_planned(a)["full_qpe_fp64"] == 0


def test_traffic_table_is_keyed_by_workload(bench):
    t = bench.measured_traffic("cfg2/c64")
    assert t is not None and t["traffic_bytes_per_launch"] > 0
    assert bench.measured_traffic("no-such-workload") is None
