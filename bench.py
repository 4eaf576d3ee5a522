"""Benchmark of the QPE hot path on one MI355X.

Workload (default): n=32 spin-orbitals (16 spatial), N=8 electrons (4+4),
m=10 ancillas, seed 2, Hartree-Fock initial state, first-order Trotter with
tau = 0.5/sum|c|, complex64 register (FP32 precision form).

A "step" is one Trotter step U of the chain: block j = U block(j-1) through
qpe_time_evolve_blocks (the reference's loop body, cpp/time_evolution_over_terms.hpp:
116-137).  --warmup steps run untimed, then exactly --steps steps are timed on
the host clock with the device synchronised on both sides; ms_per_step and
`value` (determinant-term updates per second) come from those steps.

Besides the steps, the full QPE of the reference's main (cpp/main.cc:119-137:
blocks 1..2^m-1, QIFT, P(k)) runs ONCE in complex64 and ONCE in complex128,
whatever --steps says: wall times, the FP32/FP64 speed-up, max|dP(k)| and the
relative L2 error of the FP32 QIFT output.  The CPU baseline (the FP64 OpenMP
oracle, the reference's loop structure) runs a bounded sample on the host cores.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "2022-qpesim_amd"))

import qpesim  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "QPE wall time + det-term updates/s, FP32 vs FP64 GPU at max|dP(k)|<=1e-5"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def parse(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="Trotter steps timed")
    ap.add_argument("--warmup", type=int, default=5, help="untimed Trotter steps before the timed ones")
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--m", type=int, default=None, help="override the ancilla count (development only)")
    ap.add_argument("--engine", type=int, default=0, help="0: fused groups of terms, 1: one term per op")
    ap.add_argument("--no-qpe", action="store_true", help="skip the full QPE runs (profiling of the steps only)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--fp64-only", action="store_true", help="run only the complex128 QPE (reference P(k) for --ref-pk)")
    ap.add_argument("--save-pk", default=None, help="write P(k) of the complex64 QPE to this .npy file")
    ap.add_argument("--ref-pk", default=None, help="complex128 P(k) of the same instance from an earlier run (.npy)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2, help="Trotter steps timed on the CPU oracle")
    ap.add_argument("--layout", choices=("alpha-beta", "single"), default="alpha-beta",
                    help="register layout: the reference's alpha x beta blocks, or one interleaved string of "
                         "2n spin-orbitals per determinant (C(2n,2e) amplitudes per block)")
    return ap.parse_args(argv)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def runs_planned(a):
    """How many Trotter steps and full QPEs one bench invocation performs.
    The full QPEs do not depend on --steps / --warmup."""
    qpe32 = 0 if a.no_qpe or a.fp64_only else 1
    qpe64 = 0 if a.no_qpe or a.no_fp64 or a.ref_pk else 1
    if a.fp64_only:
        qpe64 = 1
    return {"warmup_steps": 0 if a.fp64_only else a.warmup, "timed_steps": 0 if a.fp64_only else a.steps,
            "full_qpe_fp32": qpe32, "full_qpe_fp64": qpe64}


# ---- GPU legs (every device call of the bench goes through these) -------------------------

def evolve_range(sv, h, j_begin, j_end):
    """Blocks j_begin..j_end-1 of the chain; returns the device milliseconds of this call."""
    before = sv.timings()["evolve_ms"] if j_begin > 1 else 0.0
    qpesim.time_evolve_blocks(sv, h, j_begin, j_end)
    return sv.timings()["evolve_ms"] - before


def chain_steps(sv, h, j_next, count, M):
    """`count` Trotter steps of the chain from block j_next on; the chain restarts at
    block 1 (from the untouched block 0) when it reaches 2^m.  Returns (next block,
    device ms)."""
    dev_ms = 0.0
    while count > 0:
        if j_next >= M:
            j_next = 1
        k = min(count, M - j_next)
        dev_ms += evolve_range(sv, h, j_next, j_next + k)
        j_next += k
        count -= k
    return j_next, dev_ms


def full_qpe(sv, h, psi0, M):
    """The reference's main call sequence on resident block 0: evolve blocks 1..M-1
    (in 8 resumable pieces that report progress), QIFT, P(k).  Returns (wall s,
    timings, P(k))."""
    sv.set(0, psi0 / math.sqrt(M))            # input resident before the timed region
    t0 = time.perf_counter()
    bounds = [1 + ((M - 1) * k) // 8 for k in range(9)]
    for k in range(8):
        if bounds[k + 1] > bounds[k]:
            qpesim.time_evolve_blocks(sv, h, bounds[k], bounds[k + 1])
            log(f"    blocks {bounds[k]}..{bounds[k + 1] - 1} of {M - 1}: "
                f"{sv.timings()['evolve_ms']:.0f} ms device time so far")
    sv.qift()
    P = sv.identify_phase_amplitudes()
    wall = time.perf_counter() - t0
    return wall, sv.timings(), P


# ---- CPU baseline ---------------------------------------------------------------------------

def cpu_info():
    model, sockets = platform.processor() or "unknown", None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() == "Model name":
                model = v.strip()
            elif k.strip() == "Socket(s)":
                sockets = int(v.strip()) if v.strip().isdigit() else v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return model, sockets


def cpu_baseline(cfg, one, two, tau, steps, m):
    """FP64 OpenMP oracle on the host cores: `steps` Trotter steps on one block
    (extrapolated to the 2^m - 1 steps of the evolution), and the QIFT + P(k) on a
    slice of the register's determinant columns (extrapolated to all of them)."""
    os.environ.setdefault("OMP_PROC_BIND", "close")   # before the OpenMP runtime starts
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    n, e = cfg["n"], cfg["e"]
    M = 1 << m
    T = oracle.compile_terms(n, one, two)
    dg, pr = oracle.count_updates(n, e, T)
    C = math.comb(n, e)
    blk = np.zeros(C * C, np.complex128)
    blk[0] = 1.0
    t0 = time.perf_counter()
    for _ in range(steps):
        oracle.apply_unitary(n, e, T, tau, blk)
    dt = time.perf_counter() - t0
    # QIFT + P(k) on [M][cols] (the oracle's own radix-2 FFT; MKL is not available)
    cols = max(1, min(C * C, (1 << 24) // M))
    rng = np.random.default_rng(0)
    sl = (rng.standard_normal((M, cols)) + 1j * rng.standard_normal((M, cols))).astype(np.complex128)
    tq = time.perf_counter()
    oracle.qift(m, sl)
    oracle.phase_probabilities(m, sl)
    dq = time.perf_counter() - tq
    model, sockets = cpu_info()
    return {"updates_per_s": steps * (dg + pr) / dt, "seconds_per_step": dt / steps, "cores": oracle.num_threads(),
            "steps": steps, "qift_pk_s_slice": dq, "qift_cols": cols, "qift_pk_s_extrapolated": dq * (C * C) / cols,
            "evolve_s_extrapolated": (M - 1) * dt / steps, "cpu_model": model, "sockets": sockets,
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}}


def measured_traffic(workload_key):
    """HBM bytes per pass-kernel launch from the committed rocprofv3 counter summary
    for this workload (None when that workload was not counted)."""
    try:
        with open(TRAFFIC_FILE) as f:
            tab = json.load(f)
    except (OSError, ValueError):
        return None
    ent = tab.get("workloads", {}).get(workload_key)
    return ent


# ---- batching form (cfg3) -------------------------------------------------------------------

def bench_batch(a):
    """Batching form: 48 Hamiltonians of uneven size (n in {12..24}, N in {4,6,8},
    m=8) evolved in FP32 by one persistent dataflow launch (captured in a hipGraph)
    after qpe_autotune_batch; per-instance max|dP| vs FP64.  A step is one whole
    batched evolution (all 255 Trotter steps of every instance)."""
    specs = qpesim.batch_instances()
    m = a.m if a.m is not None else specs[0]["m"]
    M = 1 << m
    hams, psis, updates = [], [], []
    for sp in specs:
        one, two = qpesim.synthetic_integrals(sp["n"], sp["seed"])
        h, _ = qpesim.build_hamiltonian(sp["n"], sp["e"], one, two)
        info = qpesim.plan_info(h, qpesim.COMPLEX64)
        hams.append(h)
        psis.append(qpesim.initial_state(sp["n"], sp["e"], sp["seed"], sp["hf"]))
        updates.append(info["pair_updates"] + info["diag_updates"])
    tot_upd = (M - 1) * sum(updates)
    log(f"batch: {len(specs)} instances, m={m}, C from {min(math.comb(s['n'], s['e']) for s in specs)} "
        f"to {max(math.comb(s['n'], s['e']) for s in specs)}, {tot_upd:.3e} updates per QPE batch")

    def make(dtype, tune):
        svs = [qpesim.StateVector(m, sp["n"], sp["e"], dtype) for sp in specs]
        choices = qpesim.autotune_batch(svs, hams) if tune else []
        return svs, choices

    def run(svs, eager=False):
        for sv, psi in zip(svs, psis):
            sv.set(0, psi / math.sqrt(M))
        return qpesim.batch_time_evolve(svs, hams, eager=eager)

    t0 = time.perf_counter()
    svs32, choices = make(qpesim.COMPLEX64, True)
    log(f"  autotune {time.perf_counter() - t0:.1f}s: " +
        ", ".join(f"{c['tile_bytes']//1024}K/e{c['engine']}" for c in choices[:8]) + " ...")
    run(svs32)                                   # plans, graph capture and instantiation
    times = []
    for it in range(a.warmup + a.steps):
        ms = run(svs32)
        log(f"  {'warmup' if it < a.warmup else 'step'} {it}: batch evolve {ms:.1f} ms")
        if it >= a.warmup:
            times.append(ms)
    ev_ms = float(np.mean(times))
    tq = time.perf_counter()
    P32 = []
    for sv in svs32:
        sv.qift()
        P32.append(sv.identify_phase_amplitudes())
    post_s = time.perf_counter() - tq
    for sv in svs32:
        sv.close()
    out = {
        "metric": METRIC,
        "value": tot_upd / (ev_ms * 1e-3), "unit": "det-term updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ev_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "complex64 (fp32 state/rotations; fp64 angles, norms, P(k))",
        "data": "synthetic: 48 seeded Hamiltonians (SplitMix64/Box-Muller), Hartree-Fock initial states",
        "config": {"workload": f"batching form: 48 QPE instances, n=12..24 spin-orbitals, N=4..8, m={m}, "
                               "one persistent dataflow launch (hipGraph-captured) per batch evolution",
                   "instances": len(specs), "ancillas": m, "parallelism": "single GPU"},
        "batch_evolve_ms": ev_ms, "qift_pk_s": post_s, "updates_per_batch": tot_upd,
    }
    if not a.no_fp64:
        svs64, _ = make(qpesim.COMPLEX128, True)
        run(svs64)
        ms64 = run(svs64)
        dps = []
        for sv, p32 in zip(svs64, P32):
            sv.qift()
            dps.append(float(np.max(np.abs(sv.identify_phase_amplitudes() - p32))))
            sv.close()
        out["batch_evolve_ms_fp64"] = ms64
        out["speedup_fp32_vs_fp64"] = ms64 / ev_ms
        out["max_abs_dP_fp32_vs_fp64"] = max(dps)
        out["per_instance_max_abs_dP"] = dps
    print(json.dumps(out))


# ---- main -------------------------------------------------------------------------------------

def qift_rate(ms, elem, M, D):
    """HBM rate of the QIFT launch: every amplitude read and written once"""
    gb = 2.0 * elem * M * D / 1e9
    tbs = gb / ms if ms > 0 else 0.0
    return {"ms": ms, "bytes": gb * 1e9, "TB/s": tbs, "frac": tbs * 1e3 / HBM_PEAK_GBS}


def main(argv=None):
    a = parse(argv)
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported (no multi-GPU path)"}))
        sys.exit(2)
    if a.steps < 1 or a.warmup < 0:
        print(json.dumps({"error": "--steps must be >= 1 and --warmup >= 0"}))
        sys.exit(2)
    if a.config == "cfg3":
        bench_batch(a)
        return
    plan = runs_planned(a)
    cfg = dict(qpesim.CONFIGS[a.config])
    m = a.m if a.m is not None else cfg["m"]
    n, e = cfg["n"], cfg["e"]
    M = 1 << m
    one, two = qpesim.synthetic_integrals(n, cfg["seed"])
    h, tau = qpesim.build_hamiltonian(n, e, one, two)
    psi0 = qpesim.initial_state(n, e, cfg["seed"], cfg["hf"])
    C = math.comb(n, e)
    D = C * C
    layout = qpesim.LAYOUT_SINGLE_STRING if a.layout == "single" else qpesim.LAYOUT_ALPHA_BETA
    if layout == qpesim.LAYOUT_SINGLE_STRING:
        # the same state in the single-string register: the sector's amplitudes at their
        # interleaved ranks with the reordering sign, every other entry zero
        idx, sign = qpesim.single_string_map(n, e)
        D = math.comb(2 * n, 2 * e)
        psi_ss = np.zeros(D, np.complex128)
        psi_ss[idx] = sign * psi0
        psi0 = psi_ss
    order = cfg.get("order", 1)
    info32 = qpesim.plan_info(h, qpesim.COMPLEX64)
    # plan_info describes the first-order step; a second-order step applies every term twice (t/2 each)
    upd_per_step = order * (info32["pair_updates"] + info32["diag_updates"])
    bytes_per_step = order * info32["pass_bytes"]
    passes_per_step = order * info32["passes"]
    log(f"workload {a.config}: n={2*n} spin-orbitals N={2*e} m={m} C={C} D={D} terms={h.term_counts()} "
        f"tau={tau:.6e}; plan c64: {info32}; runs: {plan}")
    shape = (f"{M} x {C}^2 = {M*D} amplitudes, alpha x beta layout" if layout == qpesim.LAYOUT_ALPHA_BETA else
             f"{M} x C({2*n},{2*e}) = {M*D} amplitudes, single-string layout")
    workload = (f"QPE {2*n} spin-orbitals, {2*e} electrons, m={m} ancillas ({shape}), "
                f"{'second-order (symmetric)' if order == 2 else 'first-order'} Trotter, seed {cfg['seed']}")
    out = {"metric": METRIC, "value": None, "unit": "det-term updates/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "complex64 (fp32 state/rotations; fp64 angles, norms, P(k))",
           "data": "synthetic: seeded integrals (SplitMix64/Box-Muller), Hartree-Fock initial state",
           "config": {"workload": workload, "spin_orbitals": 2 * n, "electrons": 2 * e, "ancillas": m,
                      "determinants_per_block": D, "trotter_order": order, "parallelism": "single GPU",
                      "step": "one Trotter step U (block j-1 -> block j) of the QPE chain"},
           "updates_per_step": upd_per_step}

    sv32 = None
    if not a.fp64_only:
        sv32 = qpesim.StateVector(m, n, e, qpesim.COMPLEX64, layout)
        sv32.set_option(qpesim.OPT_ENGINE, a.engine)
        sv32.set_option(qpesim.OPT_TROTTER_ORDER, order)
        sv32.set(0, psi0 / math.sqrt(M))
        j, _ = chain_steps(sv32, h, 1, a.warmup, M) if a.warmup else (1, 0.0)
        log(f"  warmup: {a.warmup} Trotter steps")
        t0 = time.perf_counter()
        j, dev_ms = chain_steps(sv32, h, j, a.steps, M)
        wall = time.perf_counter() - t0      # each call returns after its device work is done
        ms_per_step = wall * 1e3 / a.steps
        log(f"  timed: {a.steps} Trotter steps, {ms_per_step:.3f} ms/step (device {dev_ms / a.steps:.3f} ms/step)")
        out["value"] = upd_per_step / (ms_per_step * 1e-3)
        out["ms_per_step"] = ms_per_step
        out["device_ms_per_step"] = dev_ms / a.steps
        # roofline of the dominant kernel (the term-rotation tile pass), from the same timed steps
        bytes_launch = bytes_per_step / passes_per_step
        avg_launch_s = dev_ms * 1e-3 / (a.steps * passes_per_step)
        achieved = bytes_launch / avg_launch_s / 1e9
        traffic = measured_traffic(f"{a.config}/c64") if layout == qpesim.LAYOUT_ALPHA_BETA else None
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS,
                           "traffic": traffic["traffic_bytes_per_launch"] if traffic else None,
                           "kernel": "pass_kernel<float, 1, *> (term-rotation tile passes: the mixed-window, "
                                     "same-spin and general op-set instantiations, averaged over a step's launches)",
                           "bytes_per_launch": bytes_launch, "launches_per_step": passes_per_step,
                           "avg_launch_us": avg_launch_s * 1e6,
                           "timing": "HIP events on the state's stream around the timed steps"}
        if layout == qpesim.LAYOUT_SINGLE_STRING:
            out["roofline"]["note"] = ("device time per step includes the two layout conversions "
                                       "(single-string block j-1 -> alpha x beta, block j back)")
        if traffic:
            out["roofline"]["traffic_source"] = traffic.get("source")
            if "dram_bytes_per_launch" in traffic:
                out["roofline"]["dram_bytes_per_launch"] = traffic["dram_bytes_per_launch"]

    y32 = P32 = None
    if plan["full_qpe_fp32"]:
        log("  full QPE, complex64")
        wall32, tm, P32 = full_qpe(sv32, h, psi0, M)
        y32 = sv32.get(0, min(sv32.size(), 1 << 24))   # leading slice of the QIFT output for the L2 comparison
        out.update({"qpe_wall_s_fp32": wall32, "evolve_ms_fp32": tm["evolve_ms"], "qift_ms_fp32": tm["qift_ms"],
                    "pk_ms_fp32": tm["pk_ms"], "updates_per_qpe": (M - 1) * upd_per_step,
                    "updates_per_s_qpe_fp32": (M - 1) * upd_per_step / (tm["evolve_ms"] * 1e-3)})
        out["qift_fp32"] = qift_rate(tm["qift_ms"], 8, M, D)
        log(f"  complex64 QPE {wall32:.2f} s (evolve {tm['evolve_ms']:.0f} ms, qift {tm['qift_ms']:.1f} ms)")
        if a.save_pk:
            np.save(a.save_pk, P32)
    if sv32 is not None:
        sv32.close()

    if plan["full_qpe_fp64"]:
        log("  full QPE, complex128")
        sv64 = qpesim.StateVector(m, n, e, qpesim.COMPLEX128, layout)
        sv64.set_option(qpesim.OPT_ENGINE, a.engine)
        sv64.set_option(qpesim.OPT_TROTTER_ORDER, order)
        wall64, tm64, P64 = full_qpe(sv64, h, psi0, M)
        y64 = sv64.get(0, min(sv64.size(), 1 << 24))
        sv64.close()
        log(f"  complex128 QPE {wall64:.2f} s (evolve {tm64['evolve_ms']:.0f} ms)")
        if a.fp64_only:
            if a.save_pk:
                np.save(a.save_pk, P64)
            print(json.dumps({"fp64_only": True, "config": a.config, "m": m, "qpe_wall_s_fp64": wall64,
                              "evolve_ms_fp64": tm64["evolve_ms"], "sum_P": float(P64.sum()),
                              "argmax_P": int(P64.argmax())}))
            return
        out["qift_fp64"] = qift_rate(tm64["qift_ms"], 16, M, D)
        out.update({"qpe_wall_s_fp64": wall64, "evolve_ms_fp64": tm64["evolve_ms"], "qift_ms_fp64": tm64["qift_ms"],
                    "value_fp64": (M - 1) * upd_per_step / (tm64["evolve_ms"] * 1e-3),
                    "sum_P_fp64": float(P64.sum()), "argmax_P": int(P64.argmax())})
        if P32 is not None:
            out["speedup_fp32_vs_fp64"] = wall64 / out["qpe_wall_s_fp32"]
            out["max_abs_dP_fp32_vs_fp64"] = float(np.max(np.abs(P32 - P64)))
            out["l2_rel_fp32_vs_fp64"] = float(np.linalg.norm(y32 - y64) / np.linalg.norm(y64))
    elif a.ref_pk and P32 is not None:
        P64 = np.load(a.ref_pk)
        out["max_abs_dP_fp32_vs_fp64"] = float(np.max(np.abs(P32 - P64)))
        out["fp64_reference"] = f"P(k) of a complex128 run of the same instance ({a.ref_pk})"

    if not a.no_cpu:
        log("  CPU baseline")
        cb = cpu_baseline(cfg, one, two, tau, a.cpu_steps, m)
        out["cpu_baseline"] = {
            "value": cb["updates_per_s"], "unit": "det-term updates/s", "cores": cb["cores"], "kind": "port",
            "sample": (f"{cb['steps']} Trotter steps U on one block of the same instance (FP64 OpenMP oracle, "
                       f"reference loop structure), extrapolated evolve {cb['evolve_s_extrapolated']:.0f} s for "
                       f"{M - 1} steps; QIFT + P(k) on {cb['qift_cols']} of {D} determinant columns "
                       f"({cb['qift_pk_s_slice']:.2f} s, extrapolated {cb['qift_pk_s_extrapolated']:.1f} s)"),
            "seconds_per_step": cb["seconds_per_step"], "evolve_s_extrapolated": cb["evolve_s_extrapolated"],
            "qift_pk_s_extrapolated": cb["qift_pk_s_extrapolated"], "cpu_model": cb["cpu_model"],
            "sockets": cb["sockets"], "omp": cb["omp"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
