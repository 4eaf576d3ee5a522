"""Benchmark of the QPE hot path on one MI355X.

Workload (default): n=32 spin-orbitals (16 spatial), N=8 electrons (4+4),
m=10 ancillas, seed 2, Hartree-Fock initial state, first-order Trotter with
tau = 0.5/sum|c|.  One step = one full QPE: blocks 1..1023 by the evolution,
then the QIFT and P(k).  The register is complex64 (FP32 precision form);
the same instance is run once in complex128 (FP64 path) for the speedup and
the max |dP(k)| / L2 error of the FP32 result.

Prints one JSON line.  `value` = FP32 determinant-term updates per second over
the timed evolution (updates = the reference's pair rotations + diagonal
multiplies, (M-1) per step of U).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "2022-qpesim_amd"))

import qpesim  # noqa: E402

HBM_PEAK_GBS = 8000.0

# HBM bytes per pass_kernel<float> launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE,
# each scaled by the same counter's reading for kernels of known size (FETCH x1.997, WRITE x1.0);
# see profiles/pass_kernel_n32N8m2_pmc_traffic.md.  Keyed by (spin-orbitals, electrons, dtype, order).
MEASURED_TRAFFIC = {(32, 8, "c64", 1): 43.22e6}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--m", type=int, default=None, help="override the ancilla count (development only)")
    ap.add_argument("--engine", type=int, default=0, help="0: fused groups of terms, 1: one term per op")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--fp64-only", action="store_true", help="run only the complex128 QPE (reference P(k) for --ref-pk)")
    ap.add_argument("--save-pk", default=None, help="write P(k) of the run to this .npy file")
    ap.add_argument("--ref-pk", default=None, help="complex128 P(k) of the same instance from an earlier run (.npy)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2, help="Trotter steps timed on the CPU oracle")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_qpe(h, cfg, m, dtype, psi0, steps, warmup, engine=0):
    n, e = cfg["n"], cfg["e"]
    M = 1 << m
    sv = qpesim.StateVector(m, n, e, dtype)
    sv.set_option(qpesim.OPT_ENGINE, engine)
    sv.set_option(qpesim.OPT_TROTTER_ORDER, cfg.get("order", 1))
    scale = 1.0 / math.sqrt(M)
    rec = []
    for it in range(warmup + steps):
        sv.set(0, psi0 * scale)                 # inputs resident before the timed region
        t0 = time.perf_counter()
        # the chain in 8 resumable pieces, each reporting its progress; device time accumulates
        bounds = [1 + ((M - 1) * k) // 8 for k in range(9)]
        for k in range(8):
            if bounds[k + 1] > bounds[k]:
                qpesim.time_evolve_blocks(sv, h, bounds[k], bounds[k + 1])
                log(f"    evolved blocks {bounds[k]}..{bounds[k + 1] - 1} of {M - 1}: "
                    f"{sv.timings()['evolve_ms']:.0f} ms device time so far")
        sv.qift()
        P = sv.identify_phase_amplitudes()
        wall = time.perf_counter() - t0
        tm = sv.timings()
        log(f"  {'warmup' if it < warmup else 'step'} {it}: dtype={'c64' if dtype else 'c128'} wall={wall:.3f}s "
            f"evolve={tm['evolve_ms']:.1f}ms qift={tm['qift_ms']:.1f}ms pk={tm['pk_ms']:.2f}ms")
        if it >= warmup:
            rec.append((wall, tm))
    y = sv.get(0, min(sv.size(), 1 << 24))   # leading slice of the QIFT output for the L2 comparison
    sv.close()
    return rec, P, y


def cpu_baseline(cfg, one, two, tau, steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    n, e = cfg["n"], cfg["e"]
    T = oracle.compile_terms(n, one, two)
    dg, pr = oracle.count_updates(n, e, T)
    C = math.comb(n, e)
    blk = np.zeros(C * C, np.complex128)
    blk[0] = 1.0
    t0 = time.perf_counter()
    for _ in range(steps):
        oracle.apply_unitary(n, e, T, tau, blk)
    dt = time.perf_counter() - t0
    return {"updates_per_s": steps * (dg + pr) / dt, "seconds_per_step": dt / steps, "cores": oracle.num_threads(),
            "steps": steps}


def bench_batch(a):
    """Batching form: 48 Hamiltonians of uneven size (n in {12..24}, N in {4,6,8},
    m=8) evolved in FP32 by one graph launch; per-instance max|dP| vs FP64."""
    specs = qpesim.batch_instances()
    m = a.m if a.m is not None else specs[0]["m"]
    M = 1 << m
    hams, psis, updates, work = [], [], [], []
    for sp in specs:
        one, two = qpesim.synthetic_integrals(sp["n"], sp["seed"])
        h, _ = qpesim.build_hamiltonian(sp["n"], sp["e"], one, two)
        info = qpesim.plan_info(h, qpesim.COMPLEX64)
        hams.append(h)
        psis.append(qpesim.initial_state(sp["n"], sp["e"], sp["seed"], sp["hf"]))
        updates.append(info["pair_updates"] + info["diag_updates"])
        work.append(info["pass_bytes"])
    tot_upd = (M - 1) * sum(updates)
    log(f"batch: {len(specs)} instances, m={m}, C from {min(math.comb(s['n'], s['e']) for s in specs)} "
        f"to {max(math.comb(s['n'], s['e']) for s in specs)}, {tot_upd:.3e} updates per QPE batch")

    def make(dtype, tune):
        svs = [qpesim.StateVector(m, sp["n"], sp["e"], dtype) for sp in specs]
        choices = []
        if tune:
            wsum = float(sum(work))
            for sv, h, w in zip(svs, hams, work):
                choices.append(qpesim.autotune(sv, h, share_cus=max(1, int(round(256 * w / wsum)))))
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
    seq_ms = 0.0
    for sv, h, psi in zip(svs32, hams, psis):       # the same instances one after another
        sv.set(0, psi / math.sqrt(M))
        qpesim.time_evolve_over_terms(sv, h)
        seq_ms += sv.timings()["evolve_ms"]
    for sv in svs32:
        sv.close()
    out = {
        "metric": "QPE wall time + det-term updates/s, FP32 vs FP64 GPU at max|dP(k)|<=1e-5",
        "value": tot_upd / (ev_ms * 1e-3), "unit": "det-term updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ev_ms + post_s * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "complex64 (fp32 state/rotations; fp64 angles, norms, P(k))",
        "data": "synthetic: 48 seeded Hamiltonians (SplitMix64/Box-Muller), Hartree-Fock initial states",
        "config": {"workload": f"batching form: 48 QPE instances, n=12..24 spin-orbitals, N=4..8, m={m}, "
                               "one hipGraph launch per batch evolution", "instances": len(specs),
                   "ancillas": m, "parallelism": "single GPU"},
        "batch_evolve_ms": ev_ms, "sequential_evolve_ms": seq_ms, "batch_speedup_vs_sequential": seq_ms / ev_ms,
        "qift_pk_s": post_s, "updates_per_batch": tot_upd,
    }
    if not a.no_fp64:
        svs64, _ = make(qpesim.COMPLEX128, False)
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


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported (no multi-GPU path)"}))
        sys.exit(2)
    if a.config == "cfg3":
        bench_batch(a)
        return
    cfg = dict(qpesim.CONFIGS[a.config])
    m = a.m if a.m is not None else cfg["m"]
    n, e = cfg["n"], cfg["e"]
    M = 1 << m
    one, two = qpesim.synthetic_integrals(n, cfg["seed"])
    h, tau = qpesim.build_hamiltonian(n, e, one, two)
    psi0 = qpesim.initial_state(n, e, cfg["seed"], cfg["hf"])
    C = math.comb(n, e)
    D = C * C

    info32 = qpesim.plan_info(h, qpesim.COMPLEX64)
    info64 = qpesim.plan_info(h, qpesim.COMPLEX128)
    order = cfg.get("order", 1)
    # plan_info describes the first-order step; a second-order step applies every term twice (t/2 each)
    upd_per_step = order * (info32["pair_updates"] + info32["diag_updates"])
    log(f"workload {a.config}: n={2*n} spin-orbitals N={2*e} m={m} C={C} D={D} terms={h.term_counts()} tau={tau:.6e}")
    log(f"plan c64: {info32}\nplan c128: {info64}")

    if a.fp64_only:
        rec64, P64, _ = run_qpe(h, cfg, m, qpesim.COMPLEX128, psi0, 1, 0, a.engine)
        if a.save_pk:
            np.save(a.save_pk, P64)
        print(json.dumps({"fp64_only": True, "config": a.config, "m": m, "qpe_wall_s_fp64": rec64[0][0],
                          "evolve_ms_fp64": rec64[0][1]["evolve_ms"], "sum_P": float(P64.sum()),
                          "argmax_P": int(P64.argmax())}))
        return
    rec32, P32, y32 = run_qpe(h, cfg, m, qpesim.COMPLEX64, psi0, a.steps, a.warmup, a.engine)
    if a.save_pk:
        np.save(a.save_pk, P32)
    ev_ms = float(np.mean([r[1]["evolve_ms"] for r in rec32]))
    q_ms = float(np.mean([r[1]["qift_ms"] for r in rec32]))
    pk_ms = float(np.mean([r[1]["pk_ms"] for r in rec32]))
    wall32 = float(np.mean([r[0] for r in rec32]))
    upd_total = (M - 1) * upd_per_step
    value = upd_total / (ev_ms * 1e-3)

    out = {
        "metric": "QPE wall time + det-term updates/s, FP32 vs FP64 GPU at max|dP(k)|<=1e-5",
        "value": value,
        "unit": "det-term updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64 (fp32 state/rotations; fp64 angles, norms, P(k))",
        "data": "synthetic: seeded integrals (SplitMix64/Box-Muller), Hartree-Fock initial state",
        "config": {"workload": f"QPE {2*n} spin-orbitals, {2*e} electrons, m={m} ancillas "
                               f"({M} x {C}^2 = {M*D} amplitudes, alpha x beta layout), "
                               f"{'second-order (symmetric)' if order == 2 else 'first-order'} Trotter, seed {cfg['seed']}",
                   "spin_orbitals": 2 * n, "electrons": 2 * e, "ancillas": m, "determinants_per_block": D,
                   "trotter_order": order,
                   "parallelism": "single GPU"},
        "qpe_wall_s_fp32": wall32,
        "evolve_ms_fp32": ev_ms, "qift_ms_fp32": q_ms, "pk_ms_fp32": pk_ms,
        "updates_per_qpe": upd_total,
    }

    # roofline of the dominant kernel (the term-rotation tile pass)
    pass_bytes = order * info32["pass_bytes"] * (M - 1)
    passes = order * info32["passes"] * (M - 1)
    avg_launch_s = ev_ms * 1e-3 / passes
    achieved = pass_bytes / passes / avg_launch_s / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": MEASURED_TRAFFIC.get((2 * n, 2 * e, "c64", 1)),
                       "kernel": "pass_kernel<float, 1> (term-rotation tile pass)",
                       "bytes_per_launch": pass_bytes / passes, "avg_launch_us": avg_launch_s * 1e6}

    if a.ref_pk:
        P64 = np.load(a.ref_pk)
        out["max_abs_dP_fp32_vs_fp64"] = float(np.max(np.abs(P32 - P64)))
        out["fp64_reference"] = f"P(k) of a complex128 run of the same instance ({a.ref_pk})"
    elif not a.no_fp64:
        rec64, P64, y64 = run_qpe(h, cfg, m, qpesim.COMPLEX128, psi0, 1, 0, a.engine)
        ev64 = rec64[0][1]["evolve_ms"]
        wall64 = rec64[0][0]
        out["qpe_wall_s_fp64"] = wall64
        out["evolve_ms_fp64"] = ev64
        out["value_fp64"] = upd_total / (ev64 * 1e-3)
        out["speedup_fp32_vs_fp64"] = wall64 / wall32
        out["max_abs_dP_fp32_vs_fp64"] = float(np.max(np.abs(P32 - P64)))
        out["l2_rel_fp32_vs_fp64"] = float(np.linalg.norm(y32 - y64) / np.linalg.norm(y64))
        out["sum_P_fp64"] = float(P64.sum())
        out["argmax_P"] = int(P64.argmax())

    if not a.no_cpu:
        cb = cpu_baseline(cfg, one, two, tau, a.cpu_steps)
        out["cpu_baseline"] = {"value": cb["updates_per_s"], "unit": "det-term updates/s", "cores": cb["cores"],
                               "kind": "port",
                               "sample": f"{cb['steps']} Trotter steps U on one block of the same instance "
                                         f"(FP64 OpenMP oracle, reference loop structure); "
                                         f"extrapolated QPE evolve time {(M-1)*cb['seconds_per_step']:.1f}s",
                               "seconds_per_step": cb["seconds_per_step"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
