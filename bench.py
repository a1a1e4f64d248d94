#!/usr/bin/env python3
"""Benchmark of the SBE RHS + Dormand-Prince propagation on one MI355X.

Headline workload (cubic3d, the largest single-GPU configuration): seeded 3-D
16-band tight-binding model on a cubic lattice (a = 4 A, R in [-4,4]^3: 729
cells), 48^3 = 110 592 k-points, linear sin^2 pulse (2 um, 10 cycles,
8 MV/cm), T2 = 8 fs, soothing width 25 meV, <= 5 Jacobi sweeps, FP64.  The
2-D 8-band 150 x 150 workload (square2d) and the 1-D two-band chain
(chain1d) are measured the same way and reported as extra keys of the line.

One step = one Dormand-Prince 5(4) step attempt over the whole device-resident
state: 6 RHS evaluations (partial phase sum + the fused stage kernel, which
forms the stage input, finishes the phase sum and applies commutator, Jacobi
and dephasing) and the error-norm reduction read back by the host
controller.  value = N_k * 6 * K / seconds (k-point RHS evaluations per
second).  The CPU baseline is the C/OpenMP oracle running the same attempts on
the host cores (-O3, and -O1 as the reference builds).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config cubic3d|square2d|chain1d]
    python bench.py --scan [--plain-sample P]   (spectra/hour, batched vs plain)
    python bench.py --surrogate                 (spectra/hour, surrogate + fallback vs plain)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "k-point RHS evals/s (FP64); HHG spectra/hour batched+surrogate vs plain"
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


def stage_flops_per_k(n: int, sweeps: int, w_last: int) -> float:
    """algorithmic FLOPs of the fused stage kernel per k-point RHS: commutator
    16 n^3, dephasing transforms 32 n^3, Jacobi 48 S n^2 (n-1), and the last
    phase-sum contraction it performs while loading H0(k) and H_E(k):
    2 operators x W_L complex multiply-adds x n^2 entries x 8 FLOP."""
    return 16 * n ** 3 + 32 * n ** 3 + 48 * sweeps * n * n * (n - 1) + 16 * w_last * n * n


def stage_bytes_per_k(n: int) -> float:
    """algorithmic HBM bytes of the stage kernel per k-point RHS, averaged over
    the six RHS of a Dormand-Prince attempt: reads x and the 10/3 stage
    derivatives it combines on load, writes d rho/dt, and once per attempt the
    new state (1/6): 5.5 complex n x n matrices."""
    return 5.5 * 16 * n * n


def pmc_traffic(path: str):
    """HBM bytes per stage-kernel launch measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE (calibrated against a device copy); the committed summary holds
    the figure and the commands that produced it"""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def sync():
    """device-wide synchronisation through the library's own HIP runtime"""
    from sbewh_amd import hiprt
    hiprt.synchronize()


def physical_cores():
    """physical cores of the host (unique (core, socket) pairs of lscpu -p)"""
    import subprocess
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
        pairs = {ln for ln in out.splitlines() if ln and not ln.startswith("#")}
        return len(pairs) or None
    except (OSError, subprocess.CalledProcessError):
        return None


def cpu_baseline(cfg, state, t, dt, budget_s: float):
    """The C/OpenMP oracle (a restatement of PropagatorCPU + the dopri5 driver)
    timing the same attempts on the host cores, at -O3 and at the reference's
    own -O1; threads pinned close (OMP_PROC_BIND is read when the OpenMP
    runtime starts, i.e. at the oracle's first load in this process)."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle.oracle_py import LIB_PATH_O1, Oracle
    o = Oracle(cfg)
    one = o.bench_attempts(state, t, dt, 1)
    count = max(2, int(budget_s / max(one, 1e-6)))
    el = o.bench_attempts(state, t, dt, count)
    threads = o.num_threads
    o.close()
    o1 = Oracle(cfg, lib_path=LIB_PATH_O1)
    el1 = o1.bench_attempts(state, t, dt, 1)
    o1.close()
    nk = o.num_k
    return {"value": nk * 6 * count / el, "unit": "k-RHS/s", "cores": threads, "kind": "port",
            "sample": f"{count} dopri5 attempts (6 RHS each) on the full {nk}-k state, C/OpenMP oracle -O3, "
                      f"{threads} threads (OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}), {el:.1f} s",
            "value_O1": nk * 6 / el1, "sample_O1": f"1 attempt at -O1 (the reference's CMAKE_CXX_FLAGS), "
                                                   f"{el1:.1f} s",
            "host_physical_cores": physical_cores()}


def spectra_of(t, j, omega0):
    """spectrum + per-harmonic yields of one run (the a13 chain, C++ host code)"""
    from sbewh_amd.spectrum import hhg_spectrum, harmonic_yields
    w, S = hhg_spectrum(t, j)
    return harmonic_yields(w, S, omega0)


def scan_bench(args):
    """64-pulse ellipticity / intensity / CEP scan over the 2-D 8-band model
    with k-grids 16 x {60^2, 90^2, 120^2, 150^2}: complete propagations
    (1024 outputs each) plus spectrum and harmonic yields, batched (one
    sbe_scan) versus the plain solver (one sbe_integrate per pulse, one after
    the other)."""
    import numpy as np
    from sbewh_amd import units
    from sbewh_amd.propagator import SBEContext
    from sbewh_amd.scan import Scan, ellipticity_scan

    cfgs = ellipticity_scan(count=args.scan_count)
    omega0 = units.omega_from_lambda_nm(cfgs[0].pulses[0].lambda_nm)
    sync()
    t0 = time.perf_counter()
    sc = Scan(cfgs)
    print(f"[batched] {len(cfgs)} pulses", file=sys.stderr, flush=True)
    r = sc.run()
    yields = [spectra_of(r["t"][i], r["j"][i], omega0) for i in range(len(cfgs))]
    sync()
    batched_s = time.perf_counter() - t0
    tuning = sc.tuning()
    sc.close()
    del sc
    # plain solver: the same pulses one after the other (all of them, or a
    # stratified sample of `--plain-sample` per grid size, extrapolated)
    by_grid = {}
    for i, c in enumerate(cfgs):
        by_grid.setdefault(c.Nk[0], []).append(i)
    plain_runs = {}
    for g, ids in sorted(by_grid.items()):
        take = ids if args.plain_sample <= 0 else ids[:args.plain_sample]
        for i in take:
            c = cfgs[i]
            sync()
            t1 = time.perf_counter()
            ctx = SBEContext(c)
            x = ctx.initial_state(ctx.fermi_level(c.occupied_bands))
            out = ctx.integrate(x, c.output_count)
            y = spectra_of(out["t"], out["j"], omega0)
            sync()
            plain_runs[i] = (time.perf_counter() - t1, out, y)
            ctx.close()
            print(f"[plain] pulse {i} on {g}^2 k: {plain_runs[i][0]:.2f} s, {out['steps']} steps",
                  file=sys.stderr, flush=True)
    plain_s = 0.0
    for g, ids in by_grid.items():
        ts = [plain_runs[i][0] for i in ids if i in plain_runs]
        plain_s += np.mean(ts) * len(ids)
    # batched == plain on the instances run both ways
    agree = max(float(np.abs(r["j"][i] - plain_runs[i][1]["j"]).max() / np.abs(plain_runs[i][1]["j"]).max())
                for i in plain_runs)
    steps_equal = all(int(r["steps"][i]) == plain_runs[i][1]["steps"] for i in plain_runs)
    n = len(cfgs)
    krhs = sum(c.num_k * int(r["rhs_calls"][i]) for i, c in enumerate(cfgs))
    out = {
        "metric": METRIC,
        "value": n / batched_s * 3600.0,
        "unit": "spectra/hour",
        "n_gpus": 1,
        "steps": n,
        "warmup": 0,
        "ms_per_step": batched_s / n * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"ellipticity scan: {n} elliptic sin2 pulses (3 um, 8 cycles, eps U[0,1], "
                               "E0 U[3,15] MV/cm, CEP U[0,2pi)) over the 2-D 8-band model, k-grids "
                               "{60,90,120,150}^2, T2 10 fs, rtol 1e-8, 1024 outputs, spectrum + yields H1-H40",
                   "total_k": int(sum(c.num_k for c in cfgs)), "outputs": cfgs[0].output_count},
        "batched": {"seconds": batched_s, "spectra_per_hour": n / batched_s * 3600.0,
                    "k_rhs_per_s": krhs / batched_s, "steps_total": int(r["steps"].sum()),
                    "rhs_calls_total": int(r["rhs_calls"].sum()), "stage_tuning": tuning},
        "plain": {"seconds": plain_s, "spectra_per_hour": n / plain_s * 3600.0,
                  "instances_timed": len(plain_runs),
                  "extrapolated": len(plain_runs) < n},
        "batched_over_plain": plain_s / batched_s,
        "batched_vs_plain_max_rel_j": agree,
        "batched_vs_plain_steps_equal": steps_equal,
        "mean_log10_yield_H1_H40": float(np.mean(np.log10(np.maximum(np.array(yields), 1e-300)))),
    }
    print(json.dumps(out), flush=True)


def surrogate_bench(args):
    """Spectra/hour of the learned surrogate with solver fallback versus the
    plain solver, on the surrogate's domain (2-D 8-band model, 15 x 15 k,
    elliptic sin2 pulses with E0, ellipticity, CEP, wavelength, cycles, T2
    drawn from the training box): the held-out queries of the committed data
    set go through Surrogate.query (network, and the solver for the queries
    outside the box or with ensemble disagreement); the plain solver is timed
    on a sample of the same queries, one sbe_integrate each."""
    import numpy as np
    from sbewh_amd.propagator import SBEContext
    from sbewh_amd.surrogate import DATA_DIR, Surrogate, config_for, log_yields, relative_log_error

    z = np.load(os.path.join(DATA_DIR, "surrogate_square2d_15x15.npz"), allow_pickle=False)
    P, Y = z["params"][:512], z["log_yields"][:512]
    s = Surrogate.load(os.path.join(DATA_DIR, "surrogate_ensemble.npz"))
    s.predict(P[:8])
    # two queries outside the training box exercise the fallback
    Q = np.concatenate([P, [[1.0, 0.3, 0.5, 2000.0, 5.0, 10.0], [0.8, 0.5, 1.0, 4400.0, 6.0, 10.0]]])
    sync()
    t0 = time.perf_counter()
    out, used, info = s.query(Q)
    sync()
    t_query = time.perf_counter() - t0
    err = relative_log_error(out[:512], Y)
    # plain solver on a sample of the queries
    plain = []
    for i in range(args.plain_sample_surrogate):
        cfg = config_for(P[i])
        sync()
        t1 = time.perf_counter()
        ctx = SBEContext(cfg)
        x = ctx.initial_state(ctx.fermi_level(cfg.occupied_bands))
        r = ctx.integrate(x, cfg.output_count)
        ly = log_yields(r["t"], r["j"], cfg.pulses[0].lambda_nm)
        sync()
        plain.append(time.perf_counter() - t1)
        ctx.close()
        print(f"[plain] query {i}: {plain[-1]:.2f} s, {r['steps']} steps, max |dlog10 Y| vs data "
              f"{np.abs(ly - Y[i]).max():.2e}", file=sys.stderr, flush=True)
    plain_per = float(np.mean(plain))
    n = len(Q)
    value = n / t_query * 3600.0
    out_json = {
        "metric": METRIC, "value": value, "unit": "spectra/hour", "n_gpus": 1, "steps": n, "warmup": 8,
        "ms_per_step": t_query / n * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16", "data": "synthetic",
        "config": {"workload": "surrogate domain: 2-D 8-band model, 15x15 k, elliptic sin2 pulses, E0 0.3-1.5 V/nm, "
                               "ellipticity 0-1, CEP 0-2pi, 2-4 um, 6-12 cycles, T2 5-20 fs; 512 held-out queries + "
                               "2 outside the training box; log10 yields H1-H40",
                   "queries": n},
        "surrogate": {"seconds": t_query, "fallback_queries": int(used.sum()),
                      "seconds_network": info["seconds_network"], "seconds_solver": info["seconds_solver"],
                      "median_rel_log10_yield_error": float(np.median(err)),
                      "p90_rel_log10_yield_error": float(np.quantile(err, 0.9))},
        "plain": {"seconds_per_spectrum": plain_per, "spectra_per_hour": 3600.0 / plain_per, "sampled": len(plain)},
        "batched_solver_spectra_per_hour": float(len(z["params"]) / z["seconds"].sum() * 3600.0),
        "surrogate_over_plain": value / (3600.0 / plain_per),
    }
    print(json.dumps(out_json), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", default="cubic3d", choices=("square2d", "chain1d", "cubic3d"),
                    help="headline workload of the k-RHS/s measurement (default: the largest, the 3-D "
                         "16-band model on 48^3 k); the other two are measured too, as extra keys")
    ap.add_argument("--headline-only", action="store_true", help="measure only the --config workload")
    ap.add_argument("--scan", action="store_true",
                    help="spectra/hour of the 64-pulse batched scan versus the plain solver")
    ap.add_argument("--scan-count", type=int, default=64)
    ap.add_argument("--plain-sample", type=int, default=2,
                    help="plain runs per grid size (0: all); the rest is extrapolated per grid size")
    ap.add_argument("--surrogate", action="store_true",
                    help="spectra/hour of the surrogate with solver fallback versus the plain solver")
    ap.add_argument("--plain-sample-surrogate", type=int, default=4)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import numpy as np
        from sbewh_amd import hiprt
        nbytes = args.copy_calibration * (1 << 20)
        a = hiprt.DeviceBuffer.from_array(np.ones(nbytes // 8))
        b = hiprt.DeviceBuffer(nbytes)
        for _ in range(3):
            b.copy_from(a, nbytes)
        hiprt.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); sharding the k-points over "
              "several GPUs is designed but not built.", file=sys.stderr)
        sys.exit(2)
    if args.scan:
        scan_bench(args)
        return
    if args.surrogate:
        surrogate_bench(args)
        return

    head = measure(args.config, args.steps, args.warmup, cpu_seconds=None if args.no_cpu else args.cpu_seconds)
    out = head["line"]
    # the other two workloads, measured the same way (no CPU leg), as extra keys
    if not args.headline_only:
        for other in ("cubic3d", "square2d", "chain1d"):
            if other != args.config:
                r = measure(other, max(args.steps, 50), args.warmup, cpu_seconds=None)
                out[other] = {k: r["line"][k] for k in ("value", "unit", "ms_per_step", "roofline", "config",
                                                         "breakdown_ms_per_rhs", "stage_tuning")}
        out["chain1d_integrate"] = measure_integrate("chain1d")
    print(json.dumps(out), flush=True)


WORKLOADS = {
    "square2d": "square2d: seeded 2-D 8-band model, 169 R, 150x150 k, elliptic sin2 3 um eps 0.3 10 MV/cm, "
                "T2 10 fs, rtol 1e-8, Jacobi <= 5 sweeps (warm-started)",
    "chain1d": "chain1d: seeded 1-D 2-band chain, 9 R, 256 k, sin2 3.2 um 8 cycles 5 MV/cm, T2 5 fs, rtol 1e-8",
    "cubic3d": "cubic3d: seeded 3-D 16-band model, 729 R, 48^3 k, sin2 2 um 10 cycles 8 MV/cm, T2 8 fs, "
               "rtol 1e-8, Jacobi <= 5 sweeps (warm-started)",
}


def measure_integrate(name, repeats=3):
    """a whole integrate_const (the model's pulse window and output count)
    with the step-size control on the host and on the resident integrator
    (one cooperative launch): wall time of the call, k-RHS/s of its RHS
    evaluations, best of `repeats` after one untimed run each"""
    import dataclasses
    import time

    from sbewh_amd import model
    from sbewh_amd.propagator import SBEContext

    cfg = model.config(name)
    res = {}
    for label, dev in (("host_controller", False), ("resident", True)):
        ctx = SBEContext(dataclasses.replace(cfg, device_controller=dev))
        x0 = ctx.initial_state(ctx.fermi_level(cfg.occupied_bands))
        ctx.integrate(x0.copy(), cfg.output_count, log_capacity=0)
        best, r = None, None
        for _ in range(repeats):
            x = x0.copy()
            t0 = time.perf_counter()
            r = ctx.integrate(x, cfg.output_count, log_capacity=0)
            dt = time.perf_counter() - t0
            best = dt if best is None or dt < best else best
        res[label] = {"seconds": best, "steps": r["steps"], "rhs_calls": r["rhs_calls"],
                      "k_rhs_per_s": ctx.num_k * r["rhs_calls"] / best, "device_controller": r["device_controller"],
                      "outputs": cfg.output_count}
        ctx.close()
    res["resident_over_host"] = res["host_controller"]["seconds"] / res["resident"]["seconds"]
    print(f"[bench] {name} integrate_const: host {res['host_controller']['seconds'] * 1e3:.1f} ms, resident "
          f"{res['resident']['seconds'] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    return res


def measure(name, steps, warmup, cpu_seconds):
    """k-RHS/s of one workload: `steps` adaptive dopri5 attempts after `warmup`,
    the stage kernel's launch time from HIP events, and its roofline"""
    from sbewh_amd import model
    from sbewh_amd.propagator import SBEContext

    cfg = model.config(name)
    ctx = SBEContext(cfg)
    mu = ctx.fermi_level(cfg.occupied_bands)
    state = ctx.initial_state(mu)
    t0, t1 = ctx.time_window
    t, dt = 0.5 * t0, 0.5   # rising edge of the pulse: E(t) != 0
    n, S, nk = cfg.model.num_wann, cfg.sweeps, ctx.num_k
    print(f"[bench] {name}: {nk} k, n = {n}, {steps} attempts", file=sys.stderr, flush=True)

    sync()
    r = ctx.bench_steps(state, t, dt, warmup, steps, kernel_timing=True)
    sync()
    secs = r["ms"] / 1e3
    value = nk * r["rhs"] / secs
    stage_avg_s = r["stage_ms"] / 1e3
    phase_avg_s = r["phase_ms"] / 1e3
    flops_launch = stage_flops_per_k(n, S, ctx.last_axis_cells) * nk
    flops_launch_table = stage_flops_per_k(n, S, 0) * nk
    bytes_launch = stage_bytes_per_k(n) * nk
    Npad = 2 if n <= 2 else 4 if n <= 4 else 8 if n <= 8 else 16
    # the lower roof of the stage kernel's algorithmic figures bounds it; the
    # kernel runs on the FP64 pipes (78.6 TFLOP/s dense, vector and matrix)
    if flops_launch / (PEAK_FP64_TFLOPS * 1e12) >= bytes_launch / (PEAK_HBM_GBS * 1e9):
        roof = {"bound": "fp64", "achieved": flops_launch_table / stage_avg_s / 1e12, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s"}
    else:
        roof = {"bound": "hbm", "achieved": bytes_launch / stage_avg_s / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    if roof["bound"] == "fp64":
        # the same launch time credited with the fused last-axis contraction too
        roof["frac_with_contraction"] = flops_launch / stage_avg_s / 1e12 / PEAK_FP64_TFLOPS
    traffic = pmc_traffic(os.path.join(ROOT, "profiles", f"stage_kernel_{name}_pmc.md"))
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "k-RHS/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": r["ms"] / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[name] + "; step = one adaptive dopri5 attempt (6 RHS) "
                                                "advancing the state",
                   "num_k": nk, "bands": n, "cells": cfg.model.num_cells, "sweeps": S,
                   "rhs_evals_timed": r["rhs"]},
        "roofline": dict(roof, traffic=traffic, kernel=f"rhs_stage_kernel<{Npad}>", launch_ms=stage_avg_s * 1e3,
                         algorithmic_flop_per_launch=flops_launch_table,
                         algorithmic_bytes_per_launch=bytes_launch),
        "breakdown_ms_per_rhs": {"stage": stage_avg_s * 1e3, "phase_sum": phase_avg_s * 1e3,
                                 "all": r["ms"] / r["rhs"]},
        "stage_tuning": ctx.stage_tuning(),
    }
    line["cpu_baseline"] = cpu_baseline(cfg, state, t, dt, cpu_seconds) if cpu_seconds else None
    ctx.close()
    return {"line": line}


if __name__ == "__main__":
    main()
