#!/usr/bin/env python3
"""Benchmark of the SBE RHS + Dormand-Prince propagation on one MI355X.

Workload (square2d): 2-D 8-band seeded tight-binding model on a square lattice
(a = 3 A, |R_i| <= 6: 169 cells), 150 x 150 k-points (22 500), elliptic sin^2
pulse (3 um, ellipticity 0.3, 10 MV/cm), T2 = 10 fs, soothing width 25 meV,
5 Jacobi sweeps, FP64.

One step = one Dormand-Prince 5(4) step attempt over the whole device-resident
state: 6 RHS evaluations (phase sum + fused commutator/Jacobi/dephasing stage
kernel), the stage algebra and the error-norm reduction read back by the host
controller.  value = N_k * 6 * K / seconds (k-point RHS evaluations per
second).  The CPU baseline is the C/OpenMP oracle running the same attempts on
the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
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


def stage_flops_per_k(n: int, sweeps: int) -> float:
    """algorithmic FLOPs of the fused stage kernel per k-point RHS: commutator
    16 n^3, dephasing transforms 32 n^3, Jacobi 48 S n^2 (n-1)."""
    return 16 * n ** 3 + 32 * n ** 3 + 48 * sweeps * n * n * (n - 1)


def stage_bytes_per_k(n: int) -> float:
    """algorithmic HBM bytes of the stage kernel per k-point: reads H0, H_E,
    rho_in, writes d rho/dt (4 complex n x n)."""
    return 4 * 16 * n * n


def pmc_traffic(path: str):
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("hbm_bytes_per_launch")
    except Exception:
        return None


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def cpu_baseline(cfg, state, t, dt, budget_s: float):
    from oracle.oracle_py import Oracle
    o = Oracle(cfg)
    one = o.bench_attempts(state, t, dt, 1)
    count = max(2, int(budget_s / max(one, 1e-6)))
    el = o.bench_attempts(state, t, dt, count)
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    return {"value": o.num_k * 6 * count / el, "unit": "k-RHS/s", "cores": cores, "kind": "port",
            "sample": f"{count} dopri5 attempts (6 RHS each) on the full 22500-k state, C/OpenMP oracle, "
                      f"{el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); multi-GPU sharding is "
              "described in DESIGN.md but not built.", file=sys.stderr)
        sys.exit(2)

    from sbewh_amd import model
    from sbewh_amd.propagator import SBEContext

    cfg = model.config("square2d")
    ctx = SBEContext(cfg)
    mu = ctx.fermi_level(cfg.occupied_bands)
    state = ctx.initial_state(mu)
    t0, t1 = ctx.time_window
    t, dt = 0.5 * t0, 0.5   # rising edge of the pulse: E(t) != 0
    n, S, nk = cfg.model.num_wann, cfg.sweeps, ctx.num_k

    sync()
    r = ctx.bench_steps(state, t, dt, args.warmup, args.steps, kernel_timing=True)
    sync()
    secs = r["ms"] / 1e3
    value = nk * r["rhs"] / secs
    stage_avg_s = r["stage_ms"] / 1e3
    phase_avg_s = r["phase_ms"] / 1e3
    flops_launch = stage_flops_per_k(n, S) * nk
    achieved = flops_launch / stage_avg_s / 1e12
    traffic = pmc_traffic(os.path.join(ROOT, "profiles", "stage_kernel_square2d_pmc.json"))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "k-RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "square2d: seeded 2-D 8-band model, 169 R, 150x150 k, elliptic sin2 3 um "
                               "eps 0.3 10 MV/cm, T2 10 fs, rtol 1e-8, Jacobi <= 5 sweeps (warm-started); "
                               "step = one adaptive dopri5 attempt (6 RHS) advancing the state",
                   "num_k": nk, "bands": n, "cells": cfg.model.num_cells, "sweeps": S,
                   "rhs_evals_timed": r["rhs"]},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "kernel": "rhs_stage_kernel<8>", "launch_ms": stage_avg_s * 1e3,
                     "algorithmic_flop_per_launch": flops_launch,
                     "algorithmic_bytes_per_launch": stage_bytes_per_k(n) * nk},
        "breakdown_ms_per_rhs": {"stage": stage_avg_s * 1e3, "phase_sum": phase_avg_s * 1e3,
                                 "all": r["ms"] / r["rhs"]},
        "stage_tuning": ctx.stage_tuning(),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, state, t, dt, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
