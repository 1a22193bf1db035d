"""Benchmark: cell-residual evaluations/s of the NS + SA residual on one MI355X.

Workload (the single-GPU configuration the metric is quoted on): NACA0012 from
the analytic 4-digit formula, hybrid grid of 40 quad wall layers (first height
for y+ ~ 1 at Re 6.5e6, growth 1.15) plus triangles out to 30 chords,
120,832 cells; Mach 0.8, AoA 1.25 deg, Re 6.5e6, Roe + Spalart-Allmaras +
Venkatakrishnan limiter, FP64.  Synthetic grid, freestream start, then a few
pseudo-time steps so the limiter and SA terms are active before timing.

One step = one full residual evaluation: the NS residual (gradients, limiter,
MUSCL + Roe, viscous flux, ordered gather) and the SA residual (spectrum,
source, convection, diffusion, gather) -- UPDATE_RESIDUALS of both solvers.
value = nTCell * steps / seconds (cell-residual evaluations per second).

Roofline: algorithmic HBM bytes of one residual evaluation,
B = 552 * nTCell + 64 * F (each array counted once, intermediates written and
read once, FP64), over the measured time of one evaluation; peak 8 TB/s.

cpu_baseline: the CPU oracle (single thread, the reference's loops) timed on
this host on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def c1_config():
    from oneflow_amd.config import CaseConfig
    return CaseConfig(ischeme=1, ilim=2, ivischeme=0, vismodel=3, mach_ref=0.8, aoa_degree=1.25,
                      reynolds=6.5e6, centropy1=0.001, centropy2=0.001, vencat_coef=0.001)


def c1_grid():
    from oneflow_amd import meshgen
    return meshgen.naca_hybrid(n_surf=512, n_quad=40, ratio=1.15, reynolds=6.5e6, n_outer=98, radius=30.0,
                               m=0.0, p=0.4, t=0.12)


def cpu_baseline(grid, cfg, budget_s=10.0):
    """Time the oracle's residual evaluation (NS + SA) on this host, one thread."""
    from oracle.pyoracle import OracleCase
    ora = OracleCase(grid, cfg)
    ora.init_freestream()
    ora.iterate(1, 2)
    ora.timestep(3)
    n = 0
    t0 = time.perf_counter()
    while True:
        ora.residual_ns()
        ora.residual_sa()
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 200:
            break
    dt = time.perf_counter() - t0
    return n, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--pre-iters", type=int, default=20, help="pseudo-time steps before timing")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    from oneflow_amd.solver import Solver

    cfg = c1_config()
    grid = c1_grid()
    print(f"[bench] grid: cells={grid.n_cells} faces={grid.n_faces} bfaces={grid.n_bfaces}", file=sys.stderr, flush=True)
    s = Solver(grid, cfg, device=0)
    s.init_freestream()
    s.iterate(1, args.pre_iters, with_norms=False)
    s.timestep(args.pre_iters + 1)
    T, F = s.n_tcell, s.n_faces

    for _ in range(args.warmup):
        s.residual_ns()
        s.residual_sa()
    s.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.residual_ns()
        s.residual_sa()
    s.sync()
    t1 = time.perf_counter()
    sec = t1 - t0
    ms_per_step = 1e3 * sec / args.steps
    value = T * args.steps / sec

    bytes_alg = 552.0 * T + 64.0 * F
    achieved = bytes_alg / (ms_per_step * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "kernel": "residual evaluation (NS+SA kernels, one stream)",
                "bytes_per_launch": bytes_alg}

    cpu = None
    if not args.no_cpu:
        n, dt = cpu_baseline(grid, cfg, args.cpu_budget)
        cpu = {"value": T * n / dt, "unit": "cell-residual evals/s", "cores": 1, "kind": "port",
               "sample": f"{n} NS+SA residual evaluations of the same {grid.n_cells}-cell grid, CPU oracle, 1 thread, {dt:.1f} s"}

    out = {
        "metric": "cell-residual evals/s + % HBM peak; time-to-1e-8 residual, learned vs plain",
        "value": value,
        "unit": "cell-residual evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "NACA0012 hybrid quad/tri grid, Roe + SA + Venkat, FP64 residual",
                   "cells": grid.n_cells, "faces": grid.n_faces, "nTCell": T,
                   "mach": cfg.mach_ref, "aoa_deg": cfg.aoa_degree, "reynolds": cfg.reynolds},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
