"""Benchmark: cell-residual evaluations/s of the NS + SA residual on one MI355X.

Workload: NACA0012 from the analytic 4-digit formula, hybrid grid of 40 quad
wall layers (first height for y+ ~ 1 at Re 6.5e6, growth 1.15) plus
triangles out to 30 chords, 120,832 cells; Mach 0.8, AoA 1.25 deg, Re 6.5e6,
Roe + Spalart-Allmaras + Venkatakrishnan limiter, FP64.  Synthetic grid,
freestream start, then a few pseudo-time steps so the limiter and SA terms
are active before timing.

One step = one full residual evaluation: the NS residual (gradients, limiter,
MUSCL + Roe, viscous flux, ordered gather) and the SA residual (spectrum,
source, convection, diffusion, gather) -- UPDATE_RESIDUALS of both solvers.
value = nTCell * steps / seconds (cell-residual evaluations per second).

roofline: the dominant kernel of the step (largest average duration among the
NS residual kernels, each timed with HIP events on the solver's stream over
--kernel-iters back-to-back launches right after the timed steps); achieved = its algorithmic bytes per launch (each array it must
read or write counted once, see kernel_bytes) / its average duration; peak
8 TB/s.  step_roofline: the whole evaluation's algorithmic bytes
552 * nTCell + 64 * F over ms_per_step.  traffic: measured HBM bytes per
launch of that kernel from the committed rocprofv3 counter summary, when one
for this workload is present.

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

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = ROOT / "profiles" / "naca0012_c1_hbm_traffic.txt"


def kernel_bytes(C, B, F, dim):
    """Algorithmic bytes per launch of the NS residual kernels: every array the
    kernel must read or write counted once (FP64 values, int32 indices).  In 2-D
    the w component and z coordinates are neither read nor written."""
    T = C + B
    nnz = 2 * F - B                  # entries of the cell -> face lists
    nq = 4 if dim == 2 else 5        # primitive components carried
    D = dim
    idx = 4 * (C + 1) + 4 * nnz      # c2f offsets + entries
    return {
        # q, centroids, volume, c2f + c2n, face centre/normal/area -> grad of nq+T, min/max
        "grad": 8 * nq * T + 8 * D * T + 8 * C + idx + 4 * nnz + 8 * (2 * D + 1) * F
                + 8 * (nq + 1) * D * C + 8 * 2 * nq * C,
        # grad, min/max, centroids, c2f, face centres -> limiter
        "limiter": 8 * nq * D * C + 8 * 2 * nq * C + 8 * D * C + idx + 8 * D * F + 8 * nq * C,
        # lc/rc, face normal/area/centre, q, centroids, grads (nq + T), limiter,
        # laminar + eddy viscosity -> inviscid and viscous face fluxes
        "face": 8 * F + 8 * (D + 1) * F + 8 * D * F + 8 * nq * T + 8 * D * T + 8 * (nq + 1) * D * T
                + 8 * nq * T + 16 * T + 2 * 8 * nq * F,
        # c2f, both face fluxes -> residual
        "gather": idx + 2 * 8 * nq * F + 8 * nq * C,
    }


def read_traffic(kernel):
    """Measured HBM bytes per launch for `kernel` from the counter summary, or None."""
    if not TRAFFIC_SUMMARY.exists():
        return None
    for line in TRAFFIC_SUMMARY.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
            return float(parts[2])
    return None


def c1_config():
    from oneflow_amd.config import CaseConfig
    return CaseConfig(ischeme=1, ilim=2, ivischeme=0, vismodel=3, mach_ref=0.8, aoa_degree=1.25,
                      reynolds=6.5e6, centropy1=0.001, centropy2=0.001, vencat_coef=0.001)


def c1_grid(numbering="multicolor"):
    from oneflow_amd import meshgen
    g = meshgen.naca_hybrid(n_surf=512, n_quad=40, ratio=1.15, reynolds=6.5e6, n_outer=98, radius=30.0,
                            m=0.0, p=0.4, t=0.12)
    if numbering == "multicolor":
        g = meshgen.renumber(g, meshgen.multicolor_order(g))
    return g


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


def _sweep_grid(case):
    return case.grid("multicolor")


def batch_main(args):
    """The batched airfoil sweep: 64 seeded NACA 4-digit cases (20k-200k cells,
    Mach 0.3-0.85, AoA -2..6 deg, Re 3e6-1e7, Roe + SA + Venkat, FP64) advanced
    together.  One step = one full pseudo-time step (NS + SA residual, LU-SGS,
    update) of every case.  Timed: graph replays of the batch (face tiles
    autotuned per case) against the same steps launched case after case."""
    from concurrent.futures import ProcessPoolExecutor

    from oneflow_amd.cases import airfoil_sweep
    from oneflow_amd.solver import Batch, Solver

    cases = airfoil_sweep(args.cases, seed=args.seed)
    t0 = time.perf_counter()
    with ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        grids = list(ex.map(_sweep_grid, cases))
    ncell = sum(g.n_cells for g in grids)
    nt = sum(g.n_tcell for g in grids)
    print(f"[bench] {len(grids)} grids, {ncell} cells, generated in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    sol = []
    for g, c in zip(grids, cases):
        s = Solver(g, c.config(), device=0)
        s.init_freestream()
        sol.append(s)
    b = Batch(sol, steps_per_graph=args.graph_steps)
    ramp = max(c.config().ncfl for c in cases)
    b.iterate(1, ramp)  # CFL ramp, launched directly
    tiles = b.autotune(iters=10)
    it = ramp + 1
    b.iterate(it, args.graph_steps)  # records the graph, first replay
    it += args.graph_steps
    t0 = time.perf_counter()
    k = b.iterate(it, args.steps)
    t1 = time.perf_counter()
    it += args.steps
    ms_batch = 1e3 * (t1 - t0) / args.steps
    # the same steps case after case, eager launches, no graph
    seq_steps = max(2, args.steps // 4)
    for s in sol:
        s.sync()
    t0 = time.perf_counter()
    for s in sol:
        s.iterate(it, seq_steps, with_norms=False)
    t1 = time.perf_counter()
    ms_seq = 1e3 * (t1 - t0) / seq_steps
    out = {
        "metric": "cell-steps/s of the batched airfoil sweep (full NS+SA pseudo-time steps)",
        "value": nt * 1e3 / ms_batch,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": ramp + args.graph_steps,
        "ms_per_step": ms_batch,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(cases)} NACA 4-digit airfoils, hybrid grids, Roe + SA + Venkat, FP64",
                   "cells": ncell, "nTCell": nt, "seed": args.seed, "graph_steps": args.graph_steps,
                   "graph_launches": k},
        "sequential_ms_per_step": ms_seq,
        "batch_speedup": ms_seq / ms_batch,
        "face_tiles": {str(t): int(n) for t, n in zip(*np.unique(tiles[:, 0], return_counts=True))},
    }
    print(json.dumps(out))
    b.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--pre-iters", type=int, default=20, help="pseudo-time steps before timing")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--numbering", default="multicolor", choices=["natural", "multicolor"])
    ap.add_argument("--kernel-iters", type=int, default=200, help="back-to-back launches per kernel timing")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30, help="size of the copy-kernel bandwidth reference")
    ap.add_argument("--workload", default="c1", choices=["c1", "batch"],
                    help="c1: single-grid residual throughput (default); batch: the 64-case sweep")
    ap.add_argument("--cases", type=int, default=64)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--graph-steps", type=int, default=10, help="steps recorded per graph replay (batch)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.workload == "batch":
        return batch_main(args)

    from oneflow_amd.solver import Solver

    cfg = c1_config()
    grid = c1_grid(args.numbering)
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

    copy_ms = s.copy_bench(args.copy_bytes, iters=20)
    copy_gbs = 2.0 * args.copy_bytes / (copy_ms * 1e-3) / 1e9
    kb = kernel_bytes(grid.n_cells, grid.n_bfaces, grid.n_faces, 2)
    step_bytes = 552.0 * T + 64.0 * F
    step_gbs = step_bytes / (ms_per_step * 1e-3) / 1e9
    # each NS residual kernel timed with HIP events on the solver's stream over
    # back-to-back launches on the state the timed steps left
    names = {"grad": "k_ns_grad", "limiter": "k_ns_limiter", "face": "k_ns_face", "gather": "k_ns_gather"}
    prof = {key: s.time_kernel(key, iters=args.kernel_iters) for key in names}
    s.residual_ns()  # leave a consistent residual behind
    s.sync()
    kernels = {}
    for key, kname in names.items():
        ms = prof[key]
        kernels[kname] = {"avg_us": 1e3 * ms, "bytes_per_launch": kb[key],
                          "achieved_gbs": kb[key] / (ms * 1e-3) / 1e9 if ms > 0 else None}
    dom = max(names, key=lambda k: prof[k])
    ach = kb[dom] / (prof[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": read_traffic(names[dom]),
                "kernel": names[dom], "avg_us": 1e3 * prof[dom], "bytes_per_launch": kb[dom],
                "launches_timed": args.kernel_iters}

    cpu = None
    if not args.no_cpu:
        n, dt = cpu_baseline(grid, cfg, args.cpu_budget)
        cpu = {"value": T * n / dt, "unit": "cell-residual evals/s", "cores": 1, "kind": "port",
               "sample": f"{n} NS+SA residual evaluations of the same {grid.n_cells}-cell grid, CPU oracle, "
                         f"1 thread of {os.cpu_count()} host cores, {dt:.1f} s"}

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
                   "numbering": args.numbering,
                   "cells": grid.n_cells, "faces": grid.n_faces, "nTCell": T,
                   "mach": cfg.mach_ref, "aoa_deg": cfg.aoa_degree, "reynolds": cfg.reynolds},
        "roofline": roofline,
        "step_roofline": {"bytes_per_step": step_bytes, "achieved": step_gbs, "frac": step_gbs / HBM_PEAK_GBS},
        "kernels": kernels,
        "copy_kernel_gbs": copy_gbs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
