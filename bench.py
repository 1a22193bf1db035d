"""Benchmark: cell-residual evaluations/s of the NS + SA residual on one MI355X.

Default workload (the largest configuration that fits one GPU): the 3-D
swept wing -- NACA0012 section, sweep 30 deg, taper 0.5, aspect ratio 4, 25
prism wall layers plus tets, 2,524,160 cells; Mach 0.84, AoA 3 deg, Re 1.2e7,
Roe + Spalart-Allmaras + Venkatakrishnan limiter, FP64, multicolour cell
numbering with interior faces in spatial order.  Synthetic grid, freestream
start, then a few pseudo-time steps so the limiter and SA terms are active
before timing.  `--workload c1` runs the 2-D NACA0012 hybrid grid (120,832
cells), `batch` the 64-case airfoil sweep, `warmstart` time to a 1e-8
residual drop, learned vs plain.

One step = one full residual evaluation: the NS residual (gradients, limiter,
MUSCL + Roe, viscous flux, ordered gather) and the SA residual (spectrum,
source, convection, diffusion, gather) -- UPDATE_RESIDUALS of both solvers.
value = nCells * steps / seconds (cell-residual evaluations per second).

roofline: the evaluation against HBM: achieved = its algorithmic bytes
552 * nTCell + 64 * nFaces (every array of the evaluation counted once:
state, geometry, wall distance, gradients and limiter written and read once,
residual; face normals, centres and cell ids) over the evaluation's average
duration, measured with HIP events on the solver's stream over back-to-back
evaluations; peak 8 TB/s.  traffic: the bytes that left L2 per evaluation
(FETCH_SIZE + WRITE_SIZE of every kernel, rocprofv3 counters, scaled by a
copy kernel) from the committed counter summary of the same workload.
`kernels` lists each kernel with its own algorithmic bytes and HIP-event
time; `pseudo_step` the whole pseudo-time step (residual, LU-SGS, update and
boundary conditions of both solvers).

The CPU leg (left out by --no-cpu) is the only part that runs the oracle:
cpu_baseline, the CPU oracle (single thread, the reference's loops) timed on
this host on a bounded sample of the same workload, and oracle_error, the
GPU's FP64 and FP32-flux residuals checked against the oracle on that sample.
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
TRAFFIC_SUMMARY = {2: ROOT / "profiles" / "naca0012_c1_hbm_traffic.txt",
                   3: ROOT / "profiles" / "swept_wing_c2_hbm_traffic.txt"}


VIS_RECORD_BYTES = {0: 32, 1: 0, 2: 160}  # per-face viscous geometry record read by face-gradient method


def kernel_bytes(C, B, F, dim, flux_bytes=8, sa_vis_method=2, ns_vis_method=0, sa_geo=0):
    """Algorithmic bytes per launch of the NS residual kernels: every array the
    kernel must read or write counted once (FP64 values, int32 indices; the
    face fluxes are FP32 in the FP32-flux variant, flux_bytes = 4).  In 2-D the
    w component and z coordinates are neither read nor written.  Static face
    terms stored at mesh upload count as what the kernels read in place of the
    metrics they are formed from: Green-Gauss weights and area-weighted normal
    (4 doubles per face, 3-D + 1), face centre relative to both cell centres
    (4 doubles, 3-D + 2), and the viscous face-gradient record of the NS and
    SA methods (ns_vis_method, sa_vis_method: STD 4 doubles, AVER none,
    TEST 20 doubles; with sa_geo = 1 the SA kernel evaluates the TEST
    geometry from the face centre and both cell centres instead)."""
    T = C + B
    nnz = 2 * F - B                  # entries of the cell -> face lists
    nq = 4 if dim == 2 else 5        # primitive components carried
    D = dim
    idx = 4 * (C + 1) + 4 * nnz      # face-list counts + entries
    gg = 8 * (4 + (D == 3)) * F      # Green-Gauss weights + area-weighted normal
    rec = 8 * (4 + 2 * (D == 3)) * F  # face centre - left / right cell centre
    return {
        # q, volume, face lists (face + neighbour), weights, centre offsets ->
        # gradients of nq + T and the limiter
        "grad": 8 * nq * T + 8 * C + idx + 4 * nnz + gg + rec + 8 * (nq + 1) * D * C + 8 * nq * C,
        # q, face lists -> the limiter's grid extrema (when no time step produced them)
        "limiter": 8 * nq * T + idx + 4 * nnz,
        # lc/rc, face normal/area, centre offsets, viscous vector, q, grads (nq + T),
        # limiter, laminar + eddy viscosity -> inviscid and viscous face fluxes
        "face": 8 * F + 8 * (D + 1) * F + rec + VIS_RECORD_BYTES[ns_vis_method] * F + 8 * nq * T + 8 * (nq + 1) * D * T
                + 8 * nq * T + 16 * T + 2 * flux_bytes * nq * F,
        # face lists, both face fluxes -> residual
        "gather": idx + 2 * flux_bytes * nq * F + 8 * nq * C,
        # SA: nu_tilde, face lists (face + neighbour), weights, volume, wall-face
        # types -> boundary ghosts of nu_tilde and its gradient
        "sa_grad": 8 * T + idx + 4 * nnz + gg + 8 * C + 4 * B + 8 * B + 8 * D * C,
        # lc/rc, normal/area, viscous record, rho + velocity, nu_tilde, mu_l,
        # volume, grad nu_tilde, wall-face types -> 5 face terms
        "sa_face": 8 * F + 8 * (D + 1) * F + (8 * D * F + 8 * D * T if sa_geo and sa_vis_method == 2
                                               else VIS_RECORD_BYTES[sa_vis_method] * F) + 8 * (D + 1) * T
                   + 8 * 3 * T + 8 * D * C + 4 * B + 40 * F,
        # face lists, 3 face terms, velocity gradient, rho, nu_tilde, mu_l, wall
        # distance, grad nu_tilde, volume, time step -> diagonal and residual
        "sa_cell": idx + 24 * F + 8 * D * D * C + 8 * 6 * C + 8 * D * C + 16 * C,
    }


def read_traffic(kernel, dim=2):
    """Measured HBM bytes per launch for `kernel` from the counter summary of
    the 2-D (NACA0012) or 3-D (swept wing) benchmark grid, or None."""
    path = TRAFFIC_SUMMARY[dim]
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
            return float(parts[2])
    return None


def c1_config():
    from oneflow_amd.config import CaseConfig
    return CaseConfig(ischeme=1, ilim=2, ivischeme=0, vismodel=3, mach_ref=0.8, aoa_degree=1.25,
                      reynolds=6.5e6, centropy1=0.001, centropy2=0.001, vencat_coef=0.001)


def _numbered(g, numbering, face_order="cell"):
    """The grid under a cell numbering; face_order "spatial" then sorts the
    interior faces by the cells' ids before the numbering (their spatial
    order), "cell" keeps them sorted by their renumbered cells."""
    from oneflow_amd import meshgen
    if numbering == "natural":
        return g
    order = meshgen.blocked_multicolor_order(g) if numbering == "blocked" else meshgen.multicolor_order(g)
    g2 = meshgen.renumber(g, order)
    if face_order == "spatial":
        g2 = meshgen.sort_interior_faces(g2, order)
    return g2


def c1_grid(numbering="multicolor", face_order="cell"):
    from oneflow_amd import meshgen
    g = meshgen.naca_hybrid(n_surf=512, n_quad=40, ratio=1.15, reynolds=6.5e6, n_outer=98, radius=30.0,
                            m=0.0, p=0.4, t=0.12)
    return _numbered(g, numbering, face_order)


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


def _sweep_grid(case, face_order="cell"):
    return case.grid("multicolor", face_order)


def _sweep_grid_spatial(case):
    return case.grid("multicolor", "spatial")


def batch_main(args):
    """The batched airfoil sweep: 64 seeded NACA 4-digit cases (20k-200k cells,
    Mach 0.3-0.85, AoA -2..6 deg, Re 3e6-1e7, Roe + SA + Venkat, FP64) advanced
    together.  One step = one full pseudo-time step (NS + SA residual, LU-SGS,
    update) of every case.  Timed: graph replays of the batch (one stream per
    case, face tiles autotuned per case) and the cases as zones of one context
    (every kernel spans all cases), against the same steps launched case after
    case.  value: the faster of the two batched forms."""
    from concurrent.futures import ProcessPoolExecutor

    from oneflow_amd.cases import airfoil_sweep
    from oneflow_amd.solver import Batch, Solver, ZoneSolver

    cases = airfoil_sweep(args.cases, seed=args.seed)
    t0 = time.perf_counter()
    with ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        grids = list(ex.map(_sweep_grid_spatial if args.face_order == "spatial" else _sweep_grid, cases))
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
    b.close()
    for s in sol:
        s.close()
    # the same cases as zones of one context: one launch sequence for all of them
    zs = ZoneSolver(grids, [c.config() for c in cases])
    zs.init_freestream()
    zs.iterate(1, ramp, with_norms=False)
    ztile = zs.autotune(iters=10)
    zs.iterate(ramp + 1, args.warmup, with_norms=False)
    zs.sync()
    t0 = time.perf_counter()
    zs.iterate(ramp + 1 + args.warmup, args.steps, with_norms=False)
    zs.sync()
    t1 = time.perf_counter()
    ms_zones = 1e3 * (t1 - t0) / args.steps
    # the zones' steps recorded into one HIP graph: one launch advances every
    # case graph_steps steps
    it = ramp + 1 + args.warmup + args.steps
    zb = Batch([zs], steps_per_graph=args.graph_steps)
    zb.iterate(it, args.graph_steps)  # records the graph, first replay
    it += args.graph_steps
    zs.sync()
    nz = max(args.graph_steps, args.steps // args.graph_steps * args.graph_steps)
    t0 = time.perf_counter()
    kz = zb.iterate(it, nz)
    zs.sync()
    t1 = time.perf_counter()
    ms_zgraph = 1e3 * (t1 - t0) / nz
    zb.close()
    zs.close()
    print(f"[bench] graph batch {ms_batch:.3f} ms/step, zones {ms_zones:.3f}, zones in one graph {ms_zgraph:.3f}, "
          f"case after case {ms_seq:.3f}", file=sys.stderr, flush=True)
    ms_best = min(ms_batch, ms_zones, ms_zgraph)
    out = {
        "metric": "cell-steps/s of the batched airfoil sweep (full NS+SA pseudo-time steps)",
        "value": ncell * 1e3 / ms_best,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": ramp + args.graph_steps,
        "ms_per_step": ms_best,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(cases)} NACA 4-digit airfoils, hybrid grids, Roe + SA + Venkat, FP64",
                   "cells": ncell, "nTCell": nt, "seed": args.seed, "face_order": args.face_order or "cell",
                   "graph_steps": args.graph_steps,
                   "graph_launches": k},
        "sequential_ms_per_step": ms_seq,
        "graph_ms_per_step": ms_batch,
        "zones_ms_per_step": ms_zones,
        "zones_graph_ms_per_step": ms_zgraph,
        "zones_graph_launches": kz,
        "batch_speedup": ms_seq / ms_batch,
        "zones_speedup": ms_seq / ms_zones,
        "face_tiles": {str(t): int(n) for t, n in zip(*np.unique(tiles[:, 0], return_counts=True))},
        "zones_face_tiles": [int(v) for v in ztile],
    }
    print(json.dumps(out))


def _every(n, fn):
    """A chunk callback that calls fn(step, running) once per n steps."""
    last = [0]

    def f(it, nl):
        if it >= last[0] + n:
            last[0] = it
            fn(it, nl)
    return f


def _sweep_grid_o1(case):
    return case.grid("multicolor")


def warmstart_main(args):
    """Time to a 1e-8 residual drop on unseen airfoils: the plain solver case
    after case from freestream, all cases as zones of one context from
    freestream, and the zones from the learned warm start.  The network is
    trained first on fields the solver converges itself (a separate seeded
    training sweep, also solved as zones).  Spatial order 1 (ilim = -1) -- the
    reference's second-order Venkat setting does not converge to 1e-8 on these
    grids.

    --phase all runs everything in one process.  For sweeps whose plain
    case-after-case baseline does not fit one run, the phases run separately
    and meet in --ws-dir: train (fields + network -> net.npy, train.json),
    seq --seq-range a:b (the plain solver on test cases a..b-1 ->
    seq_a_b.json; the baseline is a sum of independent per-case times), zones
    (plain and learned zones -> zones.json), combine (the JSON line).  Every
    phase regenerates the same seeded cases."""
    from pathlib import Path

    from oneflow_amd import warmstart as ws
    from oneflow_amd.cases import airfoil_sweep
    from oneflow_amd.mlp import Mlp
    from oneflow_amd.solver import Solver, ZoneSolver

    log = lambda m: print(f"[warmstart] {m}", file=sys.stderr, flush=True)
    phase = args.phase
    wsdir = Path(args.ws_dir)
    wsdir.mkdir(parents=True, exist_ok=True)
    train = airfoil_sweep(args.n_train, seed=args.seed + 1, min_cells=args.train_min_cells, max_cells=args.train_max_cells)
    test = airfoil_sweep(args.n_test, seed=args.seed, min_cells=20_000, max_cells=args.test_max_cells)
    if args.test_ids:
        test = [test[int(i)] for i in args.test_ids.split(",")]
    cfg_of = lambda c: c.config(ilim=-1, cfled=args.cfl)
    if phase == "combine":
        return _warmstart_combine(args, wsdir, train, test)

    def grids_of(cases):
        from concurrent.futures import ProcessPoolExecutor
        t0 = time.perf_counter()
        with ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
            g = list(ex.map(_sweep_grid_o1, cases))
        log(f"{len(g)} grids, {sum(x.n_cells for x in g)} cells, {time.perf_counter() - t0:.1f} s")
        return g

    info = {}
    if phase in ("all", "train"):
        gtrain = grids_of(train)
        # 1. training fields: the training cases as zones of one context, each
        #    converged to the tolerance (a tol drop from its peak density
        #    residual over the first 1000 steps).  Converged cases leave the run
        #    as the held-out runs do (--compact); each case's samples are read
        #    from the last context that holds it, once, when that context is left.
        zt = ZoneSolver(gtrain, [cfg_of(c) for c in train])
        zt.init_freestream()
        t0 = time.perf_counter()
        peak = np.zeros(len(train))
        for it in range(1, 1001, 100):
            zt.iterate(it, 100, with_norms=False)
            peak = np.maximum(peak, zt.zone_norms()[:, 0])
        xs, ys = {}, {}

        def take(ctx, pairs):
            fc = ws.FieldCache(ctx)
            for z, cid in pairs:
                v = ws.ZoneView(fc, z)
                xs[cid] = ws.cell_features(v, train[cid])
                ys[cid] = ws.cell_targets(v, train[cid], v.cfg)
        rebuild_t = None
        if args.compact > 0:
            rebuild_t = lambda ids: ZoneSolver([gtrain[i] for i in ids], [cfg_of(train[i]) for i in ids])
        _, tsteps, _, tstat = ws.solve_zones_to_tolerance(
            zt, list(peak), tol=args.tol, chunk=2000, max_steps=args.train_steps, first_iter=1001,
            progress=lambda i, n: log(f"  training case {i} done at step {n}"),
            on_chunk=_every(10000, lambda it, nl: log(f"  training fields: step {it}, {nl} running")),
            rebuild=rebuild_t, compact=args.compact, on_leave=take)
        t_gen = time.perf_counter() - t0
        used = [i for i in range(len(train)) if tstat[i] == ws.CONVERGED]
        log(f"training fields: steps {min(tsteps)}-{max(tsteps)}, {t_gen:.1f} s, {len(used)} of {len(train)} converged")
        X = np.concatenate([xs[i] for i in used])
        Y = np.concatenate([ys[i] for i in used])
        del xs, ys
        zt.close()
        n_train_cells = int(sum(g.n_cells for g in gtrain))
        del gtrain
        log(f"training fields: {X.shape[0]} samples in {t_gen:.1f} s")
        # 2. train the network on the GPU
        net = Mlp(seed=args.seed)
        t0 = time.perf_counter()
        losses = net.train(X, Y, batch=8192, steps=args.mlp_steps, lr=1e-3)
        t_train = time.perf_counter() - t0
        flops_step = 3 * 2 * 8192 * (16 * 256 + 4 * 256 * 256 + 256 * 64)
        log(f"trained {args.mlp_steps} steps in {t_train:.1f} s, loss {losses[:10].mean():.3e} -> {losses[-50:].mean():.3e}")
        xt = np.random.default_rng(0).normal(0, 1, (1 << 20, 9)).astype(np.float32)
        net.forward(xt)  # sizes the activation buffers
        t0 = time.perf_counter()
        net.forward(xt)
        t_fwd = time.perf_counter() - t0
        info = {"train_cases": len(train), "train_samples": int(X.shape[0]), "train_max_steps": args.train_steps,
                "train_case_steps": [int(v) for v in tsteps], "train_case_status": tstat,
                "train_cases_used": len(used), "train_cells": n_train_cells,
                "training_fields_s": t_gen, "training_s": t_train,
                "final_loss": float(losses[-50:].mean()), "first_loss": float(losses[:10].mean()),
                "train_tflops": flops_step * args.mlp_steps / t_train / 1e12, "forward_1M_ms": 1e3 * t_fwd,
                "forward_tflops": 2.0 * (1 << 20) * (16 * 256 + 4 * 256 * 256 + 256 * 64) / t_fwd / 1e12}
        del X, Y
        if phase == "train":
            np.save(wsdir / "net.npy", net.params())
            (wsdir / "train.json").write_text(json.dumps(info))
            log(f"network and training record written to {wsdir}")
            return

    # 3. unseen cases: reference levels (untimed pre-pass of the plain run)
    lo, hi = 0, len(test)
    if phase == "seq" and args.seq_range:
        lo, hi = (int(v) for v in args.seq_range.split(":"))
    cases = test[lo:hi] if phase == "seq" else test
    gtest = grids_of(cases)

    def fresh(i):
        s = Solver(gtest[i], cfg_of(cases[i]))
        s.init_freestream()
        return s
    peaks = []
    for i in range(len(cases)):
        s = fresh(i)
        peaks.append(s.iterate(1, 1000).max(axis=0))  # peak norm per equation over the first 1000 steps
        s.close()
    peaks = np.array(peaks)
    refs = [float(v) for v in peaks[:, 0]]
    prog = lambda i, n: log(f"  case {lo + i} done at step {n}")
    seq = {}
    if phase in ("all", "seq"):
        # plain solver, case after case: one context at a time, each timed alone
        t_case, st_seq, aero_seq, stat_seq = [], [], [], []
        for i in range(len(cases)):
            s = fresh(i)
            t, st, status = ws.solve_to_tolerance([s], [refs[i]], tol=args.tol, batched=False)
            t_case.append(t)
            st_seq.append(int(st[0]))
            stat_seq.append(status[0])
            aero_seq.append(s.aero().tolist() if status[0] != ws.DIVERGED else None)
            s.close()
            prog(i, st[0])
        seq = {"range": [lo, hi], "seconds_per_case": t_case, "steps": st_seq, "aero": aero_seq,
               "status": stat_seq}
        log(f"plain, case after case: {sum(t_case):.1f} s")
        if phase == "seq":
            (wsdir / f"seq_{lo}_{hi}.json").write_text(json.dumps(seq))
            return

    if phase == "zones":
        net = None
        if args.zones_run != "plain":
            net = Mlp(seed=args.seed)
            net.set_params(np.load(wsdir / "net.npy"))
    zres = _warmstart_zones(args, test, gtest, cfg_of, refs, peaks, ws, ZoneSolver, log, net)
    if phase == "zones":
        ids = [int(i) for i in args.test_ids.split(",")] if args.test_ids else list(range(len(test)))
        zres["test_ids"] = ids
        name = "zones" if args.zones_run == "both" else f"zones_{args.zones_run}"
        if args.test_ids:
            name += f"_{ids[0]}-{ids[-1]}"
        (wsdir / f"{name}.json").write_text(json.dumps(zres))
        return
    print(json.dumps(_warmstart_result(args, train, test, info, [seq], zres)))


def _warmstart_zones(args, test, gtest, cfg_of, refs, peaks, ws, ZoneSolver, log, net):
    """Plain and learned zones to the tolerance (and optionally past it)."""
    prog = lambda i, n: log(f"  zone {i} done at step {n}")
    chunk_log = _every(10000, lambda it, nl: log(f"  step {it}, {nl} running"))

    def zones():
        z = ZoneSolver(gtest, [cfg_of(c) for c in test])
        z.init_freestream()
        return z
    # converged cases leave: the unconverged ones move into a smaller context
    rebuild = None
    if args.compact > 0:
        rebuild = lambda ids: ZoneSolver([gtest[i] for i in ids], [cfg_of(test[i]) for i in ids])
    out = {}
    which = args.zones_run
    if which in ("plain", "both"):
        zp = zones()
        t_zp, st_zp, aero_zp, stat_zp = ws.solve_zones_to_tolerance(zp, refs, tol=args.tol, progress=prog,
                                                                    rebuild=rebuild, compact=args.compact,
                                                                    on_chunk=chunk_log, max_steps=args.max_steps)
        log(f"plain, zones: {t_zp:.1f} s")
        out.update({"plain_zones_s": t_zp, "plain_zones_steps": [int(v) for v in st_zp],
                    "aero_plain_zones": aero_zp, "status_plain_zones": stat_zp})
    if which in ("learned", "both"):
        zl = zones()
        zl.sync()
        t0 = time.perf_counter()
        ws.apply_predictions_zones(zl, test, net, mach_gate=args.mach_gate if args.mach_gate > 0 else None)
        zl.sync()
        t_inf = time.perf_counter() - t0
        t_zl, st_zl, aero_zl, stat_zl = ws.solve_zones_to_tolerance(zl, refs, tol=args.tol, progress=prog,
                                                                    rebuild=rebuild, compact=args.compact,
                                                                    on_chunk=chunk_log, max_steps=args.max_steps)
        log(f"learned, zones: {t_inf:.2f} s inference + {t_zl:.1f} s")
        out.update({"learned_zones_s": t_zl, "inference_s": t_inf, "learned_zones_steps": [int(v) for v in st_zl],
                    "aero_learned_zones": aero_zl, "status_learned_zones": stat_zl})
    if which != "both":
        if which == "plain":
            zp.close()
        else:
            zl.close()
        return out
    if args.tight and args.compact <= 0:
        # both zone runs continued (untimed) until every equation's norm (SA
        # included) is 100x below tol times its peak: the forces both starts reach
        lim = args.tol * 1e-2 * np.where(peaks > 0, peaks, np.inf)
        for zs, it in ((zp, max(st_zp) + 1), (zl, max(st_zl) + 1)):
            while np.any(zs.zone_norms() > lim) and it < 400_000:
                zs.iterate(it, 2000, with_norms=False)
                it += 2000
            log(f"continued to step {it - 1}")
        out["aero_plain_tight"] = zp.zone_aero().tolist()
        out["aero_learned_tight"] = zl.zone_aero().tolist()
    zp.close()
    zl.close()
    return out


def _warmstart_result(args, train, test, info, seqs, z):
    """The JSON line of the warm-start workload.  value: the learned start's
    own gain, plain zones from freestream over learned zones (inference
    included) -- the same batched solver on both sides, so batching is not
    credited to the network.  Forces are compared on the cases both runs
    converged; cases that hit max_steps or diverged are counted, not timed as
    converged."""
    seqs = [s for s in seqs if s]
    if seqs:
        t_case = [t for s in seqs for t in s["seconds_per_case"]]
        st_seq = [v for s in seqs for v in s["steps"]]
        stat_seq = [v for s in seqs for v in s.get("status", ["converged"] * len(s["steps"]))]
        aero_seq = [a for s in seqs for a in s["aero"]]
        t_seq = float(sum(t_case))
    else:  # no case-after-case run: forces against the plain zones (the same solver)
        st_seq = z["plain_zones_steps"]
        stat_seq = z.get("status_plain_zones", ["converged"] * len(test))
        aero_seq = z["aero_plain_zones"]
        t_seq = None
    t_zp, t_zl, t_inf = z["plain_zones_s"], z["learned_zones_s"], z["inference_s"]
    stat_zp = z.get("status_plain_zones", ["converged"] * len(test))
    stat_zl = z.get("status_learned_zones", ["converged"] * len(test))
    ok = [i for i in range(len(test)) if stat_seq[i] == stat_zl[i] == "converged"]
    aero_zl = np.array([z["aero_learned_zones"][i] for i in ok]).reshape(-1, 4)
    aero_sq = np.array([aero_seq[i] for i in ok]).reshape(-1, 4)

    def rel(a, b, k):
        return np.abs(a[:, k] - b[:, k]) / np.maximum(np.abs(b[:, k]), 1e-12)
    dcl, dcd = rel(aero_zl, aero_sq, 0), rel(aero_zl, aero_sq, 1)
    count = lambda st: {k: int(sum(1 for v in st if v == k)) for k in ("converged", "max_steps", "diverged")}
    out = {
        "metric": "time to a 1e-8 residual drop, learned vs plain",
        "value": t_zp / (t_inf + t_zl),
        "unit": "x speedup (plain zones from freestream / zones from the learned warm start)",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64 solver, bf16 network",
        "data": "synthetic",
        "config": {"workload": f"{len(test)} unseen NACA 4-digit airfoils (20k-{args.test_max_cells // 1000}k cells), "
                               "Roe + SA, first order",
                   "train_cases": len(train), "tol": args.tol, "cfl": args.cfl, "mach_gate": args.mach_gate,
                   "compact": args.compact},
        "seconds": {"plain_sequential": t_seq, "plain_zones": t_zp, "learned_zones": t_zl, "inference": t_inf},
        "speedup": {"zones": t_seq / t_zp if t_seq else None, "warm_start": t_zp / (t_inf + t_zl),
                    "total": t_seq / (t_inf + t_zl) if t_seq else None},
        "status": {"plain_sequential": count(stat_seq) if seqs else None, "plain_zones": count(stat_zp),
                   "learned_zones": count(stat_zl)},
        "forces_compared_against": "plain case after case" if seqs else "plain zones",
        "aero_rel_diff_learned_vs_plain": {
            "max_CL": float(dcl.max()) if len(ok) else None, "max_CD": float(dcd.max()) if len(ok) else None,
            "cases_within_1e-6": int(np.sum((dcl <= 1e-6) & (dcd <= 1e-6))), "cases_compared": len(ok),
            "cases": len(test)},
        "steps": {"plain_per_case": st_seq, "learned_zones_per_case": z["learned_zones_steps"]},
        "per_case": [{"mach": c.mach, "aoa": c.aoa, "t": c.t, "cells": c.target_cells,
                      "dCL": float(abs(a[0] - b[0]) / max(abs(b[0]), 1e-12)) if a and b else None,
                      "dCD": float(abs(a[1] - b[1]) / max(abs(b[1]), 1e-12)) if a and b else None}
                     for c, a, b in zip(test, z["aero_learned_zones"], aero_seq)],
        "training": info,
    }
    if "aero_plain_tight" in z:
        tp, tl = np.array(z["aero_plain_tight"]), np.array(z["aero_learned_tight"])
        out["aero_rel_diff_learned_vs_plain_tight"] = {"max_CL": float(rel(tl, tp, 0).max()),
                                                       "max_CD": float(rel(tl, tp, 1).max())}
    return out


def _warmstart_combine(args, wsdir, train, test):
    info = json.loads((wsdir / "train.json").read_text())
    seqs = sorted((json.loads(p.read_text()) for p in wsdir.glob("seq_*.json")), key=lambda s: s["range"][0])
    covered = [i for s in seqs for i in range(*s["range"])]
    if seqs and covered != list(range(len(test))):
        raise SystemExit(f"seq phases cover {covered[:3]}..{covered[-3:]} of {len(test)} cases; run the missing ranges")
    if (wsdir / "zones.json").exists():
        z = json.loads((wsdir / "zones.json").read_text())
    else:  # the zone runs made by separate processes, possibly over parts of the cases
        z = {}
        for run in ("plain", "learned"):
            parts = [json.loads(p.read_text()) for p in sorted(wsdir.glob(f"zones_{run}*.json"))]
            n = len(test)
            merged = {}
            for part in parts:
                for key, val in part.items():
                    if key == "test_ids":
                        continue
                    if isinstance(val, list):
                        merged.setdefault(key, [None] * n)
                        for i, v in zip(part["test_ids"], val):
                            merged[key][i] = v
                    else:  # seconds: the parts ran one after another
                        merged[key] = merged.get(key, 0.0) + val
            missing = [i for i in range(n) if merged.get(f"status_{run}_zones", [None] * n)[i] is None]
            if missing:
                raise SystemExit(f"zones_{run}: cases {missing[:5]}... not run")
            z.update(merged)
    print(json.dumps(_warmstart_result(args, train, test, info, seqs, z)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--pre-iters", type=int, default=20, help="pseudo-time steps before timing")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--numbering", default=None, choices=["natural", "multicolor", "blocked"],
                    help="cell numbering (default multicolor)")
    ap.add_argument("--face-order", default=None, choices=["cell", "spatial"],
                    help="interior faces sorted by their renumbered cells or by their spatial order "
                         "(default: cell for c1, spatial for c2)")
    ap.add_argument("--kernel-iters", type=int, default=200, help="back-to-back launches per kernel timing")
    ap.add_argument("--storage", type=int, default=None, choices=[0, 1],
                    help="device storage order: 0 the grid's (library default), 1 spatial")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30, help="size of the copy-kernel bandwidth reference")
    ap.add_argument("--launch", default="",
                    help="c1/c2: fixed launch configuration instead of the autotuner, "
                         "ns_tile,sa_tile,gather,face_split,sa_geometry,grad,sa_grad,normals (the order of "
                         "the bench line's launch record)")
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "batch", "warmstart", "mlp"],
                    help="c2: the 3-D swept wing (default, the largest single-GPU configuration); "
                         "c1: the 2-D NACA0012 grid; batch: the 64-case sweep; "
                         "warmstart: time to 1e-8, learned vs plain; mlp: the warm-start network's "
                         "forward pass and training steps alone")
    ap.add_argument("--mlp-samples", type=int, default=1 << 20, help="mlp: samples per forward pass")
    ap.add_argument("--mlp-batch", type=int, default=4096, help="mlp: training batch")
    ap.add_argument("--n-train", type=int, default=48)
    ap.add_argument("--n-test", type=int, default=16)
    ap.add_argument("--train-steps", type=int, default=200_000,
                    help="warmstart: step cap while the training fields converge to --tol")
    ap.add_argument("--train-min-cells", type=int, default=20_000, help="warmstart: smallest training grid")
    ap.add_argument("--train-max-cells", type=int, default=50_000, help="warmstart: largest training grid")
    ap.add_argument("--test-max-cells", type=int, default=200_000,
                    help="warmstart: largest held-out grid (the batched sweep's range is 20k-200k)")
    ap.add_argument("--test-ids", default="", help="warmstart: comma-separated subset of the held-out cases")
    ap.add_argument("--max-steps", type=int, default=250_000, help="warmstart: step cap of the held-out zone runs")
    ap.add_argument("--zones-run", default="both", choices=["both", "plain", "learned"],
                    help="warmstart zones phase: both zone runs, or one of them (zones_plain.json / "
                         "zones_learned.json, merged by the combine phase)")
    ap.add_argument("--mlp-steps", type=int, default=3000)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--cfl", type=float, default=10.0, help="final CFL of the warm-start cases")
    ap.add_argument("--phase", default="all", choices=["all", "train", "seq", "zones", "combine"],
                    help="warmstart: run everything, or one phase of a split run (see warmstart_main)")
    ap.add_argument("--ws-dir", default="warmstart_run", help="warmstart: where split phases meet")
    ap.add_argument("--seq-range", default="", help="warmstart seq phase: test cases a:b")
    ap.add_argument("--tight", action="store_true", help="warmstart: continue both zone runs 100x past tol")
    ap.add_argument("--mach-gate", type=float, default=0.0,
                    help="warmstart: cases above this Mach keep the freestream start (0: no gate)")
    ap.add_argument("--compact", type=float, default=0.5,
                    help="warmstart: move unconverged cases into a smaller context once they are at most this "
                         "fraction of the zones (0: never)")
    ap.add_argument("--cases", type=int, default=64)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--graph-steps", type=int, default=10, help="steps recorded per graph replay (batch)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.workload == "batch":
        return batch_main(args)
    if args.workload == "warmstart":
        return warmstart_main(args)
    if args.workload == "mlp":
        return mlp_main(args)

    return grid_main(args)


# Multiply-adds of one sample through the padded network (16 -> 256 x 5 -> 64)
MLP_MACS = 16 * 256 + 4 * 256 * 256 + 256 * 64


def mlp_main(args):
    """The warm-start network alone: `steps` forward passes of --mlp-samples
    samples in each form (one kernel, the default; one GEMM per layer), host
    timed with the PCIe copies of inputs and outputs included, and
    --mlp-steps training steps of --mlp-batch samples (forward, backward and
    Adam; one call).  Kernel durations come from rocprofv3 --kernel-trace;
    FLOP counts use the padded layer shapes."""
    from oneflow_amd.mlp import Mlp
    rng = np.random.default_rng(args.seed)
    n = args.mlp_samples
    x = rng.normal(0.0, 1.0, (n, 9)).astype(np.float32)
    net = Mlp(seed=1)
    fwd = {}
    for variant in (0, 1):
        net.set_forward_variant(variant)
        net.forward(x[:4096])
        t0 = time.perf_counter()
        for _ in range(args.steps):
            net.forward(x)
        fwd[variant] = (time.perf_counter() - t0) / args.steps
    fwd_s = fwd[1]
    nt = 1 << 16
    xt = x[:nt]
    yt = np.tanh(xt[:, :5]).astype(np.float32)
    net.train(xt, yt, batch=args.mlp_batch, steps=10)
    t0 = time.perf_counter()
    losses = net.train(xt, yt, batch=args.mlp_batch, steps=args.mlp_steps)
    train_s = (time.perf_counter() - t0) / args.mlp_steps
    net.close()
    flop_fwd = 2.0 * MLP_MACS * n
    flop_step = 3 * 2.0 * MLP_MACS * args.mlp_batch  # forward + two backward products per layer
    print(json.dumps({
        "metric": "mlp forward samples/s", "value": n / fwd_s, "unit": "samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": 1e3 * fwd_s, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": "mlp", "samples": n, "train_batch": args.mlp_batch},
        "forward_tflops_host": flop_fwd / fwd_s / 1e12,
        "forward_layered_ms_host": 1e3 * fwd[0],
        "train": {"ms_per_step": 1e3 * train_s, "tflops_host": flop_step / train_s / 1e12,
                  "loss_first": float(losses[0]), "loss_last": float(losses[-1])}}), flush=True)


def grid_main(args):
    """One grid (c1: the 2-D NACA0012; c2: the 3-D swept wing): FP64 residual
    throughput and roofline, the FP32-flux variant (c2), the whole pseudo-step,
    the error against the oracle and the CPU baseline."""
    from oneflow_amd.solver import Solver
    c2 = args.workload == "c2"
    cfg = c2_config() if c2 else c1_config()
    if args.numbering is None:
        args.numbering = "multicolor"
    if args.face_order is None:
        args.face_order = "spatial" if c2 else "cell"
    t0 = time.perf_counter()
    grid = c2_grid(numbering=args.numbering, face_order=args.face_order) if c2 else \
        c1_grid(args.numbering, args.face_order)
    log(f"grid: cells={grid.n_cells} faces={grid.n_faces} bfaces={grid.n_bfaces} ({time.perf_counter() - t0:.1f} s)")
    s = Solver(grid, cfg, device=0, storage_order=args.storage)
    r64 = residual_bench(s, grid, args)
    s.close()
    log(f"fp64: {r64['ms_per_step']:.4f} ms per evaluation (host), {r64['roofline']['avg_us']:.1f} us (events), "
        f"pseudo-step {r64['pseudo_step']['ms']:.4f} ms")
    r32 = None
    if c2:
        s = Solver(grid, cfg.replace(flux_fp32=1), device=0, storage_order=args.storage)
        r32 = residual_bench(s, grid, args, flux_bytes=4)
        s.close()
        log(f"fp32 flux: {r32['ms_per_step']:.4f} ms per evaluation")
    # The CPU leg, the only part that runs the oracle: on a bounded sample of
    # the workload it is timed (the baseline) and checks the GPU's residuals
    # on the same state (oracle_error).  --no-cpu leaves both out.
    err, cpu = None, None
    if not args.no_cpu:
        sample = c2_grid(n_span=2) if c2 else grid
        err = oracle_error(sample, cfg, fp32=c2)
        log(f"oracle error: {err}")
        n, dt = cpu_baseline(sample, cfg, args.cpu_budget)
        what = (f"a {sample.n_cells}-cell 2-layer span slice of the same wing grid" if c2
                else f"the same {grid.n_cells}-cell grid")
        cpu = {"value": sample.n_cells * n / dt, "unit": "cell-residual evals/s", "cores": 1, "kind": "port",
               "sample": f"{n} NS+SA residual evaluations of {what}, CPU oracle, 1 thread of "
                         f"{os.cpu_count()} host cores, {dt:.1f} s"}
    workload = ("swept wing NACA0012, sweep 30 deg, taper 0.5, AR 4, 25 prism layers + tets, Roe + SA + Venkat, "
                "FP64 residual" if c2 else "NACA0012 hybrid quad/tri grid, Roe + SA + Venkat, FP64 residual")
    out = {
        "metric": "cell-residual evals/s + % HBM peak; time-to-1e-8 residual, learned vs plain",
        "value": r64["value"],
        "unit": "cell-residual evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "numbering": args.numbering, "face_order": args.face_order,
                   "storage_order": "spatial" if args.storage == 1 else "grid",
                   "cells": grid.n_cells, "faces": grid.n_faces, "nTCell": grid.n_tcell,
                   "mach": cfg.mach_ref, "aoa_deg": cfg.aoa_degree, "reynolds": cfg.reynolds},
        "roofline": r64["roofline"],
        "pseudo_step": r64["pseudo_step"],
        "oracle_error": err,
        "kernels": r64["kernels"],
        "dominant_kernel": r64["dominant_kernel"],
        "copy_kernel_gbs": r64["copy_kernel_gbs"],
        "launch": r64["launch"],
        "cpu_baseline": cpu,
    }
    if r32 is not None:
        out["fp32_flux"] = {"value": r32["value"], "ms_per_step": r32["ms_per_step"], "roofline": r32["roofline"],
                            "pseudo_step": r32["pseudo_step"], "kernels": r32["kernels"]}
    print(json.dumps(out))


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def step_bytes(grid):
    """Algorithmic bytes of one NS + SA residual evaluation: 552 per cell
    (ghosts included) + 64 per face, FP64 values and int32 ids, every array
    counted once (state 48, centre and volume 32, wall distance 8, gradients
    of 5 primitives, T and nu_tilde written and read 336, limiter written and
    read 80, residual 48; per face normal, area and centre 56 and the two
    cell ids 8)."""
    return 552.0 * grid.n_tcell + 64.0 * grid.n_faces


def residual_bench(s, grid, args, flux_bytes=8):
    """Time `args.steps` NS + SA residual evaluations of one solver (after
    `args.pre_iters` pseudo-steps and `args.warmup` untimed evaluations), the
    evaluation and each residual kernel with HIP events over back-to-back
    launches, and the whole pseudo-step."""
    s.init_freestream()
    s.iterate(1, args.pre_iters, with_norms=False)
    s.timestep(args.pre_iters + 1)
    # launch configuration: the two forms of the gather and of both gradients
    # timed here for the record, then the autotuner picks the variants and
    # face tiles on this state
    variant_us = {}
    for name, kernel, part in (("gather", 2, "gather"), ("grad", 0, "grad"), ("sa_grad", 5, "sa_grad"),
                               ("timestep", 6, "timestep"), ("lusgs_ns", 6, "lusgs_ns")):
        t = {}
        for v in (0, 1):
            s.set_kernel_variant(kernel, v)
            t[v] = 1e3 * s.time_kernel(part, iters=args.kernel_iters)
        variant_us[name] = t
    if args.launch:
        # a fixed launch configuration instead of the autotuner's choice (a
        # profiled run then times what an unprofiled run chose: under a
        # kernel-trace profiler every launch of a few microseconds takes about
        # the same time, and the autotuner's choice among them is arbitrary)
        v = [int(x) for x in args.launch.split(",")]
        if len(v) != 8:
            raise SystemExit("--launch takes ns_tile,sa_tile,gather,face_split,sa_geometry,grad,sa_grad,normals")
        tiles = (v[0], v[1])
        s.set_face_tile(*tiles)
        for kernel, var in zip((2, 3, 4, 0, 5, 6), v[2:]):
            s.set_kernel_variant(kernel, var)
    else:
        tiles = s.autotune(iters=20)
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
    sec = time.perf_counter() - t0
    ms_per_step = 1e3 * sec / args.steps
    # the evaluation with HIP events on the solver's stream, back to back
    eval_ms = s.time_kernel("residual", iters=args.steps)
    copy_ms = s.copy_bench(args.copy_bytes, iters=20)
    copy_gbs = 2.0 * args.copy_bytes / (copy_ms * 1e-3) / 1e9
    kb = kernel_bytes(grid.n_cells, grid.n_bfaces, grid.n_faces, grid.dim, flux_bytes, s.cfg.iturb_visflux,
                      s.cfg.ivischeme, s.kernel_variant(4))
    names = {"grad": "k_ns_grad",
             "limiter": "k_ns_qext",
             "face": "k_ns_face",
             "gather": ("k_ns_gather", "k_ns_gather_c")[s.kernel_variant(2)],
             "sa_grad": "k_sa_grad", "sa_face": "k_sa_face", "sa_cell": "k_sa_cell"}
    # keys of the counter summary: the instantiation this solver runs
    fp32 = "_fp32" if flux_bytes == 4 else ""
    geo = "_geo" if s.kernel_variant(4) == 1 and s.cfg.iturb_visflux == 2 else ""
    tkey = {"face": "k_ns_face" + fp32, "gather": names["gather"] + fp32, "sa_face": "k_sa_face" + geo,
            "grad": "k_ns_grad" + ("_eg" if s.kernel_variant(0) == 1 else ""),
            "sa_grad": "k_sa_grad" + ("_eg" if s.kernel_variant(5) == 1 else "")}
    prof = {key: s.time_kernel(key, iters=args.kernel_iters) for key in names}
    s.residual_ns()  # leave a consistent residual behind
    s.residual_sa()
    s.sync()
    kernels = {}
    for key, kname in names.items():
        ms = prof[key]
        kernels[kname] = {"avg_us": 1e3 * ms, "bytes_per_launch": kb[key],
                          "achieved_gbs": kb[key] / (ms * 1e-3) / 1e9 if ms > 0 else None,
                          "traffic_bytes": read_traffic(tkey.get(key, kname), grid.dim)}
    dom = max(names, key=lambda k: prof[k])
    ach = kb[dom] / (prof[dom] * 1e-3) / 1e9
    dominant = {"kernel": names[dom], "avg_us": 1e3 * prof[dom], "bytes_per_launch": kb[dom], "achieved": ach,
                "frac": ach / HBM_PEAK_GBS, "launches_timed": args.kernel_iters}
    sb = step_bytes(grid)
    sgbs = sb / (eval_ms * 1e-3) / 1e9
    traffic = [kernels[k]["traffic_bytes"] for k in kernels]
    roofline = {"bound": "hbm", "achieved": sgbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": sgbs / HBM_PEAK_GBS,
                "traffic": float(sum(traffic)) if all(t is not None for t in traffic) else None,
                "traffic_kind": "bytes leaving L2 per evaluation (FETCH_SIZE + WRITE_SIZE, copy-kernel scaled), "
                                "committed counter summary of this workload",
                "kernel": "NS + SA residual evaluation (7 kernels)", "avg_us": 1e3 * eval_ms,
                "bytes_per_launch": sb, "launches_timed": args.steps}
    # the whole pseudo-time step (residual, LU-SGS, update, boundary of both
    # solvers), HIP events; this advances the state, so it runs last
    ps_ms = s.time_kernel("pseudo_step", iters=args.steps)
    pseudo = {"ms": ps_ms, "residual_ms": eval_ms, "implicit_ms": ps_ms - eval_ms,
              "cell_steps_per_s": grid.n_cells / (ps_ms * 1e-3)}
    return {"value": grid.n_cells * args.steps / sec, "ms_per_step": ms_per_step, "roofline": roofline,
            "kernels": kernels, "dominant_kernel": dominant, "pseudo_step": pseudo, "copy_kernel_gbs": copy_gbs,
            "launch": {"variant_us": variant_us, "face_tiles": list(tiles), "gather": s.kernel_variant(2),
                       "face_split": s.kernel_variant(3), "sa_geometry": s.kernel_variant(4),
                       "grad_entry_terms": s.kernel_variant(0), "sa_grad_entry_terms": s.kernel_variant(5),
                       "step_entry_normals": s.kernel_variant(6), "autotuned": not args.launch}}


def oracle_error(grid, cfg, fp32=True):
    """FP64 and FP32-flux NS + SA residuals against the CPU oracle on the same
    state (the oracle's after 10 pseudo-steps from freestream): normwise
    ||res_gpu - res_oracle|| / ||res_oracle|| and the largest per-cell error
    relative to the cell's flux scale sum_f |F_f| A_f."""
    from oneflow_amd.solver import Solver
    from oracle.pyoracle import OracleCase
    ora = OracleCase(grid, cfg)
    ora.init_freestream()
    ora.iterate(1, 10)
    ora.timestep(11)
    ora.residual_ns()
    ora.residual_sa()
    C = grid.n_cells
    ref = np.concatenate([ora.get("RES")[:, :C], ora.get("TURBRES")[None, :C]])
    scale = _flux_scale(ora, grid)
    out = {"sample": f"{C} cells, state after 10 oracle pseudo-steps"}
    for name, c in (("fp64", cfg), ("fp32_flux", cfg.replace(flux_fp32=1))):
        if name == "fp32_flux" and not fp32:
            continue
        s = Solver(grid, c)
        s.init_freestream()
        for n in ("Q", "TURBQ", "TEMPR", "VISL", "VIST", "TIMESTEP"):
            s.set(n, ora.get(n))
        s.residual_ns()
        s.residual_sa()
        got = np.concatenate([s.get("RES")[:, :C], s.get("TURBRES")[None, :C]])
        s.close()
        d = np.abs(got - ref)
        out[name] = {"norm_rel": float(np.linalg.norm(d) / np.linalg.norm(ref)),
                     "max_cell_rel_ns": float(np.max(d[:5] / np.maximum(scale, 1e-300))),
                     "bitwise_frac": float(np.mean(got == ref))}
    return out


def _flux_scale(ora, g):
    """Per cell and NS equation: sum over the cell's faces of the area times
    the magnitude of the physical flux terms of the cell's own state."""
    q = ora.get("Q")
    area = ora.get("FACE_NORMAL")[3]
    C = g.n_cells
    lc, rc = np.asarray(g.lc), np.asarray(g.rc)
    rho, u, v, w, p = q[:, :C]
    h = 1.4 / 0.4 * p / rho + 0.5 * (u * u + v * v + w * w)
    vm = np.sqrt(u * u + v * v + w * w)
    mag = np.stack([rho * vm, rho * vm * vm + p, rho * vm * vm + p, rho * vm * vm + p, rho * vm * h])
    scale = np.zeros((5, C))
    for side in (lc, rc):
        ok = (side >= 0) & (side < C)
        np.add.at(scale, (slice(None), side[ok]), mag[:, side[ok]] * area[ok])
    return scale


def c2_config():
    from oneflow_amd.config import CaseConfig
    return CaseConfig(dim=3, ischeme=1, ilim=2, ivischeme=0, vismodel=3, mach_ref=0.84, aoa_degree=3.0,
                      reynolds=1.2e7, centropy1=0.001, centropy2=0.001, vencat_coef=0.001)


def c2_grid(n_span=34, numbering="multicolor", face_order="spatial"):
    from oneflow_amd import meshgen
    g = meshgen.swept_wing(n_surf=256, n_quad=25, ratio=1.2, reynolds=1.2e7, n_outer=40, radius=20.0,
                           n_span=n_span, sweep_deg=30.0, taper=0.5, aspect=4.0, t=0.12)
    return _numbered(g, numbering, face_order)


if __name__ == "__main__":
    main()
