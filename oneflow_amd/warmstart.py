"""Learned warm start for the airfoil sweep.

Pipeline (all compute on the GPU through libofx.so):
  1. solve a training set of airfoil cases with the plain solver (batched);
  2. sample their converged cell fields into (features, targets);
  3. train the BF16 MLP (ofx_mlp_*) on them with the library's own kernels;
  4. for unseen cases, predict the field, write it as the initial state, and
     let the plain solver iterate from there to the same residual tolerance as
     from freestream -- the solver is both the fallback and the test of convergence.

Inputs per cell (9): the centroid folded into the unit disk around mid-chord,
x' = (x - 0.5) / (1 + r), y' = y / (1 + r), r = |(x - 0.5, y)|; the wall
distance as log10(d + 1e-6) / 3 + 1; the case: 25 m, p, 10 t,
2 (log10 Re - 6.5), 4 (Mach - 0.55), 10 AoA[rad].  The four airfoil
parameters are (m, p, t, Re).
Targets per cell (5): rho - 1, u - cos(AoA), v - sin(AoA), p / p_inf - 1 and
log10(nu_tilde / nu_tilde_inf) / 2 (nu_tilde floored at 1e-3 nu_tilde_inf).
"""
from __future__ import annotations

import math
import time

import numpy as np

from .cases import AirfoilCase
from .solver import Batch, Solver


def cell_features(s: Solver, case: AirfoilCase) -> np.ndarray:
    C = s.n_cells
    xc = s.get("XCC")[:, :C]
    d = s.get("DIST")[:C]
    x = xc[0] - 0.5
    y = xc[1]
    r = np.sqrt(x * x + y * y)
    f = np.empty((C, 9), np.float32)
    f[:, 0] = x / (1.0 + r)
    f[:, 1] = y / (1.0 + r)
    f[:, 2] = np.log10(d + 1e-6) / 3.0 + 1.0
    f[:, 3] = 25.0 * case.m
    f[:, 4] = case.p
    f[:, 5] = 10.0 * case.t
    f[:, 6] = 2.0 * (math.log10(case.reynolds) - 6.5)
    f[:, 7] = 4.0 * (case.mach - 0.55)
    f[:, 8] = 10.0 * math.radians(case.aoa)
    return f


def _free(case: AirfoilCase, cfg):
    a = math.radians(case.aoa)
    p_inf = 1.0 / (cfg.gama_ref * case.mach * case.mach)
    return math.cos(a), math.sin(a), p_inf, cfg.ref_sa


def cell_targets(s: Solver, case: AirfoilCase, cfg) -> np.ndarray:
    C = s.n_cells
    q = s.get("Q")[:, :C]
    nu = s.get("TURBQ")[:C]
    ca, sa, p_inf, nu_inf = _free(case, cfg)
    t = np.empty((C, 5), np.float32)
    t[:, 0] = q[0] - 1.0
    t[:, 1] = q[1] - ca
    t[:, 2] = q[2] - sa
    t[:, 3] = q[4] / p_inf - 1.0
    t[:, 4] = np.log10(np.maximum(nu, 1e-3 * nu_inf) / nu_inf) / 2.0
    return t


def _write_prediction(q, nu, case: AirfoilCase, cfg, y: np.ndarray) -> None:
    """Predicted primitive state and nu_tilde of a case's cells into q[5][k], nu[k] (views)."""
    ca, sa, p_inf, nu_inf = _free(case, cfg)
    q[0] = np.maximum(1.0 + y[:, 0], 0.05)
    q[1] = ca + y[:, 1]
    q[2] = sa + y[:, 2]
    q[3] = 0.0
    q[4] = p_inf * np.maximum(1.0 + y[:, 3], 0.05)
    nu[:] = nu_inf * np.power(10.0, 2.0 * np.clip(y[:, 4], -1.5, 2.5))


def apply_prediction(s: Solver, case: AirfoilCase, cfg, y: np.ndarray) -> None:
    """Write a predicted field as the initial state (positivity-floored), then
    the boundary tasks fill ghosts, temperature and viscosities."""
    C = s.n_cells
    q = s.get("Q")
    nu = s.get("TURBQ")
    _write_prediction(q[:, :C], nu[:C], case, cfg, y)
    s.set("Q", q)
    s.set("TURBQ", nu)
    s.ns_boundary()
    s.sa_boundary()


def apply_predictions_zones(zs, cases, net, mach_gate=None) -> None:
    """The learned initial state of every zone of a ZoneSolver: the fields are
    read once, one network forward covers all zones' cells, the state is
    written once and the boundary tasks fill the ghosts of all zones.  Cases
    above `mach_gate` (if given) keep the freestream start the plain solver
    uses (transonic cases can reach a different steady state from another
    start)."""
    xcc, dist = zs.get("XCC"), zs.get("DIST")
    q, nu = zs.get("Q"), zs.get("TURBQ")
    feats, spans = [], []
    for z, case in enumerate(cases):
        if mach_gate is not None and case.mach > mach_gate:
            continue
        cs, _ = zs.zone_slice(z)
        feats.append(cell_features(_Fields(xcc[:, cs], dist[cs]), case))
        spans.append((z, cs))
    if not feats:
        return
    y = net.forward(np.concatenate(feats))
    o = 0
    for z, cs in spans:
        n = cs.stop - cs.start
        _write_prediction(q[:, cs], nu[cs], cases[z], zs.zone_cfgs[z], y[o:o + n])
        o += n
    zs.set("Q", q)
    zs.set("TURBQ", nu)
    zs.ns_boundary()
    zs.sa_boundary()


class _Fields:
    def __init__(self, xcc, dist):
        self.n_cells = xcc.shape[1]
        self._f = {"XCC": xcc, "DIST": dist}

    def get(self, name):
        return self._f[name]


class ZoneView:
    """Zone z of a ZoneSolver seen as a one-case solver by the field readers
    above (cell fields in the zone's own layout: cells, then ghosts)."""

    def __init__(self, zs, z: int):
        self.zs, self.z = zs, z
        self.n_cells = int(zs.cell_start[z + 1] - zs.cell_start[z])
        self.cfg = zs.zone_cfgs[z]

    def get(self, name):
        return self.zs.get_zone(name, self.z)


class FieldCache:
    """A ZoneSolver whose fields are downloaded once each, for reading many
    zones of one state (ZoneView over it)."""

    def __init__(self, zs):
        self.zs = zs
        self.cell_start, self.zone_cfgs = zs.cell_start, zs.zone_cfgs
        self._f = {}

    def get_zone(self, name, z):
        if name not in self._f:
            self._f[name] = self.zs.get(name)
        a = self._f[name]
        cs, gs = self.zs.zone_slice(z)
        return np.concatenate([a[..., cs], a[..., gs]], axis=-1)


CONVERGED, MAX_STEPS, DIVERGED = "converged", "max_steps", "diverged"


def _iterate_nan_tolerant(solver, it, n):
    """solver.iterate(it, n) without norms; a NaN residual (OFX_ENAN, checked
    after the last step) is left to the caller's per-case norms."""
    from .solver import OFX_ENAN, OfxError
    try:
        solver.iterate(it, n, with_norms=False)
    except OfxError as e:
        if e.code != OFX_ENAN:
            raise


def _case_status(norm, limit, it, max_steps):
    """Status of a case after step it-1, or None while it keeps running."""
    if not np.isfinite(norm):
        return DIVERGED
    if norm <= limit:
        return CONVERGED
    if it > max_steps:
        return MAX_STEPS
    return None


def solve_zones_to_tolerance(zs, refs, tol=1e-8, chunk=500, max_steps=150_000, progress=None, rebuild=None,
                             compact=0.5, first_iter=1, on_chunk=None, on_leave=None):
    """All cases as zones of one context advance together until every
    zone's density residual norm is below tol * refs[case], its norm is not
    finite (diverged: the reference stops on NaN,
    codes/residual/src/ResidualTask.cpp:95-100) or max_steps is passed.
    With `rebuild` (case ids -> a new ZoneSolver for those cases), finished
    cases leave: once the running cases are at most `compact` of the
    context's zones, or some case diverged, their state moves into a new
    context holding only them (ghosts and viscosities recomputed by the
    boundary tasks), so finished cases stop costing work.  Returns (wall
    seconds including any rebuilds, step at which each case finished, forces
    [CL, CD, CDp, CDf] of each case at that step (None if diverged), status
    of each case: CONVERGED, MAX_STEPS or DIVERGED).  first_iter: the outer
    step the context is at (steps already taken + 1).  on_chunk(step, n_running)
    is called after every chunk of steps.  on_leave(context, [(zone, case)])
    is called once for every case, with the context that holds its last state,
    just before that context is left (at a rebuild, for the cases that
    finished in it; at the end, for all cases of the last context): the
    finished cases' states are read there without a download per case."""
    n = zs.n_zones
    steps = [0] * n
    aero = [None] * n
    status = [None] * n
    zs.sync()
    t0 = time.perf_counter()
    ramp = zs.cfg.ncfl
    it = first_iter
    cur, ids = zs, list(range(n))
    while True:
        nstep = chunk + (ramp - it if it < ramp else 0)
        _iterate_nan_tolerant(cur, it, nstep)
        it += nstep
        nr = cur.zone_norms()[:, 0]
        za = None
        diverged = False
        for z, cid in enumerate(ids):
            if status[cid] is None:
                st = _case_status(nr[z], tol * refs[cid], it, max_steps)
                if st is None:
                    continue
                status[cid] = st
                steps[cid] = it - 1
                if st == DIVERGED:
                    diverged = True
                else:
                    za = cur.zone_aero() if za is None else za
                    aero[cid] = za[z].tolist()
                if progress:
                    progress(cid, steps[cid])
        live = [cid for cid in ids if status[cid] is None]
        if on_chunk:
            on_chunk(it - 1, len(live))
        if not live:
            break
        if rebuild is not None and len(live) < len(ids) and (diverged or len(live) <= compact * len(ids)):
            new = rebuild(live)
            q, nu = cur.get("Q"), cur.get("TURBQ")
            qn, nun = new.get("Q"), new.get("TURBQ")
            for j, cid in enumerate(live):
                z = ids.index(cid)
                (cs, gs), (cn, gn) = cur.zone_slice(z), new.zone_slice(j)
                qn[:, cn], qn[:, gn] = q[:, cs], q[:, gs]
                nun[cn], nun[gn] = nu[cs], nu[gs]
            new.set("Q", qn)
            new.set("TURBQ", nun)
            new.ns_boundary()
            new.sa_boundary()
            if on_leave:
                on_leave(cur, [(z, cid) for z, cid in enumerate(ids) if status[cid] is not None])
            if cur is not zs:
                cur.close()
            cur, ids = new, live
    if on_leave:
        on_leave(cur, list(enumerate(ids)))
    if cur is not zs:
        cur.close()
    return time.perf_counter() - t0, steps, aero, status


def solve_to_tolerance(solvers, refs, tol=1e-8, chunk=500, max_steps=150_000, batched=True, steps_per_graph=10,
                       progress=None):
    """Iterate every case until its density residual norm falls below
    tol * refs[i] (an absolute level per case), is not finite (diverged) or
    max_steps is passed.  Batched: all running cases advance together in one
    recorded graph; finished cases leave the batch.  Sequential: one case
    after another.  `progress(case, steps)` is called as each case finishes.
    Returns (wall seconds, steps per case, status per case)."""
    from .solver import OFX_ENAN, OfxError
    n = len(solvers)
    steps = [0] * n
    status = [None] * n
    for s in solvers:
        s.sync()
    t0 = time.perf_counter()
    ramp = max(s.cfg.ncfl for s in solvers)
    chunk = max(steps_per_graph, chunk // steps_per_graph * steps_per_graph)
    if batched:
        live = list(range(n))
        while live:
            b = Batch([solvers[i] for i in live], steps_per_graph=steps_per_graph)
            it = steps[live[0]] + 1
            while True:
                # after the ramp, whole graphs only: no leftover step moves the
                # reduction parities, so the recorded graph stays valid
                nstep = chunk + (ramp - it if it < ramp else 0)
                try:
                    b.iterate(it, nstep)
                except OfxError as e:
                    if e.code != OFX_ENAN:
                        raise
                it += nstep
                nr = [solvers[i].norms(check=False)[0] for i in live]
                hit = []
                for i, r in zip(live, nr):
                    steps[i] = it - 1
                    st = _case_status(r, tol * refs[i], it, max_steps)
                    if st is not None:
                        status[i] = st
                        hit.append(i)
                if hit:
                    if progress:
                        for i in hit:
                            progress(i, steps[i])
                    live = [i for i in live if status[i] is None]
                    break
            b.close()
    else:
        for i, s in enumerate(solvers):
            it = 1
            while True:
                _iterate_nan_tolerant(s, it, chunk)
                it += chunk
                st = _case_status(s.norms(check=False)[0], tol * refs[i], it, max_steps)
                if st is not None:
                    status[i] = st
                    break
            steps[i] = it - 1
            if progress:
                progress(i, steps[i])
    return time.perf_counter() - t0, steps, status
