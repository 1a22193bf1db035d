"""Batched cases (ofx_batch_*): graph replays against eager stepping.

Three airfoil cases of different size, shape, Mach and AoA run twice: each
on its own context with ofx_iterate (eager), and together through one batch
whose post-ramp steps are recorded into one HIP graph and replayed.  After
every chunk, with eager host work in between (norms, downloads, single tasks
that move the reduction parities, an autotune that changes the face tiles),
every case's state must equal the eager run's bit for bit.
"""
import numpy as np
import pytest

from oneflow_amd import meshgen
from oneflow_amd.config import CaseConfig

pytestmark = pytest.mark.gpu

SPECS = [  # (n_surf, n_quad, n_outer, m, t, mach, aoa)
    (64, 8, 10, 0.00, 0.12, 0.50, 2.0),
    (96, 10, 12, 0.02, 0.10, 0.70, 1.0),
    (80, 6, 14, 0.04, 0.15, 0.30, 4.0),
]


def cases():
    out = []
    for ns, nq, no, m, t, mach, aoa in SPECS:
        g = meshgen.naca_hybrid(n_surf=ns, n_quad=nq, ratio=1.3, reynolds=5.0e6, n_outer=no, radius=20.0, m=m, p=0.4, t=t)
        g = meshgen.renumber(g, meshgen.multicolor_order(g))
        cfg = CaseConfig(ischeme=1, ilim=2, vismodel=3, mach_ref=mach, aoa_degree=aoa, reynolds=5.0e6,
                         vencat_coef=0.001, centropy1=0.001, centropy2=0.001, ncfl=6, cflst=1.0, cfled=5.0)
        out.append((g, cfg))
    return out


def fresh(cs):
    from oneflow_amd.solver import Solver
    sol = [Solver(g, c) for g, c in cs]
    for s in sol:
        s.init_freestream()
    return sol


def same(a, b, what):
    for i, (x, y) in enumerate(zip(a, b)):
        for f in ("Q", "TURBQ"):
            u, v = x.get(f), y.get(f)
            assert np.array_equal(u, v), (what, i, f, float(np.max(np.abs(u - v))))


def test_batch_graph_replays_match_eager():
    from oneflow_amd.solver import Batch
    cs = cases()
    print(f"[batch] cells {[g.n_cells for g, _ in cs]}", flush=True)
    eager, bat = fresh(cs), fresh(cs)
    b = Batch(bat, steps_per_graph=4)

    # steps 1..14: 5 ramp steps direct, 2 replays of 4 steps, 1 leftover step direct
    k = b.iterate(1, 14)
    for s in eager:
        s.iterate(1, 14, with_norms=False)
    assert k == 2
    same(eager, bat, "chunk 1")
    ne = np.array([s.norms() for s in eager])
    assert np.array_equal(ne, b.norms())

    # the odd leftover moved the parities: the graph is recorded again
    k = b.iterate(15, 9)
    for s in eager:
        s.iterate(15, 9, with_norms=False)
    assert k == 2
    same(eager, bat, "chunk 2")

    # single tasks between chunks (flip one reduction parity only)
    for s in (eager[0], bat[0]):
        s.residual_ns()
    k = b.iterate(24, 8)
    for s in eager:
        s.iterate(24, 8, with_norms=False)
    assert k == 2
    same(eager, bat, "chunk 3")

    # autotuned face tiles change launches, not results
    tiles = b.autotune(iters=5)
    print(f"[batch] tiles {tiles.tolist()}", flush=True)
    k = b.iterate(32, 12)
    for s in eager:
        s.iterate(32, 12, with_norms=False)
    assert k == 3
    same(eager, bat, "chunk 4")
    b.close()


def test_zones_in_one_graph_match_eager_zones():
    """All cases as zones of one context (every kernel spans all cases), their
    post-ramp steps recorded into one HIP graph: one graph launch advances
    every case steps_per_graph steps, bit for bit as eager zone stepping,
    which equals each case alone (tests/test_gpu_zones.py)."""
    from oneflow_amd.solver import Batch, ZoneSolver
    cs = cases()
    grids, cfgs = [g for g, _ in cs], [c for _, c in cs]
    eager, graphed = ZoneSolver(grids, cfgs), ZoneSolver(grids, cfgs)
    for z in (eager, graphed):
        z.init_freestream()
    b = Batch([graphed], steps_per_graph=6)
    k = b.iterate(1, 23)           # 5 ramp steps direct, 3 replays of 6, no leftover
    eager.iterate(1, 23, with_norms=False)
    assert k == 3
    same([eager], [graphed], "zones graph")
    assert np.array_equal(eager.zone_norms(), graphed.zone_norms())
    k = b.iterate(24, 12)
    eager.iterate(24, 12, with_norms=False)
    assert k == 2
    same([eager], [graphed], "zones graph, second chunk")
    assert np.array_equal(eager.zone_aero(), graphed.zone_aero())
    b.close()
