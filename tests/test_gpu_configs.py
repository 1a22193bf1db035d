"""Each benchmark configuration's code path on the GPU against the oracle.

* the laminar flat plate of the CPU-runnable configuration (161 x 97 nodes on
  [-0.5, 1] x [0, 0.5], wall stretching 1.1, Mach 0.2, Re 5e6, SLAU2, second
  order without limiter): GPU = oracle over 20 pseudo-steps;
* the 3-D swept wing (a small instance of the benchmark generator: prism wall
  layers + tets, Roe + SA + Venkat) in FP64: GPU = oracle; and its FP32-flux
  variant against the FP64 oracle: per cell |dres| <= 64 u sum_f |F_f|
  (the bound of tests/test_gpu_fp32.py); after 100 transient steps state,
  CL and CD within 1e-3 relative of the FP64 run;
* the batched airfoil sweep: the three smallest of the 64 seeded cases
  (about 28-30k cells each), solved as zones of one context, each zone
  against the oracle on that case alone.

Tolerances: states, residuals and forces equal the oracle's bit for bit (the
one transcendental function, pow, is the routine both sides share); the
residual norms, a reduction in another order on the GPU, agree to 1e-12.
"""
import numpy as np
import pytest

from oneflow_amd import meshgen
from oneflow_amd.config import CaseConfig
from oracle.pyoracle import OracleCase
from tests.parity_util import report
from tests.test_gpu_fp32 import U32, flux_scale

pytestmark = pytest.mark.gpu


def plate_c0():
    g = meshgen.flat_plate(161, 97, -0.5, 1.0, 0.5, 1.1)
    cfg = CaseConfig(ischeme=12, ilim=0, ivischeme=0, vismodel=1, mach_ref=0.2, reynolds=5.0e6,
                     centropy1=0.0001, centropy2=0.0001)
    return g, cfg


def wing_small(fp32=0):
    g = meshgen.swept_wing(n_surf=32, n_quad=4, ratio=1.5, reynolds=1.2e7, n_outer=4, radius=10.0, n_span=3)
    g = meshgen.renumber(g, meshgen.multicolor_order(g))
    cfg = CaseConfig(dim=3, ischeme=1, ilim=2, ivischeme=0, vismodel=3, mach_ref=0.84, aoa_degree=3.0,
                     reynolds=1.2e7, centropy1=0.001, centropy2=0.001, vencat_coef=0.001, flux_fp32=fp32)
    return g, cfg


def run_pair(g, cfg, steps):
    from oneflow_amd.solver import Solver
    gpu, ora = Solver(g, cfg), OracleCase(g, cfg)
    for s in (gpu, ora):
        s.init_freestream()
    ng = gpu.iterate(1, steps)
    no = ora.iterate(1, steps)
    return gpu, ora, ng, no


def test_flat_plate_c0():
    g, cfg = plate_c0()
    print(f"[configs] flat plate: cells={g.n_cells} faces={g.n_faces}", flush=True)
    assert g.n_cells == 160 * 96
    gpu, ora, ng, no = run_pair(g, cfg, 20)
    assert report("norms", ng, no) <= 1e-12
    assert report("Q", gpu.get("Q"), ora.get("Q")) == 0.0
    assert report("aero", gpu.aero_row(), ora.aero_row()) == 0.0
    gpu.close()


def test_swept_wing_fp64():
    g, cfg = wing_small()
    print(f"[configs] swept wing: cells={g.n_cells} faces={g.n_faces}", flush=True)
    gpu, ora, ng, no = run_pair(g, cfg, 10)
    assert report("norms", ng, no) <= 1e-12
    C = g.n_cells
    for n in ("Q", "TURBQ"):
        assert report(n, gpu.get(n)[..., :C], ora.get(n)[..., :C]) == 0.0, n
    gpu.close()


def test_swept_wing_fp32_flux_residual_bound():
    g, cfg = wing_small()
    ora = OracleCase(g, cfg)
    ora.init_freestream()
    ora.iterate(1, 20)
    ora.timestep(21)
    from oneflow_amd.solver import Solver
    gpu = Solver(g, cfg.replace(flux_fp32=1))
    gpu.init_freestream()
    for n in ("Q", "TURBQ", "TEMPR", "VISL", "VIST", "TIMESTEP"):
        gpu.set(n, ora.get(n))
    ora.residual_ns()
    gpu.residual_ns()
    C = g.n_cells
    r64, r32 = ora.get("RES")[:, :C], gpu.get("RES")[:, :C]
    d = np.abs(r32 - r64)
    cell_rel = float(np.max(d / np.maximum(flux_scale(ora, g), 1e-300)))
    norm_rel = float(np.linalg.norm(d) / np.linalg.norm(r64))
    print(f"[configs] wing fp32 flux: max |dres|/sum|F| = {cell_rel / U32:.1f} u, "
          f"||dres||/||res|| = {norm_rel:.2e}", flush=True)
    assert cell_rel <= 64 * U32
    assert norm_rel < 1e-2
    gpu.close()


def test_swept_wing_fp32_flux_run():
    from oneflow_amd.solver import Solver
    g, cfg = wing_small()
    a, b = Solver(g, cfg), Solver(g, cfg.replace(flux_fp32=1))
    for s in (a, b):
        s.init_freestream()
    a.iterate(1, 100, with_norms=False)
    b.iterate(1, 100, with_norms=False)
    C = g.n_cells
    qa, qb = a.get("Q")[:, :C], b.get("Q")[:, :C]
    eq = float(np.linalg.norm(qb - qa) / np.linalg.norm(qa))
    fa, fb = a.aero(), b.aero()
    ecl, ecd = abs(fb[0] - fa[0]) / abs(fa[0]), abs(fb[1] - fa[1]) / abs(fa[1])
    print(f"[configs] wing fp32 flux 100 steps: state {eq:.2e}, CL {ecl:.2e}, CD {ecd:.2e}", flush=True)
    # 100 steps of the transonic transient through the Venkatakrishnan limiter
    # amplify the FP32 flux rounding (measured: state 6.7e-5, CL 1.2e-4, CD
    # 3.5e-5): the stated bound for this run is 1e-3
    assert eq <= 1e-3 and ecl <= 1e-3 and ecd <= 1e-3
    a.close()
    b.close()


def test_sweep_zones_against_oracle():
    from oneflow_amd.cases import airfoil_sweep
    from oneflow_amd.solver import ZoneSolver
    cases = airfoil_sweep(64, seed=2026)
    pick = sorted(range(len(cases)), key=lambda i: cases[i].target_cells)[:3]
    grids = [cases[i].grid() for i in pick]
    cfgs = [cases[i].config() for i in pick]
    print(f"[configs] sweep cases {pick}: cells {[g.n_cells for g in grids]}", flush=True)
    zs = ZoneSolver(grids, cfgs)
    zs.init_freestream()
    zs.iterate(1, 8, with_norms=False)
    zn, za = zs.zone_norms(), zs.zone_aero()
    for z, (g, cfg) in enumerate(zip(grids, cfgs)):
        ora = OracleCase(g, cfg)
        ora.init_freestream()
        no = ora.iterate(1, 8)[-1]
        C = g.n_cells
        assert report(f"norms z{z}", zn[z], no) <= 1e-12
        for n in ("Q", "TURBQ"):
            a = zs.get_zone(n, z)[..., :C]
            b = ora.get(n)[..., :C]
            assert report(f"{n} z{z}", a, b) == 0.0, (z, n)
        assert report(f"aero z{z}", za[z], ora.aero()) == 0.0
    zs.close()
