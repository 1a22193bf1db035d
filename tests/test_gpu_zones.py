"""Several cases in one context (ofx_zones_set) against each case alone.

Three airfoil cases of different size, shape, Mach, AoA and Reynolds number,
concatenated into one context, must give every zone exactly the bits of the
same case run on its own context: state, time step, limiter, residual norms
and forces, through the CFL ramp and beyond, on natural (one-workgroup
LU-SGS sweeps) and multicolour (per-level sweeps) numberings; a 3-D pair of
boxes checks the 3-D paths and the per-zone wall distance.
"""
import numpy as np
import pytest

from oneflow_amd import meshgen
from oneflow_amd.config import CaseConfig
from oneflow_amd.mesh import BC_FARFIELD, BC_SOLID_SURFACE, BC_SYMMETRY

pytestmark = pytest.mark.gpu

SPECS = [  # (n_surf, n_quad, n_outer, m, t, mach, aoa, re)
    (64, 8, 10, 0.00, 0.12, 0.50, 2.0, 5.0e6),
    (96, 10, 12, 0.02, 0.10, 0.70, 1.0, 8.0e6),
    (80, 6, 14, 0.04, 0.15, 0.30, 4.0, 3.0e6),
]


def airfoils(numbering):
    grids, cfgs = [], []
    for ns, nq, no, m, t, mach, aoa, re in SPECS:
        g = meshgen.naca_hybrid(n_surf=ns, n_quad=nq, ratio=1.3, reynolds=re, n_outer=no, radius=20.0, m=m, p=0.4, t=t)
        if numbering == "mc":
            g = meshgen.renumber(g, meshgen.multicolor_order(g))
        grids.append(g)
        cfgs.append(CaseConfig(ischeme=1, ilim=2, vismodel=3, mach_ref=mach, aoa_degree=aoa, reynolds=re,
                               vencat_coef=0.001, centropy1=0.001, centropy2=0.001, ncfl=10, cflst=1.0, cfled=5.0))
    return grids, cfgs


def boxes():
    def bc(xc, yc, zc, n):
        if abs(zc) < 1e-12:
            return BC_SOLID_SURFACE if xc > 0.25 else BC_SYMMETRY
        return BC_FARFIELD
    grids = [meshgen.box3d(4, 3, 3, "mixed", prism_layers=1, jitter=0.15, seed=21, bc=bc),
             meshgen.box3d(3, 3, 4, "hex", jitter=0.15, seed=22, bc=bc)]
    cfgs = [CaseConfig(dim=3, ischeme=1, ilim=2, vismodel=3, mach_ref=0.5, aoa_degree=2.0, reynolds=5e4,
                       vencat_coef=0.01, ncfl=10, cflst=1.0, cfled=5.0),
            CaseConfig(dim=3, ischeme=1, ilim=2, vismodel=3, mach_ref=0.3, aoa_degree=-1.0, reynolds=2e4,
                       vencat_coef=0.01, ncfl=10, cflst=1.0, cfled=5.0)]
    return grids, cfgs


@pytest.mark.parametrize("which,variant", [("natural", 0), ("mc", 0), ("boxes3d", 0), ("mc", 1), ("boxes3d", 1)])
def test_zones_equal_single_cases(which, variant):
    # variant: the residual gather's form (kernel 2) and the per-(cell, face)
    # face terms of the gradients, time step and sweeps (kernels 0, 5, 6);
    # the single cases keep the defaults
    from oneflow_amd.solver import Solver, ZoneSolver
    grids, cfgs = boxes() if which == "boxes3d" else airfoils(which)
    print(f"\n[zones] {which} gather variant {variant}: cells {[g.n_cells for g in grids]}", flush=True)
    single = [Solver(g, c) for g, c in zip(grids, cfgs)]
    multi = ZoneSolver(grids, cfgs)
    for kernel in (2, 0, 5, 6):
        multi.set_kernel_variant(kernel, variant)
    for z, s in enumerate(single):
        assert np.array_equal(multi.get("DIST")[multi.zone_slice(z)[0]], s.get("DIST")), ("DIST", z)
    for s in single + [multi]:
        s.init_freestream()
    for chunk in (6, 9):  # across the CFL ramp (ncfl = 10) and past it
        it0 = 1 if chunk == 6 else 7
        for s in single:
            s.iterate(it0, chunk, with_norms=False)
        multi.iterate(it0, chunk, with_norms=False)
        zn, za = multi.zone_norms(), multi.zone_aero()
        for z, s in enumerate(single):
            for f in ("Q", "TURBQ", "TIMESTEP", "LIMITER"):
                a, b = multi.get_zone(f, z), s.get(f)
                if f in ("TIMESTEP", "LIMITER"):
                    nc = s.n_cells
                    a, b = a[..., :nc], b[..., :nc]
                assert np.array_equal(a, b), (which, chunk, z, f, float(np.max(np.abs(a - b))))
            assert np.array_equal(zn[z], s.norms()), (which, z, zn[z], s.norms())
            assert np.array_equal(za[z], s.aero()), (which, z, za[z], s.aero())
        # wall rows: the zones' walls in face order, zone after zone
        walls = np.concatenate([s.wall_data() for s in single])
        assert np.array_equal(multi.wall_data(), walls), which
