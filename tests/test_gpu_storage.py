"""The device storage order (ofx_set_storage_order) changes where cells and
faces live in HBM, never a result: the spatial order (cells of each connected
part in Morton order, interior faces by their cells) against the grid's own
order, after whole pseudo-steps, on a multicolour-numbered 2-D airfoil with SA,
the small 3-D swept wing, natural numbering (one-workgroup sweeps) and zones.
States, eddy viscosity, limiter, time step, forces and wall rows must agree
bit for bit; norms (sums over cells in storage order) to 1e-13.  The spatial
runs also read the face terms of the gradients, time step and sweeps from
the per-(cell, face) copies (kernel variants 0, 5 and 6), which are built in
storage order."""
import numpy as np
import pytest

from oneflow_amd import meshgen
from oneflow_amd.config import CaseConfig

pytestmark = pytest.mark.gpu


def airfoil(numbering):
    g = meshgen.naca_hybrid(n_surf=96, n_quad=8, ratio=1.3, reynolds=6.5e6, n_outer=12, radius=20.0, m=0.02, p=0.4,
                            t=0.12)
    if numbering == "mc":
        g = meshgen.renumber(g, meshgen.multicolor_order(g))
    cfg = CaseConfig(ischeme=1, ilim=2, vismodel=3, mach_ref=0.8, aoa_degree=1.25, reynolds=6.5e6,
                     centropy1=0.001, centropy2=0.001, vencat_coef=0.001, ncfl=10, cflst=1.0, cfled=5.0)
    return g, cfg


def wing():
    g = meshgen.swept_wing(n_surf=32, n_quad=4, ratio=1.5, reynolds=1.2e7, n_outer=4, radius=10.0, n_span=3)
    g = meshgen.renumber(g, meshgen.multicolor_order(g))
    cfg = CaseConfig(dim=3, ischeme=1, ilim=2, vismodel=3, mach_ref=0.84, aoa_degree=3.0, reynolds=1.2e7,
                     centropy1=0.001, centropy2=0.001, vencat_coef=0.001, ncfl=10, cflst=1.0, cfled=5.0)
    return g, cfg


def compare(a, b, C):
    for f in ("Q", "TURBQ", "VIST", "LIMITER", "TIMESTEP", "RES", "DQ", "DIST", "XCC", "FACE_NORMAL", "DQDX"):
        u, v = a.get(f), b.get(f)
        if f in ("LIMITER", "TIMESTEP", "RES", "DQ"):
            u, v = u[..., :C], v[..., :C]
        assert np.array_equal(u, v), (f, float(np.max(np.abs(u - v))))
    assert np.array_equal(a.aero_row(), b.aero_row())
    assert np.array_equal(a.wall_data(), b.wall_data())


@pytest.mark.parametrize("which", ["mc", "natural", "wing"])
def test_storage_order_changes_no_result(which):
    from oneflow_amd.solver import Solver
    g, cfg = wing() if which == "wing" else airfoil(which)
    print(f"[storage] {which}: cells={g.n_cells}", flush=True)
    a, b = Solver(g, cfg, storage_order=0), Solver(g, cfg, storage_order=1)
    for kernel in (0, 5, 6):
        b.set_kernel_variant(kernel, 1)
    for s in (a, b):
        s.init_freestream()
    na, nb = a.iterate(1, 15), b.iterate(1, 15)
    np.testing.assert_allclose(nb, na, rtol=1e-13, atol=0)
    compare(a, b, g.n_cells)
    a.close()
    b.close()


def test_storage_order_zones():
    from oneflow_amd.solver import ZoneSolver
    (g1, c1), (g2, c2) = airfoil("mc"), airfoil("natural")
    c2 = c2.replace(mach_ref=0.5, aoa_degree=3.0)
    a = ZoneSolver([g1, g2], [c1, c2], storage_order=0)
    b = ZoneSolver([g1, g2], [c1, c2], storage_order=1)
    for kernel in (0, 5, 6):
        b.set_kernel_variant(kernel, 1)
    for s in (a, b):
        s.init_freestream()
        s.iterate(1, 12, with_norms=False)
    for f in ("Q", "TURBQ", "DIST"):
        assert np.array_equal(a.get(f), b.get(f)), f
    np.testing.assert_allclose(b.zone_norms(), a.zone_norms(), rtol=1e-13, atol=0)
    assert np.array_equal(a.zone_aero(), b.zone_aero())
