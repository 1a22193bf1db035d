"""Launch variants give identical bits.

The NS gradient (one thread per cell; one wave per quantity), the limiter
and the residual gather (one lane per face; one thread per cell) have two
kernels each, the cell kernels two walks over a colouring (colour by colour;
interleaved chunks), the viscous face kernel a split form (inviscid and
viscous parts on separate waves), and the face kernels take any face tile.  Every combination must give the
same state, gradients, limiter and residual bit for bit, here after twelve
pseudo-time steps from freestream (through the CFL ramp) on 2-D and 3-D grids
and several option sets.  Variant 0 with the default tile is checked against
the oracle by the other GPU tests.
"""
import numpy as np
import pytest

from tests.test_gpu_options import GRIDS, OPTS, make

pytestmark = pytest.mark.gpu

FIELDS = ["Q", "TURBQ", "DQDX", "LIMITER", "RES", "TIMESTEP", "DQ"]


@pytest.mark.parametrize("grid", ["quad", "tri", "mixed", "wing"])
@pytest.mark.parametrize("opt", ["roe_venkat_sa_test", "roe_barth_lam_std", "slau2_o1_lam_aver_iso"])
def test_grad_variant_and_tiles_bitwise(grid, opt):
    from oneflow_amd.solver import Solver
    base, _, g = make(grid, opt)
    runs = [base]
    # (gradient, limiter, gather, cell walk, face split) variants and face tiles
    for variants, tiles in (((1, 0, 0, 0, 1), (256, 256)), ((1, 1, 1, 8, 0), (64, 128)),
                            ((0, 1, 0, 32, 1), (512, 64)), ((0, 0, 1, 0, 0), (128, 512))):
        s = Solver(g, base.cfg)
        for kernel, v in enumerate(variants):
            s.set_kernel_variant(kernel, v)
        s.set_face_tile(*tiles)
        runs.append(s)
    for s in runs:
        s.init_freestream()
        s.iterate(1, 12, with_norms=False)
        s.residual_ns()
    for n in FIELDS:
        a = runs[0].get(n)
        for k, s in enumerate(runs[1:]):
            assert np.array_equal(a, s.get(n)), (grid, opt, n, k)
    print(f"\n[variants] {grid} {opt}: {len(runs)} launch configurations agree", flush=True)


@pytest.mark.parametrize("face_order", ["cell", "spatial"])
def test_cell_walk_variant_bitwise(face_order):
    """The interleaved cell walk (kernel 3) on a multicolour numbering: the
    same bits as the colour-by-colour walk after twelve full steps."""
    from oneflow_amd import meshgen
    from oneflow_amd.config import CaseConfig
    from oneflow_amd.solver import Solver
    g = meshgen.naca_hybrid(n_surf=96, n_quad=8, ratio=1.3, reynolds=6.5e6, n_outer=12, radius=20.0,
                            m=0.02, p=0.4, t=0.12)
    order = meshgen.multicolor_order(g)
    g = meshgen.renumber(g, order)
    if face_order == "spatial":
        g = meshgen.sort_interior_faces(g, order)
    cfg = CaseConfig(ischeme=1, ilim=2, vismodel=3, mach_ref=0.8, aoa_degree=1.25, reynolds=6.5e6,
                     centropy1=0.001, centropy2=0.001, vencat_coef=0.001)
    runs = []
    for v, chunk in ((0, 0), (1, 8), (3, 3), (64, 64)):
        s = Solver(g, cfg)
        s.set_kernel_variant(3, v)
        assert s.kernel_variant(3) == chunk
        s.init_freestream()
        s.iterate(1, 12, with_norms=False)
        s.residual_ns()
        s.residual_sa()
        runs.append(s)
    for n in FIELDS + ["TURBRES"]:
        for k in range(1, len(runs)):
            assert np.array_equal(runs[0].get(n), runs[k].get(n)), (face_order, n, k)
    print(f"\n[variants] cell walk, faces in {face_order} order: both walks agree", flush=True)
