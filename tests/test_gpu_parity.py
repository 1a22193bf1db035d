"""GPU (libofx.so) vs the CPU oracle on the reference's own grids, task by task.

Every field must equal the oracle's bit for bit: both sides keep the
reference's operation order with -ffp-contract=off, and the one function
whose device and host library versions differ (pow) is the shared routine of
oneflow_amd/csrc/ofx_math.h on both.  The residual norms are a reduction in a
different order on the GPU (a deterministic tree) and are compared at 1e-12.
Whole 50-step runs must also reproduce the reference goldens (res.dat,
turbres.dat, aero.dat) under the reference harness rule.
"""
import numpy as np
import pytest

from oneflow_amd.config import REFERENCE_CASES
from oneflow_amd.mesh import read_ofl
from oracle.pyoracle import OracleCase
from oneflow_amd.restart import aero_lines, wallaero_lines
from tests.golden_util import GOLDEN, harness_equal, load_res
from tests.parity_util import report

pytestmark = pytest.mark.gpu

TOL = 0.0
NORM_TOL = 1.0e-12  # the norms' reduction order differs
CASES = ["plateuns2dslau2", "turbplateuns2droe_sa", "rae2822_roe_sa"]


def synthetic(case):
    """Small generated grids: natural and multicolour numbering (1-workgroup and
    per-level LU-SGS sweeps), Roe + SA + Venkat at a transonic condition."""
    from oneflow_amd import meshgen
    from oneflow_amd.config import CaseConfig
    g = meshgen.naca_hybrid(n_surf=96, n_quad=8, ratio=1.3, reynolds=6.5e6, n_outer=12, radius=20.0, m=0.02, p=0.4, t=0.12)
    if case.endswith("mc") or case.endswith("mcsf"):
        order = meshgen.multicolor_order(g)
        g = meshgen.renumber(g, order)
        if case.endswith("mcsf"):  # interior faces in their cells' spatial order
            g = meshgen.sort_interior_faces(g, order)
    elif case.endswith("blk"):  # blocked colouring: levels that are not runs of consecutive cells
        g = meshgen.renumber(g, meshgen.blocked_multicolor_order(g, part_cells=400))
    cfg = CaseConfig(ischeme=1, ilim=2, vismodel=3, mach_ref=0.8, aoa_degree=1.25, reynolds=6.5e6,
                     centropy1=0.001, centropy2=0.001, vencat_coef=0.001)
    return g, cfg


def make(case):
    from oneflow_amd.solver import Solver
    if case.startswith("syn"):
        g, cfg = synthetic(case)
    else:
        g = read_ofl(GOLDEN / case / "grid.ofl")
        cfg = REFERENCE_CASES[case]
    print(f"[parity] case={case} cells={g.n_cells} faces={g.n_faces}", flush=True)
    return Solver(g, cfg), OracleCase(g, cfg)


def cmp(gpu, ora, names, tol=TOL):
    for n in names:
        a, b = gpu.get(n), ora.get(n)
        if n in ("RES", "DQ", "LIMITER", "TIMESTEP", "INVSR", "VISSR", "IMPSR", "TURBRES", "TURBDQ",
                 "TURBIMPSR"):
            C = ora.n_cells
            a, b = a[..., :C], b[..., :C]
        e = report(n, a, b)
        assert e <= tol, (n, e)


@pytest.mark.parametrize("case", CASES)
def test_geometry(case):
    gpu, ora = make(case)
    cmp(gpu, ora, ["FACE_NORMAL", "FACE_CENTER", "XCC", "VOL", "DIST"], tol=0.0)


STATE = ["Q", "TURBQ", "TEMPR", "VISL", "VIST", "TIMESTEP", "INVSR", "VISSR", "RES", "TURBRES", "DQ", "TURBDQ"]


def sync(gpu, ora, names=STATE):
    """Give the GPU exactly the oracle's state so each task is checked on identical inputs."""
    for n in names:
        gpu.set(n, ora.get(n))


@pytest.mark.parametrize("case", CASES + ["syn_natural", "syn_mc", "syn_mcsf", "syn_blk"])
def test_one_step_task_by_task(case):
    gpu, ora = make(case)
    for s in (gpu, ora):
        s.init_freestream()
    cmp(gpu, ora, ["Q", "TEMPR", "VISL", "VIST", "TURBQ"])
    for it in (1, 2, 3):
        sync(gpu, ora)
        for s in (gpu, ora):
            s.timestep(it)
        cmp(gpu, ora, ["TIMESTEP", "INVSR", "VISSR"])
        sync(gpu, ora)
        for s in (gpu, ora):
            s.residual_ns()
        cmp(gpu, ora, ["DQDX", "DTDX"])
        if ora.cfg.ilim in (1, 2):
            cmp(gpu, ora, ["LIMITER"])
        cmp(gpu, ora, ["RES"])
        sync(gpu, ora)
        for s in (gpu, ora):
            s.lusgs_ns()
        cmp(gpu, ora, ["IMPSR", "DQ"])
        sync(gpu, ora)
        for s in (gpu, ora):
            s.update_ns()
        cmp(gpu, ora, ["Q"])
        sync(gpu, ora)
        for s in (gpu, ora):
            s.ns_boundary()
        cmp(gpu, ora, ["Q", "TEMPR", "VISL"])
        if ora.cfg.vismodel == 3:
            sync(gpu, ora)
            gpu.residual_ns()  # the SA source reads the NS gradients of this step's state
            ora.residual_ns()
            sync(gpu, ora)
            for s in (gpu, ora):
                s.residual_sa()
            cmp(gpu, ora, ["TURBDQDX", "TURBRES", "TURBIMPSR"])
            sync(gpu, ora)
            gpu.set("TURBIMPSR", ora.get("TURBIMPSR"))
            for s in (gpu, ora):
                s.lusgs_sa()
            cmp(gpu, ora, ["TURBDQ"])
            sync(gpu, ora)
            for s in (gpu, ora):
                s.update_sa()
                s.sa_boundary()
            cmp(gpu, ora, ["TURBQ", "VIST"])
        # advance the oracle one full step so the next pass starts from a non-trivial state
        ora.iterate(it, 1)


@pytest.mark.parametrize("case", CASES)
def test_fifty_steps_against_goldens(case):
    gpu, ora = make(case)
    gpu.init_freestream()
    ora.init_freestream()
    norms, aero, aero_rows = [], [], []
    for it in range(1, 51):
        norms.append(gpu.iterate(it, 1)[0])
        aero.append(gpu.aero())
        aero_rows.append(gpu.aero_row())
    norms = np.array(norms)
    onorms = ora.iterate(1, 50)
    e = report("norms", norms, onorms)
    assert e <= NORM_TOL, e
    cmp(gpu, ora, ["Q", "TURBQ"])
    gold = load_res(GOLDEN / case / "res.dat", 5)
    for it in range(50):
        for q in range(5):
            assert harness_equal("%.5e" % norms[it, q], "%.5e" % gold[it, q]), (case, it + 1, q)
    tp = GOLDEN / case / "turbres.dat"
    if tp.exists():
        tg = load_res(tp, 1)
        for it in range(50):
            assert harness_equal("%.5e" % norms[it, 5], "%.5e" % tg[it, 0]), (case, it + 1)
    rows = [l.split() for l in (GOLDEN / case / "aero.dat").read_text().splitlines() if l[:1].isdigit()]
    for it, row in enumerate(rows):
        for mine, ref in zip(aero[it][:2], row[3:5]):
            assert harness_equal("%.4e" % mine, "%.4e" % float(ref)), (case, it + 1, aero[it], row[3:5])
    # the whole aero.dat file (15 columns) from the GPU's rows
    from tests.test_oracle_golden import compare_aero_file
    compare_aero_file(case, aero_lines(1, aero_rows), (GOLDEN / case / "aero.dat").read_text().splitlines())
    # wallaero.dat after step 50: per wall face -cp and cf at 11 digits
    wall = gpu.wall_data()
    ew = report("wall", wall, ora.wall_data())
    assert ew == 0.0, ew
    mine = wallaero_lines(wall)
    gold_w = (GOLDEN / case / "wallaero.dat").read_text().splitlines()
    assert len(mine) == len(gold_w)
    for a, b in zip(mine[8:], gold_w[8:]):
        for x, y in zip(a.split(), b.split()):
            assert harness_equal(x, y), (case, a, b)


@pytest.mark.parametrize("case", ["syn_natural", "syn_mc", "syn_blk"])
def test_synthetic_twenty_steps(case):
    gpu, ora = make(case)
    gpu.init_freestream()
    ora.init_freestream()
    ng = gpu.iterate(1, 20)
    no = ora.iterate(1, 20)
    e = report("norms", ng, no)
    assert e <= NORM_TOL, e
    cmp(gpu, ora, ["Q", "TURBQ"])


def test_continue_from_reference_restart():
    """The reference's restart file (plate after 34950 steps) through the GPU:
    same pins as tests/test_restart.py, and the oracle's norms and forces."""
    from oneflow_amd.solver import Solver
    from tests.test_restart import check_against_goldens, continued
    norms, aero = continued(Solver)
    onorms, oaero = continued(OracleCase)
    e = report("norms", norms, onorms)
    assert e <= NORM_TOL, e
    assert report("aero", aero, oaero) == 0.0
    check_against_goldens(norms, aero)


def test_step_reproduces_reference_restart_state_bitwise():
    """One GPU pseudo-step from the reference restart file's q2 gives its q1
    bit for bit (see tests/test_restart.py)."""
    from oneflow_amd.solver import Solver
    from tests.test_restart import step_from_q2
    q, r = step_from_q2(Solver)
    assert np.array_equal(q, r.q1)
