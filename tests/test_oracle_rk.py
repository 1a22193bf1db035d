"""The multi-stage (Runge-Kutta) time integral of the oracle, pinned on the CPU.

The reference's TimeIntegral::RungeKutta (codes/solver/src/TimeIntegral.cpp:81-110)
runs, per solver, CALC_TIME_STEP and then, per stage s with
lhscoef = 1/(s+1) (codes/global/src/Ctrl.cpp:74-79), UPDATE_RESIDUALS,
CALC_LHS (ULhs::CalcLHS, codes/usolver/src/ULhs.cpp:47-62:
res *= (dt/vol) * lhscoef on the inner cells), UPDATE_FLOWFIELD with the
residual as the increment (dq aliases res under MULTI_STAGE,
codes/uns/src/UNsCom.cpp:85-88) and CALC_BOUNDARY.

The residual and the time step of the oracle are pinned by the reference
goldens (test_oracle_golden.py).  Here the stage bookkeeping is checked
against an independent numpy statement of the rules above: the per-cell
scaling, the stage coefficients, the conservative update and the norms of
the scaled residual.  A dropped volume, a stage coefficient in the wrong
order or an update from the stage-0 state fails one of them.
"""
import numpy as np
import pytest

from oneflow_amd.config import REFERENCE_CASES
from oneflow_amd.mesh import read_ofl
from oracle.pyoracle import OracleCase
from tests.golden_util import GOLDEN

GAMMA = 1.4


def prim_to_cons(p):
    rho, u, v, w, pr = p
    em = (pr / rho) / (GAMMA - 1.0) + 0.5 * (u * u + v * v + w * w)
    return np.array([rho, rho * u, rho * v, rho * w, rho * em])


def cons_to_prim(q):
    rho = np.abs(q[0])
    rd = 1.0 / rho
    u, v, w = q[1] * rd, q[2] * rd, q[3] * rd
    p = (GAMMA - 1.0) * (q[4] - 0.5 * rho * (u * u + v * v + w * w))
    return np.array([rho, u, v, w, p])


def case(name, **kw):
    g = read_ofl(GOLDEN / name / "grid.ofl")
    cfg = REFERENCE_CASES[name].replace(**kw)
    return g, cfg


def warm(o, steps=5):
    """A non-uniform state: a few LU-SGS steps from freestream."""
    o.init_freestream()
    o.iterate(1, steps)


@pytest.mark.parametrize("stages", [1, 2, 3])
def test_rk_stages_match_numpy_statement(stages):
    g, cfg = case("plateuns2dslau2", cflst=0.5, cfled=0.5)
    lu = OracleCase(g, cfg)
    warm(lu)
    q0 = lu.get("Q")
    rk = OracleCase(g, cfg.replace(time_integral=1, rk_stage=stages))
    step = OracleCase(g, cfg.replace(time_integral=1, rk_stage=stages))
    for o in (rk, step):
        o.set("Q", q0)
        o.ns_boundary()
    C = g.n_cells
    it = 6
    rk.iterate(it, 1)
    # the same stages task by task, the arithmetic of CALC_LHS and UPDATE_FLOWFIELD in numpy
    step.timestep(it)
    dt = step.get("TIMESTEP")[:C]
    vol = step.get("VOL")[:C]
    for s in range(stages):
        step.residual_ns()
        res = step.get("RES")[:, :C]
        inc = res * ((dt / vol) * (1.0 / (s + 1)))
        q = step.get("Q")
        prim = q[:, :C].copy()
        new = cons_to_prim(prim_to_cons(prim) + inc)
        assert np.all(new[0] > 0) and np.all(new[4] > 0)  # no positivity fix in this state
        q[:, :C] = new
        step.set("Q", q)
        step.ns_boundary()
        assert np.array_equal(step.get("Q")[:, :C], new)
    assert np.array_equal(rk.get("Q"), step.get("Q"))
    # the residual the step leaves is the last stage's, scaled by CALC_LHS
    assert np.array_equal(rk.get("RES")[:, :C], inc)
    n = rk.norms()
    expect = np.sqrt(np.sum(inc * inc, axis=1) / C)
    np.testing.assert_allclose(n[:5], expect, rtol=1e-13, atol=0)


def test_rk_lhs_scales_inner_cells_only():
    g, cfg = case("rae2822_roe_sa", cflst=0.5, cfled=0.5, time_integral=1, rk_stage=1)
    o = OracleCase(g, cfg)
    warm(o, 3)
    o.timestep(4)
    o.residual_ns()
    o.residual_sa()
    C = g.n_cells
    res, tres = o.get("RES"), o.get("TURBRES")
    dt, vol = o.get("TIMESTEP")[:C], o.get("VOL")[:C]
    o.lhs(0, 0.25)
    o.lhs(1, 0.5)
    r2, t2 = o.get("RES"), o.get("TURBRES")
    assert np.array_equal(r2[:, :C], res[:, :C] * ((dt / vol) * 0.25))
    assert np.array_equal(r2[:, C:], res[:, C:])
    assert np.array_equal(t2[:C], tres[:C] * ((dt / vol) * 0.5))
    assert np.array_equal(t2[C:], tres[C:])


def test_rk_sa_stage_uses_turbres_as_increment():
    g, cfg = case("turbplateuns2droe_sa", cflst=0.5, cfled=0.5)
    lu = OracleCase(g, cfg)
    warm(lu)
    q0, t0 = lu.get("Q"), lu.get("TURBQ")
    C = g.n_cells
    rk = OracleCase(g, cfg.replace(time_integral=1, rk_stage=2))
    ref = OracleCase(g, cfg.replace(time_integral=1, rk_stage=2))
    for o in (rk, ref):
        o.set("Q", q0)
        o.set("TURBQ", t0)
        o.ns_boundary()
        o.sa_boundary()
    rk.iterate(6, 1)
    # NS stages through the oracle's own tasks, then the SA stages by hand
    ref.timestep(6)
    for s in range(2):
        ref.residual_ns()
        ref.lhs(0, 1.0 / (s + 1))
        ref.update_ns()
        ref.ns_boundary()
    dt, vol = ref.get("TIMESTEP")[:C], ref.get("VOL")[:C]
    for s in range(2):
        ref.residual_sa()
        tr = ref.get("TURBRES")[:C]
        t = ref.get("TURBQ")
        t[:C] = t[:C] + tr * ((dt / vol) * (1.0 / (s + 1)))
        ref.set("TURBQ", t)
        ref.sa_boundary()
    assert np.array_equal(rk.get("Q"), ref.get("Q"))
    assert np.array_equal(rk.get("TURBQ")[:C], ref.get("TURBQ")[:C])


def test_rk_preserves_freestream():
    from oneflow_amd import meshgen
    from oneflow_amd.config import CaseConfig
    g = meshgen.naca_hybrid(n_surf=48, n_quad=4, ratio=1.3, reynolds=1e6, n_outer=6, radius=10.0,
                            m=0.0, p=0.4, t=0.12)
    # an inviscid case with farfield on every boundary keeps the uniform state
    bc = g.bc_type.copy()
    bc[:] = 4
    g.bc_type = bc
    cfg = CaseConfig(ischeme=1, ilim=2, vismodel=0, mach_ref=0.5, aoa_degree=2.0, time_integral=1, rk_stage=3)
    o = OracleCase(g, cfg)
    o.init_freestream()
    q0 = o.get("Q")
    o.iterate(1, 3)
    np.testing.assert_allclose(o.get("Q"), q0, rtol=0, atol=1e-13)


def test_rk_config_is_validated():
    g, cfg = case("plateuns2dslau2")
    with pytest.raises(RuntimeError):
        OracleCase(g, cfg.replace(time_integral=3))
    with pytest.raises(RuntimeError):
        OracleCase(g, cfg.replace(time_integral=1, rk_stage=0))
