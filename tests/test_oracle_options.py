"""Closed-form pins of the oracle options that no reference golden runs.

The reference cases use the viscous spectral-radius model 2 and the SA
production form 2.  These tests pin the other forms the path supports with
values the mathematics fixes on tiny meshes:

* viscous spectral radius, models 1, 2 and 3 (TimeStep::CalcFaceVisSpec,
  codes/ns/src/TimeStep.cpp:90-126), and the inviscid radius and local time
  step beside them (TimeStep.cpp:62-88, 128-138): on a uniform rectangular
  grid in a uniform stream every cell sees two faces of length hy with
  |vn| = |u| and two of length hx with vn = 0, so each radius is a short
  closed form in hx, hy, mu, rho, Re, c;
* the LU-SGS viscous term of models 1 and 2/3 (UNsLusgs::CalcViscousTerm,
  codes/uns/src/UNsLusgs.cpp:294-330): on two cells at rest, a density-only
  right-hand side in the first cell gives flux increments (c/2) A drho in the
  density row only, so one lower and one upper sweep leave
  dq1 = (cA/2 + v) r / (D0 D1) and dq0 = r / D0 + (cA/2 + v)^2 r / (D0^2 D1);
* the SA production forms 1, 4 and 5 (TurbCom::CalcSaProd,
  codes/turb/src/TurbCom.cpp:444-465): each replaces the vorticity magnitude
  by a function of stp = 2 work2 + work3 - 2/3 work1^2, so a cell's SA source
  under form p with velocity gradient G equals the source under the
  rotational form 2 with a pure shear gradient of magnitude S(G); S is
  sqrt(|stp|) for form 1 and the stated corrections for forms 4 and 5.
"""
import math

import numpy as np
import pytest

from oneflow_amd import meshgen
from oneflow_amd.config import CaseConfig
from oneflow_amd.mesh import BC_SOLID_SURFACE, BC_SYMMETRY, BC_FARFIELD
from oracle.pyoracle import OracleCase

GAMMA = 1.4
SMALL = 1.0e-40


def _rect_case(model, vis_ts=1, nx=5, ny=4, lx=1.0, ly=0.6, mach=0.5):
    g = meshgen.box2d(nx, ny, 0.0, lx, 0.0, ly)
    cfg = CaseConfig(dim=2, ischeme=1, ilim=0, vismodel=1, visSRModel=model, visTimeStepModel=vis_ts,
                     mach_ref=mach, aoa_degree=0.0, reynolds=3.0e4)
    o = OracleCase(g, cfg)
    o.init_freestream()
    o.ns_boundary()
    return g, cfg, o


@pytest.mark.parametrize("model", [1, 2, 3])
@pytest.mark.parametrize("vis_ts", [0, 1])
def test_spectral_radii_and_time_step_closed_form(model, vis_ts):
    nx, ny, lx, ly = 5, 4, 1.0, 0.6
    g, cfg, o = _rect_case(model, vis_ts, nx, ny, lx, ly)
    hx, hy = lx / nx, ly / ny
    it = 500  # past the CFL ramp: cfl = cfled
    o.timestep(it)
    C = g.n_cells
    q = o.get("Q")[:, :C]
    mu = o.get("VISL")[:C]
    rho, u, p = q[0], q[1], q[4]
    c = np.sqrt(GAMMA * p / rho)
    re = cfg.reynolds
    inv = hy * (np.abs(u) + c) + hx * c
    if model == 1:
        c3 = 2.0 * np.maximum(4.0 / 3.0 * mu, GAMMA * (mu / cfg.prl)) / (re * rho)
        vis = 2.0 * hy * hy * c3 + 2.0 * hx * hx * c3
    elif model == 2:
        vis = 2.0 * mu * hy / (rho * hx * re) + 2.0 * mu * hx / (rho * hy * re)
    else:
        vis = 2.0 * hy * hy / (re * rho) + 2.0 * hx * hx / (re * rho)
    np.testing.assert_allclose(o.get("INVSR")[:C], inv, rtol=1e-13)
    np.testing.assert_allclose(o.get("VISSR")[:C], vis, rtol=1e-13)
    vol = hx * hy
    dt = cfg.cfled * vol / inv
    if vis_ts:
        vts = cfg.cfled * vol / vis
        dt = dt * vts / (dt + vts)
    np.testing.assert_allclose(o.get("TIMESTEP")[:C], dt, rtol=1e-13)
    # the three models differ on this grid (a swapped model would fail)
    assert len({round(float(v), 12) for v in (
        2 * (hy * hy + hx * hx) * 2 * max(4 / 3, GAMMA / cfg.prl) * mu[0] / (re * rho[0]),
        2 * mu[0] * (hy / hx + hx / hy) / (rho[0] * re),
        2 * (hy * hy + hx * hx) / (re * rho[0]))}) == 3


@pytest.mark.parametrize("model", [1, 2, 3])
def test_lusgs_viscous_term_two_cells_at_rest(model):
    lx, ly = 1.0, 0.7
    g = meshgen.box2d(2, 1, 0.0, lx, 0.0, ly, bc=lambda xm, ym: BC_SYMMETRY)
    cfg = CaseConfig(dim=2, ischeme=1, ilim=0, vismodel=1, visSRModel=model, visTimeStepModel=1,
                     mach_ref=0.3, reynolds=500.0)
    o = OracleCase(g, cfg)
    o.init_freestream()
    T = g.n_tcell
    q = o.get("Q")
    q[1:4] = 0.0  # at rest, cells and ghosts
    o.set("Q", q)
    o.ns_boundary()
    o.timestep(500)
    res = np.zeros((5, T))
    r0 = 0.37
    res[0, 0] = r0
    o.set("RES", res)
    o.lusgs_ns()
    dq = o.get("DQ")
    D = o.get("IMPSR")[:2]
    mu = o.get("VISL")[:2]
    rho, p = q[0, 0], q[4, 0]
    c = math.sqrt(GAMMA * p / rho)
    hx = lx / 2
    A = ly  # the shared face
    if model == 1:
        c3 = 2.0 * max(4.0 / 3.0 * mu[1], GAMMA * (mu[1] / cfg.prl)) / (cfg.reynolds * rho)
        v = A * A * c3 / (hx * ly)
    else:  # models 2 and 3 share the distance-based term in the sweeps
        v = 0.5 * (2.0 * mu[1] / (rho * hx * cfg.reynolds + SMALL)) * A
    a = 0.5 * c * A + v
    want1 = a * r0 / (D[0] * D[1])
    want0 = r0 / D[0] + a * a * r0 / (D[0] ** 2 * D[1])
    np.testing.assert_allclose(dq[0, 1], want1, rtol=1e-12)
    np.testing.assert_allclose(dq[0, 0], want0, rtol=1e-12)
    assert np.max(np.abs(dq[1:, :2])) <= 1e-15 * abs(want0)
    # the diagonal: vol/dt + invsr + vissr with dt from the same radii
    inv = o.get("INVSR")[:2]
    vis = o.get("VISSR")[:2]
    vol = hx * ly
    dt = cfg.cfled * vol / inv
    vts = cfg.cfled * vol / vis
    dt = dt * vts / (dt + vts)
    np.testing.assert_allclose(D, vol / dt + inv + vis, rtol=1e-13)


def _stp(G):
    (ux, uy, uz), (vx, vy, vz), (wx, wy, wz) = G
    w1 = ux + vy + wz
    w2 = ux * ux + uy * uy + uz * uz  # as TurbCom::CalcWorkVar forms it
    w3 = (uy + vx) ** 2 + (vz + wy) ** 2 + (wx + uz) ** 2
    return 2.0 * w2 + w3 - 2.0 / 3.0 * w1 * w1


def _vort(G):
    (ux, uy, uz), (vx, vy, vz), (wx, wy, wz) = G
    return math.sqrt((wy - vz) ** 2 + (uz - wx) ** 2 + (vx - uy) ** 2)


def _effective_omega(form, G):
    stp = _stp(G)
    str0 = _vort(G)
    if form == 1:
        return math.sqrt(abs(stp))
    if form == 4:
        return str0 - 2.0 * max(0.0, str0 - math.sqrt(abs(stp)))
    return str0 * (1.0 - max(0.0, (str0 * str0 - abs(stp)) / (str0 * str0 + abs(stp) + SMALL)))


GRADS = {
    "strain": ((0.8, 0.0, 0.0), (0.0, -0.8, 0.0), (0.0, 0.0, 0.0)),   # irrotational: form 2 gives 0
    "dilatation": ((0.6, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)),  # the 2/3 work1^2 term
    "shear": ((0.0, 0.9, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)),      # work2 holds du/dy
    "mixed": ((0.3, -0.7, 0.0), (0.2, 0.1, 0.0), (0.0, 0.0, 0.0)),
}


def _sa_case(form):
    def bc(xm, ym):
        return BC_SOLID_SURFACE if abs(ym) < 1e-12 else BC_FARFIELD
    g = meshgen.box2d(4, 3, 0.0, 1.0, 0.0, 0.05, bc=bc)
    cfg = CaseConfig(dim=2, ischeme=1, ilim=0, vismodel=3, iprod_sa=form, mach_ref=0.3, reynolds=1.0e6,
                     ref_sa=3.0)
    o = OracleCase(g, cfg)
    o.init_freestream()
    o.ns_boundary()
    o.sa_boundary()
    o.timestep(500)
    o.residual_ns()
    return g, o


def _set_grad(o, g, G):
    d = o.get("DQDX")
    for a in range(3):
        for s in range(3):
            d[a, 1 + s, :] = G[s][a]
    o.set("DQDX", d)


@pytest.mark.parametrize("form", [1, 4, 5])
@pytest.mark.parametrize("name", list(GRADS))
def test_sa_production_forms_equal_rotational_form_at_effective_omega(form, name):
    G = GRADS[name]
    om = _effective_omega(form, G)
    shear = ((0.0, -om, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0))  # |dv/dx - du/dy| = om
    assert abs(_vort(shear) - om) <= 1e-15 * max(om, 1.0)
    g, o = _sa_case(form)
    _set_grad(o, g, G)
    o.residual_sa()
    r_form = o.get("TURBRES")[: g.n_cells].copy()
    s_form = o.get("TURBIMPSR")[: g.n_cells].copy()
    g2, o2 = _sa_case(2)
    _set_grad(o2, g2, shear)
    o2.residual_sa()
    r_rot = o2.get("TURBRES")[: g2.n_cells]
    s_rot = o2.get("TURBIMPSR")[: g2.n_cells]
    scale = np.max(np.abs(r_rot))
    assert scale > 0
    np.testing.assert_allclose(r_form, r_rot, rtol=0, atol=1e-13 * scale)
    np.testing.assert_allclose(s_form, s_rot, rtol=1e-13)


def test_sa_production_form1_differs_from_rotational_on_strain():
    """Form 1 must see a pure strain that the rotational form ignores."""
    G = GRADS["strain"]
    g, o = _sa_case(1)
    _set_grad(o, g, G)
    o.residual_sa()
    g2, o2 = _sa_case(2)
    _set_grad(o2, g2, G)
    o2.residual_sa()
    r1, r2 = o.get("TURBRES")[: g.n_cells], o2.get("TURBRES")[: g2.n_cells]
    assert np.max(np.abs(r1 - r2)) > 1e-6 * np.max(np.abs(r2))


@pytest.mark.parametrize("rho_scale", [1.0, 3.0])
def test_isothermal_wall_ghost_closed_form(rho_scale):
    """IsothermalVisWallBc (codes/ns/src/NsBcSolver.cpp:271-329): the ghost of a
    no-slip wall at T_w = twall_dim / tref_dim carries -u, -v, -w, the owner
    pressure p, and density max(0.5 rho_w, 2 rho_w - rho) with
    rho_w = p / (statecoef T_w) (CalcDensity, :439-444, one species).
    rho_scale 3 puts the owner density above 1.5 rho_w, where the floor
    0.5 rho_w is taken."""
    wall = lambda xm, ym: BC_SOLID_SURFACE if ym < 1e-12 else BC_FARFIELD
    g = meshgen.box2d(6, 4, bc=wall)
    cfg = CaseConfig(dim=2, ischeme=1, ilim=0, vismodel=1, isowallbc=1, mach_ref=0.5, reynolds=1e4,
                     twall_dim=250.0, tref_dim=300.0)
    o = OracleCase(g, cfg)
    o.init_freestream()
    q = o.get("Q")
    xc = o.get("XCC")
    C = g.n_cells
    q[0, :C] = rho_scale * (1.0 + 0.2 * np.sin(4 * xc[0, :C]))
    q[1, :C] = 0.8 + 0.1 * xc[1, :C]
    q[2, :C] = 0.05 * np.cos(3 * xc[0, :C])
    q[4, :C] = (1.0 + 0.1 * xc[0, :C]) / (GAMMA * 0.25)
    o.set("Q", q)
    o.ns_boundary()
    out = o.get("Q")
    statecoef = 1.0 / (GAMMA * 0.5 * 0.5)
    tw = 250.0 / 300.0
    walls = [f for f in range(g.n_bfaces) if g.bc_type[f] == BC_SOLID_SURFACE]
    assert len(walls) == 6
    took_floor = 0
    for f in walls:
        l, gh = int(g.lc[f]), C + f
        rho, u, v, w, p = q[:, l]
        rw = p / (statecoef * tw * 1.0)
        rg = max(0.5 * rw, 2.0 * rw - rho)
        took_floor += rg == 0.5 * rw
        assert out[0, gh] == rg
        assert out[1, gh] == -u and out[2, gh] == -v and out[3, gh] == -w
        assert out[4, gh] == p
    assert took_floor == (6 if rho_scale == 3.0 else 0)
