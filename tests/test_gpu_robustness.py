"""The update's positivity fixes, the reconstruction's first-order fallback and
the NaN stop on the GPU, against the oracle.

The same forced states as tests/test_oracle_robustness.py (whose expected
values are computed by hand from NsUpdate::CalcFlowField / WeekSolutionFix,
codes/ns/src/NsUpdate.cpp:40-67,192-217, and UNsUpdate::SolutionFix,
codes/uns/src/UNsUpdate.cpp:94-133), run through the C ABI: the GPU update
must equal the oracle bit for bit, including the SolutionFix neighbour
average in the serial loop's half-updated state, on natural and
multicolour numberings (whose lower-sweep levels order the GPU's fix walk).
"""
import numpy as np
import pytest

from oneflow_amd import meshgen
from oneflow_amd.config import CaseConfig
from oracle.pyoracle import OracleCase
from tests.parity_util import report
from tests.test_oracle_robustness import expected_update, prim_to_q, spike_pair

pytestmark = pytest.mark.gpu


def pair(numbering, nx=12, ny=9):
    from oneflow_amd.solver import Solver
    g = meshgen.box2d(nx, ny, tri=True, jitter=0.2, seed=1)
    if numbering == "mc":
        g = meshgen.renumber(g, meshgen.multicolor_order(g))
    cfg = CaseConfig(dim=2, ischeme=1, ilim=0, vismodel=1, mach_ref=0.5, aoa_degree=5.0, reynolds=1e4)
    o = OracleCase(g, cfg)
    o.init_freestream()
    q = o.get("Q")
    xc = o.get("XCC")
    q[0] *= 1.0 + 0.1 * np.sin(3 * xc[0]) * np.cos(2 * xc[1])
    q[4] *= 1.0 + 0.05 * np.cos(5 * xc[0] + xc[1])
    o.set("Q", q)
    s = Solver(g, cfg)
    s.set("Q", q)
    print(f"[robust] numbering={numbering} cells={g.n_cells}", flush=True)
    return g, s, o, q


def forced_increment(g, q, seed):
    """Small random increments, three cells driven to p < 0 (one recovers by
    the weak fix, two adjacent ones lose all density and need SolutionFix)."""
    rng = np.random.default_rng(seed)
    C = g.n_cells
    dq = 1e-3 * rng.standard_normal((5, C))
    weak = C // 3
    dq[:, weak] = 0.0
    dq[1, weak], dq[4, weak] = 5.0, 5.0
    i = C // 2
    # a neighbour of i through an interior face
    f = next(f for f in range(g.n_bfaces, g.n_faces) if g.lc[f] == i or g.rc[f] == i)
    j = int(g.rc[f] if g.lc[f] == i else g.lc[f])
    for c in (i, j):
        dq[:, c] = 0.0
        dq[0, c] = -prim_to_q(q[:, c])[0]
    return dq, {weak}, {i, j}


@pytest.mark.parametrize("numbering", ["natural", "mc"])
def test_fixes_bitwise(numbering):
    g, s, o, q = pair(numbering)
    dq, weak, fix = forced_increment(g, q, seed=7)
    want, fixed = expected_update(g, q, dq)
    assert set(fixed) == fix
    full = np.zeros((5, g.n_tcell))
    full[:, : g.n_cells] = dq
    for x in (s, o):
        x.set("DQ", full)
        x.update_ns()
    C = g.n_cells
    a, b = s.get("Q")[:, :C], o.get("Q")[:, :C]
    assert report("Q", a, b) == 0.0
    np.testing.assert_allclose(b, want[:, :C], rtol=1e-14, atol=1e-14)
    # the boundary pass after a fixed update: temperatures, viscosities, ghosts
    for x in (s, o):
        x.ns_boundary()
    # ghosts included (the farfield ones through the shared pow)
    for n in ("Q", "TEMPR", "VISL"):
        assert report(n, s.get(n), o.get(n)) == 0.0, n
    s.close()


def test_no_fix_update_unchanged_by_fix_path():
    """With no cell needing a fix, the update is the plain parallel one."""
    g, s, o, q = pair("mc")
    rng = np.random.default_rng(11)
    full = np.zeros((5, g.n_tcell))
    full[:, : g.n_cells] = 1e-3 * rng.standard_normal((5, g.n_cells))
    for x in (s, o):
        x.set("DQ", full)
        x.update_ns()
    assert report("Q", s.get("Q")[:, : g.n_cells], o.get("Q")[:, : g.n_cells]) == 0.0
    s.close()


def test_nan_state_stops_the_step():
    """A NaN in the state ends ofx_iterate's norms with OFX_ENAN (the
    reference stops, codes/residual/src/ResidualTask.cpp:95-100)."""
    from oneflow_amd.solver import OfxError
    g, s, o, q = pair("natural")
    q2 = q.copy()
    q2[4, 3] = np.nan
    s.set("Q", q2)
    with pytest.raises(OfxError, match="-5"):
        s.iterate(1, 1)
    s.close()


def test_reconstruction_fallback_bitwise():
    """Unlimited MUSCL on the two-spike state of
    tests/test_oracle_robustness.py, whose middle face falls back to first
    order on both sides (LimField::CalcFaceValue, codes/usolver/src/ULimiter.cpp:218-254):
    the GPU's NS residual equals the oracle's bit for bit."""
    from oneflow_amd.solver import Solver
    g, cfg, o, q, face, control = spike_pair()
    s = Solver(g, cfg)
    s.init_freestream()
    s.set("Q", q)
    s.ns_boundary()
    for x in (s, o):
        x.residual_ns()
    C = g.n_cells
    assert report("RES", s.get("RES")[:, :C], o.get("RES")[:, :C]) == 0.0
    s.close()
