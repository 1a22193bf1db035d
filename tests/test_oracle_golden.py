"""Pin the CPU oracle to the reference's own regression goldens.

test/test.py (reference) compares each output file line by line: identical
text passes, otherwise every numeric token must satisfy |a-b| <= 1e-8
(test/test.py:38-79).  The oracle's residual norms are printed with the
reference's own format (%.5e, codes/residual/src/ResidualTask.cpp:160-170) and
CL/CD with %.4e (codes/force/src/AeroForceTask.cpp:118-131) before comparison.
"""
import numpy as np
import pytest

from oneflow_amd.config import REFERENCE_CASES
from oneflow_amd.mesh import read_ofl
from oneflow_amd.restart import aero_lines, wallaero_lines
from oracle.pyoracle import OracleCase
from tests.golden_util import GOLDEN, load_res, harness_equal

CASES = ["plateuns2dslau2", "turbplateuns2droe_sa", "rae2822_roe_sa"]
WALL = {}  # wall rows of each case after step 50
ROWS = {}  # the 12 aero.dat value columns of each case, per step


@pytest.fixture(scope="module", params=CASES)
def run(request):
    case = request.param
    g = read_ofl(GOLDEN / case / "grid.ofl")
    c = OracleCase(g, REFERENCE_CASES[case])
    c.init_freestream()
    norms, aero, rows = [], [], []
    for it in range(1, 51):
        norms.append(c.iterate(it, 1)[0])
        aero.append(c.aero())
        rows.append(c.aero_row())
    WALL[case] = c.wall_data()
    ROWS[case] = np.array(rows)
    return case, np.array(norms), np.array(aero)


def test_res_dat(run):
    case, norms, _ = run
    gold = load_res(GOLDEN / case / "res.dat", 5)
    assert gold.shape == (50, 5)
    for it in range(50):
        for e in range(5):
            assert harness_equal("%.5e" % norms[it, e], "%.5e" % gold[it, e]), (case, it + 1, e, norms[it, e], gold[it, e])


def test_turbres_dat(run):
    case, norms, _ = run
    p = GOLDEN / case / "turbres.dat"
    if not p.exists():
        pytest.skip("laminar case has no turbres.dat")
    gold = load_res(p, 1)
    for it in range(50):
        assert harness_equal("%.5e" % norms[it, 5], "%.5e" % gold[it, 0]), (case, it + 1, norms[it, 5], gold[it, 0])


def test_aero_dat(run):
    case, _, aero = run
    rows = [l.split() for l in (GOLDEN / case / "aero.dat").read_text().splitlines() if l[:1].isdigit()]
    assert len(rows) == 50
    for it, row in enumerate(rows):
        cl, cd, cdp, cdf = (float(v) for v in row[3:7])
        for mine, ref in zip(aero[it], (cl, cd, cdp, cdf)):
            assert harness_equal("%.4e" % mine, "%.4e" % ref), (case, it + 1, aero[it], row[3:7])


def compare_aero_file(case, mine, gold):
    """aero.dat text of `mine` against the golden file under the harness rule:
    the header text-equal, every row token equal or within 1e-8."""
    assert len(mine) == len(gold), (case, len(mine), len(gold))
    assert mine[:17] == [g.rstrip("\r") for g in gold[:17]]
    exact = 0
    for k, (a, b) in enumerate(zip(mine[17:], gold[17:])):
        exact += a == b.rstrip("\r")
        ta, tb = a.split(), b.split()
        assert len(ta) == len(tb) == 15, (case, k)
        for x, y in zip(ta, tb):
            assert harness_equal(x, y), (case, k + 1, a, b)
    return exact


def test_aero_dat_file(run):
    """All 15 columns of aero.dat, written by oneflow_amd.restart.aero_lines
    from the oracle's rows: CL, CD, CD_PR, CD_SF, CDL2, Xcp (signed zero),
    Fx, Fy, Fz and the zero moment columns."""
    case, _, _ = run
    gold = (GOLDEN / case / "aero.dat").read_text().splitlines()
    mine = aero_lines(1, ROWS[case])
    exact = compare_aero_file(case, mine, gold)
    print(f"[aero.dat] {case}: {exact} of 50 rows text-identical")
    # the columns the harness rule would let through as near-zero must carry the reference's sign
    assert [l.split()[8] for l in mine[17:]] == [l.split()[8] for l in gold[17:]]


def test_wallaero_dat(run):
    """Per wall face x, y, z, -cp and cf after step 50 at 11 significant
    digits: a field-level pin of the wall state, its gradients and the
    viscous stress."""
    case, _, _ = run
    gold = (GOLDEN / case / "wallaero.dat").read_text().splitlines()
    mine = wallaero_lines(WALL[case])
    assert len(mine) == len(gold), (case, len(mine), len(gold))
    assert mine[:8] == [g.rstrip("\r") for g in gold[:8]]
    worst = 0.0
    for k, (a, b) in enumerate(zip(mine[8:], gold[8:])):
        ta, tb = a.split(), b.split()
        assert len(ta) == len(tb) == 5, (case, k)
        for x, y in zip(ta, tb):
            worst = max(worst, abs(float(x) - float(y)))
            assert harness_equal(x, y), (case, k, a, b)
    print(f"[wallaero] {case}: {len(mine) - 8} wall faces, worst |diff| {worst:.2e}")
