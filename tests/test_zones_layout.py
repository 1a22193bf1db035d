"""Layout of several zones in one grid (meshgen.concat_grids).

Each zone's faces, cells and nodes must come back from the concatenated grid
unchanged up to the offsets, and -- since zones share no face -- a solver
without grid-wide reductions (first order, local time steps, no limiter
extents) must advance every zone of the concatenated grid exactly as it
advances that zone alone: the oracle, run on both, agrees bit for bit.
"""
import numpy as np

from oneflow_amd import meshgen
from oneflow_amd.config import CaseConfig
from oneflow_amd.mesh import BC_FARFIELD, BC_SOLID_SURFACE
from oracle.pyoracle import OracleCase


def zones():
    def bc(xm, ym):
        return BC_SOLID_SURFACE if abs(ym) < 1e-12 else BC_FARFIELD
    return [meshgen.box2d(5, 4, jitter=0.2, seed=1, bc=bc),
            meshgen.box2d(4, 6, tri=True, jitter=0.2, seed=2, bc=bc),
            meshgen.box2d(3, 3, x0=-1.0, jitter=0.1, seed=3, bc=bc)]


def face_nodes(g, f):
    off = np.concatenate([[0], np.cumsum(g.face_node_count)])
    return g.face_nodes[off[f]:off[f + 1]]


def test_concat_layout():
    gs = zones()
    g, cs, bs = meshgen.concat_grids(gs)
    assert list(cs) == [0] + list(np.cumsum([z.n_cells for z in gs]))
    assert list(bs) == [0] + list(np.cumsum([z.n_bfaces for z in gs]))
    B = g.n_bfaces
    assert np.all(g.rc[:B] == -1) and np.all(g.rc[B:] >= 0)
    nodes0 = np.concatenate([[0], np.cumsum([z.n_nodes for z in gs])])
    ioff = B
    for k, z in enumerate(gs):
        zb = z.n_bfaces
        # boundary faces of zone k: [bs[k], bs[k+1]); interior faces after all boundary faces
        fmap = list(range(bs[k], bs[k + 1])) + list(range(ioff, ioff + z.n_faces - zb))
        ioff += z.n_faces - zb
        for f, F in enumerate(fmap):
            assert np.array_equal(face_nodes(g, F), face_nodes(z, f) + nodes0[k])
            assert g.lc[F] == z.lc[f] + cs[k]
            assert (g.rc[F] == -1) if f < zb else (g.rc[F] == z.rc[f] + cs[k])
        assert np.array_equal(g.bc_type[bs[k]:bs[k + 1]], z.bc_type)
        assert np.array_equal(g.x[nodes0[k]:nodes0[k + 1]], z.x)


def test_oracle_zones_advance_independently():
    gs = zones()
    g, cs, bs = meshgen.concat_grids(gs)
    cfg = CaseConfig(dim=2, ischeme=1, ilim=-1, vismodel=0, mach_ref=0.5, aoa_degree=3.0, reynolds=1e5,
                     ncfl=5, cflst=1.0, cfled=5.0)
    whole = OracleCase(g, cfg)
    whole.init_freestream()
    alone = [OracleCase(z, cfg) for z in gs]
    for o in alone:
        o.init_freestream()
    # a zone-dependent disturbance so that the zones do not all sit at freestream
    for o, k in [(whole, None)] + [(a, i) for i, a in enumerate(alone)]:
        q = o.get("Q")
        xc = o.get("XCC")
        q[0] *= 1.0 + 0.05 * np.sin(3.0 * xc[0] + xc[1])
        o.set("Q", q)
        o.ns_boundary()
    whole.iterate(1, 6)
    for o in alone:
        o.iterate(1, 6)
    q = whole.get("Q")
    C = g.n_cells
    for k, o in enumerate(alone):
        qa = o.get("Q")
        nc = gs[k].n_cells
        assert np.array_equal(q[:, cs[k]:cs[k + 1]], qa[:, :nc]), k
        assert np.array_equal(q[:, C + bs[k]:C + bs[k + 1]], qa[:, nc:]), k
