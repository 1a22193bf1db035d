"""The .ofl writer against the reader, on a reference grid and after a
renumbering (boundary-region name ids travel with their faces)."""
import numpy as np

from oneflow_amd import meshgen
from oneflow_amd.mesh import read_ofl, write_ofl
from tests.golden_util import GOLDEN


def _same(a, b):
    assert a.dim == b.dim and a.n_cells == b.n_cells and a.n_faces == b.n_faces and a.n_bfaces == b.n_bfaces
    for k in ("x", "y", "z", "face_node_count", "face_nodes", "lc", "rc", "bc_type", "bc_name_id"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)


def test_write_read_round_trip_reference_grid(tmp_path):
    src = GOLDEN / "rae2822_roe_sa" / "grid.ofl"
    g = read_ofl(src)
    p = tmp_path / "g.ofl"
    write_ofl(g, p)
    _same(read_ofl(p), g)
    # byte-identical to the reference's own file
    assert p.read_bytes() == src.read_bytes()


def test_renumbered_grid_keeps_region_names(tmp_path):
    g = read_ofl(GOLDEN / "turbplateuns2droe_sa" / "grid.ofl")
    # give each boundary face a region id that is not its type, then renumber
    g.bc_name_id = (np.arange(g.n_bfaces) % 7 + 100).astype(np.int32)
    order = meshgen.multicolor_order(g)
    r = meshgen.renumber(g, order)
    # each boundary face keeps its (type, region) pair: compare the multisets per owner cell
    new_of_old = np.empty(g.n_cells, np.int64)
    new_of_old[order] = np.arange(g.n_cells)
    nb = g.n_bfaces
    before = sorted(zip(new_of_old[g.lc[:nb]].tolist(), g.bc_type.tolist(), g.bc_name_id.tolist()))
    after = sorted(zip(r.lc[:nb].tolist(), r.bc_type.tolist(), r.bc_name_id.tolist()))
    assert before == after
    p = tmp_path / "r.ofl"
    write_ofl(r, p)
    _same(read_ofl(p), r)
