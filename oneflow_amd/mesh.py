"""Unstructured zone grids: the .ofl file format and the in-memory face list.

A grid is kept exactly as the reference stores one zone
(codes/geometry/src/UnsGrid.cpp:90-200, framing codes/zone/src/GridGroup.cpp:67-138
and codes/database/src/DataBook.cpp:283-298):

    int32 nZones; int32 pid[nZones]; int32 zoneType[nZones]
    per zone: int64 nBytes, then
        int32 nNodes, nFaces, nCells
        float64 x[nNodes], y[nNodes], z[nNodes]
        int32 volBcType
        int32 faceNodeCount[nFaces]; int32 faceNodes[sum(faceNodeCount)]
        int32 lc[nFaces]; int32 rc[nFaces]          (rc = -1 on boundary faces)
        int32 nBFaces; int32 bcType[nBFaces]; int32 bcNameId[nBFaces]
        int32 nIFaces  (+ interface arrays when nIFaces > 0)

Boundary faces come first.  Only single-zone grids without interfaces are on
this path.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

# codes/geometry/src/Boundary.cpp:35-46
BC_EXTRAPOLATION = 1
BC_SOLID_SURFACE = 2
BC_SYMMETRY = 3
BC_FARFIELD = 4
BC_INFLOW = 5
BC_OUTFLOW = 6


@dataclass
class Grid:
    dim: int
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    face_node_count: np.ndarray   # int32 [nFaces]
    face_nodes: np.ndarray        # int32 [sum]
    lc: np.ndarray                # int32 [nFaces]
    rc: np.ndarray                # int32 [nFaces], -1 on boundary faces
    bc_type: np.ndarray           # int32 [nBFaces]
    bc_name_id: np.ndarray = field(default=None)  # int32 [nBFaces]
    n_cells: int = 0

    @property
    def n_nodes(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.lc.shape[0])

    @property
    def n_bfaces(self) -> int:
        return int(self.bc_type.shape[0])

    @property
    def n_tcell(self) -> int:
        return self.n_cells + self.n_bfaces

    def check(self) -> None:
        """Validate the invariants every consumer relies on."""
        nb, nf, nc = self.n_bfaces, self.n_faces, self.n_cells
        assert self.face_node_count.shape == (nf,)
        assert int(self.face_node_count.sum()) == self.face_nodes.shape[0]
        assert self.face_nodes.min() >= 0 and self.face_nodes.max() < self.n_nodes
        assert np.all(self.lc >= 0) and np.all(self.lc < nc), "lc out of range"
        assert np.all(self.rc[:nb] < 0), "boundary faces must come first with rc < 0"
        assert np.all(self.rc[nb:] >= 0) and np.all(self.rc[nb:] < nc), "interior rc out of range"
        if self.dim == 2:
            assert np.all(self.face_node_count == 2)


def read_ofl(path: str | Path, dim: int = 2) -> Grid:
    """Read a single-zone .ofl grid (codes/geometry/src/UnsGrid.cpp:90-200)."""
    b = Path(path).read_bytes()
    off = 0
    nz = struct.unpack_from("<i", b, off)[0]
    off += 4
    if nz != 1:
        raise ValueError(f"{path}: only single-zone grids are supported (nZones={nz})")
    off += 8  # pid[1], zoneType[1]
    off += 8  # int64 byte count of the zone's data book
    nn, nf, nc = struct.unpack_from("<3i", b, off)
    off += 12

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(b, dtype=dtype, count=n, offset=off).copy()
        off += a.nbytes
        return a

    x = take("<f8", nn)
    y = take("<f8", nn)
    z = take("<f8", nn)
    off += 4  # volBcType
    fnc = take("<i4", nf)
    fnn = take("<i4", int(fnc.sum()))
    lc = take("<i4", nf)
    rc = take("<i4", nf)
    nb = struct.unpack_from("<i", b, off)[0]
    off += 4
    bct = take("<i4", nb)
    bcn = take("<i4", nb)
    ni = struct.unpack_from("<i", b, off)[0]
    if ni != 0:
        raise ValueError(f"{path}: interface faces are not supported")
    g = Grid(dim=dim, x=x, y=y, z=z, face_node_count=fnc, face_nodes=fnn, lc=lc, rc=rc,
             bc_type=bct, bc_name_id=bcn, n_cells=nc)
    # faces with lc < 0 are flipped by the reader (UnsGrid.cpp:160-170)
    flip = np.nonzero(g.lc < 0)[0]
    if flip.size:
        offs = np.concatenate([[0], np.cumsum(fnc)])
        for f in flip:
            g.face_nodes[offs[f]:offs[f + 1]] = g.face_nodes[offs[f]:offs[f + 1]][::-1]
            g.lc[f], g.rc[f] = g.rc[f], g.lc[f]
    g.check()
    return g


def write_ofl(grid: Grid, path: str | Path) -> None:
    """Write a single-zone .ofl grid readable by the reference (UnsGrid::WriteGrid)."""
    grid.check()
    body = bytearray()
    body += struct.pack("<3i", grid.n_nodes, grid.n_faces, grid.n_cells)
    for a in (grid.x, grid.y, grid.z):
        body += np.ascontiguousarray(a, dtype="<f8").tobytes()
    body += struct.pack("<i", -1)
    body += np.ascontiguousarray(grid.face_node_count, dtype="<i4").tobytes()
    body += np.ascontiguousarray(grid.face_nodes, dtype="<i4").tobytes()
    body += np.ascontiguousarray(grid.lc, dtype="<i4").tobytes()
    body += np.ascontiguousarray(grid.rc, dtype="<i4").tobytes()
    body += struct.pack("<i", grid.n_bfaces)
    body += np.ascontiguousarray(grid.bc_type, dtype="<i4").tobytes()
    names = grid.bc_name_id if grid.bc_name_id is not None else np.zeros(grid.n_bfaces, np.int32)
    body += np.ascontiguousarray(names, dtype="<i4").tobytes()
    body += struct.pack("<i", 0)
    head = struct.pack("<iiiq", 1, 0, 0, len(body))
    Path(path).write_bytes(head + bytes(body))
