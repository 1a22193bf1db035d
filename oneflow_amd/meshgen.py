"""Deterministic synthetic grids for the benchmark and parity workloads.

Every generator returns a `Grid` in the reference's single-zone layout
(boundary faces first, face nodes ordered so that the 2-D face normal
(y2-y1, x1-x2) points from the left cell to the right cell, as
FaceMesh::CalcFaceNormal2D computes it, codes/geometry/src/FaceMesh.cpp:112-136).

* flat_plate      -- laminar flat plate, quads, geometric wall stretching
* naca_hybrid     -- NACA 4-digit airfoil O-grid: quad wall layers plus
                     triangles out to a far-field circle

Cell numbering matters for LU-SGS (it defines the lower/upper split).  The
generators number cells in their natural structured order; `renumber()`
applies any permutation (e.g. `multicolor_order`) to a grid, and both the GPU
library and the CPU oracle see the same numbering.
"""
from __future__ import annotations

import math

import numpy as np

from .mesh import BC_FARFIELD, BC_OUTFLOW, BC_SOLID_SURFACE, BC_SYMMETRY, Grid


# ---------------------------------------------------------------- assembly

def grid_from_cells(x, y, cells, bc_of_edge, dim=2):
    """Assemble a 2-D face list from polygon cells given CCW.

    cells: list/array of node index rows (-1 padded) in counter-clockwise order.
    bc_of_edge(a, b) -> bc type for a boundary edge (a, b) in the owner's CCW order.
    """
    cells = np.asarray(cells)
    nc = cells.shape[0]
    nvert = (cells >= 0).sum(axis=1)
    # all cell edges in CCW order
    e_cell, e_a, e_b = [], [], []
    for k in range(cells.shape[1]):
        k1 = (np.arange(nc), )
        a = cells[:, k]
        nxt = np.where(k + 1 < nvert, (k + 1), 0)
        b = cells[np.arange(nc), nxt]
        valid = k < nvert
        e_cell.append(np.nonzero(valid)[0])
        e_a.append(a[valid])
        e_b.append(b[valid])
    e_cell = np.concatenate(e_cell)
    e_a = np.concatenate(e_a)
    e_b = np.concatenate(e_b)
    key_lo = np.minimum(e_a, e_b).astype(np.int64)
    key_hi = np.maximum(e_a, e_b).astype(np.int64)
    key = key_lo * (len(x) + 1) + key_hi
    order = np.argsort(key, kind="stable")
    key_s = key[order]
    # group edges
    starts = np.concatenate([[0], np.nonzero(np.diff(key_s))[0] + 1])
    counts = np.diff(np.concatenate([starts, [len(key_s)]]))
    if np.any(counts > 2):
        raise ValueError("non-manifold edge")
    bnd_idx = order[starts[counts == 1]]
    pair0 = order[starts[counts == 2]]
    pair1 = order[starts[counts == 2] + 1]
    # interior: lc = lower cell id, nodes in lc's CCW order
    c0, c1 = e_cell[pair0], e_cell[pair1]
    swap = c1 < c0
    lc_i = np.where(swap, c1, c0)
    rc_i = np.where(swap, c0, c1)
    a_i = np.where(swap, e_a[pair1], e_a[pair0])
    b_i = np.where(swap, e_b[pair1], e_b[pair0])
    srt = np.lexsort((np.maximum(lc_i, rc_i), lc_i))
    lc_i, rc_i, a_i, b_i = lc_i[srt], rc_i[srt], a_i[srt], b_i[srt]
    # boundary faces: owner CCW order; group by bc type then owner
    lc_b = e_cell[bnd_idx]
    a_b = e_a[bnd_idx]
    b_b = e_b[bnd_idx]
    bct = np.array([bc_of_edge(int(a), int(b)) for a, b in zip(a_b, b_b)], dtype=np.int32)
    srt = np.lexsort((lc_b, bct))
    lc_b, a_b, b_b, bct = lc_b[srt], a_b[srt], b_b[srt], bct[srt]
    nb = len(lc_b)
    lc = np.concatenate([lc_b, lc_i]).astype(np.int32)
    rc = np.concatenate([-np.ones(nb, np.int64), rc_i]).astype(np.int32)
    fn = np.stack([np.concatenate([a_b, a_i]), np.concatenate([b_b, b_i])], axis=1).astype(np.int32).ravel()
    nf = len(lc)
    g = Grid(dim=dim, x=np.asarray(x, np.float64), y=np.asarray(y, np.float64), z=np.zeros(len(x)),
             face_node_count=np.full(nf, 2, np.int32), face_nodes=fn, lc=lc, rc=rc, bc_type=bct,
             bc_name_id=bct.copy(), n_cells=nc)
    g.check()
    return g


def cell_areas(x, y, cells):
    cells = np.asarray(cells)
    nv = (cells >= 0).sum(axis=1)
    a = np.zeros(cells.shape[0])
    for k in range(cells.shape[1]):
        i = cells[:, k]
        j = cells[np.arange(len(cells)), np.where(k + 1 < nv, k + 1, 0)]
        valid = k < nv
        a += np.where(valid, x[i] * y[j] - x[j] * y[i], 0.0)
    return 0.5 * a


# ---------------------------------------------------------------- numbering

def cell_adjacency(g: Grid):
    nb = g.n_bfaces
    return g.lc[nb:], g.rc[nb:]


def multicolor_order(g: Grid):
    """Greedy colouring of the cell graph; cells sorted by (colour, old id).

    With this numbering every LU-SGS sweep has as many dependency levels as
    colours (a cell's lower neighbours all carry smaller colours)."""
    a, b = cell_adjacency(g)
    n = g.n_cells
    nbr = [[] for _ in range(n)]
    for u, v in zip(a.tolist(), b.tolist()):
        nbr[u].append(v)
        nbr[v].append(u)
    color = np.full(n, -1, np.int64)
    for i in range(n):
        used = {color[j] for j in nbr[i] if color[j] >= 0}
        c = 0
        while c in used:
            c += 1
        color[i] = c
    return np.lexsort((np.arange(n), color))


def _rounds_colouring(n, a, b, seed=1):
    """Colouring of the graph (n vertices, edges a-b) in rounds: each round
    colours the uncoloured vertices whose random priority beats every
    uncoloured neighbour's, with the lowest colour no coloured neighbour has
    (Jones-Plassmann).  Deterministic for a seed; vectorised."""
    w = np.random.default_rng(seed).permutation(n)
    col = np.full(n, -1, np.int64)
    used = np.zeros(n, np.int64)  # bit c: some coloured neighbour has colour c
    u = np.concatenate([a, b]).astype(np.int64)
    v = np.concatenate([b, a]).astype(np.int64)
    while True:
        open_ = col < 0
        if not open_.any():
            return col
        e = open_[u] & open_[v]
        beaten = np.zeros(n, bool)
        np.logical_or.at(beaten, u[e], w[v[e]] > w[u[e]])
        pick = np.flatnonzero(open_ & ~beaten)
        free = ~used[pick]
        low = free & -free
        col[pick] = np.round(np.log2(low.astype(np.float64))).astype(np.int64)
        chosen = np.zeros(n, bool)
        chosen[pick] = True
        e2 = chosen[u]
        np.bitwise_or.at(used, v[e2], np.left_shift(np.int64(1), col[u[e2]]))


def cell_centroids(g: Grid):
    """Approximate cell centres (mean of the cell's face centres, each face
    centre the mean of its nodes): [n_cells][3].  For ordering, not metrics."""
    cnt = g.face_node_count.astype(np.int64)
    fid = np.repeat(np.arange(g.n_faces), cnt)
    fn = g.face_nodes.astype(np.int64)
    zs = g.z if g.dim == 3 else np.zeros_like(g.x)
    fc = [np.bincount(fid, c[fn], g.n_faces) / cnt for c in (g.x, g.y, zs)]
    nb = g.n_bfaces
    owner = np.concatenate([g.lc, g.rc[nb:]]).astype(np.int64)
    face = np.concatenate([np.arange(g.n_faces), np.arange(nb, g.n_faces)])
    k = np.bincount(owner, minlength=g.n_cells)
    return np.stack([np.bincount(owner, c[face], g.n_cells) / k for c in fc], 1)


def _rcb(xyz, idx, nparts, out, base):
    """Recursive coordinate bisection of cells idx into nparts parts of equal
    size (split across the longest extent)."""
    if nparts == 1:
        out[idx] = base
        return
    ax = int(np.argmax(xyz[idx].max(0) - xyz[idx].min(0)))
    h = nparts // 2
    k = len(idx) * h // nparts
    srt = idx[np.argsort(xyz[idx, ax], kind="stable")]
    _rcb(xyz, srt[:k], h, out, base)
    _rcb(xyz, srt[k:], nparts - h, out, base + h)


def blocked_multicolor_order(g: Grid, part_cells: int = 20000, seed: int = 1):
    """A colouring numbering that keeps neighbours close in memory.

    The cells are cut into spatial parts of about part_cells cells (recursive
    coordinate bisection).  The parts are coloured so that no two touching
    parts share a colour, and the cells of each part are coloured along the
    part's inner faces.  Cells are sorted by (part colour, part, cell colour,
    old id).  A cell's lower neighbours then lie in its own part with a
    smaller cell colour or in a part of smaller part colour, so the LU-SGS
    dependency levels stay few (14 on the 2.5M-cell swept wing, against 6 for
    a global colouring), while every kernel walks one part -- a few MB of
    state -- at a time instead of sweeping the whole grid once per colour."""
    n = g.n_cells
    a, b = (np.asarray(t, np.int64) for t in cell_adjacency(g))
    nparts = max(1, -(-n // part_cells))
    part = np.empty(n, np.int64)
    _rcb(cell_centroids(g), np.arange(n), nparts, part, 0)
    cut = part[a] != part[b]
    pe = np.unique(np.stack([part[a][cut], part[b][cut]], 1), axis=0)
    pcol = _rounds_colouring(nparts, pe[:, 0], pe[:, 1], seed)
    ccol = _rounds_colouring(n, a[~cut], b[~cut], seed)
    return np.lexsort((np.arange(n), ccol, part, pcol[part]))


def renumber(g: Grid, order) -> Grid:
    """Return a copy with cells renumbered: new cell k is old cell order[k]."""
    order = np.asarray(order)
    new_of_old = np.empty(g.n_cells, np.int64)
    new_of_old[order] = np.arange(g.n_cells)
    nb = g.n_bfaces
    lc = new_of_old[g.lc]
    rc = g.rc.astype(np.int64).copy()
    rc[nb:] = new_of_old[g.rc[nb:]]
    if g.dim == 3:
        return _renumber_faces_nd(g, lc, rc)
    fn = g.face_nodes.reshape(-1, 2).copy()
    # interior faces: keep lc < rc (flip node order when swapping) and sort
    li, ri = lc[nb:].copy(), rc[nb:].copy()
    sw = li > ri
    li[sw], ri[sw] = ri[sw].copy(), li[sw].copy()
    fi = fn[nb:].copy()
    fi[sw] = fi[sw][:, ::-1]
    srt = np.lexsort((ri, li))
    fb = fn[:nb]
    lb = lc[:nb]
    srtb = np.lexsort((lb, g.bc_type))
    out = Grid(dim=g.dim, x=g.x, y=g.y, z=g.z, face_node_count=g.face_node_count.copy(),
               face_nodes=np.concatenate([fb[srtb], fi[srt]]).astype(np.int32).ravel(),
               lc=np.concatenate([lb[srtb], li[srt]]).astype(np.int32),
               rc=np.concatenate([-np.ones(nb, np.int64), ri[srt]]).astype(np.int32),
               bc_type=g.bc_type[srtb].copy(), bc_name_id=g.bc_name_id[srtb].copy(), n_cells=g.n_cells)
    out.check()
    return out


def _renumber_faces_nd(g: Grid, lc, rc) -> Grid:
    """Face re-sort after a cell renumbering for grids with any face node
    counts: interior faces keep lc < rc (node order reversed when the sides
    swap, which flips the normal), boundary faces grouped by type then owner,
    interior faces sorted by (lc, rc)."""
    nb, nf = g.n_bfaces, g.n_faces
    cnt = g.face_node_count.astype(np.int64)
    off = np.concatenate([[0], np.cumsum(cnt)])
    width = int(cnt.max())
    pad = np.full((nf, width), -1, np.int64)
    col = np.arange(width)[None, :]
    valid = col < cnt[:, None]
    pad[valid] = g.face_nodes
    li, ri = lc[nb:].copy(), rc[nb:].copy()
    sw = li > ri
    li[sw], ri[sw] = ri[sw].copy(), li[sw].copy()
    fi = pad[nb:].copy()
    if np.any(sw):
        rows = fi[sw]
        c = cnt[nb:][sw]
        rev = np.full_like(rows, -1)
        for k in np.unique(c):
            m = c == k
            rev[m, :k] = rows[m, :k][:, ::-1]
        fi[sw] = rev
    srt = np.lexsort((ri, li))
    lb = lc[:nb]
    srtb = np.lexsort((lb, g.bc_type))
    nodes = np.concatenate([pad[:nb][srtb], fi[srt]])
    fc = np.concatenate([cnt[:nb][srtb], cnt[nb:][srt]]).astype(np.int32)
    out = Grid(dim=g.dim, x=g.x, y=g.y, z=g.z, face_node_count=fc, face_nodes=nodes[nodes >= 0].astype(np.int32),
               lc=np.concatenate([lb[srtb], li[srt]]).astype(np.int32),
               rc=np.concatenate([-np.ones(nb, np.int64), ri[srt]]).astype(np.int32),
               bc_type=g.bc_type[srtb].copy(), bc_name_id=g.bc_name_id[srtb].copy(), n_cells=g.n_cells)
    out.check()
    return out


def sort_interior_faces(g: Grid, cell_key) -> Grid:
    """Return a copy whose interior faces are sorted by the (smaller, larger)
    cell_key of their two cells; boundary faces, every face's cells and node
    order (so its normal) are unchanged.  With cell_key a spatial order of the
    cells (for a colouring numbering, the cells' ids before it), the face
    kernels walk the grid once in space instead of once per colour."""
    nb = g.n_bfaces
    key = np.asarray(cell_key, np.int64)
    a, b = key[g.lc[nb:]], key[g.rc[nb:]]
    srt = np.lexsort((np.maximum(a, b), np.minimum(a, b)))
    perm = np.concatenate([np.arange(nb), nb + srt])
    cnt = g.face_node_count.astype(np.int64)
    off = np.concatenate([[0], np.cumsum(cnt)])
    lens = cnt[perm]
    first = np.concatenate([[0], np.cumsum(lens)[:-1]])
    idx = np.repeat(off[perm] - first, lens) + np.arange(int(lens.sum()))
    out = Grid(dim=g.dim, x=g.x, y=g.y, z=g.z, face_node_count=g.face_node_count[perm].copy(),
               face_nodes=g.face_nodes[idx].astype(np.int32), lc=g.lc[perm].copy(), rc=g.rc[perm].copy(),
               bc_type=g.bc_type.copy(), bc_name_id=g.bc_name_id.copy(), n_cells=g.n_cells)
    out.check()
    return out


# ---------------------------------------------------------------- flat plate

def geometric(n, h0, ratio):
    """n+1 coordinates 0, h0, h0(1+r), ... (geometric spacing)."""
    d = h0 * ratio ** np.arange(n)
    return np.concatenate([[0.0], np.cumsum(d)])


def flat_plate(ni: int = 161, nj: int = 97, x0: float = -0.5, x1: float = 1.0, height: float = 0.5,
               ratio: float = 1.1) -> Grid:
    """Laminar flat plate on [x0,x1]x[0,height]: ni x nj nodes, wall-normal spacing
    growing by `ratio` from the wall (y=0), x clustered at the leading edge x=0.
    BCs: y=0, x<0 symmetry; y=0, x>=0 adiabatic wall; x=x0 and y=height far field;
    x=x1 outflow (the region layout of test/plateuns2dslau2/grid/bcRegionMap.txt)."""
    nj_c = nj - 1
    ys = geometric(nj_c, 1.0, ratio)
    ys = ys / ys[-1] * height
    # x: cells upstream of the leading edge and along the plate, both clustered at x = 0
    n_up = max(2, (ni - 1) // 5)
    n_pl = ni - 1 - n_up
    xp = geometric(n_pl, 1.0, 1.03)
    xp = xp / xp[-1] * x1
    xu = geometric(n_up, 1.0, 1.08)
    xu = -(xu / xu[-1] * (-x0))[::-1]
    xs = np.concatenate([xu[:-1], xp])
    assert len(xs) == ni
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    x, y = X.ravel(), Y.ravel()
    nid = lambda i, j: i * nj + j
    I, J = np.meshgrid(np.arange(ni - 1), np.arange(nj - 1), indexing="ij")
    I, J = I.ravel(), J.ravel()
    cells = np.stack([nid(I, J), nid(I + 1, J), nid(I + 1, J + 1), nid(I, J + 1)], axis=1)

    def bc(a, b):
        xa, ya, xb, yb = x[a], y[a], x[b], y[b]
        if ya == 0.0 and yb == 0.0:
            return BC_SOLID_SURFACE if min(xa, xb) >= 0.0 else BC_SYMMETRY
        if xa == x1 and xb == x1:
            return BC_OUTFLOW
        return BC_FARFIELD

    return grid_from_cells(x, y, cells, bc)


# ---------------------------------------------------------------- airfoil

def naca4(m: float, p: float, t: float, n: int, closed_te: bool = True):
    """NACA 4-digit surface, n points, clockwise from the trailing edge along
    the lower surface to the leading edge and back along the upper surface
    (the trailing-edge point appears once).  m, p, t as fractions of chord."""
    a4 = -0.1036 if closed_te else -0.1015
    npts = n // 2 + 1
    beta = np.linspace(0.0, math.pi, npts)
    xc = 0.5 * (1.0 - np.cos(beta))  # 0 .. 1, cosine clustered
    yt = 5.0 * t * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc ** 2 + 0.2843 * xc ** 3 + a4 * xc ** 4)
    if m > 0.0 and p > 0.0:
        yc = np.where(xc < p, m / p ** 2 * (2 * p * xc - xc ** 2), m / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * xc - xc ** 2))
        dyc = np.where(xc < p, 2 * m / p ** 2 * (p - xc), 2 * m / (1 - p) ** 2 * (p - xc))
    else:
        yc = np.zeros_like(xc)
        dyc = np.zeros_like(xc)
    th = np.arctan(dyc)
    xu, yu = xc - yt * np.sin(th), yc + yt * np.cos(th)
    xl, yl = xc + yt * np.sin(th), yc - yt * np.cos(th)
    # lower surface from TE to LE, then upper from LE to TE (excluding duplicates)
    xs = np.concatenate([xl[::-1], xu[1:-1]])
    ys = np.concatenate([yl[::-1], yu[1:-1]])
    if closed_te:
        ys[0] = 0.5 * (yl[-1] + yu[-1])
        xs[0] = 1.0
    return xs, ys


def _smooth_normals(xs, ys, passes=20):
    n = len(xs)
    tx = np.roll(xs, -1) - np.roll(xs, 1)
    ty = np.roll(ys, -1) - np.roll(ys, 1)
    nx, ny = -ty, tx          # surface traversed clockwise -> outward normal
    L = np.hypot(nx, ny)
    nx, ny = nx / L, ny / L
    for _ in range(passes):
        nx = 0.5 * nx + 0.25 * (np.roll(nx, 1) + np.roll(nx, -1))
        ny = 0.5 * ny + 0.25 * (np.roll(ny, 1) + np.roll(ny, -1))
        L = np.hypot(nx, ny)
        nx, ny = nx / L, ny / L
    return nx, ny


def naca_hybrid(n_surf: int = 512, n_quad: int = 40, ratio: float = 1.15, reynolds: float = 6.5e6,
                yplus: float = 1.0, n_outer: int = 98, radius: float = 30.0,
                m: float = 0.0, p: float = 0.4, t: float = 0.12) -> Grid:
    """Airfoil O-grid: n_quad quad layers (first height for y+ ~ yplus at
    `reynolds`, growth `ratio`), then n_outer layers of quads each split into
    two triangles, reaching a far-field circle of `radius` chords.
    Cells: n_surf * (n_quad + 2 n_outer)."""
    xs, ys = naca4(m, p, t, n_surf)
    ns = len(xs)
    # first cell height for y+ (flat-plate skin friction estimate)
    cf = 0.026 / reynolds ** (1.0 / 7.0)
    h0 = yplus / (reynolds * math.sqrt(0.5 * cf))
    d_in = geometric(n_quad, h0, ratio)
    nx, ny = _smooth_normals(xs, ys)
    # outer layers: distance grows geometrically from the last wall layer to `radius`
    d_last = d_in[-1]
    step0 = d_in[-1] - d_in[-2]
    # solve growth g: step0 * (g^n_outer - 1)/(g-1) = radius - d_last (approximately, by bisection)
    lo, hi = 1.0001, 2.0
    while step0 * (hi ** n_outer - 1.0) / (hi - 1.0) <= radius:  # few outer layers: widen the bracket
        hi *= 2.0
    for _ in range(100):
        g = 0.5 * (lo + hi)
        s = step0 * (g ** n_outer - 1.0) / (g - 1.0)
        if s > radius:
            hi = g
        else:
            lo = g
    d_out = d_last + step0 * (g ** np.arange(1, n_outer + 1) - 1.0) / (g - 1.0)
    dists = np.concatenate([d_in, d_out])
    nl = len(dists)  # layers of nodes (including the surface)
    # march outward: each layer moves along the smoothed outward normal of the
    # previous layer curve; smoothing grows with distance so the layers relax
    # from body-fitted to circular
    X = np.zeros((nl, ns))
    Y = np.zeros((nl, ns))
    X[0], Y[0] = xs, ys
    # a smoothing pass reaches one surface spacing: scale the passes with the
    # square of the point density so the physical smoothing length is the same
    dens = max(1.0, (ns / 512.0) ** 2)
    def layer_ok(xa, ya, xb, yb):
        # both triangles (a, b, c), (a, c, d) of every cell between two layers positive
        bx, by = np.roll(xa, -1), np.roll(ya, -1)
        cx, cy = np.roll(xb, -1), np.roll(yb, -1)
        t1 = (bx - xa) * (cy - ya) - (cx - xa) * (by - ya)
        t2 = (cx - xa) * (yb - ya) - (xb - xa) * (cy - ya)
        return bool(np.all(t1 > 0) and np.all(t2 > 0))

    for j in range(1, nl):
        passes = int(dens * min(400, 4 + 2 * j + (40 * max(0.0, dists[j] - d_last) / radius) * 10))
        step = dists[j] - dists[j - 1]
        while True:
            # more smoothing where a concave stretch would fold the next layer
            nx, ny = _smooth_normals(X[j - 1], Y[j - 1], passes=passes)
            X[j] = X[j - 1] + step * nx
            Y[j] = Y[j - 1] + step * ny
            if layer_ok(X[j - 1], Y[j - 1], X[j], Y[j]) or passes > 20000:
                break
            passes *= 2
    x = X.ravel()
    y = Y.ravel()
    nid = lambda j, i: j * ns + (i % ns)
    cells = []
    for j in range(nl - 1):
        i = np.arange(ns)
        a, b, c, d = nid(j, i), nid(j, i + 1), nid(j + 1, i + 1), nid(j + 1, i)
        # surface is clockwise, layers go outward: (a, b, c, d) is counter-clockwise
        if j < n_quad:
            cells.append(np.stack([a, b, c, d], axis=1))
        else:
            t1 = np.stack([a, b, c, -np.ones_like(a)], axis=1)
            t2 = np.stack([a, c, d, -np.ones_like(a)], axis=1)
            cells.append(np.stack([t1, t2], axis=1).reshape(-1, 4))
    cells = np.concatenate(cells)
    ar = cell_areas(x, y, cells)
    if np.any(ar <= 0):
        raise ValueError(f"naca_hybrid: {int((ar <= 0).sum())} cells with non-positive area")
    jmax = nl - 1

    def bc(a, b):
        ja, jb = a // ns, b // ns
        if ja == 0 and jb == 0:
            return BC_SOLID_SURFACE
        if ja == jmax and jb == jmax:
            return BC_FARFIELD
        raise ValueError("unexpected boundary edge")

    return grid_from_cells(x, y, cells, bc)


# ---------------------------------------------------------------- 2-D box

def box2d(nx: int, ny: int, x0=0.0, x1=1.0, y0=0.0, y1=1.0, tri=False, bc=None, jitter=0.0, seed=0):
    """Rectangle of nx x ny quads (or 2 triangles each when tri), optional
    node jitter (interior nodes, fraction of the spacing, seeded).
    bc(xm, ym) -> bc type for a boundary edge with midpoint (xm, ym);
    default: far field everywhere."""
    xs, ys = np.linspace(x0, x1, nx + 1), np.linspace(y0, y1, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    if jitter > 0:
        rng = np.random.default_rng(seed)
        hx, hy = (x1 - x0) / nx, (y1 - y0) / ny
        inner = np.zeros_like(X, bool)
        inner[1:-1, 1:-1] = True
        X = X + np.where(inner, rng.uniform(-jitter, jitter, X.shape) * hx, 0.0)
        Y = Y + np.where(inner, rng.uniform(-jitter, jitter, Y.shape) * hy, 0.0)
    x, y = X.ravel(), Y.ravel()
    nid = lambda i, j: i * (ny + 1) + j
    cells = []
    for i in range(nx):
        for j in range(ny):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            if tri:
                cells.append([a, b, c, -1])
                cells.append([a, c, d, -1])
            else:
                cells.append([a, b, c, d])
    bcf = bc if bc is not None else (lambda xm, ym: BC_FARFIELD)
    return grid_from_cells(x, y, np.array(cells), lambda a, b: bcf(0.5 * (x[a] + x[b]), 0.5 * (y[a] + y[b])))


# ---------------------------------------------------------------- 3-D assembly

# Outward-oriented local faces (right-hand rule gives the outward normal) of
# positively oriented cells.  Hex: bottom 0-3 counter-clockwise seen from +z,
# top 4-7 above them.  Prism: bottom triangle 0,1,2 counter-clockwise, top 3,4,5.
HEX_FACES = ((0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (3, 7, 6, 2), (0, 4, 7, 3), (1, 2, 6, 5))
PRISM_FACES = ((0, 2, 1), (3, 4, 5), (0, 1, 4, 3), (1, 2, 5, 4), (2, 0, 3, 5))
TET_FACES = ((0, 2, 1), (0, 1, 3), (0, 3, 2), (1, 2, 3))


def grid_from_polyhedra(x, y, z, cells, bc_of_face):
    """Assemble a 3-D face list from cells given as node tuples of length 4
    (tet), 6 (prism) or 8 (hex), each positively oriented.  Interior faces get
    lc < rc with nodes ordered so the normal points from lc to rc; boundary
    faces point out of their owner.  bc_of_face(xc, yc, zc, n) -> bc type
    (face centroid and outward unit normal)."""
    x, y, z = (np.asarray(v, np.float64) for v in (x, y, z))
    table = {4: TET_FACES, 6: PRISM_FACES, 8: HEX_FACES}
    faces = {}
    for ci, cell in enumerate(cells):
        for loc in table[len(cell)]:
            nodes = tuple(int(cell[k]) for k in loc)
            key = tuple(sorted(nodes))
            if key in faces:
                owner, onodes, other = faces[key]
                if other is not None:
                    raise ValueError("non-manifold face")
                faces[key] = (owner, onodes, ci)
            else:
                faces[key] = (ci, nodes, None)
    bnd, inner = [], []
    for key, (owner, nodes, other) in faces.items():
        if other is None:
            p = np.array([[x[n], y[n], z[n]] for n in nodes])
            c = p.mean(axis=0)
            nrm = np.zeros(3)
            for i in range(len(nodes)):
                nrm += np.cross(p[i], p[(i + 1) % len(nodes)])
            nrm /= np.linalg.norm(nrm)
            bnd.append((bc_of_face(c[0], c[1], c[2], nrm), owner, nodes))
        else:
            # nodes are in `owner`'s outward order; lc must be the smaller id
            if owner < other:
                inner.append((owner, other, nodes))
            else:
                inner.append((other, owner, tuple(reversed(nodes))))
    bnd.sort(key=lambda t: (t[0], t[1]))
    inner.sort(key=lambda t: (t[0], t[1]))
    nb = len(bnd)
    fnc = [len(t[2]) for t in bnd] + [len(t[2]) for t in inner]
    fn = [n for t in bnd for n in t[2]] + [n for t in inner for n in t[2]]
    lc = [t[1] for t in bnd] + [t[0] for t in inner]
    rc = [-1] * nb + [t[1] for t in inner]
    bct = np.array([t[0] for t in bnd], np.int32)
    g = Grid(dim=3, x=x, y=y, z=z, face_node_count=np.array(fnc, np.int32), face_nodes=np.array(fn, np.int32),
             lc=np.array(lc, np.int32), rc=np.array(rc, np.int32), bc_type=bct, bc_name_id=bct.copy(),
             n_cells=len(cells))
    g.check()
    return g


def _tet_volume(p, cell):
    a, b, c, d = (p[k] for k in cell)
    return float(np.dot(np.cross(b - a, c - a), d - a)) / 6.0


def hex_to_tets(h, p):
    """Kuhn split of a hex (HEX node order) into 6 positively oriented tets;
    conforming between neighbouring hexes split the same way."""
    # corner index by (i, j, k) bits
    corner = {(0, 0, 0): h[0], (1, 0, 0): h[1], (1, 1, 0): h[2], (0, 1, 0): h[3],
              (0, 0, 1): h[4], (1, 0, 1): h[5], (1, 1, 1): h[6], (0, 1, 1): h[7]}
    tets = []
    for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        v = [0, 0, 0]
        path = [corner[tuple(v)]]
        for ax in perm:
            v[ax] = 1
            path.append(corner[tuple(v)])
        if _tet_volume(p, path) < 0:
            path[1], path[2] = path[2], path[1]
        tets.append(tuple(path))
    return tets


def hex_to_prisms(h):
    """Two prisms per hex, bottom quad split along its 0-2 diagonal."""
    return [(h[0], h[1], h[2], h[4], h[5], h[6]), (h[0], h[2], h[3], h[4], h[6], h[7])]


def box3d(nx: int, ny: int, nz: int, kind: str = "hex", extent=(1.0, 1.0, 1.0), zs=None, prism_layers=0,
          bc=None, jitter=0.0, seed=0):
    """Box of nx x ny x nz hexes, kept as hexes ("hex"), split into 6 tets
    ("tet"), 2 prisms ("prism"), or prisms in the lowest `prism_layers`
    layers and tets above ("mixed").  zs: optional node z levels.
    bc(xc, yc, zc, n) -> bc type (default far field)."""
    xs = np.linspace(0.0, extent[0], nx + 1)
    ys = np.linspace(0.0, extent[1], ny + 1)
    zz = np.asarray(zs, np.float64) if zs is not None else np.linspace(0.0, extent[2], nz + 1)
    X, Y, Z = np.meshgrid(xs, ys, zz, indexing="ij")
    if jitter > 0:
        rng = np.random.default_rng(seed)
        inner = np.zeros_like(X, bool)
        inner[1:-1, 1:-1, 1:-1] = True
        h = min(extent[0] / nx, extent[1] / ny, float(np.min(np.diff(zz))))
        for A in (X, Y, Z):
            A += np.where(inner, rng.uniform(-jitter, jitter, A.shape) * h, 0.0)
    p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    nid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k
    cells = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                h = (nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                     nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1))
                if kind == "hex":
                    cells.append(h)
                elif kind == "prism" or (kind == "mixed" and k < prism_layers):
                    cells.extend(hex_to_prisms(h))
                else:
                    cells.extend(hex_to_tets(h, p))
    bcf = bc if bc is not None else (lambda xc, yc, zc, n: BC_FARFIELD)
    return grid_from_polyhedra(p[:, 0], p[:, 1], p[:, 2], cells, bcf)


# ---------------------------------------------------------------- 3-D, vectorised

def _rows_key(a):
    """Sort each row (node tuples padded with -1) so equal faces compare equal."""
    s = np.sort(a, axis=1)
    return s


def grid_from_polyhedra_np(x, y, z, tets, prisms, bc_of_faces):
    """Vectorised 3-D assembly for large grids.  tets [n,4] and prisms [n,6]
    (positively oriented, node orders as TET_FACES / PRISM_FACES); cells are
    numbered tets first, then prisms.  bc_of_faces(nodes[k, 4]) -> bc type per
    boundary face (nodes padded with -1).  Same conventions as
    grid_from_polyhedra: interior faces lc < rc with the normal from lc to rc,
    boundary faces first, grouped by bc type then owner."""
    x, y, z = (np.asarray(v, np.float64) for v in (x, y, z))
    tets = np.asarray(tets, np.int64).reshape(-1, 4)
    prisms = np.asarray(prisms, np.int64).reshape(-1, 6)
    nt, npr = len(tets), len(prisms)
    fl, owner = [], []
    for loc in TET_FACES:
        f = np.full((nt, 4), -1, np.int64)
        f[:, :3] = tets[:, list(loc)]
        fl.append(f)
        owner.append(np.arange(nt))
    for loc in PRISM_FACES:
        f = np.full((npr, 4), -1, np.int64)
        f[:, :len(loc)] = prisms[:, list(loc)]
        fl.append(f)
        owner.append(nt + np.arange(npr))
    faces = np.concatenate(fl)
    own = np.concatenate(owner)
    key = _rows_key(faces)
    order = np.lexsort((key[:, 3], key[:, 2], key[:, 1], key[:, 0]))
    ks = key[order]
    same = np.all(ks[1:] == ks[:-1], axis=1)
    start = np.concatenate([[True], ~same])
    grp = np.cumsum(start) - 1
    cnt = np.bincount(grp)
    if np.any(cnt > 2):
        raise ValueError("non-manifold face")
    first = order[start]
    single = cnt == 1
    # boundary faces
    bidx = first[single]
    bnodes = faces[bidx]
    bown = own[bidx]
    bct = np.asarray(bc_of_faces(bnodes), np.int32)
    srt = np.lexsort((bown, bct))
    bnodes, bown, bct = bnodes[srt], bown[srt], bct[srt]
    # interior pairs
    pos = np.nonzero(~single)[0]
    starts = np.nonzero(start)[0][pos]
    a, b = order[starts], order[starts + 1]
    oa, ob = own[a], own[b]
    lcell = np.where(oa < ob, oa, ob)
    rcell = np.where(oa < ob, ob, oa)
    inodes = np.where((oa < ob)[:, None], faces[a], faces[b])
    srt = np.lexsort((rcell, lcell))
    lcell, rcell, inodes = lcell[srt], rcell[srt], inodes[srt]
    allnodes = np.concatenate([bnodes, inodes])
    fnc = (allnodes >= 0).sum(axis=1).astype(np.int32)
    fn = allnodes[allnodes >= 0].astype(np.int32)
    nb = len(bnodes)
    g = Grid(dim=3, x=x, y=y, z=z, face_node_count=fnc, face_nodes=fn,
             lc=np.concatenate([bown, lcell]).astype(np.int32),
             rc=np.concatenate([-np.ones(nb, np.int64), rcell]).astype(np.int32),
             bc_type=bct, bc_name_id=bct.copy(), n_cells=nt + npr)
    g.check()
    return g


def swept_wing(n_surf: int = 256, n_quad: int = 25, ratio: float = 1.2, reynolds: float = 1.2e7,
               n_outer: int = 40, radius: float = 20.0, n_span: int = 34, sweep_deg: float = 30.0,
               taper: float = 0.5, aspect: float = 4.0, t: float = 0.12):
    """Tapered swept wing (NACA 4-digit section, root chord 1) between two
    symmetry planes: the hybrid airfoil section grid of `naca_hybrid` extruded
    over the semi-span b = aspect (1 + taper) / 4 in n_span layers; each section
    is scaled by the local chord and shifted by the swept leading edge.  The
    n_quad wall layers become prism layers (2 prisms per extruded quad, their
    triangles on the wall-parallel rings), the outer triangles become prisms
    split into 3 tets.  Cells: n_span * n_surf * (2 n_quad + 6 n_outer).
    No wing tip: the span ends on the second symmetry plane."""
    sec = naca_hybrid(n_surf=n_surf, n_quad=n_quad, ratio=ratio, reynolds=reynolds, n_outer=n_outer,
                      radius=radius, m=0.0, p=0.4, t=t)
    ns = n_surf + 0  # nodes per layer ring
    xs2, ys2 = sec.x, sec.y
    n2 = len(xs2)
    ns = n2 // (n_quad + n_outer + 1)
    semi = aspect * (1.0 + taper) / 4.0
    zk = np.linspace(0.0, semi, n_span + 1)
    chord = 1.0 - (1.0 - taper) * zk / semi
    xle = zk * math.tan(math.radians(sweep_deg))
    X = (xle[:, None] + chord[:, None] * xs2[None, :]).ravel()
    Y = (chord[:, None] * ys2[None, :]).ravel()
    Z = np.repeat(zk, n2)
    nid = lambda k, v: k * n2 + v
    # 2-D cells of the section: quads (j < n_quad) and triangles, in naca_hybrid's order
    jj = np.arange(n_quad + n_outer)
    i = np.arange(ns)
    lay = lambda j, ii: j * ns + (ii % ns)
    prisms, tets = [], []
    P = np.stack([X, Y, Z], axis=1)
    for k in range(n_span):
        for j in range(n_quad + n_outer):
            a, b, c, d = lay(j, i), lay(j, i + 1), lay(j + 1, i + 1), lay(j + 1, i)
            if j < n_quad:
                # wall layer: the hex (a, b on ring j; d, c above them on ring j+1)
                # -> 2 prisms whose triangles lie on the rings, split along the
                # diagonal u_k - v_{k+1} of each ring quad, u < v by 2-D node id
                # (the same rule the tets use, so the rings conform)
                sw = b < a
                u, v = np.where(sw, b, a), np.where(sw, a, b)
                uo, vo = np.where(sw, c, d), np.where(sw, d, c)
                prisms.append(np.stack([nid(k, u), nid(k, v), nid(k + 1, v),
                                        nid(k, uo), nid(k, vo), nid(k + 1, vo)], axis=1))
                prisms.append(np.stack([nid(k, u), nid(k + 1, v), nid(k + 1, u),
                                        nid(k, uo), nid(k + 1, vo), nid(k + 1, uo)], axis=1))
            else:
                for tri in ((a, b, c), (a, c, d)):
                    v = np.sort(np.stack(tri, axis=1), axis=1)  # by global 2-D node id
                    v0, v1, v2 = v[:, 0], v[:, 1], v[:, 2]
                    lo, hi = lambda u: nid(k, u), lambda u: nid(k + 1, u)
                    for tt in ((lo(v0), lo(v1), lo(v2), hi(v2)), (lo(v0), lo(v1), hi(v1), hi(v2)),
                               (lo(v0), hi(v0), hi(v1), hi(v2))):
                        tets.append(np.stack(tt, axis=1))
    prisms = np.concatenate(prisms)
    tets = np.concatenate(tets)
    # orientation: positive volumes
    def fix(cells, swap):
        p = P[cells]
        if cells.shape[1] == 4:
            vol = np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0])
        else:  # prism: bottom triangle normal against the extrusion direction
            vol = np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0])
        neg = vol < 0
        cells[neg] = cells[neg][:, swap]
        return cells
    tets = fix(tets, [0, 2, 1, 3])
    prisms = fix(prisms, [0, 2, 1, 3, 5, 4])
    jmax = n_quad + n_outer

    def bcs(nodes):
        v = nodes[:, 0]
        k = v // n2
        r = (v % n2) // ns
        out = np.full(len(nodes), BC_FARFIELD, np.int32)
        valid = nodes >= 0
        kk = np.where(valid, nodes // n2, -1)
        rr = np.where(valid, (nodes % n2) // ns, -1)
        all_root = np.all(np.where(valid, kk == 0, True), axis=1)
        all_tip = np.all(np.where(valid, kk == n_span, True), axis=1)
        all_wall = np.all(np.where(valid, rr == 0, True), axis=1)
        out[all_wall] = BC_SOLID_SURFACE
        out[all_root | all_tip] = BC_SYMMETRY
        return out

    return grid_from_polyhedra_np(X, Y, Z, tets, prisms, bcs)


# ---------------------------------------------------------------- several zones

def concat_grids(grids):
    """One grid holding independent zones, in the layout ofx_zones_set expects:
    cells zone after zone, then every zone's boundary faces, then every zone's
    interior faces, each part in zone order with the zone's own order kept.
    Returns (grid, cell_start[n+1], bface_start[n+1])."""
    dim = grids[0].dim
    if any(g.dim != dim for g in grids):
        raise ValueError("concat_grids: all zones must have the same dimension")
    n = len(grids)
    node_off = np.concatenate([[0], np.cumsum([g.n_nodes for g in grids])])
    cell_off = np.concatenate([[0], np.cumsum([g.n_cells for g in grids])])
    bface_off = np.concatenate([[0], np.cumsum([g.n_bfaces for g in grids])])
    fnc_b, fnc_i, fn_b, fn_i, lc_b, lc_i, rc_i, bct, bnm = [], [], [], [], [], [], [], [], []
    for z, g in enumerate(grids):
        nb = g.n_bfaces
        off = np.concatenate([[0], np.cumsum(g.face_node_count)])
        fnc_b.append(g.face_node_count[:nb])
        fnc_i.append(g.face_node_count[nb:])
        fn_b.append(g.face_nodes[:off[nb]] + node_off[z])
        fn_i.append(g.face_nodes[off[nb]:] + node_off[z])
        lc_b.append(g.lc[:nb] + cell_off[z])
        lc_i.append(g.lc[nb:] + cell_off[z])
        rc_i.append(g.rc[nb:] + cell_off[z])
        bct.append(g.bc_type)
        bnm.append(g.bc_name_id if g.bc_name_id is not None else g.bc_type)
    B = int(bface_off[-1])
    out = Grid(dim=dim, x=np.concatenate([g.x for g in grids]), y=np.concatenate([g.y for g in grids]),
               z=np.concatenate([g.z for g in grids]),
               face_node_count=np.concatenate(fnc_b + fnc_i).astype(np.int32),
               face_nodes=np.concatenate(fn_b + fn_i).astype(np.int32),
               lc=np.concatenate(lc_b + lc_i).astype(np.int32),
               rc=np.concatenate([-np.ones(B, np.int64)] + rc_i).astype(np.int32),
               bc_type=np.concatenate(bct).astype(np.int32), bc_name_id=np.concatenate(bnm).astype(np.int32),
               n_cells=int(cell_off[-1]))
    out.check()
    return out, cell_off.astype(np.int32), bface_off.astype(np.int32)
