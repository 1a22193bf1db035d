// ofx_mesh.hip -- grid upload and geometry build on the device.
//
// Face normals/areas/centres, cell centres/volumes, ghost-cell mirrors and the
// wall distance are computed by GPU kernels that keep the reference's
// arithmetic order (codes/geometry/src/FaceMesh.cpp:112-287,
// codes/geometry/src/UnsGrid.cpp:588-887, codes/geometry/src/WallDist.cpp:136-358).
// The cell -> face lists and the LU-SGS level schedule are integer
// preprocessing done once on the host.
#include "ofx_common.h"
#include "ofx_physics.h"

#include <algorithm>
#include <numeric>

namespace ofx {

// ---------------------------------------------------------------- face metrics

// FaceMesh::CalcFaceNormal2D + CalcFaceCenter2D (codes/geometry/src/FaceMesh.cpp:112-155)
__global__ void k_face_metrics_2d(DevMesh m) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    int o = m.fnOff[f];
    int p1 = m.fnNodes[o], p2 = m.fnNodes[o + 1];
    double xfn = m.yn[p2] - m.yn[p1];
    double yfn = m.xn[p1] - m.xn[p2];
    double zfn = 0.0;
    double area = sqrt(sqr3(xfn, yfn, zfn));
    double oArea = 1.0 / (area + SMALL);
    m.xfn[f] = xfn * oArea;
    m.yfn[f] = yfn * oArea;
    m.zfn[f] = zfn * oArea;
    m.area[f] = area;
    m.xfc[f] = 0.5 * (m.xn[p1] + m.xn[p2]);
    m.yfc[f] = 0.5 * (m.yn[p1] + m.yn[p2]);
    m.zfc[f] = 0.5 * (m.zn[p1] + m.zn[p2]);
}

// FaceMesh::CalcFaceNormal3D + CalcFaceCenter3D (codes/geometry/src/FaceMesh.cpp:157-287)
__global__ void k_face_metrics_3d(DevMesh m) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    int o = m.fnOff[f], n = m.fnOff[f + 1] - o;
    const int* fv = m.fnNodes + o;
    double xfn = 0, yfn = 0, zfn = 0;
    for (int i = 0; i < n; ++i) {
        int p1 = fv[i], p2 = fv[(i + 1) % n];
        double dx1 = m.xn[p1], dy1 = m.yn[p1], dz1 = m.zn[p1];
        double dx2 = m.xn[p2], dy2 = m.yn[p2], dz2 = m.zn[p2];
        xfn += 0.5 * (dy1 * dz2 - dy2 * dz1);
        yfn += 0.5 * (dz1 * dx2 - dz2 * dx1);
        zfn += 0.5 * (dx1 * dy2 - dx2 * dy1);
    }
    double area = sqrt(sqr3(xfn, yfn, zfn));
    double oArea = 1.0 / (area + SMALL);
    xfn *= oArea; yfn *= oArea; zfn *= oArea;
    m.xfn[f] = xfn; m.yfn[f] = yfn; m.zfn[f] = zfn; m.area[f] = area;
    const double third = 1.0 / 3.0;
    double x0 = 0, y0 = 0, z0 = 0;
    for (int i = 0; i < n; ++i) { x0 += m.xn[fv[i]]; y0 += m.yn[fv[i]]; z0 += m.zn[fv[i]]; }
    double factor = 1.0 / n;
    x0 *= factor; y0 *= factor; z0 *= factor;
    double xc = 0, yc = 0, zc = 0, sarea = 0;
    for (int i = 0; i < n; ++i) {
        int p1 = fv[i], p2 = fv[(i + 1) % n];
        double dx1 = m.xn[p1] - x0, dy1 = m.yn[p1] - y0, dz1 = m.zn[p1] - z0;
        double dx2 = m.xn[p2] - x0, dy2 = m.yn[p2] - y0, dz2 = m.zn[p2] - z0;
        double x00 = third * (m.xn[p1] + m.xn[p2] + x0);
        double y00 = third * (m.yn[p1] + m.yn[p2] + y0);
        double z00 = third * (m.zn[p1] + m.zn[p2] + z0);
        double anx = dy1 * dz2 - dy2 * dz1;
        double any = dz1 * dx2 - dz2 * dx1;
        double anz = dx1 * dy2 - dx2 * dy1;
        double fa = 0.5 * sqrt(sqr3(anx, any, anz));
        double norm = anx * xfn + any * yfn + anz * zfn;
        if (norm < 0.0) fa = -fa;
        sarea += fa;
        xc += fa * x00; yc += fa * y00; zc += fa * z00;
    }
    double osarea = 1.0 / (sarea + SMALL);
    m.xfc[f] = xc * osarea; m.yfc[f] = yc * osarea; m.zfc[f] = zc * osarea;
}

// ---------------------------------------------------------------- cell metrics

// UnsGrid::CalcCellCenterVol2D as a per-cell gather over the cell's faces in
// ascending face order (codes/geometry/src/UnsGrid.cpp:588-667).
__global__ void k_cell_metrics_2d(DevMesh m) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    double xc = 0, yc = 0, zc = 0, v = 0;
    for (int k = m.c2fOff[i]; k < m.c2fOff[i + 1]; ++k) {
        int e = m.c2f[k];
        int f = e >= 0 ? e : ~e;
        double dot = (m.xfc[f] * m.xfn[f] + m.yfc[f] * m.yfn[f] + m.zfc[f] * m.zfn[f]) * m.area[f];
        if (e >= 0) {
            xc += m.xfc[f] * dot; yc += m.yfc[f] * dot; zc += m.zfc[f] * dot; v += dot;
        } else {
            xc -= m.xfc[f] * dot; yc -= m.yfc[f] * dot; zc -= m.zfc[f] * dot; v -= dot;
        }
    }
    double tmp = 1.0 / (1.5 * v + SMALL);
    xc *= tmp; yc *= tmp; zc *= tmp;
    v *= 0.5;
    if (v <= 0.0) v = -v;
    m.xcc[i] = xc; m.ycc[i] = yc; m.zcc[i] = zc; m.vol[i] = v;
}

// UnsGrid::CalcCellCenterVol3D, per-cell gather (codes/geometry/src/UnsGrid.cpp:740-860)
__global__ void k_cell_metrics_3d(DevMesh m) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    double xc = 0, yc = 0, zc = 0, v = 0;
    for (int k = m.c2fOff[i]; k < m.c2fOff[i + 1]; ++k) {
        int e = m.c2f[k];
        int f = e >= 0 ? e : ~e;
        int o = m.fnOff[f], n = m.fnOff[f + 1] - o;
        const int* fv = m.fnNodes + o;
        for (int j = 0; j < n; ++j) {
            int p2 = fv[j], p3 = fv[(j + 1) % n];
            double x21 = m.xn[p2] - m.xfc[f], y21 = m.yn[p2] - m.yfc[f], z21 = m.zn[p2] - m.zfc[f];
            double x31 = m.xn[p3] - m.xfc[f], y31 = m.yn[p3] - m.yfc[f], z31 = m.zn[p3] - m.zfc[f];
            double nx = y21 * z31 - y31 * z21;
            double ny = z21 * x31 - z31 * x21;
            double nz = x21 * y31 - x31 * y21;
            double trix = m.xfc[f] + m.xn[p2] + m.xn[p3];
            double triy = m.yfc[f] + m.yn[p2] + m.yn[p3];
            double triz = m.zfc[f] + m.zn[p2] + m.zn[p3];
            double tmp = nx * trix + ny * triy + nz * triz;
            trix *= tmp; triy *= tmp; triz *= tmp;
            if (e >= 0) { xc += trix; yc += triy; zc += triz; v += tmp; }
            else { xc -= trix; yc -= triy; zc -= triz; v -= tmp; }
        }
    }
    double tmp = 1.0 / (4.0 * v + SMALL);
    xc *= tmp; yc *= tmp; zc *= tmp;
    v /= 18.0;
    if (v <= 0.0) v = -v;
    m.xcc[i] = xc; m.ycc[i] = yc; m.zcc[i] = zc; m.vol[i] = v;
}

// ghost mirror (codes/geometry/src/UnsGrid.cpp:668-690)
// Static per-face terms (DevMesh::fgg / frec / fz).  The expressions are
// those of the gradient (CalcGradGGCellWeight, codes/usolver/src/UGrad.cpp:88-149)
// and reconstruction (LimField::CalcFaceValue, codes/usolver/src/ULimiter.cpp:181-214)
// kernels, with z terms 0 in 2-D as there.
template <int DIM>
__global__ void k_face_static(DevMesh m) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    int l = m.lc[f], r = (f < m.B) ? m.C + f : m.rc[f];
    double fx = m.xfc[f], fy = m.yfc[f], fz = (DIM == 3) ? m.zfc[f] : 0.0;
    double xl = m.xcc[l], yl = m.ycc[l], zl = (DIM == 3) ? m.zcc[l] : 0.0;
    double xr = m.xcc[r], yr = m.ycc[r], zr = (DIM == 3) ? m.zcc[r] : 0.0;
    double delt1 = sqrt(sqr3(fx - xl, fy - yl, fz - zl));
    double delt2 = sqrt(sqr3(fx - xr, fy - yr, fz - zr));
    double delta = 1.0 / (delt1 + delt2 + SMALL);
    double cl = delt2 * delta, cr = delt1 * delta;
    double ar = m.area[f];
    m.fgg[f] = dvec4{cl, cr, m.xfn[f] * ar, m.yfn[f] * ar};
    m.frec[f] = dvec4{fx - xl, fy - yl, fx - xr, fy - yr};
    if (DIM == 3) {
        m.fz[f] = dvec4{m.zfn[f] * ar, fz - zl, fz - zr, 0.0};
        m.fzr[f] = dvec2{fz - zl, fz - zr};
    }
}

// The face terms of the cell kernels once per (cell, entry) (DevMesh::ellg):
// the gradients' ones as seen from the cell, copied from fgg / frec / fz with
// the side's selection and sign applied, and the face's unit normal and area
// as stored.
template <int DIM>
__global__ void k_ell_geom(DevMesh m) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    const int n = m.c2fOff[i + 1] - m.c2fOff[i];
    const size_t S = (size_t)m.ellw * m.C;
    for (int j = 0; j < m.ellw; ++j) {
        const size_t o = (size_t)j * m.C + i;
        double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // the two records, in order
        if (j < n) {
            const int e = m.ellf[o], f = e >= 0 ? e : ~e;
            const bool own = e >= 0;
            const dvec4 g = m.fgg[f], rc4 = m.frec[f];
            const dvec4 z4 = (DIM == 3) ? m.fz[f] : dvec4{0.0, 0.0, 0.0, 0.0};
            v[0] = own ? g.x : g.y;
            v[1] = own ? g.y : g.x;
            v[2] = own ? g.z : -g.z;
            v[3] = own ? g.w : -g.w;
            v[4] = (DIM == 3) ? (own ? z4.x : -z4.x) : 0.0;
            v[5] = own ? rc4.x : rc4.z;
            v[6] = own ? rc4.y : rc4.w;
            v[7] = (DIM == 3) ? (own ? z4.y : z4.z) : 0.0;
        }
        m.ellg[o] = dvec4{v[0], v[1], v[2], v[3]};
        m.ellg[S + o] = dvec4{v[4], v[5], v[6], v[7]};
        dvec4 nrm{0.0, 0.0, 0.0, 0.0};
        if (j < n) {
            const int e = m.ellf[o], f = e >= 0 ? e : ~e;
            nrm = dvec4{m.xfn[f], m.yfn[f], m.zfn[f], m.area[f]};
        }
        m.ellg[2 * S + o] = nrm;
        double dist = 0.0, vnb = 0.0;
        if (j < n) {
            const int e = m.ellf[o], c = m.elln[o];
            const bool own = e >= 0;
            const double xo = own ? nrm.x : -nrm.x, yo = own ? nrm.y : -nrm.y, zo = own ? nrm.z : -nrm.z;
            dist = dabs(xo * (m.xcc[c] - m.xcc[i]) + yo * (m.ycc[c] - m.ycc[i]) + zo * (m.zcc[c] - m.zcc[i]));
            vnb = m.vol[c];
        }
        m.elldv[o] = dist;
        m.elldv[S + o] = vnb;
        m.elldv[2 * S + o] = v[4];
    }
}

// Viscous face-gradient geometry records (DevMesh::fv0 / fv2) from vis_geom.
__global__ void k_face_vis_static(DevMesh m) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    int l = m.lc[f], r = (f < m.B) ? m.C + f : m.rc[f];
    VisGeom G;
    vis_geom(m, f, l, r, 0, G);
    m.fv0[f] = dvec4{G.dx, G.dy, G.dz, G.ods};
    vis_geom(m, f, l, r, 2, G);
    dvec4* o = m.fv2 + 5 * (size_t)f;
    o[0] = dvec4{G.d1, G.d2, G.angle1, G.angle2};
    o[1] = dvec4{G.dxnl, G.dynl, G.dznl, G.dxnr};
    o[2] = dvec4{G.dynr, G.dznr, G.t1[0], G.t1[1]};
    o[3] = dvec4{G.t1[2], G.t2[0], G.t2[1], G.t2[2]};
    o[4] = dvec4{G.fw1, G.fw2, 0.0, 0.0};
}

__global__ void k_ghost_centres(DevMesh m) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.B) return;
    int l = m.lc[f], r = m.C + f;
    if (m.area[f] > SMALL) {
        double tmp = 2.0 * ((m.xcc[l] - m.xfc[f]) * m.xfn[f] + (m.ycc[l] - m.yfc[f]) * m.yfn[f] +
                            (m.zcc[l] - m.zfc[f]) * m.zfn[f]);
        m.xcc[r] = m.xcc[l] - m.xfn[f] * tmp;
        m.ycc[r] = m.ycc[l] - m.yfn[f] * tmp;
        m.zcc[r] = m.zcc[l] - m.zfn[f] * tmp;
    } else {
        m.xcc[r] = -m.xcc[l] + 2.0 * m.xfc[f];
        m.ycc[r] = -m.ycc[l] + 2.0 * m.yfc[f];
        m.zcc[r] = -m.zcc[l] + 2.0 * m.zfc[f];
    }
    m.vol[r] = m.vol[l];
}

// ---------------------------------------------------------------- wall distance

// gte DCPQuery point/segment, squared distance (codes/gte/include/GteDistPointSegment.h)
__device__ double seg_dist2(const double p[3], const double a[3], const double b[3]) {
    double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double df[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    double t = d[0] * df[0] + d[1] * df[1] + d[2] * df[2];
    double cl[3];
    if (t >= 0.0) { cl[0] = b[0]; cl[1] = b[1]; cl[2] = b[2]; }
    else {
        df[0] = p[0] - a[0]; df[1] = p[1] - a[1]; df[2] = p[2] - a[2];
        t = d[0] * df[0] + d[1] * df[1] + d[2] * df[2];
        if (t <= 0.0) { cl[0] = a[0]; cl[1] = a[1]; cl[2] = a[2]; }
        else {
            double sl = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (sl > 0.0) {
                t /= sl;
                cl[0] = a[0] + t * d[0]; cl[1] = a[1] + t * d[1]; cl[2] = a[2] + t * d[2];
            } else { cl[0] = a[0]; cl[1] = a[1]; cl[2] = a[2]; }
        }
    }
    df[0] = p[0] - cl[0]; df[1] = p[1] - cl[1]; df[2] = p[2] - cl[2];
    return df[0] * df[0] + df[1] * df[1] + df[2] * df[2];
}

// gte DistancePointTriangleExact, squared distance (codes/gte/include/GteDistPointTriangleExact.h)
__device__ double tri_dist2(const double P[3], const double V0[3], const double V1[3], const double V2[3]) {
    double diff[3] = {P[0] - V0[0], P[1] - V0[1], P[2] - V0[2]};
    double e0[3] = {V1[0] - V0[0], V1[1] - V0[1], V1[2] - V0[2]};
    double e1[3] = {V2[0] - V0[0], V2[1] - V0[1], V2[2] - V0[2]};
    double a00 = e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2];
    double a01 = e0[0] * e1[0] + e0[1] * e1[1] + e0[2] * e1[2];
    double a11 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    double b0 = -(diff[0] * e0[0] + diff[1] * e0[1] + diff[2] * e0[2]);
    double b1 = -(diff[0] * e1[0] + diff[1] * e1[1] + diff[2] * e1[2]);
    double det = a00 * a11 - a01 * a01;
    double t0 = a01 * b1 - a11 * b0;
    double t1 = a01 * b0 - a00 * b1;
    if (t0 + t1 <= det) {
        if (t0 < 0) {
            if (t1 < 0) {
                if (b0 < 0) { t1 = 0; if (-b0 >= a00) t0 = 1; else t0 = -b0 / a00; }
                else { t0 = 0; if (b1 >= 0) t1 = 0; else if (-b1 >= a11) t1 = 1; else t1 = -b1 / a11; }
            } else {
                t0 = 0;
                if (b1 >= 0) t1 = 0; else if (-b1 >= a11) t1 = 1; else t1 = -b1 / a11;
            }
        } else if (t1 < 0) {
            t1 = 0;
            if (b0 >= 0) t0 = 0; else if (-b0 >= a00) t0 = 1; else t0 = -b0 / a00;
        } else {
            double invDet = 1.0 / det;
            t0 *= invDet; t1 *= invDet;
        }
    } else {
        double tmp0, tmp1, numer, denom;
        if (t0 < 0) {
            tmp0 = a01 + b0; tmp1 = a11 + b1;
            if (tmp1 > tmp0) {
                numer = tmp1 - tmp0; denom = a00 - 2 * a01 + a11;
                if (numer >= denom) { t0 = 1; t1 = 0; } else { t0 = numer / denom; t1 = 1 - t0; }
            } else {
                t0 = 0;
                if (tmp1 <= 0) t1 = 1; else if (b1 >= 0) t1 = 0; else t1 = -b1 / a11;
            }
        } else if (t1 < 0) {
            tmp0 = a01 + b1; tmp1 = a00 + b0;
            if (tmp1 > tmp0) {
                numer = tmp1 - tmp0; denom = a00 - 2 * a01 + a11;
                if (numer >= denom) { t1 = 1; t0 = 0; } else { t1 = numer / denom; t0 = 1 - t1; }
            } else {
                t1 = 0;
                if (tmp1 <= 0) t0 = 1; else if (b0 >= 0) t0 = 0; else t0 = -b0 / a00;
            }
        } else {
            numer = a11 + b1 - a01 - b0;
            if (numer <= 0) { t0 = 0; t1 = 1; }
            else {
                denom = a00 - 2 * a01 + a11;
                if (numer >= denom) { t0 = 1; t1 = 0; } else { t0 = numer / denom; t1 = 1 - t0; }
            }
        }
    }
    double cp[3] = {V0[0] + t0 * e0[0] + t1 * e1[0], V0[1] + t0 * e0[1] + t1 * e1[1], V0[2] + t0 * e0[2] + t1 * e1[2]};
    double d[3] = {P[0] - cp[0], P[1] - cp[1], P[2] - cp[2]};
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

// CalcWallDist: brute force over all wall faces, min of squared distance, sqrt
// (codes/geometry/src/WallDist.cpp:136-358).  One thread per cell; wall-face
// node coordinates are staged through LDS in tiles shared by the block.  With
// several zones a cell sees only its own zone's wall faces (same faces, same
// order as a single-zone run); the block tiles the union of its zones' ranges.
constexpr int WD_TILE = 256;
__global__ void __launch_bounds__(256) k_wall_distance(DevMesh m) {
    __shared__ double sx[WD_TILE * 4], sy[WD_TILE * 4], sz[WD_TILE * 4];
    __shared__ int sn[WD_TILE];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool live = i < m.C;
    double p[3] = {0, 0, 0};
    if (live) { p[0] = m.xcc[i]; p[1] = m.ycc[i]; p[2] = m.zcc[i]; }
    double best = LARGE;
    int w0 = 0, w1 = m.nWall, my0 = 0, my1 = m.nWall;
    if (m.nz > 1) {
        int first = blockIdx.x * blockDim.x, last = min(first + (int)blockDim.x, m.C) - 1;
        w0 = m.zwall[m.zcell[first]];
        w1 = m.zwall[m.zcell[last] + 1];
        my0 = live ? m.zwall[m.zcell[i]] : 0;
        my1 = live ? m.zwall[m.zcell[i] + 1] : 0;
    }
    for (int base = w0; base < w1; base += WD_TILE) {
        int nt = min(WD_TILE, w1 - base);
        __syncthreads();
        for (int t = threadIdx.x; t < nt; t += blockDim.x) {
            int f = m.wallFaces[base + t];
            int o = m.fnOff[f], n = m.fnOff[f + 1] - o;
            sn[t] = n;
            for (int j = 0; j < 4; ++j) {
                int pj = m.fnNodes[o + (j < n ? j : 0)];
                sx[t * 4 + j] = m.xn[pj]; sy[t * 4 + j] = m.yn[pj]; sz[t * 4 + j] = m.zn[pj];
            }
        }
        __syncthreads();
        if (!live) continue;
        for (int t = 0; t < nt; ++t) {
            if (base + t < my0 || base + t >= my1) continue;
            int n = sn[t];
            double w;
            double a[3] = {sx[t * 4], sy[t * 4], sz[t * 4]};
            double b[3] = {sx[t * 4 + 1], sy[t * 4 + 1], sz[t * 4 + 1]};
            if (n <= 2) {
                w = seg_dist2(p, a, b);
            } else if (n == 3) {
                double c3[3] = {sx[t * 4 + 2], sy[t * 4 + 2], sz[t * 4 + 2]};
                w = tri_dist2(p, a, b, c3);
            } else {
                double ctr[3] = {0, 0, 0};
                for (int k = 0; k < n; ++k) { ctr[0] += sx[t * 4 + k]; ctr[1] += sy[t * 4 + k]; ctr[2] += sz[t * 4 + k]; }
                double coef = 1.0 / n;
                ctr[0] *= coef; ctr[1] *= coef; ctr[2] *= coef;
                w = LARGE;
                for (int k = 0; k < n; ++k) {
                    int k1 = (k + 1) % n;
                    double u[3] = {sx[t * 4 + k], sy[t * 4 + k], sz[t * 4 + k]};
                    double v[3] = {sx[t * 4 + k1], sy[t * 4 + k1], sz[t * 4 + k1]};
                    double dd = tri_dist2(p, u, v, ctr);
                    w = (w < dd) ? w : dd;
                }
            }
            if (best > w) best = w;
        }
    }
    if (live) m.dist[i] = sqrt(best);
}

// ---------------------------------------------------------------- host side

// Spatial storage order (host, once per mesh).  Cells: grouped by connected
// part (parts in the order of their lowest cell id, so the cells of
// disconnected zones stay in their id ranges), within a part sorted by the
// Morton code of the centroid of their faces' node averages (21 bits per axis
// in 3-D, 32 in 2-D, over the part's bounding box), ties by id.  Faces:
// boundary faces keep their positions (their ghosts are cells C + f);
// interior faces sorted by the positions of their two cells, lower first.
static uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
    v &= 0x1fffffull;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}
static uint64_t spread2(uint64_t v) {  // 32 bits -> every second bit
    v &= 0xffffffffull;
    v = (v | v << 16) & 0x0000ffff0000ffffull;
    v = (v | v << 8) & 0x00ff00ff00ff00ffull;
    v = (v | v << 4) & 0x0f0f0f0f0f0f0f0full;
    v = (v | v << 2) & 0x3333333333333333ull;
    v = (v | v << 1) & 0x5555555555555555ull;
    return v;
}

static void spatial_order(int dim, int C, int F, int B, const double* x, const double* y, const double* z,
                          const std::vector<int>& fnOff, const std::vector<int>& nodes, const std::vector<int>& lc,
                          const std::vector<int>& rc, std::vector<int>& cperm, std::vector<int>& fperm) {
    // centroid estimate: mean over the cell's faces of the face's node average
    std::vector<double> cx(C, 0.0), cy(C, 0.0), cz(C, 0.0);
    std::vector<int> cnt(C, 0);
    for (int f = 0; f < F; ++f) {
        double fx = 0, fy = 0, fz = 0;
        int n = fnOff[f + 1] - fnOff[f];
        for (int k = fnOff[f]; k < fnOff[f + 1]; ++k) { fx += x[nodes[k]]; fy += y[nodes[k]]; fz += z[nodes[k]]; }
        fx /= n; fy /= n; fz /= n;
        int cs[2] = {lc[f], f < B ? -1 : rc[f]};
        for (int c : cs)
            if (c >= 0) { cx[c] += fx; cy[c] += fy; cz[c] += fz; ++cnt[c]; }
    }
    for (int i = 0; i < C; ++i) { cx[i] /= cnt[i]; cy[i] /= cnt[i]; cz[i] /= cnt[i]; }
    // connected parts (union-find over interior faces; root = lowest id)
    std::vector<int> par(C);
    std::iota(par.begin(), par.end(), 0);
    auto find = [&](int a) { while (par[a] != a) { par[a] = par[par[a]]; a = par[a]; } return a; };
    for (int f = B; f < F; ++f) {
        int a = find(lc[f]), b = find(rc[f]);
        if (a != b) { if (a < b) par[b] = a; else par[a] = b; }
    }
    std::vector<int> part(C);
    for (int i = 0; i < C; ++i) part[i] = find(i);
    // per-part bounding boxes
    std::vector<double> lo(3 * (size_t)C, 0.0), hi(3 * (size_t)C, 0.0);
    std::vector<char> seen(C, 0);
    for (int i = 0; i < C; ++i) {
        int r = part[i];
        double v[3] = {cx[i], cy[i], cz[i]};
        for (int a = 0; a < 3; ++a) {
            double& l = lo[3 * (size_t)r + a];
            double& h = hi[3 * (size_t)r + a];
            if (!seen[r]) { l = v[a]; h = v[a]; }
            else { l = std::min(l, v[a]); h = std::max(h, v[a]); }
        }
        seen[r] = 1;
    }
    std::vector<uint64_t> key(C);
    for (int i = 0; i < C; ++i) {
        int r = part[i];
        double v[3] = {cx[i], cy[i], cz[i]};
        uint64_t q[3];
        const double scale = dim == 3 ? 2097151.0 : 4294967295.0;
        for (int a = 0; a < 3; ++a) {
            double l = lo[3 * (size_t)r + a], h = hi[3 * (size_t)r + a];
            double t = h > l ? (v[a] - l) / (h - l) : 0.0;
            q[a] = (uint64_t)std::llround(std::min(1.0, std::max(0.0, t)) * scale);
        }
        key[i] = dim == 3 ? (spread3(q[0]) | spread3(q[1]) << 1 | spread3(q[2]) << 2) : (spread2(q[0]) | spread2(q[1]) << 1);
    }
    std::vector<int> ord(C);
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](int a, int b) {
        if (part[a] != part[b]) return part[a] < part[b];
        if (key[a] != key[b]) return key[a] < key[b];
        return a < b;
    });
    cperm.assign(C, 0);
    for (int k = 0; k < C; ++k) cperm[ord[k]] = k;
    std::vector<int> ford(F - B);
    std::iota(ford.begin(), ford.end(), B);
    auto lohi = [&](int f) {
        int a = cperm[lc[f]], b = cperm[rc[f]];
        return std::make_pair(std::min(a, b), std::max(a, b));
    };
    std::sort(ford.begin(), ford.end(), [&](int a, int b) {
        auto ka = lohi(a), kb = lohi(b);
        if (ka != kb) return ka < kb;
        return a < b;
    });
    fperm.assign(F, 0);
    for (int f = 0; f < B; ++f) fperm[f] = f;
    for (int k = 0; k < F - B; ++k) fperm[ford[k]] = B + k;
}

int mesh_build(ofx_ctx* c, const double* x, const double* y, const double* z, const int* fnc,
               const int* fnn, const int* lcIn, const int* rcIn, const int* bct) {
    DevMesh& m = c->m;
    const int C = m.C, F = m.F, B = m.B, T = m.T;
    // face -> node CSR, orientation fix (codes/geometry/src/UnsGrid.cpp:160-170), NormalizeBc
    std::vector<int> fnOff(F + 1, 0);
    for (int f = 0; f < F; ++f) {
        if (fnc[f] < 2 || (m.dim == 2 && fnc[f] != 2) || (m.dim == 3 && fnc[f] > 4)) {
            set_error("ofx_mesh_upload: unsupported face node count"); return OFX_EINVAL;
        }
        fnOff[f + 1] = fnOff[f] + fnc[f];
    }
    std::vector<int> nodes(fnn, fnn + fnOff[F]);
    std::vector<int> lc(lcIn, lcIn + F), rc(rcIn, rcIn + F);
    for (int f = 0; f < F; ++f) {
        if (lc[f] < 0) {
            std::reverse(nodes.begin() + fnOff[f], nodes.begin() + fnOff[f + 1]);
            std::swap(lc[f], rc[f]);
        }
    }
    for (int k = 0; k < fnOff[F]; ++k)
        if (nodes[k] < 0 || nodes[k] >= m.nNodes) { set_error("ofx_mesh_upload: face node out of range"); return OFX_EINVAL; }
    for (int f = 0; f < B; ++f) {
        if (lc[f] < 0 || lc[f] >= C) { set_error("ofx_mesh_upload: boundary face owner out of range"); return OFX_EINVAL; }
        if (bct[f] < 1 || bct[f] > 6) { set_error("ofx_mesh_upload: unsupported boundary type (1..6 supported)"); return OFX_EINVAL; }
        rc[f] = C + f;
    }
    for (int f = B; f < F; ++f)
        if (lc[f] < 0 || lc[f] >= C || rc[f] < 0 || rc[f] >= C || lc[f] == rc[f]) {
            set_error("ofx_mesh_upload: interior face cell index out of range"); return OFX_EINVAL;
        }
    // ---- storage order: grid id -> storage position of cells and faces
    std::vector<int> cperm, fperm;
    if (c->storage_order == 1) spatial_order(m.dim, C, F, B, x, y, z, fnOff, nodes, lc, rc, cperm, fperm);
    else {
        cperm.resize(C); fperm.resize(F);
        std::iota(cperm.begin(), cperm.end(), 0);
        std::iota(fperm.begin(), fperm.end(), 0);
    }
    std::vector<int> cinv(C), finv(F);
    for (int i = 0; i < C; ++i) cinv[cperm[i]] = i;
    for (int f = 0; f < F; ++f) finv[fperm[f]] = f;
    bool ident = true;
    for (int i = 0; i < C && ident; ++i) ident = cperm[i] == i;
    for (int f = 0; f < F && ident; ++f) ident = fperm[f] == f;
    if (ident) { c->cperm.clear(); c->fperm.clear(); }
    else { c->cperm = cperm; c->fperm = fperm; }
    // faces in storage order: cells by position, ghosts of boundary face f at C + f
    {
        std::vector<int> lc2(F), rc2(F), off2(F + 1, 0), nodes2(nodes.size());
        for (int s2 = 0; s2 < F; ++s2) {
            int f = finv[s2];
            lc2[s2] = cperm[lc[f]];
            rc2[s2] = f < B ? C + f : cperm[rc[f]];
            off2[s2 + 1] = off2[s2] + (fnOff[f + 1] - fnOff[f]);
            std::copy(nodes.begin() + fnOff[f], nodes.begin() + fnOff[f + 1], nodes2.begin() + off2[s2]);
        }
        lc.swap(lc2); rc.swap(rc2); fnOff.swap(off2); nodes.swap(nodes2);
    }
    // CellTopo::CalcC2f (codes/geometry/src/CellTopo.cpp:95-119): each cell's
    // faces in ascending face id (grid numbering), as storage positions
    std::vector<int> off(T + 1, 0);
    for (int f = 0; f < B; ++f) off[lc[fperm[f]] + 1]++;
    for (int f = B; f < F; ++f) { off[lc[fperm[f]] + 1]++; off[rc[fperm[f]] + 1]++; }
    for (int i = 0; i < T; ++i) off[i + 1] += off[i];
    std::vector<int> c2f(off[T]);
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (int f = 0; f < B; ++f) { int s2 = fperm[f]; c2f[pos[lc[s2]]++] = s2; }
    for (int f = B; f < F; ++f) { int s2 = fperm[f]; c2f[pos[lc[s2]]++] = s2; c2f[pos[rc[s2]]++] = ~s2; }
    std::vector<int> c2n(off[T]);
    for (int k = 0; k < off[T]; ++k) {
        int e = c2f[k], f = e >= 0 ? e : ~e;
        c2n[k] = (e >= 0) ? rc[f] : lc[f];
    }
    int maxf = 0;
    for (int i = 0; i < C; ++i) maxf = std::max(maxf, off[i + 1] - off[i]);
    m.maxFacesPerCell = maxf;
    if (maxf > 31) { set_error("ofx_mesh_upload: cells with more than 31 faces are not supported"); return OFX_EINVAL; }
    // grid id of the neighbour of each entry (interior faces; ghosts -1)
    auto nb_id = [&](int k) { int j = c2n[k]; return j < C ? cinv[j] : -1; };
    std::vector<int> lomask(C, 0);
    for (int sc = 0; sc < C; ++sc)
        for (int k = off[sc]; k < off[sc + 1]; ++k) {
            int j = nb_id(k);
            if (j >= 0 && j < cinv[sc]) lomask[sc] |= 1 << (k - off[sc]);
        }
    // LU-SGS level schedule in the grid's numbering (exact: a cell depends
    // only on lower / upper neighbours, UNsLusgs::CanNotLowerSolve /
    // CanNotUpperSolve, codes/uns/src/UNsLusgs.cpp:174-196); the level lists
    // hold storage positions, ascending within a level
    std::vector<int> lev(C, 0), levU(C, 0);  // by grid id
    int nlo = 0, nup = 0;
    for (int i = 0; i < C; ++i) {
        int sc = cperm[i], mx = -1;
        for (int k = off[sc]; k < off[sc + 1]; ++k) {
            int j = nb_id(k);
            if (j >= 0 && j < i) mx = std::max(mx, lev[j]);
        }
        lev[i] = mx + 1;
        nlo = std::max(nlo, lev[i] + 1);
    }
    for (int i = C - 1; i >= 0; --i) {
        int sc = cperm[i], mx = -1;
        for (int k = off[sc]; k < off[sc + 1]; ++k) {
            int j = nb_id(k);
            if (j > i) mx = std::max(mx, levU[j]);
        }
        levU[i] = mx + 1;
        nup = std::max(nup, levU[i] + 1);
    }
    auto order_by = [&](const std::vector<int>& L, int nl, std::vector<int>& ord, std::vector<int>& lo) {
        lo.assign(nl + 1, 0);
        for (int i = 0; i < C; ++i) lo[L[i] + 1]++;
        for (int l = 0; l < nl; ++l) lo[l + 1] += lo[l];
        ord.assign(C, 0);
        std::vector<int> p(lo.begin(), lo.end() - 1);
        for (int sc = 0; sc < C; ++sc) ord[p[L[cinv[sc]]]++] = sc;  // storage order within a level
    };
    std::vector<int> loOrd, upOrd;
    order_by(lev, nlo, loOrd, c->sched.loOff);
    order_by(levU, nup, upOrd, c->sched.upOff);
    auto runs = [](const std::vector<int>& ord, const std::vector<int>& lo, std::vector<int>& base) {
        base.assign(lo.size() - 1, -1);
        for (size_t L = 0; L + 1 < lo.size(); ++L) {
            int b = lo[L], e = lo[L + 1];
            bool run = e > b;
            for (int k = b + 1; k < e && run; ++k) run = ord[k] == ord[b] + (k - b);
            if (run) base[L] = ord[b];
        }
    };
    runs(loOrd, c->sched.loOff, c->sched.loBase);
    runs(upOrd, c->sched.upOff, c->sched.upBase);
    m.nLevLo = nlo; m.nLevUp = nup;
    // Segments for XCD placement: cells stored sorted by lower-sweep level with
    // at most 8 levels are a colouring stored colour by colour; each colour is
    // one segment, and the interior faces (sorted by left cell) split at the
    // same cells after the boundary-face segment.
    m.nsegc = 1; m.segc[0] = 0; m.segc[1] = C;
    m.nsegf = 1; m.segf[0] = 0; m.segf[1] = F;
    bool sorted = nlo <= MAXSEG - 1;
    for (int sc = 1; sc < C && sorted; ++sc) sorted = lev[cinv[sc]] >= lev[cinv[sc - 1]];
    for (int f = B + 1; f < F && sorted; ++f) sorted = lc[f] >= lc[f - 1];
    if (sorted && nlo > 1) {
        m.nsegc = nlo;
        for (int l = 0; l <= nlo; ++l) m.segc[l] = c->sched.loOff[l];
        int ns = 0;
        m.segf[ns++] = 0;
        int f = B;
        for (int l = 0; l < nlo; ++l) {
            if (B > 0 || l > 0) m.segf[ns++] = f;
            while (f < F && lc[f] < m.segc[l + 1]) ++f;
        }
        m.segf[ns] = F;
        m.nsegf = ns;
    }
    std::vector<int> wall;
    for (int f = 0; f < B; ++f) if (bct[f] == BC_SOLID_SURFACE) wall.push_back(f);
    m.nWall = (int)wall.size();

    // ---- device allocations
    auto A = [&](void** p, size_t bytes) { return alloc_bytes(c, p, bytes); };
    size_t dF = sizeof(double) * F, dT = sizeof(double) * T;
    OFX_TRY(A((void**)&m.xn, sizeof(double) * m.nNodes));
    OFX_TRY(A((void**)&m.yn, sizeof(double) * m.nNodes));
    OFX_TRY(A((void**)&m.zn, sizeof(double) * m.nNodes));
    OFX_TRY(A((void**)&m.fnOff, sizeof(int) * (F + 1)));
    OFX_TRY(A((void**)&m.fnNodes, sizeof(int) * std::max(1, fnOff[F])));
    double** fa[] = {&m.xfn, &m.yfn, &m.zfn, &m.area, &m.xfc, &m.yfc, &m.zfc};
    for (double** p : fa) OFX_TRY(A((void**)p, dF));
    double** ca[] = {&m.xcc, &m.ycc, &m.zcc, &m.vol};
    for (double** p : ca) OFX_TRY(A((void**)p, dT));
    OFX_TRY(A((void**)&m.dist, sizeof(double) * C));
    OFX_TRY(A((void**)&m.fgg, sizeof(dvec4) * F));
    OFX_TRY(A((void**)&m.frec, sizeof(dvec4) * F));
    OFX_TRY(A((void**)&m.fz, sizeof(dvec4) * (m.dim == 3 ? F : 1)));
    OFX_TRY(A((void**)&m.fzr, sizeof(dvec2) * (m.dim == 3 ? F : 1)));
    OFX_TRY(A((void**)&m.fv0, sizeof(dvec4) * F));
    OFX_TRY(A((void**)&m.fv2, sizeof(dvec4) * 5 * (size_t)F));
    OFX_TRY(A((void**)&m.lc, sizeof(int) * F));
    OFX_TRY(A((void**)&m.rc, sizeof(int) * F));
    OFX_TRY(A((void**)&m.bct, sizeof(int) * std::max(1, B)));
    OFX_TRY(A((void**)&m.c2fOff, sizeof(int) * (T + 1)));
    OFX_TRY(A((void**)&m.c2f, sizeof(int) * std::max(1, off[T])));
    OFX_TRY(A((void**)&m.c2n, sizeof(int) * std::max(1, off[T])));
    OFX_TRY(A((void**)&m.lomask, sizeof(int) * C));
    m.ellw = maxf;
    std::vector<int> ellf((size_t)maxf * C, 0), elln((size_t)maxf * C, 0);
    for (int i = 0; i < C; ++i)
        for (int j = 0; j < maxf; ++j) {
            bool in = j < off[i + 1] - off[i];
            ellf[(size_t)j * C + i] = in ? c2f[off[i] + j] : 0;
            elln[(size_t)j * C + i] = in ? c2n[off[i] + j] : i;
        }
    OFX_TRY(A((void**)&m.ellf, sizeof(int) * std::max<size_t>(1, ellf.size())));
    OFX_TRY(A((void**)&m.elln, sizeof(int) * std::max<size_t>(1, elln.size())));
    OFX_TRY(A((void**)&m.ellg, sizeof(dvec4) * 3 * std::max<size_t>(1, ellf.size())));
    OFX_TRY(A((void**)&m.elldv, sizeof(double) * 3 * std::max<size_t>(1, ellf.size())));
    OFX_TRY(A((void**)&m.loOrder, sizeof(int) * C));
    OFX_TRY(A((void**)&m.upOrder, sizeof(int) * C));
    OFX_TRY(A((void**)&m.loOff, sizeof(int) * (nlo + 1)));
    OFX_TRY(A((void**)&m.upOff, sizeof(int) * (nup + 1)));
    OFX_TRY(A((void**)&m.wallFaces, sizeof(int) * std::max(1, m.nWall)));
    OFX_TRY(upload(c, m.xn, x, sizeof(double) * m.nNodes));
    OFX_TRY(upload(c, m.yn, y, sizeof(double) * m.nNodes));
    OFX_TRY(upload(c, m.zn, z, sizeof(double) * m.nNodes));
    OFX_TRY(upload(c, m.fnOff, fnOff.data(), sizeof(int) * (F + 1)));
    OFX_TRY(upload(c, m.fnNodes, nodes.data(), sizeof(int) * fnOff[F]));
    OFX_TRY(upload(c, m.lc, lc.data(), sizeof(int) * F));
    OFX_TRY(upload(c, m.rc, rc.data(), sizeof(int) * F));
    if (B) OFX_TRY(upload(c, m.bct, bct, sizeof(int) * B));
    OFX_TRY(upload(c, m.c2fOff, off.data(), sizeof(int) * (T + 1)));
    OFX_TRY(upload(c, m.c2f, c2f.data(), sizeof(int) * off[T]));
    OFX_TRY(upload(c, m.c2n, c2n.data(), sizeof(int) * off[T]));
    OFX_TRY(upload(c, m.lomask, lomask.data(), sizeof(int) * C));
    OFX_TRY(upload(c, m.ellf, ellf.data(), sizeof(int) * ellf.size()));
    OFX_TRY(upload(c, m.elln, elln.data(), sizeof(int) * elln.size()));
    OFX_TRY(upload(c, m.loOrder, loOrd.data(), sizeof(int) * C));
    OFX_TRY(upload(c, m.upOrder, upOrd.data(), sizeof(int) * C));
    OFX_TRY(upload(c, m.loOff, c->sched.loOff.data(), sizeof(int) * (nlo + 1)));
    OFX_TRY(upload(c, m.upOff, c->sched.upOff.data(), sizeof(int) * (nup + 1)));
    if (m.nWall) OFX_TRY(upload(c, m.wallFaces, wall.data(), sizeof(int) * m.nWall));

    // ---- geometry kernels
    const int TB = 256;
    if (m.dim == 2) k_face_metrics_2d<<<(F + TB - 1) / TB, TB, 0, c->stream>>>(m);
    else k_face_metrics_3d<<<(F + TB - 1) / TB, TB, 0, c->stream>>>(m);
    OFX_HIP(hipGetLastError());
    if (m.dim == 2) k_cell_metrics_2d<<<(C + TB - 1) / TB, TB, 0, c->stream>>>(m);
    else k_cell_metrics_3d<<<(C + TB - 1) / TB, TB, 0, c->stream>>>(m);
    OFX_HIP(hipGetLastError());
    if (B) k_ghost_centres<<<(B + TB - 1) / TB, TB, 0, c->stream>>>(m);
    OFX_HIP(hipGetLastError());
    if (m.dim == 2) k_face_static<2><<<(F + TB - 1) / TB, TB, 0, c->stream>>>(m);
    else k_face_static<3><<<(F + TB - 1) / TB, TB, 0, c->stream>>>(m);
    OFX_HIP(hipGetLastError());
    if (C) {
        if (m.dim == 2) k_ell_geom<2><<<(C + TB - 1) / TB, TB, 0, c->stream>>>(m);
        else k_ell_geom<3><<<(C + TB - 1) / TB, TB, 0, c->stream>>>(m);
        OFX_HIP(hipGetLastError());
    }
    k_face_vis_static<<<(F + TB - 1) / TB, TB, 0, c->stream>>>(m);
    OFX_HIP(hipGetLastError());
    // the wall distance is computed on first use (ensure_dist): a mesh that
    // is declared as several zones next searches each zone's walls only
    c->dist_valid = false;
    OFX_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int wall_distance(ofx_ctx* c) {
    const int TB = 256;
    if (c->m.C) k_wall_distance<<<(c->m.C + TB - 1) / TB, TB, 0, c->stream>>>(c->m);
    OFX_HIP(hipGetLastError());
    c->dist_valid = true;
    return 0;
}

}  // namespace ofx
