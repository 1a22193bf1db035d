// oracle.cpp -- CPU oracle of OneFLOW's unstructured NS + SA pseudo-time step.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  Single-threaded restatement of the
// reference loops with the same face order, cell order and per-cell
// accumulation order; compiled with -ffp-contract=off.  Each function names
// the reference function (file:line, relative to the OneFLOW tree) it follows.
//
// The end-to-end pin is the reference's own golden res.dat / turbres.dat /
// aero.dat of test/plateuns2dslau2, test/turbplateuns2droe_sa and
// test/rae2822_roe_sa (tests/test_oracle_golden.py); the options those cases
// do not reach are pinned by closed forms and invariants
// (tests/test_oracle_invariants.py).

#include "oracle.h"

#include "../oneflow_amd/csrc/ofx_math.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

typedef std::vector<double> RV;
typedef std::vector<int> IV;

// codes/basic/include/HXTypeBasic.h:33-34, codes/basic/include/Constant.h:27-57
const double SMALL = 1.0e-40;
const double LARGE = 1.0e40;
const double half = 0.5;
const double two3rd = 2.0 / 3.0;
const double fourth = 0.25;
const double PI = 3.141592653589793;

// codes/geometry/src/Boundary.cpp:35-46
const int BC_INTERFACE = -1;
const int BC_EXTRAPOLATION = 1;
const int BC_SOLID_SURFACE = 2;
const int BC_SYMMETRY = 3;
const int BC_FARFIELD = 4;
const int BC_INFLOW = 5;
const int BC_OUTFLOW = 6;
const int BC_PERIODIC = 99;

// codes/math/include/HXMath.h:186-256
inline double ABS(double a) { return (a < 0) ? -a : a; }
inline double MIN(double a, double b) { return (a < b) ? a : b; }
inline double MAX(double a, double b) { return (a > b) ? a : b; }
inline double SQR(double a) { return a * a; }
inline double SQR(double a, double b) { return a * a + b * b; }
inline double SQR(double a, double b, double c) { return a * a + b * b + c * c; }
inline double DIST(double a, double b, double c) { return std::sqrt(SQR(a, b, c)); }
inline double SIGN(double a, double b) { return a * ABS(b) / (b + SMALL); }

thread_local std::string g_err;
int fail(const std::string& m) { g_err = m; return OFX_EINVAL; }

}  // namespace

struct ofo_case {
    // ---- grid ----
    int dim = 2, nNodes = 0, nFaces = 0, nCells = 0, nBFaces = 0, nTCell = 0;
    RV xn, yn, zn;
    IV fnOff, fnNodes;
    IV lc, rc, bcType;
    RV xfn, yfn, zfn, area, xfc, yfc, zfc, vfn;
    RV xcc, ycc, zcc, vol;
    IV c2fOff, c2f;
    RV dist;
    // ---- config / constants ----
    ofx_config cfg{};
    bool has_cfg = false;
    double statecoef = 0, oreynolds = 0, const_cp = 0, oprl = 0, oprt = 0, twall = 0;
    double inflow[5] = {0, 0, 0, 0, 0};
    double sa_inflow = 0;
    double cfl = 0;
    // SA constants (codes/turb/src/TurbCom.cpp:110-345)
    double karm, karm2, okarm2, cb1, cb2, cv1, cv13, cw2, cw3, cw36, ct3, ct4, cw1, sigma, osigma, sac2, sac3, cb2s, cw1k;
    // ---- fields (SoA [eq][nTCell]) ----
    RV q[5], gama, tempr, visl, vist, res[5], dq[5], timestep, invsr, vissr, impsr;
    RV dqdx[5], dqdy[5], dqdz[5], dtdx, dtdy, dtdz, limiter[5];
    RV bcdqdx[5], bcdqdy[5], bcdqdz[5], bc_q[5];
    RV turbq, turbres, turbdq, turbimpsr, turbdqdx, turbdqdy, turbdqdz, mat_l, mat_r;
    // face flux arrays of the last residual
    RV invflux[5], visflux[5];
    // stale value carried by the reference's global nscom.prim between BC faces
    double bcprim_carry[5] = {0, 0, 0, 0, 0};
};

namespace {

// ===================================================================== grid

// FaceMesh::CalcFaceNormal2D / CalcFaceCenter2D (codes/geometry/src/FaceMesh.cpp:112-155)
void face_metrics_2d(ofo_case& c) {
    for (int f = 0; f < c.nFaces; ++f) {
        int p1 = c.fnNodes[c.fnOff[f]], p2 = c.fnNodes[c.fnOff[f] + 1];
        c.xfn[f] = c.yn[p2] - c.yn[p1];
        c.yfn[f] = c.xn[p1] - c.xn[p2];
        c.zfn[f] = 0.0;
        c.area[f] = DIST(c.xfn[f], c.yfn[f], c.zfn[f]);
    }
    for (int f = 0; f < c.nFaces; ++f) {
        double oArea = 1.0 / (c.area[f] + SMALL);
        c.xfn[f] *= oArea; c.yfn[f] *= oArea; c.zfn[f] *= oArea;
    }
    for (int f = 0; f < c.nFaces; ++f) {
        int p1 = c.fnNodes[c.fnOff[f]], p2 = c.fnNodes[c.fnOff[f] + 1];
        c.xfc[f] = half * (c.xn[p1] + c.xn[p2]);
        c.yfc[f] = half * (c.yn[p1] + c.yn[p2]);
        c.zfc[f] = half * (c.zn[p1] + c.zn[p2]);
    }
}

// FaceMesh::CalcFaceNormal3D / CalcFaceCenter3D (codes/geometry/src/FaceMesh.cpp:157-287)
void face_metrics_3d(ofo_case& c) {
    for (int f = 0; f < c.nFaces; ++f) {
        c.xfn[f] = 0; c.yfn[f] = 0; c.zfn[f] = 0;
        int n = c.fnOff[f + 1] - c.fnOff[f];
        const int* fv = &c.fnNodes[c.fnOff[f]];
        for (int i = 0; i < n; ++i) {
            int p1 = fv[i], p2 = fv[(i + 1) % n];
            double dx1 = c.xn[p1], dy1 = c.yn[p1], dz1 = c.zn[p1];
            double dx2 = c.xn[p2], dy2 = c.yn[p2], dz2 = c.zn[p2];
            c.xfn[f] += half * (dy1 * dz2 - dy2 * dz1);
            c.yfn[f] += half * (dz1 * dx2 - dz2 * dx1);
            c.zfn[f] += half * (dx1 * dy2 - dx2 * dy1);
        }
        c.area[f] = DIST(c.xfn[f], c.yfn[f], c.zfn[f]);
    }
    for (int f = 0; f < c.nFaces; ++f) {
        double oArea = 1.0 / (c.area[f] + SMALL);
        c.xfn[f] *= oArea; c.yfn[f] *= oArea; c.zfn[f] *= oArea;
    }
    const double third = 1.0 / 3.0;
    for (int f = 0; f < c.nFaces; ++f) {
        int n = c.fnOff[f + 1] - c.fnOff[f];
        const int* fv = &c.fnNodes[c.fnOff[f]];
        double x0 = 0, y0 = 0, z0 = 0;
        for (int i = 0; i < n; ++i) { x0 += c.xn[fv[i]]; y0 += c.yn[fv[i]]; z0 += c.zn[fv[i]]; }
        double factor = 1.0 / n;
        x0 *= factor; y0 *= factor; z0 *= factor;
        c.xfc[f] = 0; c.yfc[f] = 0; c.zfc[f] = 0;
        double sarea = 0;
        for (int i = 0; i < n; ++i) {
            int p1 = fv[i], p2 = fv[(i + 1) % n];
            double dx1 = c.xn[p1] - x0, dy1 = c.yn[p1] - y0, dz1 = c.zn[p1] - z0;
            double dx2 = c.xn[p2] - x0, dy2 = c.yn[p2] - y0, dz2 = c.zn[p2] - z0;
            double x00 = third * (c.xn[p1] + c.xn[p2] + x0);
            double y00 = third * (c.yn[p1] + c.yn[p2] + y0);
            double z00 = third * (c.zn[p1] + c.zn[p2] + z0);
            double anx = dy1 * dz2 - dy2 * dz1;
            double any = dz1 * dx2 - dz2 * dx1;
            double anz = dx1 * dy2 - dx2 * dy1;
            double faceArea = half * DIST(anx, any, anz);
            double norm = anx * c.xfn[f] + any * c.yfn[f] + anz * c.zfn[f];
            if (norm < 0.0) faceArea = -faceArea;
            sarea += faceArea;
            c.xfc[f] += faceArea * x00; c.yfc[f] += faceArea * y00; c.zfc[f] += faceArea * z00;
        }
        double osarea = 1.0 / (sarea + SMALL);
        c.xfc[f] *= osarea; c.yfc[f] *= osarea; c.zfc[f] *= osarea;
    }
}

// ghost-cell mirror, shared tail of UnsGrid::CalcCellCenterVol2D/3D
// (codes/geometry/src/UnsGrid.cpp:668-690, 862-887)
void ghost_centres(ofo_case& c) {
    for (int f = 0; f < c.nBFaces; ++f) {
        int l = c.lc[f], r = f + c.nCells;
        if (c.area[f] > SMALL) {
            double tmp = 2.0 * ((c.xcc[l] - c.xfc[f]) * c.xfn[f] + (c.ycc[l] - c.yfc[f]) * c.yfn[f] +
                                (c.zcc[l] - c.zfc[f]) * c.zfn[f]);
            c.xcc[r] = c.xcc[l] - c.xfn[f] * tmp;
            c.ycc[r] = c.ycc[l] - c.yfn[f] * tmp;
            c.zcc[r] = c.zcc[l] - c.zfn[f] * tmp;
        } else {
            c.xcc[r] = -c.xcc[l] + 2.0 * c.xfc[f];
            c.ycc[r] = -c.ycc[l] + 2.0 * c.yfc[f];
            c.zcc[r] = -c.zcc[l] + 2.0 * c.zfc[f];
        }
        c.vol[r] = c.vol[l];
    }
}

// UnsGrid::CalcCellCenterVol2D (codes/geometry/src/UnsGrid.cpp:588-690)
void cell_metrics_2d(ofo_case& c) {
    std::fill(c.xcc.begin(), c.xcc.end(), 0.0);
    std::fill(c.ycc.begin(), c.ycc.end(), 0.0);
    std::fill(c.zcc.begin(), c.zcc.end(), 0.0);
    std::fill(c.vol.begin(), c.vol.end(), 0.0);
    for (int f = 0; f < c.nBFaces; ++f) {
        int l = c.lc[f];
        double dot = (c.xfc[f] * c.xfn[f] + c.yfc[f] * c.yfn[f] + c.zfc[f] * c.zfn[f]) * c.area[f];
        c.xcc[l] += c.xfc[f] * dot; c.ycc[l] += c.yfc[f] * dot; c.zcc[l] += c.zfc[f] * dot;
        c.vol[l] += dot;
    }
    for (int f = c.nBFaces; f < c.nFaces; ++f) {
        int l = c.lc[f], r = c.rc[f];
        double dot = (c.xfc[f] * c.xfn[f] + c.yfc[f] * c.yfn[f] + c.zfc[f] * c.zfn[f]) * c.area[f];
        c.xcc[l] += c.xfc[f] * dot; c.ycc[l] += c.yfc[f] * dot; c.zcc[l] += c.zfc[f] * dot;
        c.vol[l] += dot;
        c.xcc[r] -= c.xfc[f] * dot; c.ycc[r] -= c.yfc[f] * dot; c.zcc[r] -= c.zfc[f] * dot;
        c.vol[r] -= dot;
    }
    for (int i = 0; i < c.nCells; ++i) {
        double tmp = 1.0 / (1.5 * c.vol[i] + SMALL);
        c.xcc[i] *= tmp; c.ycc[i] *= tmp; c.zcc[i] *= tmp;
        c.vol[i] *= half;
        if (c.vol[i] <= 0.0) c.vol[i] = -c.vol[i];
    }
    ghost_centres(c);
}

// UnsGrid::CalcCellCenterVol3D (codes/geometry/src/UnsGrid.cpp:740-887)
void cell_metrics_3d(ofo_case& c) {
    std::fill(c.xcc.begin(), c.xcc.end(), 0.0);
    std::fill(c.ycc.begin(), c.ycc.end(), 0.0);
    std::fill(c.zcc.begin(), c.zcc.end(), 0.0);
    std::fill(c.vol.begin(), c.vol.end(), 0.0);
    for (int f = 0; f < c.nFaces; ++f) {
        int l = c.lc[f], r = c.rc[f];
        int n = c.fnOff[f + 1] - c.fnOff[f];
        const int* fv = &c.fnNodes[c.fnOff[f]];
        for (int i = 0; i < n; ++i) {
            int p2 = fv[i], p3 = fv[(i + 1) % n];
            double x21 = c.xn[p2] - c.xfc[f], y21 = c.yn[p2] - c.yfc[f], z21 = c.zn[p2] - c.zfc[f];
            double x31 = c.xn[p3] - c.xfc[f], y31 = c.yn[p3] - c.yfc[f], z31 = c.zn[p3] - c.zfc[f];
            double nx = y21 * z31 - y31 * z21;
            double ny = z21 * x31 - z31 * x21;
            double nz = x21 * y31 - x31 * y21;
            double trix = c.xfc[f] + c.xn[p2] + c.xn[p3];
            double triy = c.yfc[f] + c.yn[p2] + c.yn[p3];
            double triz = c.zfc[f] + c.zn[p2] + c.zn[p3];
            double tmp = nx * trix + ny * triy + nz * triz;
            trix *= tmp; triy *= tmp; triz *= tmp;
            c.xcc[l] += trix; c.ycc[l] += triy; c.zcc[l] += triz; c.vol[l] += tmp;
            c.xcc[r] -= trix; c.ycc[r] -= triy; c.zcc[r] -= triz; c.vol[r] -= tmp;
        }
    }
    for (int i = 0; i < c.nCells; ++i) {
        double tmp = 1.0 / (4.0 * c.vol[i] + SMALL);
        c.xcc[i] *= tmp; c.ycc[i] *= tmp; c.zcc[i] *= tmp;
        c.vol[i] /= 18.0;
        if (c.vol[i] <= 0.0) c.vol[i] = -c.vol[i];
    }
    ghost_centres(c);
}

// CellTopo::CalcC2f (codes/geometry/src/CellTopo.cpp:95-119): boundary faces
// first, then interior, each in increasing face id -> ascending lists.
void build_c2f(ofo_case& c) {
    IV cnt(c.nTCell + 1, 0);
    for (int f = 0; f < c.nBFaces; ++f) cnt[c.lc[f] + 1]++;
    for (int f = c.nBFaces; f < c.nFaces; ++f) { cnt[c.lc[f] + 1]++; cnt[c.rc[f] + 1]++; }
    for (int i = 0; i < c.nTCell; ++i) cnt[i + 1] += cnt[i];
    c.c2fOff = cnt;
    c.c2f.assign(cnt[c.nTCell], 0);
    IV pos(cnt.begin(), cnt.end() - 1);
    for (int f = 0; f < c.nBFaces; ++f) c.c2f[pos[c.lc[f]]++] = f;
    for (int f = c.nBFaces; f < c.nFaces; ++f) { c.c2f[pos[c.lc[f]]++] = f; c.c2f[pos[c.rc[f]]++] = f; }
}

// gte DCPQuery point-segment (codes/gte/include/GteDistPointSegment.h), squared distance
double seg_dist2(const double p[3], const double a[3], const double b[3]) {
    double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double diff[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    double t = d[0] * diff[0] + d[1] * diff[1] + d[2] * diff[2];
    double cl[3];
    if (t >= 0.0) { cl[0] = b[0]; cl[1] = b[1]; cl[2] = b[2]; }
    else {
        diff[0] = p[0] - a[0]; diff[1] = p[1] - a[1]; diff[2] = p[2] - a[2];
        t = d[0] * diff[0] + d[1] * diff[1] + d[2] * diff[2];
        if (t <= 0.0) { cl[0] = a[0]; cl[1] = a[1]; cl[2] = a[2]; }
        else {
            double sl = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (sl > 0.0) {
                t /= sl;
                cl[0] = a[0] + t * d[0]; cl[1] = a[1] + t * d[1]; cl[2] = a[2] + t * d[2];
            } else { cl[0] = a[0]; cl[1] = a[1]; cl[2] = a[2]; }
        }
    }
    diff[0] = p[0] - cl[0]; diff[1] = p[1] - cl[1]; diff[2] = p[2] - cl[2];
    return diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2];
}

// gte DistancePointTriangleExact (codes/gte/include/GteDistPointTriangleExact.h):
// minimise |B + s E0 + t E1 - P|^2 over the triangle, region-based closed form.
double tri_dist2(const double P[3], const double V0[3], const double V1[3], const double V2[3]) {
    double diff[3] = {P[0] - V0[0], P[1] - V0[1], P[2] - V0[2]};
    double e0[3] = {V1[0] - V0[0], V1[1] - V0[1], V1[2] - V0[2]};
    double e1[3] = {V2[0] - V0[0], V2[1] - V0[1], V2[2] - V0[2]};
    auto dot = [](const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
    double a00 = dot(e0, e0), a01 = dot(e0, e1), a11 = dot(e1, e1);
    double b0 = -dot(diff, e0), b1 = -dot(diff, e1);
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
    return dot(d, d);
}

// CalcWallDist + CalcPoint2FaceDist (codes/geometry/src/WallDist.cpp:136-358)
void wall_distance(ofo_case& c) {
    std::vector<int> wf;
    for (int f = 0; f < c.nBFaces; ++f) if (c.bcType[f] == BC_SOLID_SURFACE) wf.push_back(f);
    c.dist.assign(c.nCells, LARGE);
    for (int i = 0; i < c.nCells; ++i) {
        double p[3] = {c.xcc[i], c.ycc[i], c.zcc[i]};
        for (int f : wf) {
            int n = c.fnOff[f + 1] - c.fnOff[f];
            const int* fv = &c.fnNodes[c.fnOff[f]];
            double w;
            if (n <= 2) {
                double a[3] = {c.xn[fv[0]], c.yn[fv[0]], c.zn[fv[0]]};
                double b[3] = {c.xn[fv[1]], c.yn[fv[1]], c.zn[fv[1]]};
                w = seg_dist2(p, a, b);
            } else if (n == 3) {
                double a[3] = {c.xn[fv[0]], c.yn[fv[0]], c.zn[fv[0]]};
                double b[3] = {c.xn[fv[1]], c.yn[fv[1]], c.zn[fv[1]]};
                double d[3] = {c.xn[fv[2]], c.yn[fv[2]], c.zn[fv[2]]};
                w = tri_dist2(p, a, b, d);
            } else {
                double ctr[3] = {0, 0, 0};
                for (int k = 0; k < n; ++k) { ctr[0] += c.xn[fv[k]]; ctr[1] += c.yn[fv[k]]; ctr[2] += c.zn[fv[k]]; }
                double coef = 1.0 / n;
                ctr[0] *= coef; ctr[1] *= coef; ctr[2] *= coef;
                w = LARGE;
                for (int k = 0; k < n; ++k) {
                    int p0 = fv[k], p1 = fv[(k + 1) % n];
                    double a[3] = {c.xn[p0], c.yn[p0], c.zn[p0]};
                    double b[3] = {c.xn[p1], c.yn[p1], c.zn[p1]};
                    w = MIN(w, tri_dist2(p, a, b, ctr));
                }
            }
            if (c.dist[i] > w) c.dist[i] = w;
        }
    }
    for (int i = 0; i < c.nCells; ++i) c.dist[i] = std::sqrt(c.dist[i]);
}

void alloc_fields(ofo_case& c) {
    int T = c.nTCell, B = c.nBFaces, F = c.nFaces;
    for (int e = 0; e < 5; ++e) {
        c.q[e].assign(T, 0); c.res[e].assign(T, 0); c.dq[e].assign(T, 0);
        c.dqdx[e].assign(T, 0); c.dqdy[e].assign(T, 0); c.dqdz[e].assign(T, 0);
        c.limiter[e].assign(T, 0);
        c.bcdqdx[e].assign(B, 0); c.bcdqdy[e].assign(B, 0); c.bcdqdz[e].assign(B, 0); c.bc_q[e].assign(B, 0);
        c.invflux[e].assign(F, 0); c.visflux[e].assign(F, 0);
    }
    c.gama.assign(T, 0); c.tempr.assign(T, 0); c.visl.assign(T, 0); c.vist.assign(T, 0);
    c.timestep.assign(T, 0); c.invsr.assign(T, 0); c.vissr.assign(T, 0); c.impsr.assign(T, 0);
    c.dtdx.assign(T, 0); c.dtdy.assign(T, 0); c.dtdz.assign(T, 0);
    c.turbq.assign(T, 0); c.turbres.assign(T, 0); c.turbdq.assign(T, 0); c.turbimpsr.assign(T, 0);
    c.turbdqdx.assign(T, 0); c.turbdqdy.assign(T, 0); c.turbdqdz.assign(T, 0);
    c.mat_l.assign(F, 0); c.mat_r.assign(F, 0);
}

// ============================================================ constants

// NsCom::Init + Chemical::CalcStateCoefNs/CalcRefPrim/CalcRefReynolds
// (codes/ns/src/NsCom.cpp:44-122, codes/chemical/src/Chemical.cpp:280-338)
// and TurbCom::InitConst/InitInflow (codes/turb/src/TurbCom.cpp:110-380).
void set_constants(ofo_case& c) {
    const ofx_config& g = c.cfg;
    double aoa = g.aoa_degree * PI / 180.0;
    double aos = g.sideslip_degree * PI / 180.0;
    c.statecoef = 1.0 / (g.gama_ref * SQR(g.mach_ref));
    c.oreynolds = 1.0 / g.reynolds;
    c.oprl = 1.0 / g.prl;
    c.oprt = 1.0 / g.prt;
    c.twall = g.twall_dim / g.tref_dim;
    c.const_cp = 1.0 / ((g.gama_ref - 1.0) * SQR(g.mach_ref));
    double pref = c.statecoef * 1.0 * 1.0 / 1.0;
    c.inflow[0] = 1.0;
    c.inflow[1] = 1.0 * std::cos(aoa) * std::cos(aos);
    c.inflow[2] = 1.0 * std::sin(aoa) * std::cos(aos);
    c.inflow[3] = 1.0 * std::sin(aos);
    c.inflow[4] = pref;
    c.sa_inflow = g.ref_sa;
    c.sigma = 2.0 / 3.0;
    c.osigma = 1.0 / c.sigma;
    c.karm = 0.41;
    c.karm2 = SQR(c.karm);
    c.okarm2 = 1.0 / c.karm2;
    c.cb1 = 0.1355;
    c.cb2 = 0.622;
    c.cv1 = 7.1;
    c.cv13 = c.cv1 * c.cv1 * c.cv1;
    c.cw2 = 0.3;
    c.cw3 = 2.0;
    c.cw36 = std::pow(c.cw3, 6);
    c.ct3 = 1.2;
    c.ct4 = 0.5;
    c.cw1 = c.cb1 * c.okarm2 + (1.0 + c.cb2) / c.sigma;
    c.sac2 = 0.7;
    c.sac3 = 0.9;
    c.cb2s = c.osigma * c.cb2;
    c.cw1k = c.cw1 * c.karm2;
}

// ======================================================= boundary / state

// NsCalcGamaT (codes/ns/src/NsRhs.cpp:77-104)
void calc_gama_t(ofo_case& c, int ist, int ied) {
    for (int i = ist; i < ied; ++i) {
        c.gama[i] = c.cfg.gama_ref;
        c.tempr[i] = c.q[4][i] / (c.statecoef * c.q[0][i] * 1.0);
    }
}

// CalcLaminarViscosity + Sutherland::CalcViscosity
// (codes/uns/src/UNsVisFlux.cpp:442-459, codes/model/src/Sutherland.cpp)
void calc_visl(ofo_case& c, int ist, int ied) {
    double ts = 110.4 / c.cfg.tref_dim;
    for (int i = ist; i < ied; ++i) {
        double t = c.tempr[i];
        double t3 = t * t * t;
        double v = std::sqrt(t3) * (1 + ts) / (t + ts);
        c.visl[i] = MAX(0.0, v);
    }
}

// UNsBcSolver::CalcBc + NsBcSolver (codes/uns/src/UNsBcSolver.cpp:55-168,
// codes/ns/src/NsBcSolver.cpp:51-410).  Regions are independent (each ghost
// reads only its owner), so face order is used; bc_q keeps the reference's
// carry of nscom.prim for BCs that do not set it (only wall faces read it).
void ns_bc(ofo_case& c) {
    const ofx_config& g = c.cfg;
    for (int f = 0; f < c.nBFaces; ++f) {
        int bt = c.bcType[f];
        if (bt < 0 || bt == BC_PERIODIC) continue;  // interface: no update
        int l = c.lc[f], r = c.rc[f];
        double xfn = c.xfn[f], yfn = c.yfn[f], zfn = c.zfn[f], vfn = c.vfn[f];
        double s[5], t[5];
        for (int e = 0; e < 5; ++e) s[e] = c.q[e][l];
        double gama = half * (c.gama[l] + c.gama[l]);
        double* prim = c.bcprim_carry;
        if (bt == BC_EXTRAPOLATION || bt == BC_OUTFLOW) {
            for (int e = 0; e < 5; ++e) t[e] = s[e];
        } else if (bt == BC_INFLOW) {
            for (int e = 0; e < 5; ++e) t[e] = c.inflow[e];
        } else if (bt == BC_SOLID_SURFACE && g.vismodel != 0) {
            for (int e = 0; e < 5; ++e) t[e] = s[e];
            if (g.isowallbc == 0) {
                // AdiabaticVisWallBc (NsBcSolver.cpp:350-362), static wall vfx=vfy=vfz=0
                t[1] = -t[1] + 2.0 * 0.0; t[2] = -t[2] + 2.0 * 0.0; t[3] = -t[3] + 2.0 * 0.0;
                prim[0] = s[0]; prim[1] = 0.0; prim[2] = 0.0; prim[3] = 0.0; prim[4] = s[4];
            } else {
                // IsothermalVisWallBc (NsBcSolver.cpp:271-324)
                t[1] = -t[1] + 2.0 * 0.0; t[2] = -t[2] + 2.0 * 0.0; t[3] = -t[3] + 2.0 * 0.0;
                double twall = g.twall_dim / g.tref_dim;
                double tlim = 0.1 * twall;
                double rm = s[0], pm = s[4];
                double temperature = c.tempr[l];
                double rw_face = pm / (c.statecoef * twall * 1.0);
                prim[0] = rw_face; prim[1] = 0.0; prim[2] = 0.0; prim[3] = 0.0; prim[4] = pm;
                double pg1 = pm;
                double tg1 = 2.0 * twall - temperature;
                if (tg1 < tlim) tg1 = tlim;
                double rw = pg1 / (c.statecoef * twall * 1.0);
                double rg1 = 2.0 * rw - rm;
                rg1 = MAX(0.5 * rw, rg1);
                t[0] = rg1; t[4] = pg1;
            }
        } else if (bt == BC_SYMMETRY || bt == BC_SOLID_SURFACE) {
            // SymmetryBc (NsBcSolver.cpp:364-395)
            double vn = xfn * s[1] + yfn * s[2] + zfn * s[3] - vfn;
            for (int e = 0; e < 5; ++e) t[e] = s[e];
            t[1] = s[1] - 2.0 * xfn * vn;
            t[2] = s[2] - 2.0 * yfn * vn;
            t[3] = s[3] - 2.0 * zfn * vn;
            for (int e = 0; e < 5; ++e) prim[e] = half * (s[e] + t[e]);
        } else if (bt == BC_FARFIELD) {
            // FarFieldBc (NsBcSolver.cpp:140-236)
            for (int e = 0; e < 5; ++e) t[e] = s[e];
            double rin = s[0], uin = s[1], vin = s[2], win = s[3], pin = s[4];
            double rref = c.inflow[0], uref = c.inflow[1], vref = c.inflow[2], wref = c.inflow[3], pref = c.inflow[4];
            double vnref = xfn * uref + yfn * vref + zfn * wref - vfn;
            double vnin = xfn * uin + yfn * vin + zfn * win - vfn;
            double cref = std::sqrt(ABS(g.gama_ref * pref / rref));
            double cin = std::sqrt(ABS(gama * pin / rin));
            double gamm1 = gama - 1.0;
            double velin = DIST(uin, vin, win);
            if (velin > cin) {
                if (vnin >= 0.0) {
                    for (int e = 0; e < 5; ++e) t[e] = s[e];
                } else {
                    for (int e = 0; e < 5; ++e) t[e] = c.inflow[e];
                }
            } else {
                double riemp = vnin + 2.0 * cin / gamm1;
                double riemm = vnref - 2.0 * cref / gamm1;
                double vnb = half * (riemp + riemm);
                double cb = fourth * (riemp - riemm) * gamm1;
                double vtx, vty, vtz, entr;
                if (vnb >= 0.0) {
                    entr = pin / ofxm::ofx_pow(rin, gama);
                    vtx = uin - xfn * vnin; vty = vin - yfn * vnin; vtz = win - zfn * vnin;
                } else {
                    entr = pref / ofxm::ofx_pow(rref, gama);
                    vtx = uref - xfn * vnref; vty = vref - yfn * vnref; vtz = wref - zfn * vnref;
                }
                double rb = ofxm::ofx_pow((cb * cb / (entr * gama)), 1.0 / gamm1);
                double ub = vtx + xfn * vnb;
                double vb = vty + yfn * vnb;
                double wb = vtz + zfn * vnb;
                double pb = cb * cb * rb / gama;
                t[0] = rb; t[1] = ub; t[2] = vb; t[3] = wb; t[4] = pb;
            }
        } else {
            g_err = "unsupported boundary type";
            continue;
        }
        for (int e = 0; e < 5; ++e) c.q[e][r] = t[e];
        for (int e = 0; e < 5; ++e) c.bc_q[e][f] = prim[e];
    }
}

// NsCalcBoundary (codes/ns/src/NsSolverImp.cpp:72-79)
void ns_boundary(ofo_case& c) {
    calc_gama_t(c, 0, c.nCells);
    calc_visl(c, 0, c.nCells);
    ns_bc(c);
    calc_gama_t(c, c.nCells, c.nTCell);
    calc_visl(c, c.nCells, c.nTCell);
}

// UTurbSrcFlux::CalcVist1Equ + SetGhostCellVist + TurbSrcFlux::CalcCellVist1Equ
// (codes/uturb/src/UTurbSrcFlux.cpp:101-125,217-241; codes/turb/src/TurbSrcFlux.cpp:279-290)
void sa_vist(ofo_case& c) {
    for (int i = 0; i < c.nCells; ++i) {
        double rho = c.q[0][i], nuet = c.turbq[i], visl = c.visl[i];
        double olam = rho / (visl + SMALL);
        double xsi = nuet * olam;
        double xsi3 = xsi * xsi * xsi;
        double fv1 = xsi3 / (xsi3 + c.cv13);
        double vist = fv1 * rho * nuet;
        c.vist[i] = MIN(visl * c.cfg.max_vis_ratio, vist);
    }
    for (int f = 0; f < c.nBFaces; ++f) c.vist[c.rc[f]] = c.vist[c.lc[f]];
}

// UTurbBcSolver + TurbBcSolver (codes/uturb/src/UTurbBcSolver.cpp:59-160,
// codes/turb/src/TurbBcSolver.cpp:142-253)
void sa_bc(ofo_case& c) {
    const ofx_config& g = c.cfg;
    for (int f = 0; f < c.nBFaces; ++f) {
        int bt = c.bcType[f];
        if (bt < 0 || bt == BC_PERIODIC) continue;
        int l = c.lc[f], r = c.rc[f];
        double s = c.turbq[l], t;
        if (bt == BC_INFLOW) t = c.sa_inflow;
        else if (bt == BC_SOLID_SURFACE) t = -s;
        else if (bt == BC_FARFIELD) {
            double xfn = c.xfn[f], yfn = c.yfn[f], zfn = c.zfn[f], vfn = c.vfn[f];
            double rin = c.q[0][l], uin = c.q[1][l], vin = c.q[2][l], win = c.q[3][l], pin = c.q[4][l];
            double rref = c.inflow[0], uref = c.inflow[1], vref = c.inflow[2], wref = c.inflow[3], pref = c.inflow[4];
            double vnref = xfn * uref + yfn * vref + zfn * wref - vfn;
            double vnin = xfn * uin + yfn * vin + zfn * win - vfn;
            double gama = g.gama_ref;  // reference reads the global nscom.gama == gama_ref
            double cref = std::sqrt(ABS(g.gama_ref * pref / rref));
            double cin = std::sqrt(ABS(gama * pin / rin));
            double velin = DIST(uin, vin, win);
            if (velin > cin) {
                t = (vnin >= 0.0) ? s : c.sa_inflow;
            } else {
                double gamm1 = gama - 1.0;
                double riemp = vnin + 2.0 * cin / gamm1;
                double riemm = vnref - 2.0 * cref / gamm1;
                double vnb = half * (riemp + riemm);
                t = (vnb >= 0.0) ? s : c.sa_inflow;
            }
        } else t = s;  // extrapolation, outflow, symmetry, pole
        c.turbq[r] = t;
    }
}

// ============================================================ gradients

// CalcGradGGCellWeight (codes/usolver/src/UGrad.cpp:88-149)
void grad_gg(const ofo_case& c, const RV& q, RV& dx, RV& dy, RV& dz) {
    std::fill(dx.begin(), dx.end(), 0.0);
    std::fill(dy.begin(), dy.end(), 0.0);
    std::fill(dz.begin(), dz.end(), 0.0);
    for (int f = 0; f < c.nFaces; ++f) {
        int l = c.lc[f], r = c.rc[f];
        double dxl = c.xfc[f] - c.xcc[l], dyl = c.yfc[f] - c.ycc[l], dzl = c.zfc[f] - c.zcc[l];
        double dxr = c.xfc[f] - c.xcc[r], dyr = c.yfc[f] - c.ycc[r], dzr = c.zfc[f] - c.zcc[r];
        double delt1 = DIST(dxl, dyl, dzl);
        double delt2 = DIST(dxr, dyr, dzr);
        double delta = 1.0 / (delt1 + delt2 + SMALL);
        double cl = delt2 * delta, cr = delt1 * delta;
        double value = cl * q[l] + cr * q[r];
        double fnxa = c.xfn[f] * c.area[f], fnya = c.yfn[f] * c.area[f], fnza = c.zfn[f] * c.area[f];
        dx[l] += fnxa * value; dy[l] += fnya * value; dz[l] += fnza * value;
        if (f < c.nBFaces) continue;
        dx[r] -= fnxa * value; dy[r] -= fnya * value; dz[r] -= fnza * value;
    }
    for (int i = 0; i < c.nCells; ++i) {
        double ovol = 1.0 / c.vol[i];
        dx[i] *= ovol; dy[i] *= ovol; dz[i] *= ovol;
    }
    for (int f = 0; f < c.nBFaces; ++f) {
        int l = c.lc[f], r = c.rc[f];
        dx[r] = dx[l]; dy[r] = dy[l]; dz[r] = dz[l];
    }
}

// ============================================================== limiter

// Limiter::CalcLimiter / CalcMinMaxDiff / CalcLocal*Limiter, Vencat, BarthFunction
// (codes/usolver/src/ULimiter.cpp:35-97,377-644)
double vencat(double x, double y, double eps) {
    double x2 = x * x, xy = x * y, y2 = y * y, cc = 2.0;
    return ((x2 + eps + cc * xy) / (x2 + cc * y2 + xy + eps));
}
double vencat_fn(double dq, double minv, double maxv, double eps) {
    if (dq > SMALL) return vencat(maxv, dq, eps);
    else if (dq < -SMALL) return vencat(minv, dq, eps);
    return 1.0;
}
double barth_fn(double dq, double minv, double maxv, double dot) {
    if (dot < 0.0) return 0.0;
    else if (dq > maxv) return maxv / dq;
    else if (dq < minv) return minv / dq;
    return 1.0;
}

void calc_limiter(ofo_case& c, int limflag, RV* lim, const RV* qv, const RV* gx, const RV* gy, const RV* gz, int nEqu) {
    RV minvf(c.nTCell), maxvf(c.nTCell);
    for (int e = 0; e < nEqu; ++e) {
        RV& L = lim[e];
        const RV& q = qv[e];
        for (int i = 0; i < c.nTCell; ++i) L[i] = 1.0;
        if (limflag == -1 || limflag == 0) {
            double v = (limflag == -1) ? 0.0 : 1.0;
            for (int f = 0; f < c.nFaces; ++f) { L[c.lc[f]] = v; L[c.rc[f]] = v; }
            continue;
        }
        std::fill(minvf.begin(), minvf.end(), 0.0);
        std::fill(maxvf.begin(), maxvf.end(), 0.0);
        for (int i = 0; i < c.nCells; ++i) { minvf[i] = q[i]; maxvf[i] = q[i]; }
        for (int f = 0; f < c.nBFaces; ++f) {
            if (c.bcType[f] != BC_INTERFACE) continue;
            int l = c.lc[f], r = c.rc[f];
            minvf[l] = MIN(minvf[l], q[r]); maxvf[l] = MAX(maxvf[l], q[r]);
        }
        for (int f = c.nBFaces; f < c.nFaces; ++f) {
            int l = c.lc[f], r = c.rc[f];
            minvf[l] = MIN(minvf[l], q[r]); maxvf[l] = MAX(maxvf[l], q[r]);
            minvf[r] = MIN(minvf[r], q[l]); maxvf[r] = MAX(maxvf[r], q[l]);
        }
        for (int i = 0; i < c.nCells; ++i) { minvf[i] -= q[i]; maxvf[i] -= q[i]; }
        double qmin = LARGE, qmax = -LARGE;
        for (int i = 0; i < c.nCells; ++i) { qmin = MIN(qmin, minvf[i]); qmax = MAX(qmax, maxvf[i]); }
        double eps = c.cfg.vencat_coef * MAX(qmax - qmin, 1.0);
        double eps1 = SQR(eps) + SMALL;
        for (int f = 0; f < c.nFaces; ++f) {
            int l = c.lc[f], r = c.rc[f];
            double lim1 = L[l], lim2 = L[r];
            double dx1 = c.xfc[f] - c.xcc[l], dy1 = c.yfc[f] - c.ycc[l], dz1 = c.zfc[f] - c.zcc[l];
            double dq1 = gx[e][l] * dx1 + gy[e][l] * dy1 + gz[e][l] * dz1;
            double dx2 = c.xfc[f] - c.xcc[r], dy2 = c.yfc[f] - c.ycc[r], dz2 = c.zfc[f] - c.zcc[r];
            double dq2 = gx[e][r] * dx2 + gy[e][r] * dy2 + gz[e][r] * dz2;
            if (limflag == 2) {
                lim1 = MIN(lim1, vencat_fn(dq1, minvf[l], maxvf[l], eps1));
                lim2 = MIN(lim2, vencat_fn(dq2, minvf[r], maxvf[r], eps1));
            } else {
                double ds1 = DIST(dx1, dy1, dz1);
                double dot1 = (c.xfn[f] * dx1 + c.yfn[f] * dy1 + c.zfn[f] * dz1) / (ds1 + SMALL);
                lim1 = MIN(lim1, barth_fn(dq1, minvf[l], maxvf[l], dot1));
                double ds2 = DIST(dx2, dy2, dz2);
                double dot2 = -(c.xfn[f] * dx2 + c.yfn[f] * dy2 + c.zfn[f] * dz2) / (ds2 + SMALL);
                lim2 = MIN(lim2, barth_fn(dq2, minvf[r], maxvf[r], dot2));
            }
            L[l] = lim1; L[r] = lim2;
        }
    }
}

// LimField::GetQlQr + CalcFaceValue (ireconface = 1) + NsLimField::BcQlQrFix
// (codes/usolver/src/ULimiter.cpp:131-256, codes/uns/src/UNsLimiter.cpp:67-101)
void face_values_ns(const ofo_case& c, RV* qf1, RV* qf2) {
    for (int f = 0; f < c.nFaces; ++f) {
        int l = c.lc[f], r = c.rc[f];
        double t[5];
        double dx = c.xfc[f] - c.xcc[l], dy = c.yfc[f] - c.ycc[l], dz = c.zfc[f] - c.zcc[l];
        for (int e = 0; e < 5; ++e) t[e] = c.q[e][l];
        for (int e = 0; e < 5; ++e) {
            double phi = MIN(c.limiter[e][l], c.limiter[e][r]);
            t[e] += phi * (c.dqdx[e][l] * dx + c.dqdy[e][l] * dy + c.dqdz[e][l] * dz);
        }
        bool ok = !(t[0] <= 0.0 || t[4] <= 0.0);
        for (int e = 0; e < 5; ++e) qf1[e][f] = ok ? t[e] : c.q[e][l];
        dx = c.xfc[f] - c.xcc[r]; dy = c.yfc[f] - c.ycc[r]; dz = c.zfc[f] - c.zcc[r];
        for (int e = 0; e < 5; ++e) t[e] = c.q[e][r];
        for (int e = 0; e < 5; ++e) {
            double phi = MIN(c.limiter[e][r], c.limiter[e][l]);
            t[e] += phi * (c.dqdx[e][r] * dx + c.dqdy[e][r] * dy + c.dqdz[e][r] * dz);
        }
        ok = !(t[0] <= 0.0 || t[4] <= 0.0);
        for (int e = 0; e < 5; ++e) qf2[e][f] = ok ? t[e] : c.q[e][r];
    }
    for (int f = 0; f < c.nBFaces; ++f) {
        int bt = c.bcType[f];
        if (bt == BC_INTERFACE || bt == BC_PERIODIC) continue;
        int l = c.lc[f], r = c.rc[f];
        for (int e = 0; e < 5; ++e) {
            double tmp = half * (c.q[e][l] + c.q[e][r]);
            qf1[e][f] = tmp; qf2[e][f] = tmp;
        }
        if (bt == BC_SOLID_SURFACE)
            for (int e = 0; e < 5; ++e) { qf1[e][f] = c.bc_q[e][f]; qf2[e][f] = c.bc_q[e][f]; }
    }
}

// ========================================================= inviscid flux

// PrimToQ (codes/ns/src/NsInvFlux.cpp:1349-1370) with CalcInternalEnergy (:1392-1406)
void prim_to_q(const double* p, double gama, double* q) {
    double v2 = SQR(p[1], p[2], p[3]);
    double em = (p[4] / p[0]) / (gama - 1.0) + half * v2;
    q[0] = p[0]; q[1] = p[0] * p[1]; q[2] = p[0] * p[2]; q[3] = p[0] * p[3]; q[4] = p[0] * em;
}

// NsInvFlux::Roe + ModifyAbsoluteEigenvalue + CalcTotalEnthalpyChange
// (codes/ns/src/NsInvFlux.cpp:167-308,1329-1347)
void roe_flux(const ofx_config& g, const double* p1, const double* p2, double gama1, double gama2,
              double xfn, double yfn, double zfn, double vfn, double* flux) {
    double rl = p1[0], ul = p1[1], vl = p1[2], wl = p1[3], pl = p1[4];
    double rr = p2[0], ur = p2[1], vr = p2[2], wr = p2[3], pr = p2[4];
    double v2l = SQR(ul, vl, wl), v2r = SQR(ur, vr, wr);
    double hint_l = (gama1 / (gama1 - 1.0)) * (pl / rl);
    double hint_r = (gama2 / (gama2 - 1.0)) * (pr / rr);
    double hl = hint_l + half * v2l, hr = hint_r + half * v2r;
    double vnl = xfn * ul + yfn * vl + zfn * wl - vfn;
    double vnr = xfn * ur + yfn * vr + zfn * wr - vfn;
    double rvnl = rl * vnl, rvnr = rr * vnr;
    double ratio = std::sqrt(rr / rl);
    double coef = 1.0 / (1.0 + ratio);
    double rm = std::sqrt(rl * rr);
    double um = (ul + ur * ratio) * coef;
    double vm = (vl + vr * ratio) * coef;
    double wm = (wl + wr * ratio) * coef;
    double hm = (hl + hr * ratio) * coef;
    double pm = (pl + pr * ratio) * coef;
    double gama = (gama1 + gama2 * ratio) * coef;
    double v2 = SQR(um, vm, wm);
    double vnflow = xfn * um + yfn * vm + zfn * wm;
    double vnrel = vnflow - vfn;
    double gamm1 = gama - 1.0;
    double c2 = gamm1 * (hm - half * v2);
    double cm = std::sqrt(ABS(c2));
    double aeig1 = ABS(vnrel), aeig2 = ABS(vnrel + cm), aeig3 = ABS(vnrel - cm);
    double f1[5] = {rvnl, rvnl * ul + xfn * pl, rvnl * vl + yfn * pl, rvnl * wl + zfn * pl, rvnl * hl + vfn * pl};
    double f2[5] = {rvnr, rvnr * ur + xfn * pr, rvnr * vr + yfn * pr, rvnr * wr + zfn * pr, rvnr * hr + vfn * pr};
    for (int e = 0; e < 5; ++e) flux[e] = f1[e] + f2[e];
    double q1[5], q2[5], dq[5];
    prim_to_q(p1, gama1, q1);
    prim_to_q(p2, gama2, q2);
    for (int e = 0; e < 5; ++e) dq[e] = q2[e] - q1[e];
    double meig1 = aeig1, meig2 = aeig2, meig3 = aeig3;
    if (g.ieigenfix == 1) {
        double maxE = MAX(aeig2, aeig3);
        double m1 = maxE * g.centropy1, m2 = maxE * g.centropy2;
        meig1 = MAX(m1, aeig1); meig2 = MAX(m2, aeig2); meig3 = MAX(m2, aeig3);
    }
    double xi1 = (2.0 * meig1 - meig2 - meig3) / (2.0 * c2);
    double xi2 = (meig2 - meig3) / (2.0 * cm);
    double dc = vnflow * dq[0] - xfn * dq[1] - yfn * dq[2] - zfn * dq[3];
    double c2dc = c2 * dc;
    // CalcTotalEnthalpyChange on prim = (rm, um, vm, wm, pm)
    double ae = gama - 1.0;
    double af = half * ae * v2;
    double dh = -ae * (um * dq[1] + vm * dq[2] + wm * dq[3] - dq[4]);
    dh += af * dq[0];
    double term1 = dh * xi1 + dc * xi2;
    double term2 = c2dc * xi1 + dh * xi2;
    flux[0] -= (meig1 * dq[0] - term1);
    flux[1] -= (meig1 * dq[1] - um * term1 + xfn * term2);
    flux[2] -= (meig1 * dq[2] - vm * term1 + yfn * term2);
    flux[3] -= (meig1 * dq[3] - wm * term1 + zfn * term2);
    flux[4] -= (meig1 * dq[4] - hm * term1 + vnflow * term2);
    (void)rm; (void)pm;
    for (int e = 0; e < 5; ++e) flux[e] *= half;
}

// FMSplit2, FPSplit5 (codes/ns/src/NsInvFlux.cpp:42-65)
double fm_split2(double mach, double ipn) { return ipn * fourth * SQR(mach + ipn); }
double fp_split5(double mach, double alpha, double ipn) {
    double macha = ABS(mach);
    if (macha >= 1.0) return half * (1.0 + ipn * mach / (macha + SMALL));
    return fm_split2(mach, ipn) * ((2.0 * ipn - mach) - ipn * 16.0 * alpha * mach * fm_split2(mach, -ipn));
}

// NsInvFlux::Slau2 (codes/ns/src/NsInvFlux.cpp:1246-1314)
void slau2_flux(const double* p1, const double* p2, double gama1, double gama2,
                double xfn, double yfn, double zfn, double vfn, double* flux) {
    double rl = p1[0], ul = p1[1], vl = p1[2], wl = p1[3], pl = p1[4];
    double rr = p2[0], ur = p2[1], vr = p2[2], wr = p2[3], pr = p2[4];
    double v2l = SQR(ul, vl, wl), v2r = SQR(ur, vr, wr);
    double hint_l = (gama1 / (gama1 - 1.0)) * (pl / rl);
    double hint_r = (gama2 / (gama2 - 1.0)) * (pr / rr);
    double hl = hint_l + half * v2l, hr = hint_r + half * v2r;
    double vnl = xfn * ul + yfn * vl + zfn * wl - vfn;
    double vnr = xfn * ur + yfn * vr + zfn * wr - vfn;
    double orl = 1.0 / (rl + SMALL), orr = 1.0 / (rr + SMALL);
    double c2l = gama1 * pl * orl, c2r = gama2 * pr * orr;
    double cl = std::sqrt(c2l), cr = std::sqrt(c2r);
    double cm = half * (cl + cr);
    double t = rr / rl;
    double vna = (ABS(vnl) + t * ABS(vnr)) / (1 + t);
    double ml = vnl / cm, mr = vnr / cm;
    double gg = -MAX(MIN(ml, 0.0), -1.0) * MIN(MAX(mr, 0.0), 1.0);
    double vnp = (1 - gg) * vna + gg * ABS(vnl);
    double vnn = (1 - gg) * vna + gg * ABS(vnr);
    double va = std::sqrt(half * (v2l + v2r));
    double m12 = MIN(1.0, va / cm);
    double ka = SQR(1 - m12);
    double ms = half * (rl * (vnl + vnp) + rr * (vnr - vnn) - ka * (pr - pl) / cm);
    double fp5ml = fp_split5(ml, 0.0, 1.0);
    double fp5mr = fp_split5(mr, 0.0, -1.0);
    double ps = 0;
    ps += half * (pl + pr);
    ps += half * (fp5ml - fp5mr) * (pl - pr);
    ps += va * (fp5ml + fp5mr - 1) * std::sqrt(rl * rr) * cm;
    double mp = half * (ms + ABS(ms)), mn = half * (ms - ABS(ms));
    flux[0] = (mp + mn);
    flux[1] = (mp * ul + mn * ur) + xfn * ps;
    flux[2] = (mp * vl + mn * vr) + yfn * ps;
    flux[3] = (mp * wl + mn * wr) + zfn * ps;
    flux[4] = (mp * hl + mn * hr) + vfn * ps;
}

// AddF2CField (codes/usolver/src/UCom.cpp:144-171)
void add_f2c(const ofo_case& c, RV* cell, const RV* face, int nEqu) {
    for (int f = 0; f < c.nBFaces; ++f) {
        int l = c.lc[f];
        for (int e = 0; e < nEqu; ++e) cell[e][l] -= face[e][f];
    }
    for (int f = c.nBFaces; f < c.nFaces; ++f) {
        int l = c.lc[f], r = c.rc[f];
        for (int e = 0; e < nEqu; ++e) { cell[e][l] -= face[e][f]; cell[e][r] += face[e][f]; }
    }
}

// UNsInvFlux::CalcFlux (codes/uns/src/UNsInvFlux.cpp:123-192)
void ns_inv_flux(ofo_case& c) {
    // uns_grad: GG of the 5 primitives + StoreBcGrad (codes/grad/src/Grad.cpp:53-156)
    for (int e = 0; e < 5; ++e) grad_gg(c, c.q[e], c.dqdx[e], c.dqdy[e], c.dqdz[e]);
    for (int f = 0; f < c.nBFaces; ++f) {
        int l = c.lc[f];
        for (int e = 0; e < 5; ++e) { c.bcdqdx[e][f] = c.dqdx[e][l]; c.bcdqdy[e][f] = c.dqdy[e][l]; c.bcdqdz[e][f] = c.dqdz[e][l]; }
    }
    calc_limiter(c, c.cfg.ilim, c.limiter, c.q, c.dqdx, c.dqdy, c.dqdz, 5);
    RV qf1[5], qf2[5];
    for (int e = 0; e < 5; ++e) { qf1[e].assign(c.nFaces, 0); qf2[e].assign(c.nFaces, 0); }
    face_values_ns(c, qf1, qf2);
    for (int f = 0; f < c.nFaces; ++f) {
        int l = c.lc[f], r = c.rc[f];
        double p1[5], p2[5], fl[5];
        for (int e = 0; e < 5; ++e) { p1[e] = qf1[e][f]; p2[e] = qf2[e][f]; }
        if (c.cfg.ischeme == 12)
            slau2_flux(p1, p2, c.gama[l], c.gama[r], c.xfn[f], c.yfn[f], c.zfn[f], c.vfn[f], fl);
        else
            roe_flux(c.cfg, p1, p2, c.gama[l], c.gama[r], c.xfn[f], c.yfn[f], c.zfn[f], c.vfn[f], fl);
        for (int e = 0; e < 5; ++e) c.invflux[e][f] = c.area[f] * fl[e];
    }
    add_f2c(c, c.res, c.invflux, 5);
}

// ========================================================== viscous flux

// GCom::CalcTangent (codes/global/src/Com.cpp:76-121)
void tangents(double xfn, double yfn, double zfn, double* t1, double* t2) {
    if (xfn != 0.0) { t1[0] = yfn; t1[1] = -xfn; t1[2] = 0.0; }
    else if (yfn != 0.0) { t1[0] = -yfn; t1[1] = xfn; t1[2] = 0.0; }
    else if (zfn != 0.0) { t1[0] = 0.0; t1[1] = -zfn; t1[2] = yfn; }
    else { t1[0] = t1[1] = t1[2] = 0; t2[0] = t2[1] = t2[2] = 0; return; }
    double dtmp = 1.0 / DIST(t1[0], t1[1], t1[2]);
    t1[0] *= dtmp; t1[1] *= dtmp; t1[2] *= dtmp;
    t2[0] = yfn * t1[2] - zfn * t1[1];
    t2[1] = zfn * t1[0] - xfn * t1[2];
    t2[2] = xfn * t1[1] - yfn * t1[0];
}

// Face-gradient of nv scalars at face f by the reference's VisGrad methods
// (codes/usolver/src/VisGrad.cpp:44-365); q1/q2 cell values, g1/g2 cell
// gradients [nv][3].  Outputs face value qf, face gradient gf[nv][3] and
// normal derivative dqdn.
struct FaceGeom {
    double xfn, yfn, zfn, t1[3], t2[3];
    double fw1, fw2;                      // CalcFaceWeight
    double dx, dy, dz, ods;               // CalcGradCoef
    double d1, d2, dxnl, dynl, dznl, dxnr, dynr, dznr, angle1, angle2;  // PrepareCellGeom
};

void face_geom(const ofo_case& c, int f, int method, FaceGeom& G) {
    int l = c.lc[f], r = c.rc[f];
    G.xfn = c.xfn[f]; G.yfn = c.yfn[f]; G.zfn = c.zfn[f];
    tangents(G.xfn, G.yfn, G.zfn, G.t1, G.t2);
    double dxl = c.xfc[f] - c.xcc[l], dyl = c.yfc[f] - c.ycc[l], dzl = c.zfc[f] - c.zcc[l];
    double dxr = c.xfc[f] - c.xcc[r], dyr = c.yfc[f] - c.ycc[r], dzr = c.zfc[f] - c.zcc[r];
    double delt1 = DIST(dxl, dyl, dzl), delt2 = DIST(dxr, dyr, dzr);
    double delta = 1.0 / (delt1 + delt2 + SMALL);
    G.fw1 = delt2 * delta; G.fw2 = delt1 * delta;
    if (method == 0) {
        G.dx = c.xcc[r] - c.xcc[l]; G.dy = c.ycc[r] - c.ycc[l]; G.dz = c.zcc[r] - c.zcc[l];
        G.ods = 1.0 / DIST(G.dx, G.dy, G.dz);
        G.dx *= G.ods; G.dy *= G.ods; G.dz *= G.ods;
    } else if (method == 2) {
        double ax = c.xcc[l] - c.xfc[f], ay = c.ycc[l] - c.yfc[f], az = c.zcc[l] - c.zfc[f];
        double bx = c.xcc[r] - c.xfc[f], by = c.ycc[r] - c.yfc[f], bz = c.zcc[r] - c.zfc[f];
        G.d1 = G.xfn * ax + G.yfn * ay + G.zfn * az;
        G.d2 = G.xfn * bx + G.yfn * by + G.zfn * bz;
        G.dxnl = G.xfn * G.d1 - ax; G.dynl = G.yfn * G.d1 - ay; G.dznl = G.zfn * G.d1 - az;
        G.dxnr = G.xfn * G.d2 - bx; G.dynr = G.yfn * G.d2 - by; G.dznr = G.zfn * G.d2 - bz;
        auto angle = [](double dx, double dy, double dz, double d) {
            double res = d / (DIST(dx, dy, dz) + SMALL);
            if (res > 1.0) res = 1.0;
            if (res < -1.0) res = -1.0;
            return std::asin(res) * 180.0 / PI;
        };
        G.angle1 = angle(ax, ay, az, -G.d1);
        G.angle2 = angle(bx, by, bz, G.d2);
    }
}

// One scalar through the selected VisGrad method.
void vis_grad_scalar(const ofo_case& c, int f, int method, const FaceGeom& G, double q1, double q2,
                     const double* g1, const double* g2, double& qf, double* gf, double& dqdn) {
    // AverGrad (always done in PrepareFaceValue)
    gf[0] = half * (g1[0] + g2[0]); gf[1] = half * (g1[1] + g2[1]); gf[2] = half * (g1[2] + g2[2]);
    dqdn = 0.0;
    qf = half * (q1 + q2);
    if (method == 0) {
        // StdMethod: CorrectGrad + CalcNormalGrad
        double fOld = gf[0] * G.dx + gf[1] * G.dy + gf[2] * G.dz;
        double fnew = (q2 - q1) * G.ods;
        double df = fnew - fOld;
        gf[0] += df * G.dx; gf[1] += df * G.dy; gf[2] += df * G.dz;
        dqdn = G.xfn * gf[0] + G.yfn * gf[1] + G.zfn * gf[2];
    } else if (method == 1) {
        dqdn = G.xfn * gf[0] + G.yfn * gf[1] + G.zfn * gf[2];
    } else {
        // TestMethod: CalcTestMethod + ModifyFaceGrad
        double q11 = q1, q22 = q2;
        if (G.angle1 > 20.0 && G.angle2 > 20.0) {
            q11 += g1[0] * G.dxnl + g1[1] * G.dynl + g1[2] * G.dznl;
            q22 += g2[0] * G.dxnr + g2[1] * G.dynr + g2[2] * G.dznr;
            qf = half * (q1 + q2);
        }
        bool ok = G.angle1 > 0.0 && G.angle2 > 0.0 && ABS(G.d1) > SMALL && ABS(G.d2) > SMALL;
        if (ok) {
            double dtmp = SQR(G.d1, G.d2);
            double c1 = SQR(G.d1) / dtmp, c2 = 1.0 - c1;
            if (f < c.nBFaces && c.bcType[f] != BC_INTERFACE) { c1 = 1.0; c2 = 0.0; }
            double dqdn1 = (q11 - qf) / G.d1;
            double dqdn2 = (q22 - qf) / G.d2;
            dqdn = c1 * dqdn1 + c2 * dqdn2;
        }
        gf[0] = G.fw1 * g1[0] + G.fw2 * g2[0];
        gf[1] = G.fw1 * g1[1] + G.fw2 * g2[1];
        gf[2] = G.fw1 * g1[2] + G.fw2 * g2[2];
        double dt1 = G.t1[0] * gf[0] + G.t1[1] * gf[1] + G.t1[2] * gf[2];
        double dt2 = G.t2[0] * gf[0] + G.t2[1] * gf[1] + G.t2[2] * gf[2];
        gf[0] = G.xfn * dqdn + G.t1[0] * dt1 + G.t2[0] * dt2;
        gf[1] = G.yfn * dqdn + G.t1[1] * dt1 + G.t2[1] * dt2;
        gf[2] = G.zfn * dqdn + G.t1[2] * dt1 + G.t2[2] * dt2;
    }
}

// UNsVisFlux::CalcFlux (codes/uns/src/UNsVisFlux.cpp:90-440)
void ns_vis_flux(ofo_case& c) {
    if (c.cfg.vismodel == 0) return;
    grad_gg(c, c.tempr, c.dtdx, c.dtdy, c.dtdz);  // ut_grad (codes/uns/src/UNsGrad.cpp:75-100)
    int method = c.cfg.ivischeme;
    for (int f = 0; f < c.nFaces; ++f) {
        int l = c.lc[f], r = c.rc[f];
        FaceGeom G;
        face_geom(c, f, method, G);
        double gq[5][3], qfv[5], dqdn[5];
        for (int e = 0; e < 5; ++e) {
            double g1[3] = {c.dqdx[e][l], c.dqdy[e][l], c.dqdz[e][l]};
            double g2[3] = {c.dqdx[e][r], c.dqdy[e][r], c.dqdz[e][r]};
            vis_grad_scalar(c, f, method, G, c.q[e][l], c.q[e][r], g1, g2, qfv[e], gq[e], dqdn[e]);
        }
        double gT[3], Tf, dtdn;
        {
            double g1[3] = {c.dtdx[l], c.dtdy[l], c.dtdz[l]};
            double g2[3] = {c.dtdx[r], c.dtdy[r], c.dtdz[r]};
            vis_grad_scalar(c, f, method, G, c.tempr[l], c.tempr[r], g1, g2, Tf, gT, dtdn);
        }
        double visl = half * (c.visl[l] + c.visl[r]);
        double vist = half * (c.vist[l] + c.vist[r]);
        double vis = visl + vist;
        double dudx = gq[1][0], dudy = gq[1][1], dudz = gq[1][2];
        double dvdx = gq[2][0], dvdy = gq[2][1], dvdz = gq[2][2];
        double dwdx = gq[3][0], dwdy = gq[3][1], dwdz = gq[3][2];
        double um = qfv[1], vm = qfv[2], wm = qfv[3];
        // AddHeatFlux
        double qNormal = 0.0;
        double kcp = (visl * c.oprl + vist * c.oprt) * c.const_cp;
        qNormal += G.xfn * 0.0 + G.yfn * 0.0 + G.zfn * 0.0;
        qNormal += kcp * dtdn;
        // CalcStress
        double divv2p3 = two3rd * (dudx + dvdy + dwdz);
        double txx = vis * (2.0 * dudx - divv2p3);
        double tyy = vis * (2.0 * dvdy - divv2p3);
        double tzz = vis * (2.0 * dwdz - divv2p3);
        double txy = vis * (dudy + dvdx);
        double txz = vis * (dudz + dwdx);
        double tyz = vis * (dvdz + dwdy);
        double fv[5];
        fv[0] = 0.0;
        fv[1] = G.xfn * txx + G.yfn * txy + G.zfn * txz;
        fv[2] = G.xfn * txy + G.yfn * tyy + G.zfn * tyz;
        fv[3] = G.xfn * txz + G.yfn * tyz + G.zfn * tzz;
        fv[4] = um * fv[1] + vm * fv[2] + wm * fv[3] + qNormal;
        double coeff = -c.oreynolds * c.area[f];
        for (int e = 0; e < 5; ++e) c.visflux[e][f] = coeff * fv[e];
    }
    add_f2c(c, c.res, c.visflux, 5);
}

// ============================================================ time step

// UTimeStep::CalcTimeStep (codes/uns/src/UTimeStep.cpp:190-416) +
// TimeStep (codes/ns/src/TimeStep.cpp:43-133)
void ns_timestep(ofo_case& c, int iter) {
    const ofx_config& g = c.cfg;
    c.cfl = g.cfled;
    if (iter < g.ncfl) {
        double delta = (g.cfled - g.cflst) / g.ncfl;
        c.cfl = g.cflst + delta * iter;
    }
    for (int i = 0; i < c.nCells; ++i) c.invsr[i] = 0.0;
    for (int f = 0; f < c.nFaces; ++f) {
        int l = c.lc[f], r = c.rc[f];
        double rl = c.q[0][l], ul = c.q[1][l], vl = c.q[2][l], wl = c.q[3][l], pl = c.q[4][l];
        double rr = c.q[0][r], ur = c.q[1][r], vr = c.q[2][r], wr = c.q[3][r], pr = c.q[4][r];
        double gama1 = c.gama[l], gama2 = c.gama[r];
        double xfn = c.xfn[f], yfn = c.yfn[f], zfn = c.zfn[f], vfn = c.vfn[f];
        (void)std::sqrt(gama1 * pl / rl); (void)std::sqrt(gama2 * pr / rr);
        double vnl = xfn * ul + yfn * vl + zfn * wl - vfn;
        double vnr = xfn * ur + yfn * vr + zfn * wr - vfn;
        double gama = half * (gama1 + gama2);
        double pm = half * (pl + pr), rm = half * (rl + rr);
        double cm = std::sqrt(gama * pm / rm);
        double vn = half * (vnl + vnr);
        double sr = half * c.area[f] * (ABS(vn) + cm);
        c.invsr[l] += sr; c.invsr[r] += sr;
    }
    for (int i = 0; i < c.nCells; ++i) c.vissr[i] = 0.0;
    if (g.vismodel > 0) {
        for (int f = 0; f < c.nFaces; ++f) {
            int l = c.lc[f], r = c.rc[f];
            double visl = half * (c.visl[l] + c.visl[r]);
            double vist = half * (c.vist[l] + c.vist[r]);
            double sr = 0.0;
            if (g.visSRModel == 1) {
                double density = half * (c.q[0][l] + c.q[0][r]);
                double c1 = 4.0 / 3.0 * (visl + vist);
                double c2 = g.gama_ref * (visl * c.oprl + vist * c.oprt);
                double c3 = 2.0 * MAX(c1, c2) / (g.reynolds * density);
                sr = SQR(c.area[f]) * c3;
            } else if (g.visSRModel == 2) {
                double dist = ABS(c.xfn[f] * (c.xcc[r] - c.xcc[l]) + c.yfn[f] * (c.ycc[r] - c.ycc[l]) +
                                  c.zfn[f] * (c.zcc[r] - c.zcc[l]));
                double viscosity = visl + vist;
                double density = half * (c.q[0][l] + c.q[0][r]);
                double c1 = 2.0 * viscosity / (density * dist * g.reynolds + SMALL);
                sr = half * c1 * c.area[f];
            } else if (g.visSRModel == 3) {
                double density = half * (c.q[0][l] + c.q[0][r]);
                sr = SQR(c.area[f]) / (g.reynolds * density + SMALL);
            }
            c.vissr[l] += sr; c.vissr[r] += sr;
        }
    }
    for (int i = 0; i < c.nCells; ++i) c.timestep[i] = c.cfl * c.vol[i] / c.invsr[i];
    if (g.vismodel > 0 && g.visTimeStepModel > 0) {
        for (int i = 0; i < c.nCells; ++i) {
            double vts = c.cfl * c.vol[i] / c.vissr[i];
            c.timestep[i] *= vts / (c.timestep[i] + vts);
        }
    }
    double mts = LARGE;
    for (int i = 0; i < c.nCells; ++i) mts = MIN(c.timestep[i], mts);
    if (g.max_time_ratio > 0) {
        double mp = g.max_time_ratio * mts;
        for (int i = 0; i < c.nCells; ++i) c.timestep[i] = MIN(c.timestep[i], mp);
    }
}

// ================================================================ LU-SGS

// UNsLusgs::Solve for face f seen from cell i (codes/uns/src/UNsLusgs.cpp:189-300,
// codes/ns/src/NsLusgs.cpp:121-228, codes/global/src/Com.cpp:40-65)
void ns_lusgs_face(const ofo_case& c, int i, int f, double* rhs0) {
    int l = c.lc[f], r = c.rc[f];
    double xfn = c.xfn[f], yfn = c.yfn[f], zfn = c.zfn[f], vfn = c.vfn[f], farea = c.area[f];
    if (r == i) { std::swap(l, r); xfn = -xfn; yfn = -yfn; zfn = -zfn; vfn = -vfn; }
    // PrepareData
    double primj[5], dqj[5];
    for (int e = 0; e < 5; ++e) { primj[e] = c.q[e][r]; dqj[e] = c.dq[e][r]; }
    double rl = c.q[0][l], ul = c.q[1][l], vl = c.q[2][l], wl = c.q[3][l], pl = c.q[4][l];
    double rr = c.q[0][r], ur = c.q[1][r], vr = c.q[2][r], wr = c.q[3][r], pr = c.q[4][r];
    double gl = c.gama[l], gr = c.gama[r];
    double hl = gl / (gl - 1.0) * pl / rl + half * SQR(ul, vl, wl);
    double hr = gr / (gr - 1.0) * pr / rr + half * SQR(ur, vr, wr);
    double tmp0 = std::sqrt(rr / rl);
    double tmp1 = 1.0 / (1.0 + tmp0);
    double rm = std::sqrt(rr * rl);
    double um = (ul + ur * tmp0) * tmp1;
    double vm = (vl + vr * tmp0) * tmp1;
    double wm = (wl + wr * tmp0) * tmp1;
    double hm = (hl + hr * tmp0) * tmp1;
    double pm = rm * (hm - half * SQR(um, vm, wm)) * (gl - 1.0) / gl;
    double gama = c.gama[r];
    double visl = c.visl[r], vist = c.vist[r];
    // GetStandardFluxIncrement(-1) + CalcFaceEigenValue(primF)
    double c2f = ABS(gama * pm / rm);
    double cmf = std::sqrt(c2f);
    double vnf = xfn * um + yfn * vm + zfn * wm;
    double lmdF = ABS(vnf - vfn) + cmf;
    double c2 = ABS(gama * primj[4] / primj[0]);
    double cm = std::sqrt(c2);
    double vn_fluid = xfn * primj[1] + yfn * primj[2] + zfn * primj[3];
    double vn_rel = vn_fluid - vfn;
    double sgn = -1.0;
    double lmd1 = half * (vn_rel + sgn * lmdF);
    double lmd2 = half * ((vn_rel + cm) + sgn * lmdF);
    double lmd3 = half * ((vn_rel - cm) + sgn * lmdF);
    double x1 = (lmd1 + lmd1 - lmd2 - lmd3) / (c2 + c2);
    double x2 = (lmd2 - lmd3) / (cm + cm);
    double dc = vn_fluid * dqj[0] - xfn * dqj[1] - yfn * dqj[2] - zfn * dqj[3];
    double c2dc = c2 * dc;
    // CalcDH (codes/ns/src/NsLusgs.cpp:340-368)
    double v2 = SQR(primj[1], primj[2], primj[3]);
    double ae = gama - 1.0;
    double af = half * ae * v2;
    double dh = -ae * (primj[1] * dqj[1] + primj[2] * dqj[2] + primj[3] * dqj[3] - dqj[4]);
    double enth = (gama / (gama - 1.0)) * (primj[4] / primj[0]);
    dh += af * dqj[0];
    double hmj = enth + half * v2;
    double term1 = dh * x1 + dc * x2;
    double term2 = c2dc * x1 + dh * x2;
    double dfj[5];
    dfj[0] = lmd1 * dqj[0] - term1;
    dfj[1] = lmd1 * dqj[1] - primj[1] * term1 + xfn * term2;
    dfj[2] = lmd1 * dqj[2] - primj[2] * term1 + yfn * term2;
    dfj[3] = lmd1 * dqj[3] - primj[3] * term1 + zfn * term2;
    dfj[4] = lmd1 * dqj[4] - hmj * term1 + vn_fluid * term2;
    for (int e = 0; e < 5; ++e) dfj[e] *= farea;
    // CalcViscousTerm (codes/uns/src/UNsLusgs.cpp:295-335)
    if (c.cfg.vismodel != 0) {
        double visrad;
        if (c.cfg.visSRModel == 1) {
            double density = half * (c.q[0][l] + c.q[0][r]);
            double c1 = 4.0 / 3.0 * (visl + vist);
            double cc2 = c.cfg.gama_ref * (visl * c.oprl + vist * c.oprt);
            double c3 = 2.0 * MAX(c1, cc2) / (c.cfg.reynolds * density);
            visrad = SQR(farea) * c3 / c.vol[r];
        } else {
            double dist = ABS(xfn * (c.xcc[r] - c.xcc[l]) + yfn * (c.ycc[r] - c.ycc[l]) + zfn * (c.zcc[r] - c.zcc[l]));
            double viscosity = visl + vist;
            double density = half * (c.q[0][l] + c.q[0][r]);
            double c1 = 2.0 * viscosity / (density * dist * c.cfg.reynolds + SMALL);
            visrad = half * c1 * farea;
        }
        for (int e = 0; e < 5; ++e) rhs0[e] -= visrad * dqj[e];
    }
    for (int e = 0; e < 5; ++e) rhs0[e] += dfj[e];
}

// UNsSpectrum::CalcImplicitSpectrum (codes/uns/src/UNsSpectrum.cpp:52-94), sp2 = 1
void ns_impsr(ofo_case& c) {
    for (int i = 0; i < c.nCells; ++i) c.impsr[i] = (1.0 / c.timestep[i]) * c.vol[i];
    for (int i = 0; i < c.nCells; ++i) c.impsr[i] += c.invsr[i];
    for (int i = 0; i < c.nCells; ++i) c.impsr[i] += c.vissr[i];
}

// UNsLusgs::LowerSweep / UpperSweep (codes/uns/src/UNsLusgs.cpp:75-138)
void ns_lusgs(ofo_case& c) {
    ns_impsr(c);
    for (int i = 0; i < c.nCells; ++i) {
        double rhs0[5] = {0, 0, 0, 0, 0};
        for (int k = c.c2fOff[i]; k < c.c2fOff[i + 1]; ++k) {
            int f = c.c2f[k];
            if (c.lc[f] > i || c.rc[f] > i) continue;
            ns_lusgs_face(c, i, f, rhs0);
        }
        double rad = c.impsr[i];
        for (int e = 0; e < 5; ++e) { double t = c.res[e][i] - rhs0[e]; t /= rad; c.dq[e][i] = t; }
    }
    for (int i = c.nCells - 1; i >= 0; --i) {
        double rhs0[5] = {0, 0, 0, 0, 0};
        for (int k = c.c2fOff[i]; k < c.c2fOff[i + 1]; ++k) {
            int f = c.c2f[k];
            if (c.lc[f] < i || c.rc[f] < i) continue;
            ns_lusgs_face(c, i, f, rhs0);
        }
        double rad = c.impsr[i];
        for (int e = 0; e < 5; ++e) { double t = -rhs0[e]; t /= rad; c.dq[e][i] += t; }
    }
}

// UNsUpdate::UpdateFlowField + NsUpdate::CalcFlowField/WeekSolutionFix + SolutionFix
// (codes/uns/src/UNsUpdate.cpp:46-146, codes/ns/src/NsUpdate.cpp:40-67,192-217)
// Under the multi-stage scheme the increment is the residual field itself:
// UNsField::Init binds dq to res when time_integral is MULTI_STAGE
// (codes/uns/src/UNsCom.cpp:85-88).
void ns_update(ofo_case& c) {
    const RV* inc = c.cfg.time_integral == 1 ? c.res : c.dq;
    for (int i = 0; i < c.nCells; ++i) {
        double prim0[5], q0[5], q[5], dq[5], prim[5];
        for (int e = 0; e < 5; ++e) { prim0[e] = c.q[e][i]; dq[e] = inc[e][i]; }
        double gama = c.gama[i];
        prim_to_q(prim0, gama, q0);
        for (int e = 0; e < 5; ++e) q[e] = q0[e] + dq[e];
        // QToPrim (codes/ns/src/NsInvFlux.cpp:1372-1390)
        double density = ABS(q[0]);
        double rd = 1.0 / density;
        double um = q[1] * rd, vm = q[2] * rd, wm = q[3] * rd, rem = q[4];
        double v2 = SQR(um, vm, wm);
        double reint = rem - half * density * v2;
        prim[0] = density; prim[1] = um; prim[2] = vm; prim[3] = wm; prim[4] = (gama - 1.0) * reint;
        if (prim[0] <= 0.0 || prim[4] <= 0.0) {
            const double coefs[3] = {0.5, 0.1, 0.01};
            bool ok = false;
            for (int k = 0; k < 3; ++k) {
                for (int e = 0; e < 5; ++e) prim[e] += coefs[k] * dq[e];
                if (!(prim[0] < 0.0 || prim[4] < 0.0)) { ok = true; break; }
            }
            if (!ok) {
                for (int e = 0; e < 5; ++e) prim[e] = 0;
                double sumV = 0.0;
                for (int k = c.c2fOff[i]; k < c.c2fOff[i + 1]; ++k) {
                    int f = c.c2f[k];
                    int nei = (i == c.lc[f]) ? c.rc[f] : c.lc[f];
                    sumV += 1.0;
                    for (int e = 0; e < 5; ++e) {
                        double v = c.q[e][nei];
                        if (e == 0 || e == 4) v = ABS(v);
                        prim[e] += v * 1.0;
                    }
                }
                double rVol = 1.0 / sumV;
                for (int e = 0; e < 5; ++e) prim[e] *= rVol;
            }
        }
        for (int e = 0; e < 5; ++e) c.q[e][i] = prim[e];
    }
}

// ==================================================================== SA

// UTurbSpectrum::CalcSpectrum1Equ + TurbSpectrum::CalcFaceSpectrum1Equ
// (codes/uturb/src/UTurbSpectrum.cpp:157-240, codes/turb/src/TurbSpectrum.cpp:75-134)
void sa_spectrum(ofo_case& c) {
    std::fill(c.turbimpsr.begin(), c.turbimpsr.end(), 0.0);
    std::fill(c.mat_l.begin(), c.mat_l.end(), 0.0);
    std::fill(c.mat_r.begin(), c.mat_r.end(), 0.0);
    for (int i = 0; i < c.nCells; ++i) {
        double ts = c.timestep[i] * c.cfg.turb_cfl_ratio;
        c.turbimpsr[i] += (1.0 / ts) * c.vol[i];
    }
    for (int f = 0; f < c.nFaces; ++f) {
        int l = c.lc[f], r = c.rc[f];
        double xfn = c.xfn[f], yfn = c.yfn[f], zfn = c.zfn[f], vfn = c.vfn[f], farea = c.area[f];
        double rl = c.q[0][l], rr = c.q[0][r];
        double vnl = xfn * c.q[1][l] + yfn * c.q[2][l] + zfn * c.q[3][l] - vfn;
        double vnr = xfn * c.q[1][r] + yfn * c.q[2][r] + zfn * c.q[3][r] - vfn;
        double absVn = ABS(half * (vnl + vnr));
        double rad1 = 0.0, rad2 = 0.0, m1 = 0.0, m2 = 0.0;
        rad1 += farea * half * absVn;
        rad2 += farea * half * absVn;
        m1 += farea * half * (-vnl - absVn);
        m2 += farea * half * (vnr - absVn);
        double orl = 1.0 / (rl + SMALL), orr = 1.0 / (rr + SMALL);
        double nuet1 = c.visl[l] * orl + c.turbq[l];
        double nuet2 = c.visl[r] * orr + c.turbq[r];
        double nueff = ABS(half * (nuet1 + nuet2));
        double cvol = half * (c.vol[l] + c.vol[r]);
        double ovol = 1.0 / cvol;
        double areaS2 = SQR(farea);
        rad1 += c.oreynolds * ovol * (c.osigma * nueff * areaS2);
        rad2 += c.oreynolds * ovol * (c.osigma * nueff * areaS2);
        m1 += -c.oreynolds * ovol * (c.osigma * nueff * areaS2);
        m2 += -c.oreynolds * ovol * (c.osigma * nueff * areaS2);
        c.turbimpsr[l] += rad1; c.turbimpsr[r] += rad2;
        c.mat_l[f] += m1; c.mat_r[f] += m2;
    }
}

// TurbCom::CalcSrcSa with CalcWorkVar/CalcSaProd (codes/turb/src/TurbCom.cpp:429-465,771-872)
void sa_source_cell(const ofo_case& c, int i, double& res_sa, double& spec_sa) {
    const double one = 1.0, two = 2.0, three = 3.0, six = 6.0, ten = 10.0, zero = 0.0;
    double dudx = c.dqdx[1][i], dudy = c.dqdy[1][i], dudz = c.dqdz[1][i];
    double dvdx = c.dqdx[2][i], dvdy = c.dqdy[2][i], dvdz = c.dqdz[2][i];
    double dwdx = c.dqdx[3][i], dwdy = c.dqdy[3][i], dwdz = c.dqdz[3][i];
    double rho = c.q[0][i], nuet = c.turbq[i], visl = c.visl[i];
    double len_scale = c.dist[i];
    double olam = rho / (visl + SMALL);
    double d2 = SQR(len_scale);
    double od2 = one / d2;
    double work1 = dudx + dvdy + dwdz;
    double work2 = SQR(dudx, dudy, dudz);
    double work3 = SQR(dudy + dvdx, dvdz + dwdy, dwdx + dudz);
    double work4 = DIST(dwdy - dvdz, dudz - dwdx, dvdx - dudy);
    if (c.cfg.iprod_sa == 1) {
        double stp = two * work2 + work3 - two3rd * work1 * work1;
        work4 = std::sqrt(ABS(stp));
    } else if (c.cfg.iprod_sa == 4) {
        double stp = two * work2 + work3 - two3rd * work1 * work1;
        double str = work4;
        work4 = str - two * MAX(zero, str - std::sqrt(ABS(stp)));
    } else if (c.cfg.iprod_sa == 5) {
        double stp = two * work2 + work3 - two3rd * work1 * work1;
        double str = work4;
        stp = ABS(stp);
        work4 = str * (one - MAX(zero, (str * str - stp) / (str * str + stp + SMALL)));
    }
    double str = work4;
    double xsi = nuet * olam + SMALL;
    double xsi3 = xsi * xsi * xsi;
    double fv1 = xsi3 / (xsi3 + c.cv13);
    double fv2 = one - xsi / (one + xsi * fv1);
    double rs = nuet * c.okarm2 * od2 * c.oreynolds;
    double nuetRs = nuet * rs;
    double sBar = rs * fv2;
    double omega = str;
    double stdv;
    if (sBar >= -c.sac2 * omega) stdv = omega + sBar;
    else {
        double term1 = (c.sac2 * c.sac2 * omega + c.sac3 * sBar);
        double term2 = (c.sac3 - 2 * c.sac2) * omega - sBar;
        stdv = omega + omega * term1 / term2;
    }
    double ostd = one / (stdv + SMALL);
    double r = rs * ostd;
    r = MIN(r, ten);
    double r2 = r * r;
    double r5 = r2 * r2 * r;
    double gg = r + c.cw2 * (r5 * r - r);
    double gg2 = gg * gg;
    double g6 = gg2 * gg2 * gg2;
    double or6 = one / six;
    double fw = gg * ofxm::ofx_pow((one + c.cw36) / (g6 + c.cw36), or6);
    double ft2 = 0.0;  // ft2_flag is never set in the reference (zero-initialised global)
    double grd2 = SQR(c.turbdqdx[i], c.turbdqdy[i], c.turbdqdz[i]);
    double prod = c.cb1 * (one - ft2) * stdv * nuet;
    double diff = c.cb2s * grd2 * c.oreynolds;
    double dest = (c.cw1k * fw - c.cb1 * ft2) * nuetRs;
    diff = MIN(c.cfg.rprod * prod, diff);
    double srcTerm = prod - dest;
    srcTerm += diff;
    double p2 = fv1 / (xsi + SMALL);
    double p3 = one + xsi * fv1;
    double dfv2dk = (three * c.cv13 * (p2 * p2) - one) / (p3 * p3);
    double dsdnu = c.oreynolds * c.okarm2 * od2 * (fv2 + xsi * dfv2dk);
    double dfwdg = fw / (gg + SMALL) * (one - g6 / (g6 + c.cw36));
    double dgdr = one + c.cw2 * (six * r5 - one);
    double drdnu = c.oreynolds * c.okarm2 * od2 * ostd * (one - nuet * ostd * dsdnu);
    double dfwdnu = dfwdg * dgdr * drdnu;
    double prod0 = c.cb1 * (one - ft2) * stdv;
    double dest0 = (c.cw1 * fw - c.cb1 * ft2 * c.okarm2) * nuet * od2 * c.oreynolds;
    double prodp = c.cb1 * ((one - ft2) * dsdnu);
    double destp = (c.cw1 * (fw + dfwdnu * nuet) - c.cb1 * ft2 * c.okarm2) * od2 * c.oreynolds;
    double prde0 = prod0 - dest0;
    double prdep = prodp - destp;
    double part1 = -half * (prde0 - ABS(prde0));
    double part2 = -half * (prdep - ABS(prdep)) * nuet;
    spec_sa = part1 + part2;
    res_sa = srcTerm;
}

// TurbCalcRHS (codes/turb/src/TurbRhs.cpp:57-71): BC, spectrum, source,
// first-order convection, diffusion.
void sa_residual(ofo_case& c) {
    for (int i = 0; i < c.nTCell; ++i) c.turbres[i] = 0.0;  // ZERO_RESIDUALS/LOAD_RESIDUALS
    sa_bc(c);
    sa_spectrum(c);
    // UTurbSrcFlux::Init -> CalcGrad (turbq) (codes/uturb/src/UTurbSrcFlux.cpp:56-60)
    grad_gg(c, c.turbq, c.turbdqdx, c.turbdqdy, c.turbdqdz);
    for (int i = 0; i < c.nCells; ++i) {
        double res_sa, spec_sa;
        sa_source_cell(c, i, res_sa, spec_sa);
        c.turbimpsr[i] += spec_sa * c.vol[i];
        c.turbres[i] += res_sa * c.vol[i];
    }
    // UTurbInvFlux::CalcFlux with turb_ilim = tns_ilim = -1 (first order)
    // (codes/uturb/src/UTurbInvFlux.cpp:130-202, codes/turb/src/TurbInvFlux.cpp:78-98)
    RV fl(c.nFaces), fv(c.nFaces);
    for (int f = 0; f < c.nFaces; ++f) {
        int l = c.lc[f], r = c.rc[f];
        double nl = c.turbq[l], nr = c.turbq[r];
        double ul = c.q[1][l], vl = c.q[2][l], wl = c.q[3][l];
        double ur = c.q[1][r], vr = c.q[2][r], wr = c.q[3][r];
        int bt = (f < c.nBFaces) ? c.bcType[f] : 0;
        if (f < c.nBFaces && bt != BC_INTERFACE && bt != BC_PERIODIC) {
            double tn = half * (c.turbq[l] + c.turbq[r]);
            nl = tn; nr = tn;
            if (bt == BC_SOLID_SURFACE) {
                ul = ur = c.bc_q[1][f]; vl = vr = c.bc_q[2][f]; wl = wr = c.bc_q[3][f];
            } else {
                double tu = half * (c.q[1][l] + c.q[1][r]);
                double tv = half * (c.q[2][l] + c.q[2][r]);
                double tw = half * (c.q[3][l] + c.q[3][r]);
                ul = ur = tu; vl = vr = tv; wl = wr = tw;
            }
        }
        double xfn = c.xfn[f], yfn = c.yfn[f], zfn = c.zfn[f], vfn = c.vfn[f];
        double vnl = xfn * ul + yfn * vl + zfn * wl - vfn;
        double vnr = xfn * ur + yfn * vr + zfn * wr - vfn;
        double vnl_p = half * (vnl + ABS(vnl));
        double vnr_n = half * (vnr - ABS(vnr));
        double coef_l = vnl_p * 1.0, coef_r = vnr_n * 1.0;  // rho_coef = 1
        fl[f] = c.area[f] * (coef_l * nl + coef_r * nr);
    }
    add_f2c(c, &c.turbres, &fl, 1);
    // UTurbVisFlux::CalcVisFlux1Equ (codes/uturb/src/UTurbVisFlux.cpp:247-369)
    int method = c.cfg.iturb_visflux;
    for (int f = 0; f < c.nFaces; ++f) {
        int l = c.lc[f], r = c.rc[f];
        FaceGeom G;
        face_geom(c, f, method, G);
        double g1[3] = {c.turbdqdx[l], c.turbdqdy[l], c.turbdqdz[l]};
        double g2[3] = {c.turbdqdx[r], c.turbdqdy[r], c.turbdqdz[r]};
        double qf, gf[3], dqdn;
        vis_grad_scalar(c, f, method, G, c.turbq[l], c.turbq[r], g1, g2, qf, gf, dqdn);
        double orl = 1.0 / (c.q[0][l] + SMALL), orr = 1.0 / (c.q[0][r] + SMALL);
        double visl_nv = half * (c.visl[l] * orl + c.visl[r] * orr);
        double vist_nv = half * (c.turbq[l] + c.turbq[r]);
        double comVis = (visl_nv + vist_nv) * c.osigma;
        fv[f] = -c.oreynolds * comVis * dqdn * c.area[f];
    }
    add_f2c(c, &c.turbres, &fv, 1);
}

// UTurbLusgs::LowerSweep/UpperSweep + TurbLusgs (codes/uturb/src/UTurbLusgs.cpp:90-304,
// codes/turb/src/TurbLusgs.cpp:123-254)
void sa_lusgs(ofo_case& c) {
    for (int i = 0; i < c.nCells; ++i) {
        double rhs0 = 0.0;
        for (int k = c.c2fOff[i]; k < c.c2fOff[i + 1]; ++k) {
            int f = c.c2f[k];
            if (c.lc[f] > i || c.rc[f] > i) continue;
            bool swap = (c.rc[f] == i);
            int j = swap ? c.lc[f] : c.rc[f];
            double m = swap ? c.mat_l[f] : c.mat_r[f];
            rhs0 += m * c.turbdq[j];
        }
        double t = c.turbres[i] - rhs0;
        t /= c.turbimpsr[i];
        c.turbdq[i] = t;
    }
    for (int i = c.nCells - 1; i >= 0; --i) {
        double rhs0 = 0.0;
        for (int k = c.c2fOff[i]; k < c.c2fOff[i + 1]; ++k) {
            int f = c.c2f[k];
            if (c.lc[f] < i || c.rc[f] < i) continue;
            bool swap = (c.rc[f] == i);
            int j = swap ? c.lc[f] : c.rc[f];
            double m = swap ? c.mat_l[f] : c.mat_r[f];
            rhs0 += m * c.turbdq[j];
        }
        double t = -rhs0;
        t /= c.turbimpsr[i];
        c.turbdq[i] += t;
    }
}

// TurbUpdate::CalcFlowField1Equ, crelax = 1 (codes/turb/src/TurbUpdate.cpp:58-69);
// the increment is turbres under the multi-stage scheme (codes/uturb/src/UTurbCom.cpp:79-82)
void sa_update(ofo_case& c) {
    const RV& inc = c.cfg.time_integral == 1 ? c.turbres : c.turbdq;
    for (int i = 0; i < c.nCells; ++i) {
        double qOld = c.turbq[i], dqNew = inc[i];
        double qNew = qOld + dqNew;
        qNew = (1.0 - 1.0) * qOld + 1.0 * qNew;
        c.turbq[i] = qNew;
    }
}

// ULhs::CalcLHS (codes/usolver/src/ULhs.cpp:47-62): the residual of each
// equation scaled by the local time step over the cell volume and lhscoef,
// inner cells only.
void calc_lhs(ofo_case& c, int solver, double lhscoef) {
    RV* r = solver == 0 ? c.res : &c.turbres;
    int nEqu = solver == 0 ? 5 : 1;
    for (int e = 0; e < nEqu; ++e)
        for (int i = 0; i < c.nCells; ++i) {
            double dt = c.timestep[i];
            r[e][i] *= (dt / c.vol[i]) * lhscoef;
        }
}

bool has_nan(const RV& v, int n) {
    for (int i = 0; i < n; ++i) if (std::isnan(v[i])) return true;
    return false;
}

}  // namespace

// ===================================================================== ABI

extern "C" {

const char* ofo_last_error(void) { return g_err.c_str(); }

ofo_case* ofo_create(int dim, int nNodes, const double* x, const double* y, const double* z,
                     int nFaces, int nCells, int nBFaces, const int* faceNodeCount, const int* faceNodes,
                     const int* lc, const int* rc, const int* bcType) {
    if (nFaces <= 0 || nCells <= 0 || nBFaces < 0 || nBFaces > nFaces || (dim != 2 && dim != 3)) {
        g_err = "ofo_create: bad sizes"; return nullptr;
    }
    ofo_case* c = new ofo_case();
    c->dim = dim; c->nNodes = nNodes; c->nFaces = nFaces; c->nCells = nCells; c->nBFaces = nBFaces;
    c->nTCell = nCells + nBFaces;
    c->xn.assign(x, x + nNodes); c->yn.assign(y, y + nNodes); c->zn.assign(z, z + nNodes);
    c->fnOff.assign(nFaces + 1, 0);
    for (int f = 0; f < nFaces; ++f) c->fnOff[f + 1] = c->fnOff[f] + faceNodeCount[f];
    c->fnNodes.assign(faceNodes, faceNodes + c->fnOff[nFaces]);
    c->lc.assign(lc, lc + nFaces); c->rc.assign(rc, rc + nFaces);
    c->bcType.assign(bcType, bcType + nBFaces);
    // UnsGrid::ReadGridFaceTopology (reverse faces with lc < 0) + NormalizeBc
    // (codes/geometry/src/UnsGrid.cpp:123-176)
    for (int f = 0; f < nFaces; ++f) {
        if (c->lc[f] < 0) {
            std::reverse(c->fnNodes.begin() + c->fnOff[f], c->fnNodes.begin() + c->fnOff[f + 1]);
            std::swap(c->lc[f], c->rc[f]);
        }
    }
    for (int f = 0; f < nBFaces; ++f) c->rc[f] = f + nCells;
    for (int f = nBFaces; f < nFaces; ++f) {
        if (c->lc[f] < 0 || c->lc[f] >= nCells || c->rc[f] < 0 || c->rc[f] >= nCells) {
            g_err = "ofo_create: interior face with bad cell index"; delete c; return nullptr;
        }
    }
    c->xfn.assign(nFaces, 0); c->yfn.assign(nFaces, 0); c->zfn.assign(nFaces, 0); c->area.assign(nFaces, 0);
    c->xfc.assign(nFaces, 0); c->yfc.assign(nFaces, 0); c->zfc.assign(nFaces, 0); c->vfn.assign(nFaces, 0);
    c->xcc.assign(c->nTCell, 0); c->ycc.assign(c->nTCell, 0); c->zcc.assign(c->nTCell, 0); c->vol.assign(c->nTCell, 0);
    if (dim == 2) { face_metrics_2d(*c); cell_metrics_2d(*c); }
    else { face_metrics_3d(*c); cell_metrics_3d(*c); }
    build_c2f(*c);
    wall_distance(*c);
    alloc_fields(*c);
    return c;
}

void ofo_destroy(ofo_case* c) { delete c; }

int ofo_mesh_info(ofo_case* c, int* out5) {
    out5[0] = c->nCells; out5[1] = c->nFaces; out5[2] = c->nBFaces; out5[3] = c->nTCell; out5[4] = 0;
    return 0;
}

int ofo_config_set(ofo_case* c, const ofx_config* cfg) {
    if (!(cfg->aref > 0.0)) return fail("aref must be positive");
    if (cfg->time_integral != 1 && cfg->time_integral != 2) return fail("time_integral must be 1 or 2");
    if (cfg->time_integral == 1 && cfg->rk_stage < 1) return fail("rk_stage must be >= 1");
    if (cfg->ischeme != 1 && cfg->ischeme != 12) return fail("ischeme must be 1 (Roe) or 12 (SLAU2)");
    if (cfg->ilim < -1 || cfg->ilim > 2) return fail("ilim must be -1..2");
    if (cfg->vismodel != 0 && cfg->vismodel != 1 && cfg->vismodel != 3) return fail("vismodel must be 0, 1 or 3");
    if (cfg->turb_ilim != -1 || cfg->tns_ilim != -1) return fail("only turb_ilim = tns_ilim = -1 supported");
    if (cfg->timestepModel != 0) return fail("only local time stepping supported");
    c->cfg = *cfg;
    c->has_cfg = true;
    set_constants(*c);
    return 0;
}

// INIT_FLOWFIELD (freestream) + INIT_FINAL for NS then SA
int ofo_init_freestream(ofo_case* c) {
    if (!c->has_cfg) return fail("config not set");
    for (int e = 0; e < 5; ++e) std::fill(c->q[e].begin(), c->q[e].end(), c->inflow[e]);
    std::fill(c->turbq.begin(), c->turbq.end(), c->sa_inflow);
    std::fill(c->vist.begin(), c->vist.end(), 0.0);
    ns_boundary(*c);
    if (c->cfg.vismodel == 3) { sa_vist(*c); sa_bc(*c); }
    return 0;
}

static RV* field_ptr(ofo_case* c, int field, int& neq, int& n) {
    n = c->nTCell; neq = 1;
    switch (field) {
        case OFX_F_Q: neq = 5; return c->q;
        case OFX_F_TURBQ: return &c->turbq;
        case OFX_F_RES: neq = 5; return c->res;
        case OFX_F_TURBRES: return &c->turbres;
        case OFX_F_LIMITER: neq = 5; return c->limiter;
        case OFX_F_TIMESTEP: return &c->timestep;
        case OFX_F_INVSR: return &c->invsr;
        case OFX_F_VISSR: return &c->vissr;
        case OFX_F_DQ: neq = 5; return c->dq;
        case OFX_F_TURBDQ: return &c->turbdq;
        case OFX_F_TEMPR: return &c->tempr;
        case OFX_F_VISL: return &c->visl;
        case OFX_F_VIST: return &c->vist;
        case OFX_F_DIST: n = c->nCells; return &c->dist;
        case OFX_F_VOL: return &c->vol;
        case OFX_F_TURBIMPSR: return &c->turbimpsr;
        case OFX_F_IMPSR: return &c->impsr;
        default: return nullptr;
    }
}

int ofo_get_field(ofo_case* c, int field, double* h) {
    int T = c->nTCell, F = c->nFaces;
    if (field == OFX_F_DQDX) {
        for (int e = 0; e < 5; ++e) {
            std::copy(c->dqdx[e].begin(), c->dqdx[e].end(), h + (0 * 5 + e) * T);
            std::copy(c->dqdy[e].begin(), c->dqdy[e].end(), h + (1 * 5 + e) * T);
            std::copy(c->dqdz[e].begin(), c->dqdz[e].end(), h + (2 * 5 + e) * T);
        }
        return 0;
    }
    if (field == OFX_F_DTDX || field == OFX_F_TURBDQDX || field == OFX_F_XCC) {
        const RV* a = field == OFX_F_DTDX ? &c->dtdx : field == OFX_F_TURBDQDX ? &c->turbdqdx : &c->xcc;
        const RV* b = field == OFX_F_DTDX ? &c->dtdy : field == OFX_F_TURBDQDX ? &c->turbdqdy : &c->ycc;
        const RV* d = field == OFX_F_DTDX ? &c->dtdz : field == OFX_F_TURBDQDX ? &c->turbdqdz : &c->zcc;
        std::copy(a->begin(), a->end(), h); std::copy(b->begin(), b->end(), h + T); std::copy(d->begin(), d->end(), h + 2 * T);
        return 0;
    }
    if (field == OFX_F_FACE_NORMAL) {
        std::copy(c->xfn.begin(), c->xfn.end(), h); std::copy(c->yfn.begin(), c->yfn.end(), h + F);
        std::copy(c->zfn.begin(), c->zfn.end(), h + 2 * F); std::copy(c->area.begin(), c->area.end(), h + 3 * F);
        return 0;
    }
    if (field == OFX_F_FACE_CENTER) {
        std::copy(c->xfc.begin(), c->xfc.end(), h); std::copy(c->yfc.begin(), c->yfc.end(), h + F);
        std::copy(c->zfc.begin(), c->zfc.end(), h + 2 * F);
        return 0;
    }
    int neq, n;
    RV* p = field_ptr(c, field, neq, n);
    if (!p) return fail("unknown field");
    for (int e = 0; e < neq; ++e) std::copy(p[e].begin(), p[e].begin() + n, h + (size_t)e * n);
    return 0;
}

int ofo_set_field(ofo_case* c, int field, const double* h) {
    if (field == OFX_F_DQDX) {
        const int T = c->nTCell;
        for (int e = 0; e < 5; ++e) {
            std::copy(h + (0 * 5 + e) * T, h + (0 * 5 + e + 1) * T, c->dqdx[e].begin());
            std::copy(h + (1 * 5 + e) * T, h + (1 * 5 + e + 1) * T, c->dqdy[e].begin());
            std::copy(h + (2 * 5 + e) * T, h + (2 * 5 + e + 1) * T, c->dqdz[e].begin());
        }
        return 0;
    }
    int neq, n;
    RV* p = field_ptr(c, field, neq, n);
    if (!p || field == OFX_F_DIST || field == OFX_F_VOL) return fail("field not settable");
    for (int e = 0; e < neq; ++e) std::copy(h + (size_t)e * n, h + (size_t)(e + 1) * n, p[e].begin());
    return 0;
}

int ofo_ns_boundary(ofo_case* c) { ns_boundary(*c); return 0; }
int ofo_sa_boundary(ofo_case* c) {
    if (c->cfg.vismodel != 3) return 0;
    sa_vist(*c); sa_bc(*c); return 0;
}
int ofo_timestep(ofo_case* c, int iter) { ns_timestep(*c, iter); return 0; }

int ofo_residual_ns(ofo_case* c) {
    for (int e = 0; e < 5; ++e) std::fill(c->res[e].begin(), c->res[e].end(), 0.0);
    ns_inv_flux(*c);
    ns_vis_flux(*c);
    for (int e = 0; e < 5; ++e) if (has_nan(c->res[e], c->nCells)) { g_err = "NotANumber"; return OFX_ENAN; }
    return 0;
}

int ofo_residual_sa(ofo_case* c) {
    if (c->cfg.vismodel != 3) return 0;
    sa_residual(*c);
    if (has_nan(c->turbres, c->nCells)) { g_err = "NotANumber"; return OFX_ENAN; }
    return 0;
}

int ofo_lusgs_ns(ofo_case* c) {
    for (int e = 0; e < 5; ++e) std::fill(c->dq[e].begin(), c->dq[e].end(), 0.0);
    ns_lusgs(*c);
    return 0;
}
int ofo_lusgs_sa(ofo_case* c) {
    if (c->cfg.vismodel != 3) return 0;
    std::fill(c->turbdq.begin(), c->turbdq.end(), 0.0);
    sa_lusgs(*c);
    return 0;
}
int ofo_update_ns(ofo_case* c) { ns_update(*c); return 0; }
int ofo_update_sa(ofo_case* c) { if (c->cfg.vismodel == 3) sa_update(*c); return 0; }

int ofo_norms_get(ofo_case* c, ofx_norms* out) {
    for (int e = 0; e < 5; ++e) {
        double s = 0;
        for (int i = 0; i < c->nCells; ++i) s += SQR(c->res[e][i]);
        out->ns[e] = std::sqrt(s / c->nCells);
    }
    double s = 0;
    if (c->cfg.vismodel == 3) for (int i = 0; i < c->nCells; ++i) s += SQR(c->turbres[i]);
    out->sa = std::sqrt(s / c->nCells);
    return 0;
}

// One StrongIter per step: NS then SA, each TimeIntegral::Lusgs
int ofo_lhs(ofo_case* c, int solver, double lhscoef) {
    if (solver != 0 && solver != 1) return fail("lhs: solver must be 0 or 1");
    if (solver == 1 && c->cfg.vismodel != 3) return 0;
    calc_lhs(*c, solver, lhscoef);
    return 0;
}

// TimeIntegral::RungeKutta of the NS solver, then of the SA solver
// (codes/solver/src/TimeIntegral.cpp:81-110, one grid level): CALC_TIME_STEP,
// then per stage with lhscoef = rk_coef[stage] = 1/(stage+1)
// (codes/global/src/Ctrl.cpp:74-79) UPDATE_RESIDUALS, CALC_LHS,
// UPDATE_FLOWFIELD, CALC_BOUNDARY.  LOAD_Q and LOAD_RESIDUALS copy into
// multigrid storage that a one-level run never reads back.  The SA solver
// registers no CALC_TIME_STEP (system/turb/function/funcMap.txt).
static int rk_step(ofo_case* c, int it) {
    ofo_timestep(c, it);
    for (int s = 0; s < c->cfg.rk_stage; ++s) {
        int rc = ofo_residual_ns(c);
        if (rc) return rc;
        calc_lhs(*c, 0, 1.0 / (s + 1));
        ns_update(*c);
        ns_boundary(*c);
    }
    if (c->cfg.vismodel == 3) {
        for (int s = 0; s < c->cfg.rk_stage; ++s) {
            int rc = ofo_residual_sa(c);
            if (rc) return rc;
            calc_lhs(*c, 1, 1.0 / (s + 1));
            sa_update(*c);
            sa_vist(*c); sa_bc(*c);
        }
    }
    return 0;
}

int ofo_iterate(ofo_case* c, int first_iter, int n, ofx_norms* norms_out) {
    for (int k = 0; k < n; ++k) {
        int it = first_iter + k;
        if (c->cfg.time_integral == 1) {
            int rc = rk_step(c, it);
            if (rc) return rc;
            if (norms_out) ofo_norms_get(c, norms_out + k);
            continue;
        }
        ofo_timestep(c, it);
        int rc = ofo_residual_ns(c);
        if (rc) return rc;
        ofo_lusgs_ns(c);
        ofo_update_ns(c);
        ns_boundary(*c);
        if (c->cfg.vismodel == 3) {
            rc = ofo_residual_sa(c);
            if (rc) return rc;
            ofo_lusgs_sa(c);
            ofo_update_sa(c);
            sa_vist(*c); sa_bc(*c);
        }
        if (norms_out) ofo_norms_get(c, norms_out + k);
    }
    return 0;
}

// Pressure and viscous force of wall face f and its pressure coefficient
// (CalcAeroForce's face loop, codes/force/src/AeroForceTask.cpp:365-424, with
// Stress::CalcForce, codes/solver/src/Stress.cpp).
static void wall_face(const ofo_case* c, int f, double* pf, double* vf, double& cp) {
    double orey = 1.0 / c->cfg.reynolds;
    int l = c->lc[f], r = c->rc[f];
    double area = c->area[f], fnx = c->xfn[f], fny = c->yfn[f], fnz = c->zfn[f];
    double fanx = fnx * area, fany = fny * area, fanz = fnz * area;
    double dx = c->xfc[f] - c->xcc[l], dy = c->yfc[f] - c->ycc[l], dz = c->zfc[f] - c->zcc[l];
    double wp = c->q[4][l] + c->bcdqdx[4][f] * dx + c->bcdqdy[4][f] * dy + c->bcdqdz[4][f] * dz;
    if (wp < 0.0) wp = c->q[4][l];
    cp = 2.0 * (wp - c->inflow[4]);
    pf[0] = fanx * cp; pf[1] = fany * cp; pf[2] = fanz * cp;
    vf[0] = vf[1] = vf[2] = 0.0;
    if (c->cfg.vismodel > 0) {
        double g[3][3];
        for (int k = 0; k < 3; ++k) { g[k][0] = c->bcdqdx[1 + k][f]; g[k][1] = c->bcdqdy[1 + k][f]; g[k][2] = c->bcdqdz[1 + k][f]; }
        double ex = c->xcc[r] - c->xcc[l], ey = c->ycc[r] - c->ycc[l], ez = c->zcc[r] - c->zcc[l];
        double ods = 1.0 / DIST(ex, ey, ez);
        ex *= ods; ey *= ods; ez *= ods;
        for (int k = 0; k < 3; ++k) {
            double fOld = g[k][0] * ex + g[k][1] * ey + g[k][2] * ez;
            double fnew = (c->q[1 + k][r] - c->q[1 + k][l]) * ods;
            double df = fnew - fOld;
            g[k][0] += df * ex; g[k][1] += df * ey; g[k][2] += df * ez;
        }
        double vis = c->visl[l];
        double dudx = g[0][0], dudy = g[0][1], dudz = g[0][2];
        double dvdx = g[1][0], dvdy = g[1][1], dvdz = g[1][2];
        double dwdx = g[2][0], dwdy = g[2][1], dwdz = g[2][2];
        double divv2p3 = two3rd * (dudx + dvdy + dwdz);
        double txx = vis * (2.0 * dudx - divv2p3), tyy = vis * (2.0 * dvdy - divv2p3), tzz = vis * (2.0 * dwdz - divv2p3);
        double txy = vis * (dudy + dvdx), txz = vis * (dudz + dwdx), tyz = vis * (dvdz + dwdy);
        double coef = -2.0 * area * orey;
        vf[0] = coef * (fnx * txx + fny * txy + fnz * txz);
        vf[1] = coef * (fnx * txy + fny * tyy + fnz * tyz);
        vf[2] = coef * (fnx * txz + fny * tyz + fnz * tzz);
    }
}

// CalcAeroForce + AeroForceInfo::CalcCoef (codes/force/src/AeroForceTask.cpp:186-440,
// codes/force/src/Force.cpp:148-270, codes/solver/src/Stress.cpp), aref = 1.
// AerodynamicForceTask::Run + Dump (codes/force/src/AeroForceTask.cpp:59-150):
// the wall forces summed in face order (AeroForce::AddForce adds total, pres,
// vis and power, codes/force/src/Force.cpp:202-208 -- not the moments), then
// AeroForceInfo::CalcCoef (codes/force/src/Force.cpp:259-276) -> the value
// columns of one aero.dat row: CL, CD, CD_PR, CD_SF, CDL2, Xcp, Fx, Fy, Fz,
// Cmx, Cmy, Cmz.
int ofo_aero_row(ofo_case* c, double* out12) {
    double aoa = c->cfg.aoa_degree * PI / 180.0, aos = c->cfg.sideslip_degree * PI / 180.0;
    double sina = std::sin(aoa), cosa = std::cos(aoa), sinb = std::sin(aos), cosb = std::cos(aos);
    double tot[3] = {0, 0, 0}, pres[3] = {0, 0, 0}, mom[3] = {0, 0, 0};
    for (int f = 0; f < c->nBFaces; ++f) {
        if (c->bcType[f] != BC_SOLID_SURFACE) continue;
        double pf[3], vf[3], cp;
        wall_face(c, f, pf, vf, cp);
        double face_total[3];
        for (int k = 0; k < 3; ++k) face_total[k] = pf[k] + vf[k];
        for (int k = 0; k < 3; ++k) { tot[k] += face_total[k]; pres[k] += pf[k]; }
    }
    double aref = c->cfg.aref, cmoment = c->cfg.aref * 1.0;  // lref = 1 in every case script
    double cf[3], cpres[3], cmom[3];
    for (int k = 0; k < 3; ++k) { cf[k] = tot[k] / aref; cpres[k] = pres[k] / aref; cmom[k] = mom[k] / cmoment; }
    const double AR = 9.5;
    double cl = -sina * cf[0] + cosa * cf[1];
    double cd = cosa * cosb * cf[0] + sina * cosb * cf[1] + sinb * cf[2];
    double cd_pres = cosa * cosb * cpres[0] + sina * cosb * cpres[1] + sinb * cpres[2];
    double cd_vis = cd - cd_pres;
    double cdl = cd - SQR(cl) / (PI * AR);
    double pres_center = SIGN(1.0, cf[1]) * cmom[2] / (ABS(cf[1]) + SMALL);
    double row[12] = {cl, cd, cd_pres, cd_vis, cdl, pres_center, cf[0], cf[1], cf[2], cmom[0], cmom[1], cmom[2]};
    for (int k = 0; k < 12; ++k) out12[k] = row[k];
    return 0;
}

int ofo_aero(ofo_case* c, double* out4) {
    double row[12];
    ofo_aero_row(c, row);
    for (int k = 0; k < 4; ++k) out4[k] = row[k];
    return 0;
}

// DUMP_PRESSURE_COEFF rows (the idump_pres branch of CalcAeroForce,
// codes/force/src/AeroForceTask.cpp:426-440, AeroCom::CalcCF,
// codes/force/src/Force.cpp:160-165): per wall face in face order
// x, y, z of the face centre, -cp, cf -> out[5 * k .. 5 * k + 4].
int ofo_wall_data(ofo_case* c, double* out) {
    double aoa = c->cfg.aoa_degree * PI / 180.0;
    double sina = std::sin(aoa), cosa = std::cos(aoa);
    int k = 0;
    for (int f = 0; f < c->nBFaces; ++f) {
        if (c->bcType[f] != BC_SOLID_SURFACE) continue;
        double pf[3], vf[3], cp;
        wall_face(c, f, pf, vf, cp);
        double dir = cosa * vf[0] + sina * vf[1];
        double cf = DIST(vf[0], vf[1], vf[2]) / c->area[f] * SIGN(1.0, dir);
        double* o = out + 5 * (size_t)k++;
        o[0] = c->xfc[f]; o[1] = c->yfc[f]; o[2] = c->zfc[f]; o[3] = -cp; o[4] = cf;
    }
    return 0;
}

int ofo_wall_count(ofo_case* c) {
    int n = 0;
    for (int f = 0; f < c->nBFaces; ++f) n += c->bcType[f] == BC_SOLID_SURFACE;
    return n;
}

int ofo_face_flux(ofo_case* c, double* inv, double* vis) {
    int F = c->nFaces;
    for (int e = 0; e < 5; ++e) {
        std::copy(c->invflux[e].begin(), c->invflux[e].end(), inv + (size_t)e * F);
        std::copy(c->visflux[e].begin(), c->visflux[e].end(), vis + (size_t)e * F);
    }
    return 0;
}

int ofo_pow(const double* x, const double* y, int n, double* out) {
    for (int i = 0; i < n; ++i) out[i] = ofxm::ofx_pow(x[i], y[i]);
    return 0;
}

}  // extern "C"
