// ofx_physics.h -- per-face / per-cell physics of the NS + SA path as device
// functions.  Each keeps the operation order of the reference function it is
// named after, so that with -ffp-contract=off the FP64 results are the
// reference's bit for bit wherever only + - * / sqrt are involved.
#pragma once

#include "ofx_common.h"

namespace ofx {

// Sutherland::CalcViscosity (codes/model/src/Sutherland.cpp), CalcLaminarViscosity clamp
__device__ __forceinline__ double sutherland(double t, double ts) {
    double t3 = t * t * t;
    double v = sqrt(t3) * (1 + ts) / (t + ts);
    return dmax(0.0, v);
}

// NsCalcGamaT: T = p / (statecoef * rho * oamw), oamw = 1 (codes/ns/src/NsRhs.cpp:95-102)
__device__ __forceinline__ double temperature(double rho, double p, double statecoef) {
    return p / (statecoef * rho * 1.0);
}

// Real-type generic helpers for the face fluxes (R = double: the FP64 path,
// R = float: the FP32-flux variant); same expressions as dmin/dmax/dabs/sqr3.
template <typename R> __device__ __forceinline__ R tmin(R a, R b) { return (a < b) ? a : b; }
template <typename R> __device__ __forceinline__ R tmax(R a, R b) { return (a > b) ? a : b; }
template <typename R> __device__ __forceinline__ R tabs(R a) { return (a < R(0)) ? -a : a; }
template <typename R> __device__ __forceinline__ R tsq3(R a, R b, R c) { return a * a + b * b + c * c; }

// PrimToQ + CalcInternalEnergy (codes/ns/src/NsInvFlux.cpp:1349-1370,1392-1406)
template <typename R>
__device__ __forceinline__ void prim_to_q(const R* p, R gama, R* q) {
    R v2 = tsq3(p[1], p[2], p[3]);
    R em = (p[4] / p[0]) / (gama - R(1.0)) + R(0.5) * v2;
    q[0] = p[0]; q[1] = p[0] * p[1]; q[2] = p[0] * p[2]; q[3] = p[0] * p[3]; q[4] = p[0] * em;
}

// NsInvFlux::Roe with ModifyAbsoluteEigenvalue and CalcTotalEnthalpyChange
// (codes/ns/src/NsInvFlux.cpp:167-308,1329-1347); static faces (vfn = 0);
// returns the flux per unit area.
template <typename R>
__device__ __forceinline__ void roe_flux(int ieigenfix, R centropy1, R centropy2, const R* p1, const R* p2,
                                         R gama1, R gama2, R xfn, R yfn, R zfn, R* flux) {
    const R vfn = R(0.0);
    R rl = p1[0], ul = p1[1], vl = p1[2], wl = p1[3], pl = p1[4];
    R rr = p2[0], ur = p2[1], vr = p2[2], wr = p2[3], pr = p2[4];
    R v2l = tsq3(ul, vl, wl), v2r = tsq3(ur, vr, wr);
    R hint_l = (gama1 / (gama1 - R(1.0))) * (pl / rl);
    R hint_r = (gama2 / (gama2 - R(1.0))) * (pr / rr);
    R hl = hint_l + R(0.5) * v2l, hr = hint_r + R(0.5) * v2r;
    R vnl = xfn * ul + yfn * vl + zfn * wl - vfn;
    R vnr = xfn * ur + yfn * vr + zfn * wr - vfn;
    R rvnl = rl * vnl, rvnr = rr * vnr;
    R ratio = sqrt(rr / rl);
    R coef = R(1.0) / (R(1.0) + ratio);
    R um = (ul + ur * ratio) * coef;
    R vm = (vl + vr * ratio) * coef;
    R wm = (wl + wr * ratio) * coef;
    R hm = (hl + hr * ratio) * coef;
    R gama = (gama1 + gama2 * ratio) * coef;
    R v2 = tsq3(um, vm, wm);
    R vnflow = xfn * um + yfn * vm + zfn * wm;
    R vnrel = vnflow - vfn;
    R gamm1 = gama - R(1.0);
    R c2 = gamm1 * (hm - R(0.5) * v2);
    R cm = sqrt(tabs(c2));
    R aeig1 = tabs(vnrel), aeig2 = tabs(vnrel + cm), aeig3 = tabs(vnrel - cm);
    flux[0] = rvnl + rvnr;
    flux[1] = (rvnl * ul + xfn * pl) + (rvnr * ur + xfn * pr);
    flux[2] = (rvnl * vl + yfn * pl) + (rvnr * vr + yfn * pr);
    flux[3] = (rvnl * wl + zfn * pl) + (rvnr * wr + zfn * pr);
    flux[4] = (rvnl * hl + vfn * pl) + (rvnr * hr + vfn * pr);
    R q1[5], q2[5], dq[5];
    prim_to_q(p1, gama1, q1);
    prim_to_q(p2, gama2, q2);
    for (int e = 0; e < 5; ++e) dq[e] = q2[e] - q1[e];
    R meig1 = aeig1, meig2 = aeig2, meig3 = aeig3;
    if (ieigenfix == 1) {
        R maxE = tmax(aeig2, aeig3);
        R m1 = maxE * centropy1, m2 = maxE * centropy2;
        meig1 = tmax(m1, aeig1); meig2 = tmax(m2, aeig2); meig3 = tmax(m2, aeig3);
    }
    R xi1 = (R(2.0) * meig1 - meig2 - meig3) / (R(2.0) * c2);
    R xi2 = (meig2 - meig3) / (R(2.0) * cm);
    R dc = vnflow * dq[0] - xfn * dq[1] - yfn * dq[2] - zfn * dq[3];
    R c2dc = c2 * dc;
    R ae = gama - R(1.0);
    R af = R(0.5) * ae * v2;
    R dh = -ae * (um * dq[1] + vm * dq[2] + wm * dq[3] - dq[4]);
    dh += af * dq[0];
    R term1 = dh * xi1 + dc * xi2;
    R term2 = c2dc * xi1 + dh * xi2;
    flux[0] -= (meig1 * dq[0] - term1);
    flux[1] -= (meig1 * dq[1] - um * term1 + xfn * term2);
    flux[2] -= (meig1 * dq[2] - vm * term1 + yfn * term2);
    flux[3] -= (meig1 * dq[3] - wm * term1 + zfn * term2);
    flux[4] -= (meig1 * dq[4] - hm * term1 + vnflow * term2);
    for (int e = 0; e < 5; ++e) flux[e] *= R(0.5);
}

// FMSplit2 / FPSplit5 (codes/ns/src/NsInvFlux.cpp:42-65)
template <typename R>
__device__ __forceinline__ R fm_split2(R mach, R ipn) { return ipn * R(0.25) * ((mach + ipn) * (mach + ipn)); }
template <typename R>
__device__ __forceinline__ R fp_split5(R mach, R alpha, R ipn) {
    R macha = tabs(mach);
    if (macha >= R(1.0)) return R(0.5) * (R(1.0) + ipn * mach / (macha + R(SMALL)));
    return fm_split2(mach, ipn) * ((R(2.0) * ipn - mach) - ipn * R(16.0) * alpha * mach * fm_split2(mach, -ipn));
}

// NsInvFlux::Slau2 (codes/ns/src/NsInvFlux.cpp:1246-1314), static faces
template <typename R>
__device__ __forceinline__ void slau2_flux(const R* p1, const R* p2, R gama1, R gama2,
                                           R xfn, R yfn, R zfn, R* flux) {
    const R vfn = R(0.0);
    R rl = p1[0], ul = p1[1], vl = p1[2], wl = p1[3], pl = p1[4];
    R rr = p2[0], ur = p2[1], vr = p2[2], wr = p2[3], pr = p2[4];
    R v2l = tsq3(ul, vl, wl), v2r = tsq3(ur, vr, wr);
    R hint_l = (gama1 / (gama1 - R(1.0))) * (pl / rl);
    R hint_r = (gama2 / (gama2 - R(1.0))) * (pr / rr);
    R hl = hint_l + R(0.5) * v2l, hr = hint_r + R(0.5) * v2r;
    R vnl = xfn * ul + yfn * vl + zfn * wl - vfn;
    R vnr = xfn * ur + yfn * vr + zfn * wr - vfn;
    R orl = R(1.0) / (rl + R(SMALL)), orr = R(1.0) / (rr + R(SMALL));
    R c2l = gama1 * pl * orl, c2r = gama2 * pr * orr;
    R cl = sqrt(c2l), cr = sqrt(c2r);
    R cm = R(0.5) * (cl + cr);
    R t = rr / rl;
    R vna = (tabs(vnl) + t * tabs(vnr)) / (R(1.0) + t);
    R ml = vnl / cm, mr = vnr / cm;
    R g = -tmax(tmin(ml, R(0.0)), R(-1.0)) * tmin(tmax(mr, R(0.0)), R(1.0));
    R vnp = (R(1.0) - g) * vna + g * tabs(vnl);
    R vnn = (R(1.0) - g) * vna + g * tabs(vnr);
    R va = sqrt(R(0.5) * (v2l + v2r));
    R m12 = tmin(R(1.0), va / cm);
    R ka = (R(1.0) - m12) * (R(1.0) - m12);
    R ms = R(0.5) * (rl * (vnl + vnp) + rr * (vnr - vnn) - ka * (pr - pl) / cm);
    R fp5ml = fp_split5(ml, R(0.0), R(1.0));
    R fp5mr = fp_split5(mr, R(0.0), R(-1.0));
    R ps = R(0.0);
    ps += R(0.5) * (pl + pr);
    ps += R(0.5) * (fp5ml - fp5mr) * (pl - pr);
    ps += va * (fp5ml + fp5mr - R(1.0)) * sqrt(rl * rr) * cm;
    R mp = R(0.5) * (ms + tabs(ms)), mn = R(0.5) * (ms - tabs(ms));
    flux[0] = (mp + mn);
    flux[1] = (mp * ul + mn * ur) + xfn * ps;
    flux[2] = (mp * vl + mn * vr) + yfn * ps;
    flux[3] = (mp * wl + mn * wr) + zfn * ps;
    flux[4] = (mp * hl + mn * hr) + vfn * ps;
}

// GCom::CalcTangent (codes/global/src/Com.cpp:76-121)
__device__ __forceinline__ void tangents(double xfn, double yfn, double zfn, double* t1, double* t2) {
    if (xfn != 0.0) { t1[0] = yfn; t1[1] = -xfn; t1[2] = 0.0; }
    else if (yfn != 0.0) { t1[0] = -yfn; t1[1] = xfn; t1[2] = 0.0; }
    else if (zfn != 0.0) { t1[0] = 0.0; t1[1] = -zfn; t1[2] = yfn; }
    else { t1[0] = t1[1] = t1[2] = 0; t2[0] = t2[1] = t2[2] = 0; return; }
    double dtmp = 1.0 / sqrt(sqr3(t1[0], t1[1], t1[2]));
    t1[0] *= dtmp; t1[1] *= dtmp; t1[2] *= dtmp;
    t2[0] = yfn * t1[2] - zfn * t1[1];
    t2[1] = zfn * t1[0] - xfn * t1[2];
    t2[2] = xfn * t1[1] - yfn * t1[0];
}

// Geometry of one face for the VisGrad face-gradient methods
// (VisGradGeom::CalcFaceWeight / CalcGradCoef / PrepareCellGeom,
//  codes/usolver/src/VisGrad.cpp:44-103)
struct VisGeom {
    double xfn, yfn, zfn, t1[3], t2[3];
    double fw1, fw2;
    double dx, dy, dz, ods;
    double d1, d2, dxnl, dynl, dznl, dxnr, dynr, dznr, angle1, angle2;
    bool bc_nonint;  // boundary face that is not an interface (VisGrad::CalcC1C2)
};

__device__ __forceinline__ double vis_angle(double dx, double dy, double dz, double dist) {
    double res = dist / (sqrt(sqr3(dx, dy, dz)) + SMALL);
    if (res > 1.0) res = 1.0;
    if (res < -1.0) res = -1.0;
    return asin(res) * 180.0 / PI;
}

__device__ __forceinline__ void vis_geom(const DevMesh& m, int f, int l, int r, int method, VisGeom& G) {
    G.xfn = m.xfn[f]; G.yfn = m.yfn[f]; G.zfn = m.zfn[f];
    G.bc_nonint = f < m.B;
    tangents(G.xfn, G.yfn, G.zfn, G.t1, G.t2);
    double xfc = m.xfc[f], yfc = m.yfc[f], zfc = m.zfc[f];
    double xl = m.xcc[l], yl = m.ycc[l], zl = m.zcc[l];
    double xr = m.xcc[r], yr = m.ycc[r], zr = m.zcc[r];
    double delt1 = sqrt(sqr3(xfc - xl, yfc - yl, zfc - zl));
    double delt2 = sqrt(sqr3(xfc - xr, yfc - yr, zfc - zr));
    double delta = 1.0 / (delt1 + delt2 + SMALL);
    G.fw1 = delt2 * delta; G.fw2 = delt1 * delta;
    if (method == 0) {
        G.dx = xr - xl; G.dy = yr - yl; G.dz = zr - zl;
        G.ods = 1.0 / sqrt(sqr3(G.dx, G.dy, G.dz));
        G.dx *= G.ods; G.dy *= G.ods; G.dz *= G.ods;
    } else if (method == 2) {
        double ax = xl - xfc, ay = yl - yfc, az = zl - zfc;
        double bx = xr - xfc, by = yr - yfc, bz = zr - zfc;
        G.d1 = G.xfn * ax + G.yfn * ay + G.zfn * az;
        G.d2 = G.xfn * bx + G.yfn * by + G.zfn * bz;
        G.dxnl = G.xfn * G.d1 - ax; G.dynl = G.yfn * G.d1 - ay; G.dznl = G.zfn * G.d1 - az;
        G.dxnr = G.xfn * G.d2 - bx; G.dynr = G.yfn * G.d2 - by; G.dznr = G.zfn * G.d2 - bz;
        G.angle1 = vis_angle(ax, ay, az, -G.d1);
        G.angle2 = vis_angle(bx, by, bz, G.d2);
    }
}

// The raw records of face f for face-gradient method `method` (run-time,
// wave-uniform): rec[0] the STD record (method 0), rec[0..4] the TEST
// records (method 2), zeros otherwise.  For kernels that issue these loads
// early, before any arithmetic; vis_geom_rec unpacks them as vis_geom_pre
// (below) loads them.
__device__ __forceinline__ void vis_rec_load(const DevMesh& m, int method, int f, dvec4* rec) {
    for (int t = 0; t < 5; ++t) rec[t] = dvec4{0.0, 0.0, 0.0, 0.0};
    if (method == 0) {
        rec[0] = m.fv0[f];
    } else if (method == 2) {
        const dvec4* r = m.fv2 + 5 * (size_t)f;
        for (int t = 0; t < 5; ++t) rec[t] = r[t];
    }
}
template <int METHOD>
__device__ __forceinline__ void vis_geom_rec(double xfn, double yfn, double zfn, bool bc_nonint, const dvec4* rec,
                                             VisGeom& G) {
    G.xfn = xfn; G.yfn = yfn; G.zfn = zfn;
    G.bc_nonint = bc_nonint;
    if (METHOD == 0) {
        dvec4 a = rec[0];
        G.dx = a.x; G.dy = a.y; G.dz = a.z; G.ods = a.w;
    } else if (METHOD == 2) {
        dvec4 a = rec[0], b = rec[1], c = rec[2], e = rec[3], w = rec[4];
        G.fw1 = w.x; G.fw2 = w.y;
        G.d1 = a.x; G.d2 = a.y; G.angle1 = a.z; G.angle2 = a.w;
        G.dxnl = b.x; G.dynl = b.y; G.dznl = b.z; G.dxnr = b.w;
        G.dynr = c.x; G.dznr = c.y; G.t1[0] = c.z; G.t1[1] = c.w;
        G.t1[2] = e.x; G.t2[0] = e.y; G.t2[1] = e.z; G.t2[2] = e.w;
    }
}
// The same geometry read from the per-face records the mesh upload stores
// (DevMesh::fv0, fv2), which vis_geom itself computed.  METHOD is the
// caller's compile-time choice, so no member is merged across branches.
template <int METHOD>
__device__ __forceinline__ void vis_geom_pre(const DevMesh& m, int f, VisGeom& G) {
    G.xfn = m.xfn[f]; G.yfn = m.yfn[f]; G.zfn = m.zfn[f];
    G.bc_nonint = f < m.B;
    if (METHOD == 0) {
        dvec4 a = m.fv0[f];
        G.dx = a.x; G.dy = a.y; G.dz = a.z; G.ods = a.w;
    } else if (METHOD == 2) {
        const dvec4* r = m.fv2 + 5 * (size_t)f;
        dvec4 a = r[0], b = r[1], c = r[2], e = r[3], w = r[4];
        G.fw1 = w.x; G.fw2 = w.y;
        G.d1 = a.x; G.d2 = a.y; G.angle1 = a.z; G.angle2 = a.w;
        G.dxnl = b.x; G.dynl = b.y; G.dznl = b.z; G.dxnr = b.w;
        G.dynr = c.x; G.dznr = c.y; G.t1[0] = c.z; G.t1[1] = c.w;
        G.t1[2] = e.x; G.t2[0] = e.y; G.t2[1] = e.z; G.t2[2] = e.w;
    }
}

// One scalar through VisGrad (AverGrad, then STD / AVER / TEST method;
// codes/usolver/src/VisGrad.cpp:130-365).  gf = face gradient, dqdn = normal derivative.
template <int METHOD>
__device__ __forceinline__ void vis_grad(const VisGeom& G, double q1, double q2, const double* g1,
                                         const double* g2, double& qf, double* gf, double& dqdn) {
    gf[0] = 0.5 * (g1[0] + g2[0]); gf[1] = 0.5 * (g1[1] + g2[1]); gf[2] = 0.5 * (g1[2] + g2[2]);
    dqdn = 0.0;
    qf = 0.5 * (q1 + q2);
    if (METHOD == 0) {
        double fOld = gf[0] * G.dx + gf[1] * G.dy + gf[2] * G.dz;
        double fnew = (q2 - q1) * G.ods;
        double df = fnew - fOld;
        gf[0] += df * G.dx; gf[1] += df * G.dy; gf[2] += df * G.dz;
        dqdn = G.xfn * gf[0] + G.yfn * gf[1] + G.zfn * gf[2];
    } else if (METHOD == 1) {
        dqdn = G.xfn * gf[0] + G.yfn * gf[1] + G.zfn * gf[2];
    } else {
        double q11 = q1, q22 = q2;
        if (G.angle1 > 20.0 && G.angle2 > 20.0) {
            q11 += g1[0] * G.dxnl + g1[1] * G.dynl + g1[2] * G.dznl;
            q22 += g2[0] * G.dxnr + g2[1] * G.dynr + g2[2] * G.dznr;
        }
        bool ok = G.angle1 > 0.0 && G.angle2 > 0.0 && dabs(G.d1) > SMALL && dabs(G.d2) > SMALL;
        if (ok) {
            double dtmp = G.d1 * G.d1 + G.d2 * G.d2;
            double c1 = (G.d1 * G.d1) / dtmp, c2 = 1.0 - c1;
            if (G.bc_nonint) { c1 = 1.0; c2 = 0.0; }
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

// Venkatakrishnan / Barth limiter functions (codes/usolver/src/ULimiter.cpp:35-97)
__device__ __forceinline__ double vencat(double x, double y, double eps) {
    double x2 = x * x, xy = x * y, y2 = y * y, cc = 2.0;
    return ((x2 + eps + cc * xy) / (x2 + cc * y2 + xy + eps));
}
__device__ __forceinline__ double vencat_fn(double dq, double minv, double maxv, double eps) {
    // one evaluation on the selected extremum (the same operations as
    // evaluating the taken branch; both branches would be if-converted)
    double phi = vencat(dq > SMALL ? maxv : minv, dq, eps);
    return (dq > SMALL || dq < -SMALL) ? phi : 1.0;
}
__device__ __forceinline__ double barth_fn(double dq, double minv, double maxv, double dot) {
    // one division by the selected extremum (the taken branch's operation)
    double r = (dq > maxv ? maxv : minv) / dq;
    if (dot < 0.0) return 0.0;
    return (dq > maxv || dq < minv) ? r : 1.0;
}

}  // namespace ofx
