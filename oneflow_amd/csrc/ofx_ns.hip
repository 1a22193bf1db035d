// ofx_ns.hip -- Navier-Stokes kernels: boundary conditions, local time step,
// residual (Green-Gauss gradients, limiter, MUSCL + Roe/SLAU2, viscous flux,
// ordered atomic-free gather), level-scheduled LU-SGS and the update.
//
// Data layout: SoA doubles, cell arrays [eq][T] (T = C + B, ghosts last),
// face arrays [eq][F], faces boundary-first.  Gathers walk each cell's face
// list in ascending face id, which is the order the reference's serial
// scatter AddF2CField (codes/usolver/src/UCom.cpp:144-171) adds into a cell,
// so the FP64 residual is the reference's bit for bit and no atomics are used.
#include "ofx_physics.h"

namespace ofx {

constexpr int TB = 256;

static inline int nblk(int n) { return (n + TB - 1) / TB; }

// ============================================================ boundary

// NsCalcGamaT + CalcLaminarViscosity on interior cells (codes/ns/src/NsSolverImp.cpp:72-79)
__global__ void __launch_bounds__(TB) k_ns_cellstate(DevMesh m, DevFields d, Consts k0) {
    int i = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    const Consts k = zoned(k0, m, cell_zone(m, i));
    double t = temperature(d.q[i], d.q[4 * m.T + i], k.statecoef);
    d.tempr[i] = t;
    d.visl[i] = sutherland(t, k.ts_suth);
}

// UNsBcSolver::CalcBc + NsBcSolver for one boundary face given the owner's
// primitive state s and temperature towner (codes/uns/src/UNsBcSolver.cpp:55-168,
// codes/ns/src/NsBcSolver.cpp:51-410); writes the ghost state, then the ghost
// gamma/T/viscosity (NsCalcGamaT / CalcLaminarViscosity on F_GHOST).
__device__ __forceinline__ void ns_bc_face(const DevMesh& m, const DevFields& d, const Consts& k, int f,
                                           const double* s, double towner) {
    const int T = m.T;
    int bt = m.bct[f];
    int r = m.C + f;
    double xfn = m.xfn[f], yfn = m.yfn[f], zfn = m.zfn[f];
    const double vfn = 0.0;
    double t[5];
    double gama = 0.5 * (k.gama + k.gama);
    if (bt == BC_EXTRAPOLATION || bt == BC_OUTFLOW) {
        for (int e = 0; e < 5; ++e) t[e] = s[e];
    } else if (bt == BC_INFLOW) {
        for (int e = 0; e < 5; ++e) t[e] = k.inflow[e];
    } else if (bt == BC_SOLID_SURFACE && k.vismodel != 0) {
        for (int e = 0; e < 5; ++e) t[e] = s[e];
        t[1] = -t[1] + 2.0 * 0.0; t[2] = -t[2] + 2.0 * 0.0; t[3] = -t[3] + 2.0 * 0.0;
        if (k.isowallbc != 0) {
            // IsothermalVisWallBc (codes/ns/src/NsBcSolver.cpp:271-324)
            double tlim = 0.1 * k.twall;
            double rm = s[0], pm = s[4];
            double pg1 = pm;
            double tg1 = 2.0 * k.twall - towner;
            if (tg1 < tlim) tg1 = tlim;
            double rw = pg1 / (k.statecoef * k.twall * 1.0);
            double rg1 = 2.0 * rw - rm;
            rg1 = dmax(0.5 * rw, rg1);
            t[0] = rg1; t[4] = pg1;
        }
    } else if (bt == BC_SYMMETRY || bt == BC_SOLID_SURFACE) {
        double vn = xfn * s[1] + yfn * s[2] + zfn * s[3] - vfn;
        for (int e = 0; e < 5; ++e) t[e] = s[e];
        t[1] = s[1] - 2.0 * xfn * vn;
        t[2] = s[2] - 2.0 * yfn * vn;
        t[3] = s[3] - 2.0 * zfn * vn;
    } else {  // BC_FARFIELD (codes/ns/src/NsBcSolver.cpp:140-236)
        double rin = s[0], uin = s[1], vin = s[2], win = s[3], pin = s[4];
        double rref = k.inflow[0], uref = k.inflow[1], vref = k.inflow[2], wref = k.inflow[3], pref = k.inflow[4];
        double vnref = xfn * uref + yfn * vref + zfn * wref - vfn;
        double vnin = xfn * uin + yfn * vin + zfn * win - vfn;
        double cref = sqrt(dabs(k.gama * pref / rref));
        double cin = sqrt(dabs(gama * pin / rin));
        double gamm1 = gama - 1.0;
        double velin = sqrt(sqr3(uin, vin, win));
        if (velin > cin) {
            if (vnin >= 0.0) { for (int e = 0; e < 5; ++e) t[e] = s[e]; }
            else { for (int e = 0; e < 5; ++e) t[e] = k.inflow[e]; }
        } else {
            double riemp = vnin + 2.0 * cin / gamm1;
            double riemm = vnref - 2.0 * cref / gamm1;
            double vnb = 0.5 * (riemp + riemm);
            double cb = 0.25 * (riemp - riemm) * gamm1;
            double vtx, vty, vtz, entr;
            if (vnb >= 0.0) {
                entr = pin / ofxm::ofx_pow(rin, gama);
                vtx = uin - xfn * vnin; vty = vin - yfn * vnin; vtz = win - zfn * vnin;
            } else {
                entr = pref / ofxm::ofx_pow(rref, gama);
                vtx = uref - xfn * vnref; vty = vref - yfn * vnref; vtz = wref - zfn * vnref;
            }
            double rb = ofxm::ofx_pow((cb * cb / (entr * gama)), 1.0 / gamm1);
            t[0] = rb; t[1] = vtx + xfn * vnb; t[2] = vty + yfn * vnb; t[3] = vtz + zfn * vnb;
            t[4] = cb * cb * rb / gama;
        }
    }
    for (int e = 0; e < 5; ++e) d.q[e * T + r] = t[e];
    double tt = temperature(t[0], t[4], k.statecoef);
    d.tempr[r] = tt;
    d.visl[r] = sutherland(tt, k.ts_suth);
}

__global__ void __launch_bounds__(TB) k_ns_bc(DevMesh m, DevFields d, Consts k0) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.B) return;
    const Consts k = zoned(k0, m, face_zone(m, f));
    int l = m.lc[f];
    double s[5];
    for (int e = 0; e < 5; ++e) s[e] = d.q[e * m.T + l];
    ns_bc_face(m, d, k, f, s, d.tempr[l]);
}

int ns_boundary(ofx_ctx* c) {
    ++c->qgen;  // ghost states change
    k_ns_cellstate<<<nblk(c->m.C), TB, 0, c->stream>>>(c->m, c->d, c->k);
    OFX_HIP(hipGetLastError());
    if (c->m.B) k_ns_bc<<<nblk(c->m.B), TB, 0, c->stream>>>(c->m, c->d, c->k);
    OFX_HIP(hipGetLastError());
    return 0;
}

// ============================================================ time step



// UTimeStep::CalcInvSpectrumField / CalcVisSpectrumField / CalcLocalTimeStep
// (codes/uns/src/UTimeStep.cpp:190-416, codes/ns/src/TimeStep.cpp:43-133):
// per cell, the face spectral radii are summed in ascending face order.
// QEXT: the same pass also reduces the extrema of the neighbour differences
// that fix the Venkatakrishnan epsilon (Limiter::CalcMinMaxDiff, the global
// part, codes/usolver/src/ULimiter.cpp:597-644): per primitive the largest
// q_i - min(q_i, q_j) and max(q_i, q_j) - q_i over interior faces, into the
// limiter's slots of parity par_lim; the residual's gradient kernel reads them.
// EN: the faces' normals and areas come from the per-entry copies
// (DevMesh::ellg plane 2).
template <int DIM, int MAXF, bool QEXT, bool EN = false>
__global__ void __launch_bounds__(TB) k_ns_timestep(DevMesh m, DevFields d, Consts k0, int par, int par_lim) {
    int i = cell_block(m, TB) * TB + threadIdx.x;
    const int T = m.T;
    const int zone = (i < m.C) ? cell_zone(m, i) : -1;
    const Consts k = zoned(k0, m, zone);
    const bool visc = k.vismodel > 0;
    double dtv = LARGE;
    double ext[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < m.C) {
        // every load of the cell's faces is issued before the sums; each face
        // is then evaluated with its own left / right roles, in face order.
        // In 2-D w and the z components are zero and left out.
        const int n = m.c2fOff[i + 1] - m.c2fOff[i];
        int ent[MAXF], nb[MAXF];
        cell_faces<MAXF>(m, i, ent, nb);
        double qi[5], vli = 0.0, vti = 0.0;
        for (int e = 0; e < 5; ++e) qi[e] = (DIM == 2 && e == 3) ? 0.0 : d.q[e * T + i];
        if (visc) { vli = d.visl[i]; vti = d.vist[i]; }
        double qn[MAXF][5], nx[MAXF], ny[MAXF], nz[MAXF], ar[MAXF], vln[MAXF], vtn[MAXF];
#pragma unroll
        for (int j = 0; j < MAXF; ++j) {
            if (j < n) {
                int e = ent[j], f = e >= 0 ? e : ~e, c = nb[j];
                for (int q = 0; q < 5; ++q) qn[j][q] = (DIM == 2 && q == 3) ? 0.0 : d.q[q * T + c];
                if (EN) {
                    const dvec4 r4 = m.ellg[(size_t)(2 * m.ellw + j) * m.C + i];
                    nx[j] = r4.x; ny[j] = r4.y; nz[j] = (DIM == 3) ? r4.z : 0.0;
                    ar[j] = r4.w;
                } else {
                    nx[j] = m.xfn[f]; ny[j] = m.yfn[f]; nz[j] = (DIM == 3) ? m.zfn[f] : 0.0;
                    ar[j] = m.area[f];
                }
                if (visc) { vln[j] = d.visl[c]; vtn[j] = d.vist[c]; }
            }
        }
        if (QEXT) {
            double mn[5], mx[5];
            for (int e = 0; e < 5; ++e) { mn[e] = qi[e]; mx[e] = qi[e]; }
#pragma unroll
            for (int j = 0; j < MAXF; ++j) {
                int f = ent[j] >= 0 ? ent[j] : ~ent[j];
                if (j < n && f >= m.B)
                    for (int e = 0; e < 5; ++e) { mn[e] = dmin(mn[e], qn[j][e]); mx[e] = dmax(mx[e], qn[j][e]); }
            }
            for (int e = 0; e < 5; ++e) { mn[e] -= qi[e]; mx[e] -= qi[e]; ext[e] = 0.0 - mn[e]; ext[5 + e] = mx[e]; }
        }
        double invsr = 0.0, vissr = 0.0;
#pragma unroll
        for (int j = 0; j < MAXF; ++j) {
            if (j < n) {
                bool own = ent[j] >= 0;  // this cell is the face's left cell
                const double* ql = own ? qi : qn[j];
                const double* qr = own ? qn[j] : qi;
                double rl = ql[0], ul = ql[1], vl = ql[2], wl = ql[3], pl = ql[4];
                double rr = qr[0], ur = qr[1], vr = qr[2], wr = qr[3], pr = qr[4];
                double xfn = nx[j], yfn = ny[j], zfn = nz[j], area = ar[j];
                double vnl = xfn * ul + yfn * vl + zfn * wl - 0.0;
                double vnr = xfn * ur + yfn * vr + zfn * wr - 0.0;
                double gama = 0.5 * (k.gama + k.gama);
                double pm = 0.5 * (pl + pr), rm = 0.5 * (rl + rr);
                double cm = sqrt(gama * pm / rm);
                double vn = 0.5 * (vnl + vnr);
                invsr += 0.5 * area * (dabs(vn) + cm);
                if (visc) {
                    double visl = own ? 0.5 * (vli + vln[j]) : 0.5 * (vln[j] + vli);
                    double vist = own ? 0.5 * (vti + vtn[j]) : 0.5 * (vtn[j] + vti);
                    double sr = 0.0;
                    if (k.visSRModel == 1) {
                        double density = 0.5 * (rl + rr);
                        double c1 = 4.0 / 3.0 * (visl + vist);
                        double c2 = k.gama * (visl * k.oprl + vist * k.oprt);
                        double c3 = 2.0 * dmax(c1, c2) / (k.reynolds * density);
                        sr = (area * area) * c3;
                    } else if (k.visSRModel == 2) {
                        double dist;
                        if (EN) {
                            dist = m.elldv[(size_t)j * m.C + i];  // the same bits (DevMesh::elldv)
                        } else {
                            int l = own ? i : nb[j], r = own ? nb[j] : i;
                            double zl = (DIM == 3) ? m.zcc[l] : 0.0, zr = (DIM == 3) ? m.zcc[r] : 0.0;
                            dist = dabs(xfn * (m.xcc[r] - m.xcc[l]) + yfn * (m.ycc[r] - m.ycc[l]) + zfn * (zr - zl));
                        }
                        double viscosity = visl + vist;
                        double density = 0.5 * (rl + rr);
                        double c1 = 2.0 * viscosity / (density * dist * k.reynolds + SMALL);
                        sr = 0.5 * c1 * area;
                    } else if (k.visSRModel == 3) {
                        double density = 0.5 * (rl + rr);
                        sr = (area * area) / (k.reynolds * density + SMALL);
                    }
                    vissr += sr;
                }
            }
        }
        double vol = m.vol[i];
        double dt = k.cfl * vol / invsr;
        if (visc && k.visTimeStepModel > 0) {
            double vts = k.cfl * vol / vissr;
            dt *= vts / (dt + vts);
        }
        d.invsr[i] = invsr;
        d.vissr[i] = vissr;
        d.timestep[i] = dt;
        dtv = dt;
    }
    if (m.nz > 1) zslot_reduce<1, false>(&dtv, zone, d.slots, par, RQ_N, RQ_DTMIN, m.nz);
    else slot_reduce<1, false>(&dtv, d.slots, par, RQ_N, RQ_DTMIN);
    if (QEXT) {
        if (m.nz > 1) zslot_reduce<10, true>(ext, zone, d.slots, par_lim, RQ_N, RQ_QMIN, m.nz);
        else slot_reduce<10, true>(ext, d.slots, par_lim, RQ_N, RQ_QMIN);
    }
}

// The limiter's extrema alone (the QEXT part of k_ns_timestep), for a residual
// evaluated on a state no time step has seen.
template <int DIM, int MAXF>
__global__ void __launch_bounds__(TB) k_ns_qext(DevMesh m, DevFields d, int par_lim) {
    int i = cell_block(m, TB) * TB + threadIdx.x;
    const int T = m.T;
    const int zone = (i < m.C) ? cell_zone(m, i) : -1;
    double ext[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < m.C) {
        const int n = m.c2fOff[i + 1] - m.c2fOff[i];
        int nb[MAXF];
        cell_nbs<MAXF>(m, i, nb);
        double qi[5], qn[MAXF][5];
        for (int e = 0; e < 5; ++e) qi[e] = (DIM == 2 && e == 3) ? 0.0 : d.q[e * T + i];
#pragma unroll
        for (int j = 0; j < MAXF; ++j)
            for (int e = 0; e < 5; ++e) qn[j][e] = (DIM == 2 && e == 3) ? 0.0 : d.q[e * T + nb[j]];
        double mn[5], mx[5];
        for (int e = 0; e < 5; ++e) { mn[e] = qi[e]; mx[e] = qi[e]; }
#pragma unroll
        for (int j = 0; j < MAXF; ++j) {
            if (j < n && nb[j] < m.C)  // interior faces (a boundary face's neighbour is its ghost, C + f)
                for (int e = 0; e < 5; ++e) { mn[e] = dmin(mn[e], qn[j][e]); mx[e] = dmax(mx[e], qn[j][e]); }
        }
        for (int e = 0; e < 5; ++e) { mn[e] -= qi[e]; mx[e] -= qi[e]; ext[e] = 0.0 - mn[e]; ext[5 + e] = mx[e]; }
    }
    if (m.nz > 1) zslot_reduce<10, true>(ext, zone, d.slots, par_lim, RQ_N, RQ_QMIN, m.nz);
    else slot_reduce<10, true>(ext, d.slots, par_lim, RQ_N, RQ_QMIN);
}


// Time step after UTimeStep::ModifyTimeStep (codes/uns/src/UTimeStep.cpp:368-381)
// and the LU-SGS diagonal of UNsSpectrum::CalcImplicitSpectrum
// (codes/uns/src/UNsSpectrum.cpp:52-94, sp2 = 1), evaluated where consumed.
__device__ __forceinline__ double clamped_dt(const DevFields& d, const Consts& k, int i, double mindt) {
    double dt = d.timestep[i];
    if (k.max_time_ratio > 0) {
        double mp = k.max_time_ratio * mindt;
        dt = (dt < mp) ? dt : mp;
    }
    return dt;
}
__device__ __forceinline__ double ns_impsr(const DevMesh& m, const DevFields& d, const Consts& k, int i, double mindt) {
    double imp = (1.0 / clamped_dt(d, k, i, mindt)) * m.vol[i];
    imp += d.invsr[i];
    imp += d.vissr[i];
    return imp;
}

// Materialise the clamped time step and the LU-SGS diagonal as fields
// (for download; the solver evaluates them where they are consumed).
__global__ void __launch_bounds__(TB) k_ns_dtclamp(DevMesh m, DevFields d, Consts k, int par) {
    double mindt = slot_result<false>(d.slots, par, m.nz * RQ_N, RQ_DTMIN);
    int i = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    mindt = zone_mindt(m, d, i, mindt);
    double dt = clamped_dt(d, k, i, mindt);
    double imp = ns_impsr(m, d, k, i, mindt);
    d.timestep[i] = dt;
    d.impsr[i] = imp;
}

int materialize_timestep(ofx_ctx* c) {
    k_ns_dtclamp<<<nblk(c->m.C), TB, 0, c->stream>>>(c->m, c->d, c->k, c->par_dt);
    OFX_HIP(hipGetLastError());
    return 0;
}

int ns_timestep(ofx_ctx* c, int iter) {
    const ofx_config& g = c->cfg;
    double cfl = g.cfled;
    if (iter < g.ncfl) {
        double delta = (g.cfled - g.cflst) / g.ncfl;
        cfl = g.cflst + delta * iter;
    }
    c->k.cfl = cfl;
    c->par_dt ^= 1;
    const int mf = c->m.maxFacesPerCell;
    const DevMesh& m = c->m;
    // the Venkatakrishnan extrema of this state come with the time step
    const bool qext = c->k.ilim == 2;
    if (qext) c->par_lim ^= 1;
    const int pd = c->par_dt, pl = c->par_lim;
    const bool en = c->normal_variant == 1;
#define OFX_TS(D, MF)                                                                                             \
    (qext ? (en ? (k_ns_timestep<D, MF, true, true><<<nblk(m.C), TB, 0, c->stream>>>(m, c->d, c->k, pd, pl), 0)   \
                : (k_ns_timestep<D, MF, true><<<nblk(m.C), TB, 0, c->stream>>>(m, c->d, c->k, pd, pl), 0))        \
          : (en ? (k_ns_timestep<D, MF, false, true><<<nblk(m.C), TB, 0, c->stream>>>(m, c->d, c->k, pd, pl), 0)  \
                : (k_ns_timestep<D, MF, false><<<nblk(m.C), TB, 0, c->stream>>>(m, c->d, c->k, pd, pl), 0)))
    if (m.dim == 2) {
        if (mf <= 4) OFX_TS(2, 4);
        else if (mf <= 8) OFX_TS(2, 8);
        else { set_error("cells with more than 8 faces are not supported"); return OFX_EINVAL; }
    } else {
        if (mf <= 6) OFX_TS(3, 6);
        else if (mf <= 16) OFX_TS(3, 16);
        else { set_error("cells with more than 16 faces are not supported"); return OFX_EINVAL; }
    }
#undef OFX_TS
    if (c->m.nz > 1) {
        k_zone_extract<false><<<c->m.nz, 64, 0, c->stream>>>(c->d.slots, c->par_dt, RQ_N, RQ_DTMIN, 1, c->m.nz, c->d.zred);
        if (qext)
            k_zone_extract<true><<<c->m.nz, 256, 0, c->stream>>>(c->d.slots, c->par_lim, RQ_N, RQ_QMIN, 10, c->m.nz,
                                                                 c->d.zred);
    }
    OFX_HIP(hipGetLastError());
    if (qext) c->qext_gen = c->qgen;
    return 0;
}

// ============================================================ residual

// Green-Gauss gradient with inverse-distance face weights of rho,u,v,w,p and T
// (CalcGradGGCellWeight, codes/usolver/src/UGrad.cpp:88-149) as a per-cell
// gather, and with LIM = 1 (Barth) or 2 (Venkatakrishnan) the cell's limiter
// in the same pass (Limiter::CalcLimiter with CalcMinMaxDiff and
// CalcBarthLimiter / CalcVencatLimiter, codes/usolver/src/ULimiter.cpp:53-97,
// 377-644): the neighbour min/max differences are the ones the gradient's face
// loop already holds, the limiter of a face needs only the cell's own
// gradient, and the Venkatakrishnan epsilon comes from the grid extrema that
// the time-step kernel (or k_ns_qext) reduced on this state into the slots of
// parity `par`.  The face loop is unrolled to MAXF: all loads of the cell's
// faces are issued before the contributions are added in ascending face
// order.  DIM == 2 skips the z components and w, identically zero in 2-D.
// EG: the face terms come from the per-entry copies (DevMesh::ellg), loaded
// with the neighbour ids rather than after the face ids; a boundary entry is
// one whose neighbour is a ghost (id >= C).
template <int DIM, int MAXF, int LIM, bool EG = false>
__global__ void __launch_bounds__(TB) k_ns_grad(DevMesh m, DevFields d, Consts k0, int par) {
    // Every lane stays active to the end (the extrema are read per wave):
    // lanes past the last cell evaluate it and store nothing.
    const int i0 = cell_block(m, TB) * TB + threadIdx.x;
    const bool live = i0 < m.C;
    const int i = live ? i0 : m.C - 1;
    const int T = m.T;
    const Consts k = zoned(k0, m, cell_zone(m, i));
    double ext[10];
    if (LIM == 2 && m.nz == 1) wave_slot_result<10, true>(d.slots, par, m.nz * RQ_N, RQ_QMIN, ext);
    const int n = m.c2fOff[i + 1] - m.c2fOff[i];
    int ent[MAXF], nb[MAXF];
    if (EG && LIM != 1) cell_nbs<MAXF>(m, i, nb);
    else cell_faces<MAXF>(m, i, ent, nb);
    double qi[5];
    for (int e = 0; e < 5; ++e) qi[e] = (DIM == 2 && e == 3) ? 0.0 : d.q[e * T + i];
    // EG: gw = {weight of this cell, weight of the neighbour, outward normal
    // x A, y A}, gz = outward z A; otherwise the face's record as stored.
    // dx, dy, dz: face centre - this cell's centre.
    dvec4 gw[MAXF];
    double gz[MAXF], qn[MAXF][5], dx[MAXF], dy[MAXF], dz[MAXF];
    const size_t S = (size_t)m.ellw * m.C;
    int imask = 0;  // EG: bit j set for an interior entry
#pragma unroll
    for (int j = 0; j < MAXF; ++j) {
        if (j < n) {
            const int c = nb[j];
            if (EG) {
                imask |= (c < m.C) << j;
                const size_t o = (size_t)j * m.C + i;
                gw[j] = m.ellg[o];
                if (DIM == 3 || LIM != 0) {
                    const dvec4 b4 = m.ellg[S + o];
                    gz[j] = (DIM == 3) ? b4.x : 0.0;
                    dx[j] = b4.y;
                    dy[j] = b4.z;
                    dz[j] = (DIM == 3) ? b4.w : 0.0;
                } else {
                    gz[j] = 0.0;
                }
            } else {
                const int e = ent[j], f = e >= 0 ? e : ~e;
                gw[j] = m.fgg[f];  // as stored: {cl, cr, nx A, ny A} of the face
                gz[j] = (DIM == 3) ? m.fz[f].x : 0.0;
                if (LIM != 0) {
                    dvec4 rc4 = m.frec[f];
                    dx[j] = e >= 0 ? rc4.x : rc4.z;
                    dy[j] = e >= 0 ? rc4.y : rc4.w;
                    dz[j] = (DIM == 3) ? (e >= 0 ? m.fz[f].y : m.fz[f].z) : 0.0;
                }
            }
            for (int e2 = 0; e2 < 5; ++e2) qn[j][e2] = (DIM == 2 && e2 == 3) ? 0.0 : d.q[e2 * T + c];
        }
    }
    double mn[5], mx[5];
    for (int e = 0; e < 5; ++e) { mn[e] = qi[e]; mx[e] = qi[e]; }
    double ti = temperature(qi[0], qi[4], k.statecoef);
    double acc[6][3];
    for (int s = 0; s < 6; ++s) acc[s][0] = acc[s][1] = acc[s][2] = 0.0;
    // EG: with the weights and the normal taken as seen from this cell, the
    // sums are those of the face records: for the right cell, cl*vn + cr*vi
    // is w_self*vi + w_nb*vn (a commuted sum) and acc - n*v is acc + (-n)*v,
    // both the same bits.
#pragma unroll
    for (int j = 0; j < MAXF; ++j) {
        if (j < n) {
            double tn = temperature(qn[j][0], qn[j][4], k.statecoef);
            double fnxa = gw[j].z, fnya = gw[j].w, fnza = gz[j];
            if (EG) {
                double ws = gw[j].x, wn = gw[j].y;
                for (int s = 0; s < 6; ++s) {
                    double vi = s < 5 ? qi[s] : ti, vn = s < 5 ? qn[j][s] : tn;
                    double value = ws * vi + wn * vn;
                    acc[s][0] += fnxa * value; acc[s][1] += fnya * value; acc[s][2] += fnza * value;
                }
            } else {
                bool own = ent[j] >= 0;  // this cell is the face's left cell
                double cl = gw[j].x, cr = gw[j].y;
                for (int s = 0; s < 6; ++s) {
                    double vi = s < 5 ? qi[s] : ti, vn = s < 5 ? qn[j][s] : tn;
                    double value = own ? (cl * vi + cr * vn) : (cl * vn + cr * vi);
                    if (own) { acc[s][0] += fnxa * value; acc[s][1] += fnya * value; acc[s][2] += fnza * value; }
                    else { acc[s][0] -= fnxa * value; acc[s][1] -= fnya * value; acc[s][2] -= fnza * value; }
                }
            }
            // interior faces only (a boundary face's neighbour is its ghost, id C + f)
            const bool interior = EG ? ((imask >> j) & 1) != 0 : (ent[j] >= 0 ? ent[j] : ~ent[j]) >= m.B;
            if (LIM != 0 && interior)
                for (int s = 0; s < 5; ++s) { mn[s] = dmin(mn[s], qn[j][s]); mx[s] = dmax(mx[s], qn[j][s]); }
        }
    }
    // In 2-D the z components and w's gradient stay at the zeros the arrays
    // were allocated with (w = 0 and z-extent 0 make them exactly 0).
    double ovol = 1.0 / m.vol[i];
    double g[6][3];
    for (int s = 0; s < 6; ++s)
        for (int a = 0; a < 3; ++a) g[s][a] = acc[s][a] * ovol;
    if (live)
        for (int s = 0; s < 6; ++s) {
            if (DIM == 2 && s == 3) continue;
            for (int a = 0; a < DIM; ++a) d.grad[(a * 6 + s) * T + i] = g[s][a];
        }
    if (LIM == 0) return;
    double dot[MAXF];
    if (LIM == 1) {
#pragma unroll
        for (int j = 0; j < MAXF; ++j) {
            if (j < n) {
                int e = ent[j], f = e >= 0 ? e : ~e;
                double ds = sqrt(sqr3(dx[j], dy[j], dz[j]));
                double nzf = (DIM == 3) ? m.zfn[f] : 0.0;
                dot[j] = (m.xfn[f] * dx[j] + m.yfn[f] * dy[j] + nzf * dz[j]) / (ds + SMALL);
                if (e < 0) dot[j] = -dot[j];
            }
        }
    }
    double lims[5];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        double lim = 1.0;
        const bool skip = (LIM == 2 && DIM == 2 && s == 3);  // Venkat on w in 2-D: every face gives 1
        if (!skip) {
            // w in 2-D (Barth): gradient and extents are exactly zero, as stored
            const bool w2 = DIM == 2 && s == 3;
            double gx = w2 ? 0.0 : g[s][0], gy = w2 ? 0.0 : g[s][1], gzs = (DIM == 3) ? g[s][2] : 0.0;
            double lo_d = w2 ? 0.0 : mn[s] - qi[s], hi_d = w2 ? 0.0 : mx[s] - qi[s];
            double eps1 = 0.0;
            if (LIM == 2) {
                double lo, hi;
                if (m.nz > 1) {
                    const double* zr = d.zred + (size_t)m.zcell[i] * RQ_N;
                    lo = zr[RQ_QMIN + s];
                    hi = zr[RQ_QMAX + s];
                } else {
                    lo = ext[s];
                    hi = ext[5 + s];
                }
                double eps = k.vencat_coef * dmax(hi - (-lo), 1.0);
                eps1 = eps * eps + SMALL;
            }
#pragma unroll
            for (int j = 0; j < MAXF; ++j) {
                if (j < n) {
                    double dq = gx * dx[j] + gy * dy[j] + gzs * dz[j];
                    double phi = (LIM == 2) ? vencat_fn(dq, lo_d, hi_d, eps1) : barth_fn(dq, lo_d, hi_d, dot[j]);
                    lim = dmin(lim, phi);
                }
            }
        }
        lims[s] = dmin(1.0, lim);
    }
    if (live)
        for (int s = 0; s < 5; ++s) d.limiter[s * T + i] = lims[s];
}

// The residual gather with one thread per cell (face loads of the cell all
// issued first, then the ordered sums: inviscid chain, then viscous).
// L0 (> 0 inside a full pseudo-time step): cells [0, L0) are the first level
// of the lower LU-SGS sweep (no lower-numbered neighbour), whose increment is
// dq = (res - 0) / D; the gather writes it with the residual and the sweep
// starts at its second level.
template <int MAXF, bool VISC, bool F32, bool FUSE>
__device__ __forceinline__ void ns_gather_cell(const DevMesh& m, const DevFields& d, const Consts& k0, int par,
                                               int L0) {
    const int T = m.T, F = m.F;
    // Every lane stays active to the end (the fused path reduces the time-step
    // slots per wave): lanes past the last cell evaluate it and store nothing.
    const int i0 = cell_block(m, TB) * TB + threadIdx.x;
    const bool live = i0 < m.C;
    const int i = live ? i0 : m.C - 1;
    double mindt1 = 0.0;
    if (FUSE) wave_slot_result<1, false>(d.slots, par, m.nz * RQ_N, RQ_DTMIN, &mindt1);
    double rad = 1.0;
    if (FUSE && i < L0) {
        const Consts k = zoned(k0, m, cell_zone(m, i));
        rad = ns_impsr(m, d, k, i, zone_mindt(m, d, i, mindt1));
    }
    const int n = m.c2fOff[i + 1] - m.c2fOff[i];
    int ent[MAXF], nb[MAXF];
    cell_faces<MAXF>(m, i, ent, nb);
    const float* inv32 = reinterpret_cast<const float*>(d.invflux);
    const float* vis32 = reinterpret_cast<const float*>(d.visflux);
    // Every entry is loaded, padding included (it names face 0), so that no
    // load sits in a divergent branch, and all loads come before the first
    // store (which the compiler cannot move them across): the first n
    // entries are summed.  The viscous flux of the continuity equation is
    // coeff * 0.0, a signed zero, on every face, and a sum that starts at +0
    // and only adds is never -0, so adding it changes no bit: it is not read.
    double fi[5][MAXF], fv[5][MAXF];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
#pragma unroll
        for (int j = 0; j < MAXF; ++j) {
            int e = ent[j], f = e >= 0 ? e : ~e;
            double x = F32 ? (double)inv32[s * F + f] : d.invflux[s * F + f];
            fi[s][j] = (e >= 0) ? -x : x;
            if (VISC && s > 0) {
                double y = F32 ? (double)vis32[s * F + f] : d.visflux[s * F + f];
                fv[s][j] = (e >= 0) ? -y : y;
            }
        }
    }
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        double sum = 0.0;
#pragma unroll
        for (int j = 0; j < MAXF; ++j)
            sum = sum + live_or_nz(j < n, fi[s][j]);
        if (VISC && s > 0) {
#pragma unroll
            for (int j = 0; j < MAXF; ++j)
                sum = sum + live_or_nz(j < n, fv[s][j]);
        }
        if (live) d.res[s * T + i] = sum;
        if (FUSE && live && i < L0) { double t = sum - 0.0; t /= rad; d.dq[s * T + i] = t; }
    }
}

template <int MAXF, bool VISC, bool F32>
__global__ void __launch_bounds__(TB) k_ns_gather_c(DevMesh m, DevFields d) {
    ns_gather_cell<MAXF, VISC, F32, false>(m, d, Consts{}, 0, 0);
}

template <int MAXF, bool VISC, bool F32>
__global__ void __launch_bounds__(TB) k_ns_gather_cf(DevMesh m, DevFields d, Consts k0, int par, int L0) {
    ns_gather_cell<MAXF, VISC, F32, true>(m, d, k0, par, L0);
}

// Viscous face flux (UNsVisFlux::CalcFlux, codes/uns/src/UNsVisFlux.cpp:127-440)
// for face-gradient method VM: face gradients of u, v, (w,) T, Newtonian
// stress and heat flux, stored as visflux[5][F] (rows 1..4; row 0 is zero and not stored).
template <int DIM, bool F32, int VM>
__device__ __forceinline__ void ns_face_visc(const DevMesh& m, const DevFields& d, const Consts& k, int f, int l,
                                             int r, int rg, const double* ql, const double* qr, double xfn,
                                             double yfn, double zfn, double area) {
    const int T = m.T;
    VisGeom G;
    vis_geom_pre<VM>(m, f, G);
    double tl = temperature(ql[0], ql[4], k.statecoef);
    double tr = temperature(qr[0], qr[4], k.statecoef);
    double gq[6][3], qfv[6], dqdn[6];
#pragma unroll
    for (int s = 1; s < 6; ++s) {
        if (s == 4) continue;  // pressure does not enter the viscous flux
        if (DIM == 2 && s == 3) {
            gq[3][0] = gq[3][1] = gq[3][2] = 0.0; qfv[3] = 0.0; dqdn[3] = 0.0;
            continue;
        }
        double g1[3] = {d.grad[(0 * 6 + s) * T + l], d.grad[(1 * 6 + s) * T + l], (DIM == 3) ? d.grad[(2 * 6 + s) * T + l] : 0.0};
        double g2[3] = {d.grad[(0 * 6 + s) * T + rg], d.grad[(1 * 6 + s) * T + rg], (DIM == 3) ? d.grad[(2 * 6 + s) * T + rg] : 0.0};
        double v1 = s < 5 ? ql[s] : tl, v2 = s < 5 ? qr[s] : tr;
        vis_grad<VM>(G, v1, v2, g1, g2, qfv[s], gq[s], dqdn[s]);
    }
    double visl = 0.5 * (sutherland(tl, k.ts_suth) + sutherland(tr, k.ts_suth));
    double vist = 0.5 * (d.vist[l] + d.vist[r]);
    double vis = visl + vist;
    double dudx = gq[1][0], dudy = gq[1][1], dudz = gq[1][2];
    double dvdx = gq[2][0], dvdy = gq[2][1], dvdz = gq[2][2];
    double dwdx = gq[3][0], dwdy = gq[3][1], dwdz = gq[3][2];
    double qNormal = 0.0;
    double kcp = (visl * k.oprl + vist * k.oprt) * k.const_cp;
    qNormal += xfn * 0.0 + yfn * 0.0 + zfn * 0.0;
    qNormal += kcp * dqdn[5];
    double divv2p3 = (2.0 / 3.0) * (dudx + dvdy + dwdz);
    double txx = vis * (2.0 * dudx - divv2p3);
    double tyy = vis * (2.0 * dvdy - divv2p3);
    double tzz = vis * (2.0 * dwdz - divv2p3);
    double txy = vis * (dudy + dvdx);
    double txz = vis * (dudz + dwdx);
    double tyz = vis * (dvdz + dwdy);
    double fv1 = xfn * txx + yfn * txy + zfn * txz;
    double fv2 = xfn * txy + yfn * tyy + zfn * tyz;
    double fv3 = xfn * txz + yfn * tyz + zfn * tzz;
    double fv4 = qfv[1] * fv1 + qfv[2] * fv2 + qfv[3] * fv3 + qNormal;
    double coeff = -k.oreynolds * area;
    // Row 0, the continuity equation's coeff * 0.0, is a signed zero that the
    // gathers do not read (adding it changes no bit), so it is not stored.
    if (F32) {
        float* out = reinterpret_cast<float*>(d.visflux);
        out[1 * m.F + f] = (float)(coeff * fv1);
        out[2 * m.F + f] = (float)(coeff * fv2);
        out[3 * m.F + f] = (float)(coeff * fv3);
        out[4 * m.F + f] = (float)(coeff * fv4);
    } else {
        d.visflux[1 * m.F + f] = coeff * fv1;
        d.visflux[2 * m.F + f] = coeff * fv2;
        d.visflux[3 * m.F + f] = coeff * fv3;
        d.visflux[4 * m.F + f] = coeff * fv4;
    }
}

// Per-face inviscid (MUSCL + Roe/SLAU2) and viscous fluxes
// (UNsInvFlux::CalcInvFlux, codes/uns/src/UNsInvFlux.cpp:141-192;
//  LimField::CalcFaceValue + NsLimField::BcQlQrFix, codes/usolver/src/ULimiter.cpp:181-256,
//  codes/uns/src/UNsLimiter.cpp:67-101; UNsVisFlux::CalcVisFlux,
//  codes/uns/src/UNsVisFlux.cpp:127-440).  LIM: -1 first order, 0 unlimited,
// 1 limiter array.  Ghost gradients equal their owner's (Grad::CalcGrad tail).
// F32: the FP32-flux variant -- the inviscid flux is evaluated in FP32 from
// the FP64 reconstruction and both face fluxes are stored as FP32 (the
// gather accumulates them in FP64).
// SPLIT (viscous runs): a block of two waves covers 64 faces, the first wave
// computing their inviscid fluxes and the second their viscous fluxes (each
// wave loads what its part needs): half the work and registers per wave,
// twice the waves in flight.
template <int DIM, int LIM, bool SLAU2, bool VISC, bool F32, bool SPLIT = false>
__global__ void __launch_bounds__(512) k_ns_face(DevMesh m, DevFields d, Consts k0, int tb) {
    const int role = SPLIT ? (int)(threadIdx.x >> 6) : 0;  // wave-uniform: 0 inviscid, 1 viscous
    int f = SPLIT ? face_block(m, 64) * 64 + (int)(threadIdx.x & 63) : face_block(m, tb) * tb + (int)threadIdx.x;
    if (f >= m.F) return;
    const Consts k = zoned(k0, m, face_zone(m, f));
    const int T = m.T;
    int l = m.lc[f], r = m.rc[f];
    bool bnd = f < m.B;
    int rg = bnd ? l : r;  // cell holding r's gradient
    double xfn = m.xfn[f], yfn = m.yfn[f], zfn = (DIM == 3) ? m.zfn[f] : 0.0, area = m.area[f];
    double ql[5], qr[5], p1[5], p2[5];
    for (int e = 0; e < 5; ++e) {
        ql[e] = (DIM == 2 && e == 3) ? 0.0 : d.q[e * T + l];
        qr[e] = (DIM == 2 && e == 3) ? 0.0 : d.q[e * T + r];
    }
    if (!SPLIT || role == 0) {
    if (bnd) {
        for (int e = 0; e < 5; ++e) { double tmp = 0.5 * (ql[e] + qr[e]); p1[e] = tmp; p2[e] = tmp; }
        if (m.bct[f] == BC_SOLID_SURFACE && k.vismodel != 0) {
            // bc_q of AdiabaticVisWallBc / IsothermalVisWallBc: owner rho (or wall density), zero velocity, owner p
            double rw = (k.isowallbc == 0) ? ql[0] : ql[4] / (k.statecoef * k.twall * 1.0);
            p1[0] = rw; p1[1] = 0.0; p1[2] = 0.0; p1[3] = 0.0; p1[4] = ql[4];
            for (int e = 0; e < 5; ++e) p2[e] = p1[e];
        } else if (m.bct[f] == BC_SOLID_SURFACE || m.bct[f] == BC_SYMMETRY) {
            // bc_q of SymmetryBc is only read on walls; symmetry faces use the average above
        }
    } else if (LIM == -1) {
        for (int e = 0; e < 5; ++e) { p1[e] = ql[e]; p2[e] = qr[e]; }
    } else {
        dvec4 rc4 = m.frec[f];
        double dxl = rc4.x, dyl = rc4.y, dxr = rc4.z, dyr = rc4.w, dzl = 0.0, dzr = 0.0;
        if (DIM == 3) { dvec2 z2 = m.fzr[f]; dzl = z2.x; dzr = z2.y; }
        for (int e = 0; e < 5; ++e) {
            if (DIM == 2 && e == 3) {
                // w and its gradient are zero in 2-D: the reconstruction leaves w = 0
                p1[e] = ql[e] + 0.0; p2[e] = qr[e] + 0.0;
                continue;
            }
            double gzl = (DIM == 3) ? d.grad[(2 * 6 + e) * T + l] : 0.0;
            double gzr = (DIM == 3) ? d.grad[(2 * 6 + e) * T + r] : 0.0;
            double gl = d.grad[(0 * 6 + e) * T + l] * dxl + d.grad[(1 * 6 + e) * T + l] * dyl + gzl * dzl;
            double gr = d.grad[(0 * 6 + e) * T + r] * dxr + d.grad[(1 * 6 + e) * T + r] * dyr + gzr * dzr;
            if (LIM == 0) {
                p1[e] = ql[e] + gl;
                p2[e] = qr[e] + gr;
            } else {
                double phil = d.limiter[e * T + l], phir = d.limiter[e * T + r];
                p1[e] = ql[e] + dmin(phil, phir) * gl;
                p2[e] = qr[e] + dmin(phir, phil) * gr;
            }
        }
        if (p1[0] <= 0.0 || p1[4] <= 0.0) for (int e = 0; e < 5; ++e) p1[e] = ql[e];
        if (p2[0] <= 0.0 || p2[4] <= 0.0) for (int e = 0; e < 5; ++e) p2[e] = qr[e];
    }
    if (F32) {
        float a1[5], a2[5], fl[5];
        for (int e = 0; e < 5; ++e) { a1[e] = (float)p1[e]; a2[e] = (float)p2[e]; }
        float gf = (float)k.gama, nx = (float)xfn, ny = (float)yfn, nz = (float)zfn;
        if (SLAU2) slau2_flux<float>(a1, a2, gf, gf, nx, ny, nz, fl);
        else roe_flux<float>(k.ieigenfix, (float)k.centropy1, (float)k.centropy2, a1, a2, gf, gf, nx, ny, nz, fl);
        float* out = reinterpret_cast<float*>(d.invflux);
        float af = (float)area;
        for (int e = 0; e < 5; ++e) out[e * m.F + f] = af * fl[e];
    } else {
        double fl[5];
        if (SLAU2) slau2_flux<double>(p1, p2, k.gama, k.gama, xfn, yfn, zfn, fl);
        else roe_flux<double>(k.ieigenfix, k.centropy1, k.centropy2, p1, p2, k.gama, k.gama, xfn, yfn, zfn, fl);
        for (int e = 0; e < 5; ++e) d.invflux[e * m.F + f] = area * fl[e];
    }
    }  // inviscid part

    if (!VISC || (SPLIT && role == 0)) return;
    // the face-gradient method is uniform: one dispatch, each branch a
    // separate instantiation of the viscous flux
    if (k.ivischeme == 0) ns_face_visc<DIM, F32, 0>(m, d, k, f, l, r, rg, ql, qr, xfn, yfn, zfn, area);
    else if (k.ivischeme == 1) ns_face_visc<DIM, F32, 1>(m, d, k, f, l, r, rg, ql, qr, xfn, yfn, zfn, area);
    else ns_face_visc<DIM, F32, 2>(m, d, k, f, l, r, rg, ql, qr, xfn, yfn, zfn, area);
}

// Residual gather: res = -sum(inviscid) then -sum(viscous), each over the
// cell's faces in ascending id, sign by side (AddF2CField called twice);
// LPC lanes per cell, ordered shuffle sums.
template <int LPC, bool VISC, bool F32>
__global__ void __launch_bounds__(TB) k_ns_gather(DevMesh m, DevFields d) {
    const int T = m.T, F = m.F;
    int gid = cell_block(m, blockDim.x / LPC) * blockDim.x + threadIdx.x;
    int i = gid / LPC, j = gid % LPC;
    bool live = i < m.C;
    int n = 0;
    if (live) n = m.c2fOff[i + 1] - m.c2fOff[i];
    // entry j of the cell's face list, loaded without waiting for the count
    const int ej = (live && j < m.ellw) ? m.ellf[(size_t)j * m.C + i] : 0;
    bool has = live && j < n;
    double fi[5] = {0, 0, 0, 0, 0}, fv[5] = {0, 0, 0, 0, 0};
    if (has) {
        int e = ej;
        int f = e >= 0 ? e : ~e;
        const float* inv32 = reinterpret_cast<const float*>(d.invflux);
        const float* vis32 = reinterpret_cast<const float*>(d.visflux);
        for (int s = 0; s < 5; ++s) {
            double x = F32 ? (double)inv32[s * F + f] : d.invflux[s * F + f];
            fi[s] = (e >= 0) ? -x : x;
        }
        if (VISC)  // the continuity equation's viscous flux is a signed zero (see ns_gather_cell)
            for (int s = 1; s < 5; ++s) {
                double x = F32 ? (double)vis32[s * F + f] : d.visflux[s * F + f];
                fv[s] = (e >= 0) ? -x : x;
            }
    }
    for (int s = 0; s < 5; ++s) {
        // continue the inviscid chain with the viscous terms: ((0 + i0) + i1 ...) + v0 + v1 ...
        int base = (int)(threadIdx.x & 63) / LPC * LPC;
        double sum = 0.0;
#pragma unroll
        for (int jj = 0; jj < LPC; ++jj) { double y = __shfl(fi[s], base + jj, 64); if (jj < n) sum = sum + y; }
        if (VISC && s > 0) {
#pragma unroll
            for (int jj = 0; jj < LPC; ++jj) { double y = __shfl(fv[s], base + jj, 64); if (jj < n) sum = sum + y; }
        }
        if (live && j == 0) d.res[s * T + i] = sum;
    }
}

template <int DIM, int LIM, bool SLAU2>
static void launch_face(ofx_ctx* c, bool visc) {
    const int tb = c->face_tb;  // face tile, chosen by ofx_autotune
    int nb = (c->m.F + tb - 1) / tb;
    if (visc && c->face_split && !c->k.flux_fp32) {
        k_ns_face<DIM, LIM, SLAU2, true, false, true><<<(c->m.F + 63) / 64, 128, 0, c->stream>>>(c->m, c->d, c->k, 64);
        return;
    }
    if (c->k.flux_fp32) {
        if (visc) k_ns_face<DIM, LIM, SLAU2, true, true><<<nb, tb, 0, c->stream>>>(c->m, c->d, c->k, tb);
        else k_ns_face<DIM, LIM, SLAU2, false, true><<<nb, tb, 0, c->stream>>>(c->m, c->d, c->k, tb);
    } else {
        if (visc) k_ns_face<DIM, LIM, SLAU2, true, false><<<nb, tb, 0, c->stream>>>(c->m, c->d, c->k, tb);
        else k_ns_face<DIM, LIM, SLAU2, false, false><<<nb, tb, 0, c->stream>>>(c->m, c->d, c->k, tb);
    }
}

template <int DIM>
static void launch_face_dim(ofx_ctx* c, bool visc) {
    const Consts& k = c->k;
    bool slau2 = k.ischeme == 12;
    if (k.ilim == -1) { if (slau2) launch_face<DIM, -1, true>(c, visc); else launch_face<DIM, -1, false>(c, visc); }
    else if (k.ilim == 0) { if (slau2) launch_face<DIM, 0, true>(c, visc); else launch_face<DIM, 0, false>(c, visc); }
    else { if (slau2) launch_face<DIM, 1, true>(c, visc); else launch_face<DIM, 1, false>(c, visc); }
}

// The limiter's grid extrema for the current state, unless the time step of
// this state produced them (k_ns_timestep<.., QEXT>); `force` launches the
// pass regardless (timing).
template <int DIM, int MAXF>
static void launch_qext(ofx_ctx* c, bool force) {
    if (c->k.ilim != 2 || (!force && c->qext_gen == c->qgen)) return;
    c->par_lim ^= 1;
    k_ns_qext<DIM, MAXF><<<nblk(c->m.C), TB, 0, c->stream>>>(c->m, c->d, c->par_lim);
    if (c->m.nz > 1)
        k_zone_extract<true><<<c->m.nz, 256, 0, c->stream>>>(c->d.slots, c->par_lim, RQ_N, RQ_QMIN, 10, c->m.nz,
                                                             c->d.zred);
    c->qext_gen = c->qgen;
}

// NS residual kernels; `parts` selects them (bit 0 gradient + limiter, bit 1
// the limiter's grid extrema, bit 2 face fluxes, bit 3 gather) so each can be
// timed alone.  MAXF bounds the faces per cell (unrolled loops); the
// lanes-per-cell gather uses LPC, the next power of two (its group
// reductions are butterflies).
template <int DIM, int MAXF>
static void launch_cells(ofx_ctx* c, unsigned parts, bool reuse_ext) {
    constexpr int LPC = MAXF <= 4 ? 4 : (MAXF <= 8 ? 8 : 16);
    const Consts& k = c->k;
    const int C1 = c->m.C;         // threads: one per cell
    const int C = c->m.C * LPC;    // threads: LPC lanes per cell
    bool visc = k.vismodel != 0;
    if (parts & 2u) launch_qext<DIM, MAXF>(c, true);
    if (parts & 1u) {
        const bool eg = c->grad_variant == 1;
        if (k.ilim == 2) {
            if (!(parts & 2u)) launch_qext<DIM, MAXF>(c, !reuse_ext);
            if (eg) k_ns_grad<DIM, MAXF, 2, true><<<nblk(C1), TB, 0, c->stream>>>(c->m, c->d, k, c->par_lim);
            else k_ns_grad<DIM, MAXF, 2><<<nblk(C1), TB, 0, c->stream>>>(c->m, c->d, k, c->par_lim);
        } else if (k.ilim == 1) {
            if (eg) k_ns_grad<DIM, MAXF, 1, true><<<nblk(C1), TB, 0, c->stream>>>(c->m, c->d, k, 0);
            else k_ns_grad<DIM, MAXF, 1><<<nblk(C1), TB, 0, c->stream>>>(c->m, c->d, k, 0);
        } else {
            if (eg) k_ns_grad<DIM, MAXF, 0, true><<<nblk(C1), TB, 0, c->stream>>>(c->m, c->d, k, 0);
            else k_ns_grad<DIM, MAXF, 0><<<nblk(C1), TB, 0, c->stream>>>(c->m, c->d, k, 0);
        }
    }
    if (parts & 4u) launch_face_dim<DIM>(c, visc);
    if ((parts & 8u) && c->gather_variant == 1) {
        const int L0 = c->fuse_l0;
        if (L0 > 0) {
            if (k.flux_fp32) {
                if (visc) k_ns_gather_cf<MAXF, true, true><<<nblk(C1), TB, 0, c->stream>>>(c->m, c->d, k, c->par_dt, L0);
                else k_ns_gather_cf<MAXF, false, true><<<nblk(C1), TB, 0, c->stream>>>(c->m, c->d, k, c->par_dt, L0);
            } else {
                if (visc) k_ns_gather_cf<MAXF, true, false><<<nblk(C1), TB, 0, c->stream>>>(c->m, c->d, k, c->par_dt, L0);
                else k_ns_gather_cf<MAXF, false, false><<<nblk(C1), TB, 0, c->stream>>>(c->m, c->d, k, c->par_dt, L0);
            }
        } else if (k.flux_fp32) {
            if (visc) k_ns_gather_c<MAXF, true, true><<<nblk(C1), TB, 0, c->stream>>>(c->m, c->d);
            else k_ns_gather_c<MAXF, false, true><<<nblk(C1), TB, 0, c->stream>>>(c->m, c->d);
        } else {
            if (visc) k_ns_gather_c<MAXF, true, false><<<nblk(C1), TB, 0, c->stream>>>(c->m, c->d);
            else k_ns_gather_c<MAXF, false, false><<<nblk(C1), TB, 0, c->stream>>>(c->m, c->d);
        }
    } else if (parts & 8u) {
        if (k.flux_fp32) {
            if (visc) k_ns_gather<LPC, true, true><<<nblk(C), TB, 0, c->stream>>>(c->m, c->d);
            else k_ns_gather<LPC, false, true><<<nblk(C), TB, 0, c->stream>>>(c->m, c->d);
        } else {
            if (visc) k_ns_gather<LPC, true, false><<<nblk(C), TB, 0, c->stream>>>(c->m, c->d);
            else k_ns_gather<LPC, false, false><<<nblk(C), TB, 0, c->stream>>>(c->m, c->d);
        }
    }
}

// reuse_ext: the limiter's extrema of this state may come from its time step
// (inside a pseudo-step); otherwise the residual reduces them itself, so a
// standalone UPDATE_RESIDUALS does the whole limiter's work.
int ns_residual_parts(ofx_ctx* c, unsigned parts, bool reuse_ext) {
    int mf = c->m.maxFacesPerCell;
    if (c->m.dim == 2) {
        if (mf <= 4) launch_cells<2, 4>(c, parts, reuse_ext);
        else if (mf <= 8) launch_cells<2, 8>(c, parts, reuse_ext);
        else { set_error("cells with more than 8 faces are not supported"); return OFX_EINVAL; }
    } else {
        if (mf <= 6) launch_cells<3, 6>(c, parts, reuse_ext);
        else if (mf <= 8) launch_cells<3, 8>(c, parts, reuse_ext);
        else if (mf <= 16) launch_cells<3, 16>(c, parts, reuse_ext);
        else { set_error("cells with more than 16 faces are not supported"); return OFX_EINVAL; }
    }
    OFX_HIP(hipGetLastError());
    return 0;
}

int ns_residual(ofx_ctx* c, bool reuse_ext) { return ns_residual_parts(c, 0xFu & ~2u, reuse_ext); }

// Ghost-cell gradients / limiter for field download (Grad::CalcGrad tail,
// codes/usolver/src/UGrad.cpp:142-148)
__global__ void k_ghost_grad(DevMesh m, DevFields d) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.B) return;
    int l = m.lc[f], r = m.C + f;
    for (int a = 0; a < 18; ++a) d.grad[a * m.T + r] = d.grad[a * m.T + l];
    for (int a = 0; a < 3; ++a) d.tgrad[a * m.T + r] = d.tgrad[a * m.T + l];
}

int fill_ghost_grad(ofx_ctx* c) {
    if (c->m.B) k_ghost_grad<<<nblk(c->m.B), TB, 0, c->stream>>>(c->m, c->d);
    OFX_HIP(hipGetLastError());
    return 0;
}

// ============================================================ LU-SGS

// UNsLusgs::Solve for face f seen from cell i: GCom::SetGeometry orientation,
// PrepareData/PrepareDataFacePrim, NsLusgs::GetStandardFluxIncrement(-1) with
// CalcFaceEigenValue and CalcDH, then CalcViscousTerm and AddFluxIncrement
// (codes/uns/src/UNsLusgs.cpp:189-335, codes/ns/src/NsLusgs.cpp:121-228,340-368).
// The face term of UNsLusgs::Solve for cell i = l (after GCom::SetGeometry
// has put i on the left): the split flux increment of the neighbour r and,
// with viscosity, its viscous radius.  Own state (rl..pl) and neighbour data
// are passed in, loaded by the caller.
// VM: the viscous term of the face, 0 none (inviscid), 1 spectral-radius
// model 1 (neighbour volume), 2 models 2 and 3 (centre distance) -- a
// compile-time choice, so a face evaluates only its own model.
template <int VM>
__device__ __forceinline__ void ns_lusgs_face(const Consts& k, double xfn, double yfn, double zfn, double vfn,
                                              double farea, double rl, double ul, double vl, double wl, double pl,
                                              const double* pj, const double* dqj, double visl, double vist,
                                              double volr, double dist, double* rhs0) {
    double rr = pj[0], ur = pj[1], vr = pj[2], wr = pj[3], pr = pj[4];
    double gl = k.gama, gr = k.gama;
    double hl = gl / (gl - 1.0) * pl / rl + 0.5 * sqr3(ul, vl, wl);
    double hr = gr / (gr - 1.0) * pr / rr + 0.5 * sqr3(ur, vr, wr);
    double tmp0 = sqrt(rr / rl);
    double tmp1 = 1.0 / (1.0 + tmp0);
    double rm = sqrt(rr * rl);
    double um = (ul + ur * tmp0) * tmp1;
    double vm = (vl + vr * tmp0) * tmp1;
    double wm = (wl + wr * tmp0) * tmp1;
    double hm = (hl + hr * tmp0) * tmp1;
    double pm = rm * (hm - 0.5 * sqr3(um, vm, wm)) * (gl - 1.0) / gl;
    double gama = k.gama;
    double c2f = dabs(gama * pm / rm);
    double cmf = sqrt(c2f);
    double vnf = xfn * um + yfn * vm + zfn * wm;
    double lmdF = dabs(vnf - vfn) + cmf;
    double c2 = dabs(gama * pj[4] / pj[0]);
    double cm = sqrt(c2);
    double vn_fluid = xfn * pj[1] + yfn * pj[2] + zfn * pj[3];
    double vn_rel = vn_fluid - vfn;
    double lmd1 = 0.5 * (vn_rel + -1.0 * lmdF);
    double lmd2 = 0.5 * ((vn_rel + cm) + -1.0 * lmdF);
    double lmd3 = 0.5 * ((vn_rel - cm) + -1.0 * lmdF);
    double x1 = (lmd1 + lmd1 - lmd2 - lmd3) / (c2 + c2);
    double x2 = (lmd2 - lmd3) / (cm + cm);
    double dc = vn_fluid * dqj[0] - xfn * dqj[1] - yfn * dqj[2] - zfn * dqj[3];
    double c2dc = c2 * dc;
    double v2 = sqr3(pj[1], pj[2], pj[3]);
    double ae = gama - 1.0;
    double af = 0.5 * ae * v2;
    double dh = -ae * (pj[1] * dqj[1] + pj[2] * dqj[2] + pj[3] * dqj[3] - dqj[4]);
    double enth = (gama / (gama - 1.0)) * (pj[4] / pj[0]);
    dh += af * dqj[0];
    double hmj = enth + 0.5 * v2;
    double term1 = dh * x1 + dc * x2;
    double term2 = c2dc * x1 + dh * x2;
    double dfj[5];
    dfj[0] = lmd1 * dqj[0] - term1;
    dfj[1] = lmd1 * dqj[1] - pj[1] * term1 + xfn * term2;
    dfj[2] = lmd1 * dqj[2] - pj[2] * term1 + yfn * term2;
    dfj[3] = lmd1 * dqj[3] - pj[3] * term1 + zfn * term2;
    dfj[4] = lmd1 * dqj[4] - hmj * term1 + vn_fluid * term2;
    for (int e = 0; e < 5; ++e) dfj[e] *= farea;
    if (VM != 0) {
        double visrad;
        if (VM == 1) {
            double density = 0.5 * (rl + rr);
            double c1 = 4.0 / 3.0 * (visl + vist);
            double cc2 = k.gama * (visl * k.oprl + vist * k.oprt);
            double c3 = 2.0 * dmax(c1, cc2) / (k.reynolds * density);
            visrad = (farea * farea) * c3 / volr;
        } else {
            double viscosity = visl + vist;
            double density = 0.5 * (rl + rr);
            double c1 = 2.0 * viscosity / (density * dist * k.reynolds + SMALL);
            visrad = 0.5 * c1 * farea;
        }
        for (int e = 0; e < 5; ++e) rhs0[e] -= visrad * dqj[e];
    }
    for (int e = 0; e < 5; ++e) rhs0[e] += dfj[e];
}

// One cell of the lower (UPPER=false) or upper sweep.  Boundary faces are
// skipped in both sweeps: their ghost increment is zero, so the reference's
// upper-sweep contribution through them is exactly +0.  The cell's faces are
// unrolled to MAXF: all neighbour loads are issued first, then the face
// terms are added in ascending face order.  EN: the faces' normals and
// areas come from the per-entry copies (DevMesh::ellg plane 2).
template <bool UPPER, int MAXF, int VM, bool EN>
__device__ __forceinline__ void ns_lusgs_cell(const DevMesh& m, const DevFields& d, const Consts& k0, int i, double mindt1,
                                              bool store = true) {
    const int T = m.T;
    const Consts k = zoned(k0, m, cell_zone(m, i));
    const double mindt = zone_mindt(m, d, i, mindt1);
    constexpr bool visc = VM != 0, sr1 = VM == 1;
    const int n = m.c2fOff[i + 1] - m.c2fOff[i];
    int ent[MAXF], nb[MAXF];
    cell_faces<MAXF>(m, i, ent, nb);
    const int lomask = m.lomask[i];
    bool act[MAXF];
#pragma unroll
    for (int j = 0; j < MAXF; ++j) {
        int f = ent[j] >= 0 ? ent[j] : ~ent[j];
        const bool lower = (lomask >> j) & 1;  // neighbour's id below the cell's
        act[j] = j < n && f >= m.B && (UPPER ? !lower : lower);
    }
    double rl = d.q[i], ul = d.q[T + i], vl = d.q[2 * T + i], wl = d.q[3 * T + i], pl = d.q[4 * T + i];
    double xi = 0.0, yi = 0.0, zi = 0.0;
    if (visc && !sr1 && !EN) { xi = m.xcc[i]; yi = m.ycc[i]; zi = m.zcc[i]; }
    double nx[MAXF], ny[MAXF], nz[MAXF], ar[MAXF], pj[MAXF][5], dqj[MAXF][5], vl_[MAXF], vt_[MAXF];
    // SR model 1: rx = vol of r; otherwise the centroid of r (EN: rx = the
    // per-entry distance along the normal)
    double rx[MAXF], ry[MAXF], rz[MAXF];
#pragma unroll
    for (int j = 0; j < MAXF; ++j) {
        if (act[j]) {
            int f = ent[j] >= 0 ? ent[j] : ~ent[j], r = nb[j];
            if (EN) {
                const dvec4 r4 = m.ellg[(size_t)(2 * m.ellw + j) * m.C + i];
                nx[j] = r4.x; ny[j] = r4.y; nz[j] = r4.z; ar[j] = r4.w;
            } else {
                nx[j] = m.xfn[f]; ny[j] = m.yfn[f]; nz[j] = m.zfn[f]; ar[j] = m.area[f];
            }
            for (int e = 0; e < 5; ++e) { pj[j][e] = d.q[e * T + r]; dqj[j][e] = d.dq[e * T + r]; }
            if (visc) {
                vl_[j] = d.visl[r]; vt_[j] = d.vist[r];
                if (EN) rx[j] = m.elldv[(size_t)((sr1 ? m.ellw : 0) + j) * m.C + i];  // volume / distance
                else if (sr1) rx[j] = m.vol[r];
                else { rx[j] = m.xcc[r]; ry[j] = m.ycc[r]; rz[j] = m.zcc[r]; }
            }
        }
    }
    double rhs0[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < MAXF; ++j) {
        if (act[j]) {
            // GCom::SetGeometry: the normal points out of i
            bool own = ent[j] >= 0;
            double xfn = own ? nx[j] : -nx[j], yfn = own ? ny[j] : -ny[j], zfn = own ? nz[j] : -nz[j];
            double vfn = own ? 0.0 : -0.0;
            double volr = 0.0, dist = 0.0;
            if (visc) {
                if (sr1) volr = rx[j];
                else if (EN) dist = rx[j];
                else dist = dabs(xfn * (rx[j] - xi) + yfn * (ry[j] - yi) + zfn * (rz[j] - zi));
            }
            ns_lusgs_face<VM>(k, xfn, yfn, zfn, vfn, ar[j], rl, ul, vl, wl, pl, pj[j], dqj[j], vl_[j], vt_[j], volr,
                              dist, rhs0);
        }
    }
    double rad = ns_impsr(m, d, k, i, mindt);
    if (!store) return;
    if (!UPPER) {
        for (int e = 0; e < 5; ++e) { double t = d.res[e * T + i] - rhs0[e]; t /= rad; d.dq[e * T + i] = t; }
    } else {
        for (int e = 0; e < 5; ++e) { double t = -rhs0[e]; t /= rad; d.dq[e * T + i] += t; }
    }
}

// Whole sweep in one workgroup: levels in order, a workgroup barrier between
// levels (every cell of a level depends only on earlier levels).
template <bool UPPER, int MAXF, int VM, bool EN>
__global__ void __launch_bounds__(1024) k_ns_sweep_1wg(DevMesh m, DevFields d, Consts k, int par) {
    double mindt = slot_result<false>(d.slots, par, m.nz * RQ_N, RQ_DTMIN);
    const int* off = UPPER ? m.upOff : m.loOff;
    const int* ord = UPPER ? m.upOrder : m.loOrder;
    int nl = UPPER ? m.nLevUp : m.nLevLo;
    for (int L = 0; L < nl; ++L) {
        for (int idx = off[L] + threadIdx.x; idx < off[L + 1]; idx += blockDim.x)
            ns_lusgs_cell<UPPER, MAXF, VM, EN>(m, d, k, ord[idx], mindt);
        __syncthreads();
    }
}

// One level per launch (wide levels on large meshes).
// base >= 0: the level's cells are base, base + 1, ... (no order lookup).
template <bool UPPER, int MAXF, int VM, bool EN>
__global__ void __launch_bounds__(TB) k_ns_sweep_level(DevMesh m, DevFields d, Consts k, int b, int e, int base,
                                                       int par) {
    // Every lane stays active (the minimum time step is reduced per wave, its
    // slot loads overlapping the cell's own): lanes past the level's end
    // evaluate its last cell and store nothing.
    const int idx0 = b + blockIdx.x * TB + threadIdx.x;
    const bool live = idx0 < e;
    const int idx = live ? idx0 : e - 1;
    double mindt;
    wave_slot_result<1, false>(d.slots, par, m.nz * RQ_N, RQ_DTMIN, &mindt);
    int i = base >= 0 ? base + (idx - b) : (UPPER ? m.upOrder : m.loOrder)[idx];
    ns_lusgs_cell<UPPER, MAXF, VM, EN>(m, d, k, i, mindt, live);
}

// Launch policy: few wide levels (multicolour numbering) -> one grid-wide
// launch per level; many narrow levels (natural numbering) -> one workgroup
// walking all levels.
bool use_level_launches(const std::vector<int>& off) {
    int nl = (int)off.size() - 1;
    return nl <= 64;
}

template <int MAXF, int VM, bool EN>
static void launch_lusgs(ofx_ctx* c, int lo_first) {
    const DevMesh& m = c->m;
    for (int pass = 0; pass < 2; ++pass) {
        bool up = pass == 1;
        const std::vector<int>& off = up ? c->sched.upOff : c->sched.loOff;
        if (!use_level_launches(off)) {
            if (up) k_ns_sweep_1wg<true, MAXF, VM, EN><<<1, 1024, 0, c->stream>>>(m, c->d, c->k, c->par_dt);
            else k_ns_sweep_1wg<false, MAXF, VM, EN><<<1, 1024, 0, c->stream>>>(m, c->d, c->k, c->par_dt);
        } else {
            const std::vector<int>& base = up ? c->sched.upBase : c->sched.loBase;
            // upper level 0 holds the cells without a higher neighbour: their
            // increment is dq += (-0)/D, which leaves dq unchanged bit for bit
            for (size_t L = up ? 1 : (size_t)lo_first; L + 1 < off.size(); ++L) {
                int b = off[L], e = off[L + 1];
                if (e <= b) continue;
                if (up)
                    k_ns_sweep_level<true, MAXF, VM, EN><<<nblk(e - b), TB, 0, c->stream>>>(m, c->d, c->k, b, e, base[L],
                                                                                       c->par_dt);
                else
                    k_ns_sweep_level<false, MAXF, VM, EN><<<nblk(e - b), TB, 0, c->stream>>>(m, c->d, c->k, b, e, base[L],
                                                                                        c->par_dt);
            }
        }
    }
}

template <int MAXF, bool EN>
static void launch_lusgs_vm(ofx_ctx* c, int lo_first) {
    if (c->k.vismodel == 0) launch_lusgs<MAXF, 0, EN>(c, lo_first);
    else if (c->k.visSRModel == 1) launch_lusgs<MAXF, 1, EN>(c, lo_first);
    else launch_lusgs<MAXF, 2, EN>(c, lo_first);
}
template <int MAXF>
static void launch_lusgs_vm(ofx_ctx* c, int lo_first) {
    if (c->normal_variant == 1) launch_lusgs_vm<MAXF, true>(c, lo_first);
    else launch_lusgs_vm<MAXF, false>(c, lo_first);
}

// The first level of the lower sweep that may be fused into the residual
// gathers: its cells must be the run 0..loOff[1]-1 and the sweeps launched
// per level.  Returns loOff[1], or 0.
int fusable_l0(const ofx_ctx* c) {
    const HostSched& h = c->sched;
    if (!use_level_launches(h.loOff) || h.loBase.empty() || h.loBase[0] != 0) return 0;
    return h.loOff[1];
}

int ns_lusgs(ofx_ctx* c, int lo_first) {
    int mf = c->m.maxFacesPerCell;
    if (mf <= 4) launch_lusgs_vm<4>(c, lo_first);
    else if (mf <= 6) launch_lusgs_vm<6>(c, lo_first);
    else if (mf <= 8) launch_lusgs_vm<8>(c, lo_first);
    else if (mf <= 16) launch_lusgs_vm<16>(c, lo_first);
    else { set_error("cells with more than 16 faces are not supported"); return OFX_EINVAL; }
    OFX_HIP(hipGetLastError());
    return 0;
}

// ============================================================ update

// UNsUpdate::UpdateFlowField + NsUpdate::CalcFlowField / WeekSolutionFix /
// SolutionFix (codes/uns/src/UNsUpdate.cpp:46-146, codes/ns/src/NsUpdate.cpp:40-67,192-217).
__device__ __forceinline__ bool ns_update_cell(const DevFields& d, const Consts& k, int T, int i, double* prim) {
    double prim0[5], q0[5], q[5], dq[5];
    for (int e = 0; e < 5; ++e) { prim0[e] = d.q[e * T + i]; dq[e] = d.dq[e * T + i]; }
    double gama = k.gama;
    prim_to_q(prim0, gama, q0);
    for (int e = 0; e < 5; ++e) q[e] = q0[e] + dq[e];
    double density = dabs(q[0]);
    double rd = 1.0 / density;
    double um = q[1] * rd, vm = q[2] * rd, wm = q[3] * rd, rem = q[4];
    double v2 = sqr3(um, vm, wm);
    double reint = rem - 0.5 * density * v2;
    prim[0] = density; prim[1] = um; prim[2] = vm; prim[3] = wm; prim[4] = (gama - 1.0) * reint;
    if (prim[0] <= 0.0 || prim[4] <= 0.0) {
        const double coefs[3] = {0.5, 0.1, 0.01};
        for (int kk = 0; kk < 3; ++kk) {
            for (int e = 0; e < 5; ++e) prim[e] += coefs[kk] * dq[e];
            if (!(prim[0] < 0.0 || prim[4] < 0.0)) return true;
        }
        return false;  // needs SolutionFix
    }
    return true;
}

// Pass 1: does any cell need SolutionFix?  (The neighbour average of
// SolutionFix reads neighbours in the serial loop's half-updated state, so
// such an update runs serially in cell order.)
__global__ void __launch_bounds__(TB) k_ns_update_check(DevMesh m, DevFields d, Consts k, int par) {
    int i = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    double flag = 0.0;
    if (i < m.C) {
        double prim[5];
        if (!ns_update_cell(d, k, m.T, i, prim)) flag = 1.0;
    }
    slot_reduce<1, true>(&flag, d.slots, par, m.nz * RQ_N, RQ_FIX);
}

__device__ void ns_solution_fix(const DevMesh& m, const DevFields& d, int i, double* prim) {
    const int T = m.T;
    for (int e = 0; e < 5; ++e) prim[e] = 0.0;
    double sumV = 0.0;
    for (int kk = m.c2fOff[i]; kk < m.c2fOff[i + 1]; ++kk) {
        int e = m.c2f[kk];
        int f = e >= 0 ? e : ~e;
        int nei = e >= 0 ? m.rc[f] : m.lc[f];
        sumV += 1.0;
        for (int s = 0; s < 5; ++s) {
            double v = d.q[s * T + nei];
            if (s == 0 || s == 4) v = dabs(v);
            prim[s] += v * 1.0;
        }
    }
    double rVol = 1.0 / sumV;
    for (int s = 0; s < 5; ++s) prim[s] *= rVol;
}

// NsCalcBoundary of one updated cell: its temperature and laminar viscosity,
// and the ghost states of its boundary faces (boundary faces come first in
// the cell's list; entries of faces where it is the right cell are ~f < 0).
__device__ __forceinline__ void ns_cell_boundary(const DevMesh& m, const DevFields& d, const Consts& k, int i,
                                                 const double* prim) {
    const Consts kz = zoned(k, m, cell_zone(m, i));
    double t = temperature(prim[0], prim[4], kz.statecoef);
    d.tempr[i] = t;
    d.visl[i] = sutherland(t, kz.ts_suth);
    if (!has_bface(m, i)) return;
    for (int kk = m.c2fOff[i]; kk < m.c2fOff[i + 1]; ++kk) {
        int f = m.c2f[kk];
        if (f < 0 || f >= m.B) break;
        ns_bc_face(m, d, kz, f, prim, t);
    }
}

// Pass 2: UPDATE_FLOWFIELD; with BC also the following CALC_BOUNDARY
// (NsCalcBoundary): T / viscosity of the cell and its boundary faces' ghosts.
// When some cell needs SolutionFix, the reference's serial loop gives its
// neighbour average the half-updated state: lower-numbered neighbours
// updated, higher-numbered ones not.  The lower-sweep levels of the LU-SGS
// schedule have exactly that order (a cell's lower-numbered neighbours lie in
// earlier levels, its higher-numbered ones in later levels), so block 0 walks
// them in order, a barrier between levels, updating each level's cells in
// parallel in place; the boundary pass follows once every cell is final.
template <bool BC>
__global__ void __launch_bounds__(TB) k_ns_update_apply(DevMesh m, DevFields d, Consts k, int par) {
    const int T = m.T;
    double anyfix = slot_result<true>(d.slots, par, m.nz * RQ_N, RQ_FIX);
    if (anyfix == 0.0) {
        int i = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
        if (i >= m.C) return;
        double prim[5];
        ns_update_cell(d, k, T, i, prim);
        for (int e = 0; e < 5; ++e) d.q[e * T + i] = prim[e];
        if (BC) ns_cell_boundary(m, d, k, i, prim);
        return;
    }
    if (blockIdx.x != 0) return;
    for (int L = 0; L < m.nLevLo; ++L) {
        for (int idx = m.loOff[L] + threadIdx.x; idx < m.loOff[L + 1]; idx += blockDim.x) {
            const int i = m.loOrder[idx];
            double prim[5];
            if (!ns_update_cell(d, k, T, i, prim)) ns_solution_fix(m, d, i, prim);
            for (int e = 0; e < 5; ++e) d.q[e * T + i] = prim[e];
        }
        __syncthreads();
    }
    if (BC) {
        for (int i = threadIdx.x; i < m.C; i += blockDim.x) {
            double prim[5];
            for (int e = 0; e < 5; ++e) prim[e] = d.q[e * T + i];
            ns_cell_boundary(m, d, k, i, prim);
        }
    }
}

int ns_update(ofx_ctx* c) {
    const DevFields d = update_fields(c);
    ++c->qgen;
    c->par_fix ^= 1;
    k_ns_update_check<<<nblk(c->m.C), TB, 0, c->stream>>>(c->m, d, c->k, c->par_fix);
    k_ns_update_apply<false><<<nblk(c->m.C), TB, 0, c->stream>>>(c->m, d, c->k, c->par_fix);
    OFX_HIP(hipGetLastError());
    return 0;
}

int ns_update_fused(ofx_ctx* c) {
    const DevFields d = update_fields(c);
    ++c->qgen;
    c->par_fix ^= 1;
    k_ns_update_check<<<nblk(c->m.C), TB, 0, c->stream>>>(c->m, d, c->k, c->par_fix);
    k_ns_update_apply<true><<<nblk(c->m.C), TB, 0, c->stream>>>(c->m, d, c->k, c->par_fix);
    OFX_HIP(hipGetLastError());
    return 0;
}

// CALC_LHS, ULhs::CalcLHS (codes/usolver/src/ULhs.cpp:47-62):
// r[e][c] *= (dt / vol) * lhscoef over the inner cells, with the time step as
// ModifyTimeStep leaves it (the clamp is applied here, where it is consumed).
template <int NEQ>
__global__ void __launch_bounds__(TB) k_lhs(DevMesh m, DevFields d, Consts k, double* r, double lhscoef, int par) {
    double mindt = slot_result<false>(d.slots, par, m.nz * RQ_N, RQ_DTMIN);
    int i = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    mindt = zone_mindt(m, d, i, mindt);
    const double dt = clamped_dt(d, k, i, mindt);
    const double s = (dt / m.vol[i]) * lhscoef;
#pragma unroll
    for (int e = 0; e < NEQ; ++e) r[(size_t)e * m.T + i] *= s;
}

int lhs(ofx_ctx* c, int solver, double lhscoef) {
    if (solver == 0)
        k_lhs<5><<<nblk(c->m.C), TB, 0, c->stream>>>(c->m, c->d, c->k, c->d.res, lhscoef, c->par_dt);
    else if (c->k.vismodel == 3)
        k_lhs<1><<<nblk(c->m.C), TB, 0, c->stream>>>(c->m, c->d, c->k, c->d.turbres, lhscoef, c->par_dt);
    OFX_HIP(hipGetLastError());
    return 0;
}

// ============================================================ norms / forces

// DUMP_RESIDUAL norm sqrt(sum res^2 / nCells) (codes/residual/src/Residual.cpp:52-72):
// fixed-shape two-level sum (block partials, then one block in order), so the
// result is identical run to run.
constexpr int NORM_BLOCKS = 512;
__global__ void __launch_bounds__(TB) k_norm_partial(DevMesh m, DevFields d, double* part, int c0, int c1) {
    __shared__ double sm[6][TB];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int i = c0 + blockIdx.x * blockDim.x + threadIdx.x; i < c1; i += gridDim.x * blockDim.x) {
        for (int e = 0; e < 5; ++e) { double v = d.res[e * m.T + i]; acc[e] += v * v; }
        double v = d.turbres[i];
        acc[5] += v * v;
    }
    for (int e = 0; e < 6; ++e) sm[e][threadIdx.x] = acc[e];
    __syncthreads();
    for (int s = TB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int e = 0; e < 6; ++e) sm[e][threadIdx.x] += sm[e][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 6) part[threadIdx.x * NORM_BLOCKS + blockIdx.x] = sm[threadIdx.x][0];
}

__global__ void k_norm_final(const double* part, double* out, int nCells) {
    int e = threadIdx.x;
    if (e >= 6) return;
    double s = 0.0;
    for (int b = 0; b < NORM_BLOCKS; ++b) s += part[e * NORM_BLOCKS + b];
    out[e] = sqrt(s / nCells);
}

// Norms over cells [c0, c1) (one zone of a multi-zone context, or all cells).
int norms_range(ofx_ctx* c, int c0, int c1, ofx_norms* out) {
    double* part = c->d.norms + 8;
    k_norm_partial<<<NORM_BLOCKS, TB, 0, c->stream>>>(c->m, c->d, part, c0, c1);
    k_norm_final<<<1, 64, 0, c->stream>>>(part, c->d.norms, c1 - c0);
    OFX_HIP(hipGetLastError());
    double h[6];
    OFX_HIP(hipMemcpyAsync(h, c->d.norms, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    OFX_HIP(hipStreamSynchronize(c->stream));
    for (int e = 0; e < 5; ++e) out->ns[e] = h[e];
    out->sa = (c->k.vismodel == 3) ? h[5] : 0.0;
    return 0;
}

int norms(ofx_ctx* c, ofx_norms* out) { return norms_range(c, 0, c->m.C, out); }

// CalcAeroForce per wall face (codes/force/src/AeroForceTask.cpp:186-440), the
// sum in face order on one thread as the reference loops.
// Pressure and viscous force of wall face f and its pressure coefficient
// (CalcAeroForce's face loop, codes/force/src/AeroForceTask.cpp:365-424, with
// Stress::CalcForce, codes/solver/src/Stress.cpp).  bcdq* = owner gradient.
__device__ __forceinline__ void wall_face_force(const DevMesh& m, const DevFields& d, const Consts& k, int f,
                                                double* pf, double* vf, double& cp) {
    const int T = m.T;
    int l = m.lc[f], r = m.C + f;
    double area = m.area[f], fnx = m.xfn[f], fny = m.yfn[f], fnz = m.zfn[f];
    double fanx = fnx * area, fany = fny * area, fanz = fnz * area;
    double dx = m.xfc[f] - m.xcc[l], dy = m.yfc[f] - m.ycc[l], dz = m.zfc[f] - m.zcc[l];
    double gp[3] = {d.grad[(0 * 6 + 4) * T + l], d.grad[(1 * 6 + 4) * T + l], d.grad[(2 * 6 + 4) * T + l]};
    double wp = d.q[4 * T + l] + gp[0] * dx + gp[1] * dy + gp[2] * dz;
    if (wp < 0.0) wp = d.q[4 * T + l];
    cp = 2.0 * (wp - k.inflow[4]);
    pf[0] = fanx * cp; pf[1] = fany * cp; pf[2] = fanz * cp;
    vf[0] = vf[1] = vf[2] = 0.0;
    if (k.vismodel > 0) {
        double g[3][3];
        for (int s = 0; s < 3; ++s)
            for (int a = 0; a < 3; ++a) g[s][a] = d.grad[(a * 6 + 1 + s) * T + l];
        double ex = m.xcc[r] - m.xcc[l], ey = m.ycc[r] - m.ycc[l], ez = m.zcc[r] - m.zcc[l];
        double ods = 1.0 / sqrt(sqr3(ex, ey, ez));
        ex *= ods; ey *= ods; ez *= ods;
        for (int s = 0; s < 3; ++s) {
            double fOld = g[s][0] * ex + g[s][1] * ey + g[s][2] * ez;
            double fnew = (d.q[(1 + s) * T + r] - d.q[(1 + s) * T + l]) * ods;
            double df = fnew - fOld;
            g[s][0] += df * ex; g[s][1] += df * ey; g[s][2] += df * ez;
        }
        double vis = d.visl[l];
        double dudx = g[0][0], dudy = g[0][1], dudz = g[0][2];
        double dvdx = g[1][0], dvdy = g[1][1], dvdz = g[1][2];
        double dwdx = g[2][0], dwdy = g[2][1], dwdz = g[2][2];
        double divv2p3 = (2.0 / 3.0) * (dudx + dvdy + dwdz);
        double txx = vis * (2.0 * dudx - divv2p3), tyy = vis * (2.0 * dvdy - divv2p3), tzz = vis * (2.0 * dwdz - divv2p3);
        double txy = vis * (dudy + dvdx), txz = vis * (dudz + dwdx), tyz = vis * (dvdz + dwdy);
        double coef = -2.0 * area * k.oreynolds;
        vf[0] = coef * (fnx * txx + fny * txy + fnz * txz);
        vf[1] = coef * (fnx * txy + fny * tyy + fnz * tyz);
        vf[2] = coef * (fnx * txz + fny * tyz + fnz * tzz);
    }
}

// Force sums over the walls of boundary faces [f0, f1), in face order (one thread).
__global__ void k_aero(DevMesh m, DevFields d, Consts k, double* out, int f0, int f1) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double tot[3] = {0, 0, 0}, pres[3] = {0, 0, 0};
    for (int f = f0; f < f1; ++f) {
        if (m.bct[f] != BC_SOLID_SURFACE) continue;
        double pf[3], vf[3], cp;
        wall_face_force(m, d, k, f, pf, vf, cp);
        for (int a = 0; a < 3; ++a) { tot[a] += pf[a] + vf[a]; pres[a] += pf[a]; }
    }
    for (int a = 0; a < 3; ++a) { out[a] = tot[a]; out[3 + a] = pres[a]; }
}

// DUMP_PRESSURE_COEFF rows (idump_pres branch of CalcAeroForce,
// codes/force/src/AeroForceTask.cpp:426-440; AeroCom::CalcCF,
// codes/force/src/Force.cpp:160-165): wall k (face order) -> x, y, z of the
// face centre, -cp, cf at out[5k..5k+4].  One thread per wall face; zone
// constants and angles by the face's zone (single zone: k, cosa, sina).
__global__ void k_wall_data(DevMesh m, DevFields d, Consts k0, double cosa0, double sina0, double* out) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= m.nWall) return;
    int f = m.wallFaces[w];
    int zone = face_zone(m, f);
    const Consts k = zoned(k0, m, zone);
    double cosa = m.nz > 1 ? m.zk[zone].cosa : cosa0, sina = m.nz > 1 ? m.zk[zone].sina : sina0;
    double pf[3], vf[3], cp;
    wall_face_force(m, d, k, f, pf, vf, cp);
    double dir = cosa * vf[0] + sina * vf[1];
    double sgn = 1.0 * dabs(dir) / (dir + SMALL);
    double cf = sqrt(sqr3(vf[0], vf[1], vf[2])) / m.area[f] * sgn;
    double* o = out + 5 * (size_t)w;
    o[0] = m.xfc[f]; o[1] = m.yfc[f]; o[2] = m.zfc[f]; o[3] = -cp; o[4] = cf;
}

int wall_data(ofx_ctx* c, double* host_out) {
    const int n = c->m.nWall;
    if (n == 0) return 0;
    double* buf = nullptr;
    OFX_HIP(hipMalloc((void**)&buf, sizeof(double) * 5 * (size_t)n));
    double aoa = c->cfg.aoa_degree * PI / 180.0;
    k_wall_data<<<(n + 255) / 256, 256, 0, c->stream>>>(c->m, c->d, c->k, std::cos(aoa), std::sin(aoa), buf);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host_out, buf, sizeof(double) * 5 * (size_t)n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); return OFX_EHIP; }
    return 0;
}

// Forces of zone z (its boundary faces [f0, f1), its constants and angles).
// Summed wall forces (total and pressure part) of zone z, and its configuration.
static int zone_forces(ofx_ctx* c, int z, double* h6, const ofx_config** cfg) {
    double* buf = c->d.norms + 8;
    int f0 = 0, f1 = c->m.B;
    Consts k = c->k;
    const ofx_config* g = &c->cfg;
    if (c->m.nz > 1) {
        f0 = c->zb[z];
        f1 = c->zb[z + 1];
        const ZoneK& zk = c->zkh[z];
        k.statecoef = zk.statecoef; k.reynolds = zk.reynolds; k.oreynolds = zk.oreynolds; k.const_cp = zk.const_cp;
        for (int e = 0; e < 5; ++e) k.inflow[e] = zk.inflow[e];
        g = &c->zcfg[z];
    }
    k_aero<<<1, 64, 0, c->stream>>>(c->m, c->d, k, buf, f0, f1);
    OFX_HIP(hipGetLastError());
    OFX_HIP(hipMemcpyAsync(h6, buf, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    OFX_HIP(hipStreamSynchronize(c->stream));
    *cfg = g;
    return 0;
}

// AeroForceInfo::CalcCoef (codes/force/src/Force.cpp:259-276) with AeroCom::CalcCL /
// CalcCD (codes/force/src/Force.cpp:150-158) -> the 12 aero.dat value columns.
// The moments the reference never accumulates are zero.
static void aero_coefficients(const double* tot, const double* pres, const ofx_config& g, double* out12) {
    double aoa = g.aoa_degree * PI / 180.0, aos = g.sideslip_degree * PI / 180.0;
    double sina = std::sin(aoa), cosa = std::cos(aoa), sinb = std::sin(aos), cosb = std::cos(aos);
    double cf[3], cp[3], cmom[3];
    for (int a = 0; a < 3; ++a) { cf[a] = tot[a] / g.aref; cp[a] = pres[a] / g.aref; cmom[a] = 0.0 / (g.aref * 1.0); }
    const double AR = 9.5;
    double cl = -sina * cf[0] + cosa * cf[1];
    double cd = cosa * cosb * cf[0] + sina * cosb * cf[1] + sinb * cf[2];
    double cd_pres = cosa * cosb * cp[0] + sina * cosb * cp[1] + sinb * cp[2];
    double cd_vis = cd - cd_pres;
    double cdl = cd - (cl * cl) / (PI * AR);
    double sgn = 1.0 * std::fabs(cf[1]) / (cf[1] + SMALL);  // SIGN(1.0, cf.y)
    double pres_center = sgn * cmom[2] / (std::fabs(cf[1]) + SMALL);
    const double row[12] = {cl, cd, cd_pres, cd_vis, cdl, pres_center, cf[0], cf[1], cf[2], cmom[0], cmom[1], cmom[2]};
    for (int i = 0; i < 12; ++i) out12[i] = row[i];
}

int aero_row_zone(ofx_ctx* c, int z, double* out12) {
    double h[6];
    const ofx_config* g = nullptr;
    OFX_TRY(zone_forces(c, z, h, &g));
    aero_coefficients(h, h + 3, *g, out12);
    return 0;
}

int aero_zone(ofx_ctx* c, int z, double* out4) {
    double row[12];
    OFX_TRY(aero_row_zone(c, z, row));
    for (int i = 0; i < 4; ++i) out4[i] = row[i];
    return 0;
}

int aero(ofx_ctx* c, double* out4) { return aero_zone(c, 0, out4); }

}  // namespace ofx
