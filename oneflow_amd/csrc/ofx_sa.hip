// ofx_sa.hip -- Spalart-Allmaras kernels: eddy viscosity and ghost cells,
// residual (spectrum + face matrices, source, first-order convection,
// diffusion, ordered gather), level-scheduled scalar LU-SGS and update.
#include "ofx_physics.h"

namespace ofx {

bool use_level_launches(const std::vector<int>& off);

constexpr int TBS = 256;
static inline int nblks(int n) { return (n + TBS - 1) / TBS; }

// TurbSrcFlux::CalcCellVist1Equ (codes/turb/src/TurbSrcFlux.cpp:279-290)
__device__ __forceinline__ double sa_vist_cell(const DevFields& d, const Consts& k, int T, int i) {
    double rho = d.q[i], nuet = d.turbq[i], visl = d.visl[i];
    double olam = rho / (visl + SMALL);
    double xsi = nuet * olam;
    double xsi3 = xsi * xsi * xsi;
    double fv1 = xsi3 / (xsi3 + k.cv13);
    double vist = fv1 * rho * nuet;
    double cap = visl * k.max_vis_ratio;
    return (cap < vist) ? cap : vist;
}

// UTurbSrcFlux::CalcVist1Equ + SetGhostCellVist (codes/uturb/src/UTurbSrcFlux.cpp:101-125,217-241)
__global__ void __launch_bounds__(TBS) k_sa_vist(DevMesh m, DevFields d, Consts k) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.T) return;
    int src = i < m.C ? i : m.lc[i - m.C];
    d.vist[i] = sa_vist_cell(d, k, m.T, src);
}

// UTurbBcSolver + TurbBcSolver for one boundary face: ghost nu_tilde from the
// owner's nu_tilde s and NS state (codes/uturb/src/UTurbBcSolver.cpp:59-160,
// codes/turb/src/TurbBcSolver.cpp:142-253)
__device__ __forceinline__ double sa_bc_value(const DevMesh& m, const DevFields& d, const Consts& k, int f, int l,
                                              double s) {
    const int T = m.T;
    int bt = m.bct[f];
    if (bt == BC_INFLOW) return k.sa_inflow;
    if (bt == BC_SOLID_SURFACE) return -s;
    if (bt != BC_FARFIELD) return s;  // extrapolation, outflow, symmetry
    double xfn = m.xfn[f], yfn = m.yfn[f], zfn = m.zfn[f];
    const double vfn = 0.0;
    double rin = d.q[l], uin = d.q[T + l], vin = d.q[2 * T + l], win = d.q[3 * T + l], pin = d.q[4 * T + l];
    double rref = k.inflow[0], uref = k.inflow[1], vref = k.inflow[2], wref = k.inflow[3], pref = k.inflow[4];
    double vnref = xfn * uref + yfn * vref + zfn * wref - vfn;
    double vnin = xfn * uin + yfn * vin + zfn * win - vfn;
    double gama = k.gama;
    double cref = sqrt(dabs(k.gama * pref / rref));
    double cin = sqrt(dabs(gama * pin / rin));
    double velin = sqrt(sqr3(uin, vin, win));
    if (velin > cin) return (vnin >= 0.0) ? s : k.sa_inflow;
    double gamm1 = gama - 1.0;
    double riemp = vnin + 2.0 * cin / gamm1;
    double riemm = vnref - 2.0 * cref / gamm1;
    double vnb = 0.5 * (riemp + riemm);
    return (vnb >= 0.0) ? s : k.sa_inflow;
}

__global__ void __launch_bounds__(TBS) k_sa_bc(DevMesh m, DevFields d, Consts k0) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.B) return;
    const Consts k = zoned(k0, m, face_zone(m, f));
    int l = m.lc[f];
    d.turbq[m.C + f] = sa_bc_value(m, d, k, f, l, d.turbq[l]);
}

int sa_boundary(ofx_ctx* c) {
    if (c->k.vismodel != 3) return 0;
    k_sa_vist<<<nblks(c->m.T), TBS, 0, c->stream>>>(c->m, c->d, c->k);
    if (c->m.B) k_sa_bc<<<nblks(c->m.B), TBS, 0, c->stream>>>(c->m, c->d, c->k);
    OFX_HIP(hipGetLastError());
    return 0;
}

// TurbCalcBc (start of TurbCalcRHS, codes/turb/src/TurbRhs.cpp:57-62) fused
// with the Green-Gauss gradient of nu_tilde (UTurbGrad,
// codes/uturb/src/UTurbGrad.cpp:51-74): each cell sets the ghosts of its own
// boundary faces (they depend only on the owner) and uses them.
// EG: the face terms come from the per-entry copies (DevMesh::ellg, as seen
// from the cell; the same sums bit for bit, see k_ns_grad), a boundary
// entry being one whose neighbour is a ghost (id C + f).
template <int DIM, int MAXF, bool EG = false>
__global__ void __launch_bounds__(TBS) k_sa_grad(DevMesh m, DevFields d, Consts k0) {
    int i = cell_block(m, TBS) * TBS + threadIdx.x;
    if (i >= m.C) return;
    const Consts k = zoned(k0, m, cell_zone(m, i));
    const int T = m.T;
    const int n = m.c2fOff[i + 1] - m.c2fOff[i];
    int ent[MAXF], nb[MAXF];
    if (EG) cell_nbs<MAXF>(m, i, nb);
    else cell_faces<MAXF>(m, i, ent, nb);
    double nui = d.turbq[i];
    // All entries are loaded (padding: face 0 or the cell's entry 0, the
    // cell itself) and masked in the sums, so that no load sits in a
    // divergent branch; only the ghosts of the cell's own boundary faces are
    // evaluated and written.
    dvec4 gw[MAXF];
    double gz[MAXF], vn[MAXF];
    const size_t S = (size_t)m.ellw * m.C;
#pragma unroll
    for (int j = 0; j < MAXF; ++j) {
        const int c = nb[j];
        if (EG) {
            // a padding entry reads entry 0 again (cached), not its zeros
            const size_t o = (size_t)(j < n ? j : 0) * m.C + i;
            gw[j] = m.ellg[o];
            gz[j] = (DIM == 3) ? m.elldv[2 * S + o] : 0.0;
        } else {
            int e = ent[j], f = e >= 0 ? e : ~e;
            gw[j] = m.fgg[f];
            gz[j] = (DIM == 3) ? m.fz[f].x : 0.0;
        }
        vn[j] = d.turbq[c];
    }
#pragma unroll
    for (int j = 0; j < MAXF; ++j) {
        const int f = EG ? nb[j] - m.C : (ent[j] >= 0 ? ent[j] : ~ent[j]);
        if (j < n && (EG ? nb[j] >= m.C : f < m.B)) {
            vn[j] = sa_bc_value(m, d, k, f, i, nui);
            d.turbq[nb[j]] = vn[j];
        }
    }
    double ax = 0, ay = 0, az = 0;
#pragma unroll
    for (int j = 0; j < MAXF; ++j) {
        const bool lv = j < n;
        if (EG) {
            double value = gw[j].x * nui + gw[j].y * vn[j];
            double tx = gw[j].z * value, ty = gw[j].w * value, tz = gz[j] * value;
            ax = ax + live_or_nz(lv, tx); ay = ay + live_or_nz(lv, ty); az = az + live_or_nz(lv, tz);
        } else {
            bool own = ent[j] >= 0;
            double cl = gw[j].x, cr = gw[j].y;
            double value = own ? (cl * nui + cr * vn[j]) : (cl * vn[j] + cr * nui);
            double tx = gw[j].z * value, ty = gw[j].w * value, tz = gz[j] * value;
            if (own) { ax = ax + live_or_nz(lv, tx); ay = ay + live_or_nz(lv, ty); az = az + live_or_nz(lv, tz); }
            else { ax = ax - live_or_pz(lv, tx); ay = ay - live_or_pz(lv, ty); az = az - live_or_pz(lv, tz); }
        }
    }
    double ovol = 1.0 / m.vol[i];
    d.tgrad[i] = ax * ovol;
    d.tgrad[T + i] = ay * ovol;
    d.tgrad[2 * T + i] = az * ovol;
}

// Normal derivative of nu_tilde at face f by face-gradient method VM
// (UTurbVisFlux::CalcFaceVisFlux1Equ via VisGrad, codes/uturb/src/UTurbVisFlux.cpp:296-369).
template <int VM>
__device__ __forceinline__ double sa_face_dqdn(const DevMesh& m, int f, double xfn, double yfn, double zfn,
                                               const dvec4* rec, double nl, double nr, const double* g1,
                                               const double* g2) {
    VisGeom G;
    vis_geom_rec<VM>(xfn, yfn, zfn, f < m.B, rec, G);
    double qf, gf[3], dqdn;
    vis_grad<VM>(G, nl, nr, g1, g2, qf, gf, dqdn);
    return dqdn;
}

// Per face: SA spectral radius and off-diagonal face matrices
// (TurbSpectrum::CalcFaceSpectrum1Equ, codes/turb/src/TurbSpectrum.cpp:75-134),
// first-order upwind convection (UTurbInvFlux + TurbInvFlux::RoeFlux with
// turb_ilim = tns_ilim = -1, codes/uturb/src/UTurbInvFlux.cpp:130-202,
// codes/turb/src/TurbInvFlux.cpp:78-98) and diffusion
// (UTurbVisFlux::CalcFaceVisFlux1Equ, codes/uturb/src/UTurbVisFlux.cpp:247-369).
// saface rows: 0 radius, 1 matrix_l, 2 matrix_r, 3 convection, 4 diffusion.
// GEO 1 (TEST method only): the face-gradient geometry is evaluated from the
// cell and face centres (vis_geom, the expressions the mesh upload used for
// the stored 160-byte record, so the same bits) instead of read from the record.
template <int DIM, int GEO>
__global__ void __launch_bounds__(512) k_sa_face(DevMesh m, DevFields d, Consts k0, int tb) {
    int f = face_block(m, tb) * tb + threadIdx.x;
    if (f >= m.F) return;
    const Consts k = zoned(k0, m, face_zone(m, f));
    const int T = m.T, F = m.F;
    int l = m.lc[f], r = m.rc[f];
    bool bnd = f < m.B;
    int rg = bnd ? l : r;
    double xfn = m.xfn[f], yfn = m.yfn[f], zfn = (DIM == 3) ? m.zfn[f] : 0.0, farea = m.area[f];
    const double vfn = 0.0;
    double rl = d.q[l], rr = d.q[r];
    double ul = d.q[T + l], vl = d.q[2 * T + l], wl = (DIM == 3) ? d.q[3 * T + l] : 0.0;
    double ur = d.q[T + r], vr = d.q[2 * T + r], wr = (DIM == 3) ? d.q[3 * T + r] : 0.0;
    double nl = d.turbq[l], nr = d.turbq[r];
    double visll = d.visl[l], vislr = d.visl[r];
    // The diffusion term's loads: in 2-D issued here with the others, before
    // any arithmetic (measured faster on the 2-D grid); in 3-D where they are
    // used (the early issue measured slower on the 3-D grid).
    const double gzf = m.zfn[f];
    dvec4 rec[5];
    double g1[3], g2[3];
    auto diffusion_loads = [&]() {
        if (GEO == 0) vis_rec_load(m, k0.iturb_visflux, f, rec);
        g1[0] = d.tgrad[l]; g1[1] = d.tgrad[T + l]; g1[2] = (DIM == 3) ? d.tgrad[2 * T + l] : 0.0;
        g2[0] = d.tgrad[rg]; g2[1] = d.tgrad[T + rg]; g2[2] = (DIM == 3) ? d.tgrad[2 * T + rg] : 0.0;
    };
    if (DIM == 2) diffusion_loads();
    const int bt = bnd ? m.bct[f] : 0;
    // spectrum
    double vnl = xfn * ul + yfn * vl + zfn * wl - vfn;
    double vnr = xfn * ur + yfn * vr + zfn * wr - vfn;
    double absVn = dabs(0.5 * (vnl + vnr));
    double rad = 0.0, m1 = 0.0, m2 = 0.0;
    rad += farea * 0.5 * absVn;
    m1 += farea * 0.5 * (-vnl - absVn);
    m2 += farea * 0.5 * (vnr - absVn);
    double orl = 1.0 / (rl + SMALL), orr = 1.0 / (rr + SMALL);
    double nuet1 = visll * orl + nl;
    double nuet2 = vislr * orr + nr;
    double nueff = dabs(0.5 * (nuet1 + nuet2));
    double cvol = 0.5 * (m.vol[l] + m.vol[r]);
    double ovol = 1.0 / cvol;
    double areaS2 = farea * farea;
    double dif = k.oreynolds * ovol * (k.osigma * nueff * areaS2);
    rad += dif;
    m1 += -dif;
    m2 += -dif;
    // convection on first-order face states (boundary faces: averaged / wall values)
    double cnl = nl, cnr = nr;
    double cul = ul, cvl = vl, cwl = wl, cur = ur, cvr = vr, cwr = wr;
    if (bnd) {
        double tn = 0.5 * (nl + nr);
        cnl = tn; cnr = tn;
        if (bt == BC_SOLID_SURFACE && k.vismodel != 0) {
            cul = cur = 0.0; cvl = cvr = 0.0; cwl = cwr = 0.0;
        } else {
            double tu = 0.5 * (ul + ur), tv = 0.5 * (vl + vr), tw = 0.5 * (wl + wr);
            cul = cur = tu; cvl = cvr = tv; cwl = cwr = tw;
        }
    }
    double cvnl = xfn * cul + yfn * cvl + zfn * cwl - vfn;
    double cvnr = xfn * cur + yfn * cvr + zfn * cwr - vfn;
    double vnl_p = 0.5 * (cvnl + dabs(cvnl));
    double vnr_n = 0.5 * (cvnr - dabs(cvnr));
    double conv = farea * (vnl_p * 1.0 * cnl + vnr_n * 1.0 * cnr);
    // diffusion (face-gradient method uniform: one dispatch, separate instantiations)
    if (DIM == 3) diffusion_loads();
    double dqdn;
    if (GEO == 1) {
        VisGeom G;
        vis_geom(m, f, l, r, 2, G);
        double qf, gf[3];
        vis_grad<2>(G, nl, nr, g1, g2, qf, gf, dqdn);
    } else if (k0.iturb_visflux == 0) dqdn = sa_face_dqdn<0>(m, f, xfn, yfn, gzf, rec, nl, nr, g1, g2);
    else if (k0.iturb_visflux == 1) dqdn = sa_face_dqdn<1>(m, f, xfn, yfn, gzf, rec, nl, nr, g1, g2);
    else dqdn = sa_face_dqdn<2>(m, f, xfn, yfn, gzf, rec, nl, nr, g1, g2);
    double visl_nv = 0.5 * (visll * orl + vislr * orr);
    double vist_nv = 0.5 * (nl + nr);
    double comVis = (visl_nv + vist_nv) * k.osigma;
    double diff = -k.oreynolds * comVis * dqdn * farea;
    d.saface[0 * F + f] = rad;
    d.saface[1 * F + f] = m1;
    d.saface[2 * F + f] = m2;
    d.saface[3 * F + f] = conv;
    d.saface[4 * F + f] = diff;
}

// TurbCom::CalcSrcSa with CalcWorkVar / CalcSaProd (codes/turb/src/TurbCom.cpp:429-465,771-872)
__device__ __forceinline__ void sa_source(const DevMesh& m, const DevFields& d, const Consts& k, int i,
                                          double& res_sa, double& spec_sa) {
    const int T = m.T;
    const double one = 1.0, two = 2.0, three = 3.0, six = 6.0, ten = 10.0, zero = 0.0;
    double dudx = d.grad[(0 * 6 + 1) * T + i], dudy = d.grad[(1 * 6 + 1) * T + i], dudz = d.grad[(2 * 6 + 1) * T + i];
    double dvdx = d.grad[(0 * 6 + 2) * T + i], dvdy = d.grad[(1 * 6 + 2) * T + i], dvdz = d.grad[(2 * 6 + 2) * T + i];
    double dwdx = d.grad[(0 * 6 + 3) * T + i], dwdy = d.grad[(1 * 6 + 3) * T + i], dwdz = d.grad[(2 * 6 + 3) * T + i];
    double rho = d.q[i], nuet = d.turbq[i], visl = d.visl[i];
    double len_scale = m.dist[i];
    double olam = rho / (visl + SMALL);
    double d2 = len_scale * len_scale;
    double od2 = one / d2;
    double work1 = dudx + dvdy + dwdz;
    double work2 = sqr3(dudx, dudy, dudz);
    double work3 = sqr3(dudy + dvdx, dvdz + dwdy, dwdx + dudz);
    double work4 = sqrt(sqr3(dwdy - dvdz, dudz - dwdx, dvdx - dudy));
    if (k.iprod_sa == 1) {
        double stp = two * work2 + work3 - (2.0 / 3.0) * work1 * work1;
        work4 = sqrt(dabs(stp));
    } else if (k.iprod_sa == 4) {
        double stp = two * work2 + work3 - (2.0 / 3.0) * work1 * work1;
        double str = work4;
        work4 = str - two * dmax(zero, str - sqrt(dabs(stp)));
    } else if (k.iprod_sa == 5) {
        double stp = two * work2 + work3 - (2.0 / 3.0) * work1 * work1;
        double str = work4;
        stp = dabs(stp);
        work4 = str * (one - dmax(zero, (str * str - stp) / (str * str + stp + SMALL)));
    }
    double str = work4;
    double xsi = nuet * olam + SMALL;
    double xsi3 = xsi * xsi * xsi;
    double fv1 = xsi3 / (xsi3 + k.cv13);
    double fv2 = one - xsi / (one + xsi * fv1);
    double rs = nuet * k.okarm2 * od2 * k.oreynolds;
    double nuetRs = nuet * rs;
    double sBar = rs * fv2;
    double omega = str;
    double stdv;
    if (sBar >= -k.sac2 * omega) stdv = omega + sBar;
    else {
        double term1 = (k.sac2 * k.sac2 * omega + k.sac3 * sBar);
        double term2 = (k.sac3 - 2 * k.sac2) * omega - sBar;
        stdv = omega + omega * term1 / term2;
    }
    double ostd = one / (stdv + SMALL);
    double r = rs * ostd;
    r = dmin(r, ten);
    // The reference's pow(x, n) for the integer powers n = 5, 6, 2 is
    // evaluated as products (within a few ulps of pow; pow itself is a long
    // log / exp sequence here, a large share of this kernel's FP64 work).
    double r2 = r * r;
    double r5 = r2 * r2 * r;
    double gg = r + k.cw2 * (r5 * r - r);
    double gg2 = gg * gg;
    double g6 = gg2 * gg2 * gg2;
    double or6 = one / six;
    double fw = gg * ofxm::ofx_pow((one + k.cw36) / (g6 + k.cw36), or6);
    const double ft2 = 0.0;
    double grd2 = sqr3(d.tgrad[i], d.tgrad[T + i], d.tgrad[2 * T + i]);
    double prod = k.cb1 * (one - ft2) * stdv * nuet;
    double diff = k.cb2s * grd2 * k.oreynolds;
    double dest = (k.cw1k * fw - k.cb1 * ft2) * nuetRs;
    diff = dmin(k.rprod * prod, diff);
    double srcTerm = prod - dest;
    srcTerm += diff;
    double p2 = fv1 / (xsi + SMALL);
    double p3 = one + xsi * fv1;
    double dfv2dk = (three * k.cv13 * (p2 * p2) - one) / (p3 * p3);
    double dsdnu = k.oreynolds * k.okarm2 * od2 * (fv2 + xsi * dfv2dk);
    double dfwdg = fw / (gg + SMALL) * (one - g6 / (g6 + k.cw36));
    double dgdr = one + k.cw2 * (six * r5 - one);
    double drdnu = k.oreynolds * k.okarm2 * od2 * ostd * (one - nuet * ostd * dsdnu);
    double dfwdnu = dfwdg * dgdr * drdnu;
    double prod0 = k.cb1 * (one - ft2) * stdv;
    double dest0 = (k.cw1 * fw - k.cb1 * ft2 * k.okarm2) * nuet * od2 * k.oreynolds;
    double prodp = k.cb1 * ((one - ft2) * dsdnu);
    double destp = (k.cw1 * (fw + dfwdnu * nuet) - k.cb1 * ft2 * k.okarm2) * od2 * k.oreynolds;
    double prde0 = prod0 - dest0;
    double prdep = prodp - destp;
    double part1 = -0.5 * (prde0 - dabs(prde0));
    double part2 = -0.5 * (prdep - dabs(prdep)) * nuet;
    spec_sa = part1 + part2;
    res_sa = srcTerm;
}

// Per cell: implicit diagonal (unsteady term, face radii in face order, then
// source Jacobian) and residual (source, then convection, then diffusion, each
// gathered in ascending face order) -- TurbCalcRHS order
// (codes/turb/src/TurbRhs.cpp:57-71).
template <int MAXF>
__global__ void __launch_bounds__(TBS) k_sa_cell(DevMesh m, DevFields d, Consts k0, int par, int L0) {
    // Every lane stays active to the end (the time-step slots are reduced per
    // wave): lanes past the last cell evaluate the last cell and store nothing.
    const int i0 = cell_block(m, TBS) * TBS + threadIdx.x;
    const bool live = i0 < m.C;
    const int i = live ? i0 : m.C - 1;
    double mindt;
    wave_slot_result<1, false>(d.slots, par, m.nz * RQ_N, RQ_DTMIN, &mindt);
    mindt = zone_mindt(m, d, i, mindt);
    const Consts k = zoned(k0, m, cell_zone(m, i));
    const int F = m.F;
    const int n = m.c2fOff[i + 1] - m.c2fOff[i];
    int ent[MAXF];
#pragma unroll
    for (int j = 0; j < MAXF; ++j) ent[j] = (j < m.ellw) ? m.ellf[(size_t)j * m.C + i] : 0;
    // all entries loaded (padding names face 0), masked in the sums
    double rad[MAXF], cv[MAXF], df[MAXF];
#pragma unroll
    for (int j = 0; j < MAXF; ++j) {
        int f = ent[j] >= 0 ? ent[j] : ~ent[j];
        rad[j] = d.saface[f];
        cv[j] = d.saface[3 * F + f];
        df[j] = d.saface[4 * F + f];
    }
    double vol = m.vol[i];
    double dt = d.timestep[i];
    if (k.max_time_ratio > 0) {
        double mp = k.max_time_ratio * mindt;
        dt = (dt < mp) ? dt : mp;
    }
    double ts = dt * k.turb_cfl_ratio;
    double imp = 0.0;
    imp += (1.0 / ts) * vol;
#pragma unroll
    for (int j = 0; j < MAXF; ++j) imp = imp + live_or_nz(j < n, rad[j]);
    double res_sa, spec_sa;
    sa_source(m, d, k, i, res_sa, spec_sa);
    imp += spec_sa * vol;
    double res = 0.0;
    res += res_sa * vol;
#pragma unroll
    for (int j = 0; j < MAXF; ++j) {
        const bool lv = j < n;
        res = (ent[j] >= 0) ? res - live_or_pz(lv, cv[j]) : res + live_or_nz(lv, cv[j]);
    }
#pragma unroll
    for (int j = 0; j < MAXF; ++j) {
        const bool lv = j < n;
        res = (ent[j] >= 0) ? res - live_or_pz(lv, df[j]) : res + live_or_nz(lv, df[j]);
    }
    if (!live) return;
    d.turbimpsr[i] = imp;
    d.turbres[i] = res;
    // cells [0, L0): the first level of the lower sweep, turbdq = (res - 0) / D
    if (i < L0) { double t = res - 0.0; t /= imp; d.turbdq[i] = t; }
}

// parts: bit 0 gradient (+ ghosts), bit 1 face terms, bit 2 cell terms.
template <int DIM, int MAXF>
static void launch_sa(ofx_ctx* c, unsigned parts) {
    const DevMesh& m = c->m;
    if (parts & 1u) {
        if (c->sagrad_variant == 1) k_sa_grad<DIM, MAXF, true><<<nblks(m.C), TBS, 0, c->stream>>>(m, c->d, c->k);
        else k_sa_grad<DIM, MAXF><<<nblks(m.C), TBS, 0, c->stream>>>(m, c->d, c->k);
    }
    if (parts & 2u) {
        const int nb = (m.F + c->saface_tb - 1) / c->saface_tb;
        if (c->sa_geom_variant == 1 && c->k.iturb_visflux == 2)
            k_sa_face<DIM, 1><<<nb, c->saface_tb, 0, c->stream>>>(m, c->d, c->k, c->saface_tb);
        else
            k_sa_face<DIM, 0><<<nb, c->saface_tb, 0, c->stream>>>(m, c->d, c->k, c->saface_tb);
    }
    if (parts & 4u) k_sa_cell<MAXF><<<nblks(m.C), TBS, 0, c->stream>>>(m, c->d, c->k, c->par_dt, c->fuse_l0);
}

int sa_residual_parts(ofx_ctx* c, unsigned parts) {
    OFX_TRY(ensure_dist(c));
    if (c->k.vismodel != 3) return 0;
    int mf = c->m.maxFacesPerCell;
    if (c->m.dim == 2) {
        if (mf <= 4) launch_sa<2, 4>(c, parts);
        else if (mf <= 8) launch_sa<2, 8>(c, parts);
        else { set_error("cells with more than 8 faces are not supported"); return OFX_EINVAL; }
    } else {
        if (mf <= 6) launch_sa<3, 6>(c, parts);
        else if (mf <= 12) launch_sa<3, 12>(c, parts);
        else { set_error("cells with more than 12 faces are not supported"); return OFX_EINVAL; }
    }
    OFX_HIP(hipGetLastError());
    return 0;
}

int sa_residual(ofx_ctx* c) { return sa_residual_parts(c, 7u); }

// UTurbLusgs lower / upper cell (codes/uturb/src/UTurbLusgs.cpp:90-304,
// codes/turb/src/TurbLusgs.cpp:123-254): matrix_l when the cell is the face's
// right cell (GCom::SetGeometry swap), matrix_r otherwise.
// The cell's faces are unrolled to MAXF: loads first, then the ordered sum.
template <bool UPPER, int MAXF>
__device__ __forceinline__ void sa_lusgs_cell(const DevMesh& m, const DevFields& d, int i) {
    const int F = m.F;
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
    double mat[MAXF], dqn[MAXF];
#pragma unroll
    for (int j = 0; j < MAXF; ++j) {
        if (act[j]) {
            int f = ent[j] >= 0 ? ent[j] : ~ent[j];
            mat[j] = (ent[j] >= 0) ? d.saface[2 * F + f] : d.saface[1 * F + f];
            dqn[j] = d.turbdq[nb[j]];
        }
    }
    double rhs0 = 0.0;
#pragma unroll
    for (int j = 0; j < MAXF; ++j)
        if (act[j]) rhs0 += mat[j] * dqn[j];
    double rad = d.turbimpsr[i];
    if (!UPPER) { double t = d.turbres[i] - rhs0; t /= rad; d.turbdq[i] = t; }
    else { double t = -rhs0; t /= rad; d.turbdq[i] += t; }
}

template <bool UPPER, int MAXF>
__global__ void __launch_bounds__(1024) k_sa_sweep_1wg(DevMesh m, DevFields d) {
    const int* off = UPPER ? m.upOff : m.loOff;
    const int* ord = UPPER ? m.upOrder : m.loOrder;
    int nl = UPPER ? m.nLevUp : m.nLevLo;
    for (int L = 0; L < nl; ++L) {
        for (int idx = off[L] + threadIdx.x; idx < off[L + 1]; idx += blockDim.x)
            sa_lusgs_cell<UPPER, MAXF>(m, d, ord[idx]);
        __syncthreads();
    }
}

template <bool UPPER, int MAXF>
__global__ void __launch_bounds__(TBS) k_sa_sweep_level(DevMesh m, DevFields d, int b, int e, int base) {
    int idx = b + blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= e) return;
    int i = base >= 0 ? base + (idx - b) : (UPPER ? m.upOrder : m.loOrder)[idx];
    sa_lusgs_cell<UPPER, MAXF>(m, d, i);
}

template <int MAXF>
static void launch_sa_lusgs(ofx_ctx* c, int lo_first) {
    const DevMesh& m = c->m;
    for (int pass = 0; pass < 2; ++pass) {
        bool up = pass == 1;
        const std::vector<int>& off = up ? c->sched.upOff : c->sched.loOff;
        if (!use_level_launches(off)) {
            if (up) k_sa_sweep_1wg<true, MAXF><<<1, 1024, 0, c->stream>>>(m, c->d);
            else k_sa_sweep_1wg<false, MAXF><<<1, 1024, 0, c->stream>>>(m, c->d);
        } else {
            const std::vector<int>& base = up ? c->sched.upBase : c->sched.loBase;
            // upper level 0: no higher neighbour, turbdq += (-0)/D changes nothing
            for (size_t L = up ? 1 : (size_t)lo_first; L + 1 < off.size(); ++L) {
                int b = off[L], e = off[L + 1];
                if (e <= b) continue;
                if (up) k_sa_sweep_level<true, MAXF><<<nblks(e - b), TBS, 0, c->stream>>>(m, c->d, b, e, base[L]);
                else k_sa_sweep_level<false, MAXF><<<nblks(e - b), TBS, 0, c->stream>>>(m, c->d, b, e, base[L]);
            }
        }
    }
}

int sa_lusgs(ofx_ctx* c, int lo_first) {
    if (c->k.vismodel != 3) return 0;
    int mf = c->m.maxFacesPerCell;
    if (mf <= 4) launch_sa_lusgs<4>(c, lo_first);
    else if (mf <= 6) launch_sa_lusgs<6>(c, lo_first);
    else if (mf <= 8) launch_sa_lusgs<8>(c, lo_first);
    else if (mf <= 16) launch_sa_lusgs<16>(c, lo_first);
    else { set_error("cells with more than 16 faces are not supported"); return OFX_EINVAL; }
    OFX_HIP(hipGetLastError());
    return 0;
}

// TurbUpdate::CalcFlowField1Equ, crelax = 1 (codes/turb/src/TurbUpdate.cpp:58-69)
__global__ void __launch_bounds__(TBS) k_sa_update(DevMesh m, DevFields d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    double qOld = d.turbq[i], dqNew = d.turbdq[i];
    double qNew = qOld + dqNew;
    qNew = (1.0 - 1.0) * qOld + 1.0 * qNew;
    d.turbq[i] = qNew;
}

// UPDATE_FLOWFIELD + CALC_BOUNDARY of the SA solver: new nu_tilde, eddy
// viscosity, and the ghosts of the cell's boundary faces (vist copied from
// the owner, nu_tilde from the SA boundary conditions).
__global__ void __launch_bounds__(TBS) k_sa_update_fused(DevMesh m, DevFields d, Consts k0) {
    int i = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    const Consts k = zoned(k0, m, cell_zone(m, i));
    double qOld = d.turbq[i], dqNew = d.turbdq[i];
    double qNew = qOld + dqNew;
    qNew = (1.0 - 1.0) * qOld + 1.0 * qNew;
    d.turbq[i] = qNew;
    double vist = sa_vist_cell(d, k, m.T, i);
    d.vist[i] = vist;
    if (!has_bface(m, i)) return;
    for (int kk = m.c2fOff[i]; kk < m.c2fOff[i + 1]; ++kk) {
        int f = m.c2f[kk];
        if (f < 0 || f >= m.B) break;  // boundary faces first; right-cell entries are ~f < 0
        d.vist[m.C + f] = vist;
        d.turbq[m.C + f] = sa_bc_value(m, d, k, f, i, qNew);
    }
}

int sa_update_fused(ofx_ctx* c) {
    if (c->k.vismodel != 3) return 0;
    k_sa_update_fused<<<nblks(c->m.C), TBS, 0, c->stream>>>(c->m, update_fields(c), c->k);
    OFX_HIP(hipGetLastError());
    return 0;
}

int sa_update(ofx_ctx* c) {
    if (c->k.vismodel != 3) return 0;
    k_sa_update<<<nblks(c->m.C), TBS, 0, c->stream>>>(c->m, update_fields(c));
    OFX_HIP(hipGetLastError());
    return 0;
}

}  // namespace ofx
