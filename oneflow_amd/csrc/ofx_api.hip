// ofx_api.hip -- the C ABI of include/ofx.h and the host-side task sequence.
//
// The host side mirrors the reference's task runtime for one zone: one call
// per registered task (UPDATE_RESIDUALS, LUSGS_*_SWEEP, UPDATE_FLOWFIELD, ...)
// and ofx_iterate() for the whole TimeIntegral::Lusgs sequence of the NS then
// the SA solver (codes/solver/src/TimeIntegral.cpp:113-130,
// codes/multigrid/src/Multigrid.cpp:339-375).
#include "ofx_common.h"

#include <algorithm>

#include <cmath>
#include <cstring>
#include <new>

namespace ofx {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int alloc_bytes(ofx_ctx* c, void** p, size_t bytes) {
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
        return OFX_ENOMEM;
    }
    c->allocs.push_back(*p);
    e = hipMemsetAsync(*p, 0, bytes, c->stream);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); return OFX_EHIP; }
    return 0;
}

int upload(ofx_ctx* c, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return 0;
    OFX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    OFX_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

__global__ void k_fill_field(double* p, long n, double v) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

static int alloc_fields(ofx_ctx* c) {
    const size_t T = c->m.T, F = c->m.F, C = c->m.C;
    DevFields& d = c->d;
    auto A = [&](double** p, size_t n) { return alloc_bytes(c, (void**)p, n * sizeof(double)); };
    OFX_TRY(A(&d.q, 5 * T));
    OFX_TRY(A(&d.tempr, T));
    OFX_TRY(A(&d.visl, T));
    OFX_TRY(A(&d.vist, T));
    OFX_TRY(A(&d.turbq, T));
    OFX_TRY(A(&d.res, 5 * T));
    OFX_TRY(A(&d.turbres, T));
    OFX_TRY(A(&d.dq, 5 * T));
    OFX_TRY(A(&d.turbdq, T));
    OFX_TRY(A(&d.grad, 18 * T));
    OFX_TRY(A(&d.tgrad, 3 * T));
    OFX_TRY(A(&d.limiter, 5 * T));
    OFX_TRY(A(&d.timestep, T));
    OFX_TRY(A(&d.invsr, T));
    OFX_TRY(A(&d.vissr, T));
    OFX_TRY(A(&d.impsr, T));
    OFX_TRY(A(&d.turbimpsr, T));
    OFX_TRY(A(&d.invflux, 5 * F));
    OFX_TRY(A(&d.visflux, 5 * F));
    OFX_TRY(A(&d.saface, 5 * F));
    OFX_TRY(A(&d.red, 16));
    OFX_TRY(A(&d.norms, 8 + 6 * 512));
    OFX_TRY(alloc_bytes(c, (void**)&d.slots, 2 * 12 * NSLOT * sizeof(unsigned long long)));
    // identity values in both parity buffers: min-reduced quantity 0 starts at LARGE, the others at +0
    for (int p = 0; p < 2; ++p)
        k_fill_field<<<1, 64, 0, c->stream>>>(reinterpret_cast<double*>(d.slots) + p * 12 * NSLOT, NSLOT, LARGE);
    OFX_HIP(hipGetLastError());
    return 0;
}

// Case constants (codes/ns/src/NsCom.cpp:44-122, codes/chemical/src/Chemical.cpp:280-338,
// codes/turb/src/TurbCom.cpp:110-380).
static void set_consts(ofx_ctx* c) {
    const ofx_config& g = c->cfg;
    Consts& k = c->k;
    double aoa = g.aoa_degree * PI / 180.0;
    double aos = g.sideslip_degree * PI / 180.0;
    k.gama = g.gama_ref;
    k.statecoef = 1.0 / (g.gama_ref * (g.mach_ref * g.mach_ref));
    k.reynolds = g.reynolds;
    k.oreynolds = 1.0 / g.reynolds;
    k.oprl = 1.0 / g.prl;
    k.oprt = 1.0 / g.prt;
    k.twall = g.twall_dim / g.tref_dim;
    k.const_cp = 1.0 / ((g.gama_ref - 1.0) * (g.mach_ref * g.mach_ref));
    k.ts_suth = 110.4 / g.tref_dim;
    k.inflow[0] = 1.0;
    k.inflow[1] = 1.0 * std::cos(aoa) * std::cos(aos);
    k.inflow[2] = 1.0 * std::sin(aoa) * std::cos(aos);
    k.inflow[3] = 1.0 * std::sin(aos);
    k.inflow[4] = k.statecoef * 1.0 * 1.0 / 1.0;
    k.cfl = g.cflst;
    k.sa_inflow = g.ref_sa;
    double sigma = 2.0 / 3.0;
    k.osigma = 1.0 / sigma;
    double karm = 0.41;
    k.karm2 = karm * karm;
    k.okarm2 = 1.0 / k.karm2;
    k.cb1 = 0.1355;
    k.cb2 = 0.622;
    double cv1 = 7.1;
    k.cv13 = cv1 * cv1 * cv1;
    k.cw2 = 0.3;
    k.cw36 = std::pow(2.0, 6);
    k.cw1 = k.cb1 * k.okarm2 + (1.0 + k.cb2) / sigma;
    k.sac2 = 0.7;
    k.sac3 = 0.9;
    k.cb2s = k.osigma * k.cb2;
    k.cw1k = k.cw1 * k.karm2;
    k.max_vis_ratio = g.max_vis_ratio;
    k.rprod = g.rprod;
    k.turb_cfl_ratio = g.turb_cfl_ratio;
    k.vencat_coef = g.vencat_coef;
    k.centropy1 = g.centropy1;
    k.centropy2 = g.centropy2;
    k.max_time_ratio = g.max_time_ratio;
    k.ischeme = g.ischeme;
    k.ilim = g.ilim;
    k.ieigenfix = g.ieigenfix;
    k.ivischeme = g.ivischeme;
    k.vismodel = g.vismodel;
    k.visSRModel = g.visSRModel;
    k.visTimeStepModel = g.visTimeStepModel;
    k.isowallbc = g.isowallbc;
    k.iturb_visflux = g.iturb_visflux;
    k.iprod_sa = g.iprod_sa;
    k.flux_fp32 = g.flux_fp32;
}

int init_freestream(ofx_ctx* c) {
    OFX_TRY(ensure_dist(c));
    ++c->qgen;
    const long T = c->m.T;
    if (c->m.nz > 1) {
        // each zone's cells and ghosts take that zone's freestream
        const long C = c->m.C;
        for (int z = 0; z < c->m.nz; ++z) {
            long c0 = c->zc[z], c1 = c->zc[z + 1], g0 = C + c->zb[z], g1 = C + c->zb[z + 1];
            for (int e = 0; e < 5; ++e) {
                double v = c->zkh[z].inflow[e];
                if (c1 > c0) k_fill_field<<<(int)((c1 - c0 + 255) / 256), 256, 0, c->stream>>>(c->d.q + e * T + c0, c1 - c0, v);
                if (g1 > g0) k_fill_field<<<(int)((g1 - g0 + 255) / 256), 256, 0, c->stream>>>(c->d.q + e * T + g0, g1 - g0, v);
            }
        }
    } else {
        for (int e = 0; e < 5; ++e)
            k_fill_field<<<(int)((T + 255) / 256), 256, 0, c->stream>>>(c->d.q + e * T, T, c->k.inflow[e]);
    }
    k_fill_field<<<(int)((T + 255) / 256), 256, 0, c->stream>>>(c->d.turbq, T, c->k.sa_inflow);
    k_fill_field<<<(int)((T + 255) / 256), 256, 0, c->stream>>>(c->d.vist, T, 0.0);
    OFX_HIP(hipGetLastError());
    OFX_TRY(ns_boundary(c));
    OFX_TRY(sa_boundary(c));
    return 0;
}

// The case-dependent constants of one configuration (as set_consts).
ZoneK zone_consts(const ofx_config& g) {
    ZoneK z;
    double aoa = g.aoa_degree * PI / 180.0, aos = g.sideslip_degree * PI / 180.0;
    z.statecoef = 1.0 / (g.gama_ref * (g.mach_ref * g.mach_ref));
    z.reynolds = g.reynolds;
    z.oreynolds = 1.0 / g.reynolds;
    z.const_cp = 1.0 / ((g.gama_ref - 1.0) * (g.mach_ref * g.mach_ref));
    z.inflow[0] = 1.0;
    z.inflow[1] = 1.0 * std::cos(aoa) * std::cos(aos);
    z.inflow[2] = 1.0 * std::sin(aoa) * std::cos(aos);
    z.inflow[3] = 1.0 * std::sin(aos);
    z.inflow[4] = z.statecoef * 1.0 * 1.0 / 1.0;
    z.cosa = std::cos(aoa);
    z.sina = std::sin(aoa);
    return z;
}

}  // namespace ofx

using namespace ofx;

#define CHECK_CTX(c)                                                           \
    do {                                                                       \
        if (!(c)) { set_error("null context"); return OFX_EINVAL; }            \
    } while (0)
#define CHECK_READY(c)                                                         \
    do {                                                                       \
        CHECK_CTX(c);                                                          \
        if (!(c)->has_mesh || !(c)->has_cfg) {                                 \
            set_error("mesh and config must be set first"); return OFX_ESTATE; \
        }                                                                      \
    } while (0)

extern "C" {

const char* ofx_last_error(void) { return g_err.c_str(); }

int ofx_ctx_create(ofx_ctx** out, int device) {
    if (!out) { set_error("null out"); return OFX_EINVAL; }
    *out = nullptr;
    int n = 0;
    OFX_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) { set_error("no such HIP device"); return OFX_EINVAL; }
    ofx_ctx* c = new (std::nothrow) ofx_ctx();
    if (!c) { set_error("out of host memory"); return OFX_ENOMEM; }
    c->device = device;
    OFX_HIP(hipSetDevice(device));
    OFX_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    *out = c;
    return 0;
}

int ofx_ctx_destroy(ofx_ctx* c) {
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int ofx_mesh_upload(ofx_ctx* c, int dim, int nNodes, const double* x, const double* y, const double* z,
                    int nFaces, int nCells, int nBFaces, const int* faceNodeCount, const int* faceNodes,
                    const int* lc, const int* rc, const int* bcType) {
    CHECK_CTX(c);
    if (c->has_mesh) { set_error("context already holds a mesh"); return OFX_ESTATE; }
    if ((dim != 2 && dim != 3) || nNodes <= 0 || nFaces <= 0 || nCells <= 0 || nBFaces < 0 || nBFaces > nFaces ||
        !x || !y || !z || !faceNodeCount || !faceNodes || !lc || !rc || (nBFaces && !bcType)) {
        set_error("ofx_mesh_upload: bad arguments"); return OFX_EINVAL;
    }
    OFX_HIP(hipSetDevice(c->device));
    c->m.dim = dim; c->m.nNodes = nNodes; c->m.C = nCells; c->m.F = nFaces; c->m.B = nBFaces;
    c->m.T = nCells + nBFaces;
    OFX_TRY(mesh_build(c, x, y, z, faceNodeCount, faceNodes, lc, rc, bcType));
    OFX_TRY(alloc_fields(c));
    OFX_HIP(hipStreamSynchronize(c->stream));
    c->has_mesh = true;
    ++c->gen;
    return 0;
}

int ofx_mesh_info(ofx_ctx* c, int* out5) {
    CHECK_CTX(c);
    if (!c->has_mesh) { set_error("no mesh"); return OFX_ESTATE; }
    out5[0] = c->m.C; out5[1] = c->m.F; out5[2] = c->m.B; out5[3] = c->m.T;
    out5[4] = std::max(c->m.nLevLo, c->m.nLevUp);
    return 0;
}

int ofx_config_set(ofx_ctx* c, const ofx_config* cfg) {
    CHECK_CTX(c);
    if (!cfg) { set_error("null config"); return OFX_EINVAL; }
    if (cfg->ischeme != 1 && cfg->ischeme != 12) { set_error("ischeme must be 1 (Roe) or 12 (SLAU2)"); return OFX_EINVAL; }
    if (cfg->ilim < -1 || cfg->ilim > 2) { set_error("ilim must be -1..2"); return OFX_EINVAL; }
    if (cfg->vismodel != 0 && cfg->vismodel != 1 && cfg->vismodel != 3) { set_error("vismodel must be 0, 1 or 3"); return OFX_EINVAL; }
    if (cfg->ivischeme < 0 || cfg->ivischeme > 2 || cfg->iturb_visflux < 0 || cfg->iturb_visflux > 2) {
        set_error("face-gradient methods 0..2 supported"); return OFX_EINVAL;
    }
    if (cfg->turb_ilim != -1 || cfg->tns_ilim != -1) { set_error("only turb_ilim = tns_ilim = -1 supported"); return OFX_EINVAL; }
    if (cfg->timestepModel != 0) { set_error("only local time stepping supported"); return OFX_EINVAL; }
    if (cfg->flux_fp32 != 0 && cfg->flux_fp32 != 1) { set_error("flux_fp32 must be 0 or 1"); return OFX_EINVAL; }
    if (!(cfg->aref > 0.0)) { set_error("aref must be positive"); return OFX_EINVAL; }
    if (cfg->time_integral != 1 && cfg->time_integral != 2) {
        set_error("time_integral must be 1 (multi-stage) or 2 (LU-SGS)"); return OFX_EINVAL;
    }
    if (cfg->time_integral == 1 && cfg->rk_stage < 1) { set_error("rk_stage must be >= 1"); return OFX_EINVAL; }
    c->cfg = *cfg;
    set_consts(c);
    c->has_cfg = true;
    c->qext_gen = -1;
    ++c->gen;
    return 0;
}

int ofx_init_freestream(ofx_ctx* c) {
    CHECK_READY(c);
    OFX_TRY(init_freestream(c));
    OFX_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

static int field_view(ofx_ctx* c, int field, double** p, size_t* n) {
    const size_t T = c->m.T, F = c->m.F, C = c->m.C;
    DevFields& d = c->d;
    DevMesh& m = c->m;
    switch (field) {
        case OFX_F_Q: *p = d.q; *n = 5 * T; return 0;
        case OFX_F_TURBQ: *p = d.turbq; *n = T; return 0;
        case OFX_F_RES: *p = d.res; *n = 5 * T; return 0;
        case OFX_F_TURBRES: *p = d.turbres; *n = T; return 0;
        case OFX_F_LIMITER: *p = d.limiter; *n = 5 * T; return 0;
        case OFX_F_TIMESTEP: *p = d.timestep; *n = T; return 0;
        case OFX_F_INVSR: *p = d.invsr; *n = T; return 0;
        case OFX_F_VISSR: *p = d.vissr; *n = T; return 0;
        case OFX_F_DQ: *p = d.dq; *n = 5 * T; return 0;
        case OFX_F_TURBDQ: *p = d.turbdq; *n = T; return 0;
        case OFX_F_TEMPR: *p = d.tempr; *n = T; return 0;
        case OFX_F_VISL: *p = d.visl; *n = T; return 0;
        case OFX_F_VIST: *p = d.vist; *n = T; return 0;
        case OFX_F_DIST: *p = m.dist; *n = C; return 0;
        case OFX_F_VOL: *p = m.vol; *n = T; return 0;
        case OFX_F_TURBIMPSR: *p = d.turbimpsr; *n = T; return 0;
        case OFX_F_IMPSR: *p = d.impsr; *n = T; return 0;
        default: break;
    }
    (void)F;
    return -1;
}

}  // extern "C"

namespace {
// A field's device blocks: `nb` arrays of `len` doubles, each indexed by cell
// (cells then ghosts; kind 0), by cell without ghosts (kind 1) or by face
// (kind 2).  The grid-order host copy is the blocks one after another.
struct FieldBlocks {
    std::vector<double*> p;
    size_t len = 0;
    int kind = 0;
};

int field_blocks(ofx_ctx* c, int field, FieldBlocks& fb) {
    const size_t T = c->m.T, F = c->m.F, C = c->m.C;
    const DevMesh& m = c->m;
    double* g = c->d.grad;
    fb.p.clear();
    fb.kind = 0;
    fb.len = T;
    switch (field) {
        case OFX_F_DQDX:  // [3][5][T]: x, y, z of rho,u,v,w,p
            for (int a = 0; a < 3; ++a)
                for (int e = 0; e < 5; ++e) fb.p.push_back(g + (a * 6 + e) * T);
            return 0;
        case OFX_F_DTDX:
            for (int a = 0; a < 3; ++a) fb.p.push_back(g + (a * 6 + 5) * T);
            return 0;
        case OFX_F_TURBDQDX:
            for (int a = 0; a < 3; ++a) fb.p.push_back(c->d.tgrad + a * T);
            return 0;
        case OFX_F_XCC:
            fb.p = {m.xcc, m.ycc, m.zcc};
            return 0;
        case OFX_F_FACE_NORMAL:
            fb.p = {m.xfn, m.yfn, m.zfn, m.area};
            fb.kind = 2; fb.len = F;
            return 0;
        case OFX_F_FACE_CENTER:
            fb.p = {m.xfc, m.yfc, m.zfc};
            fb.kind = 2; fb.len = F;
            return 0;
        default: break;
    }
    double* p;
    size_t n;
    if (field_view(c, field, &p, &n)) return -1;
    if (field == OFX_F_DIST) { fb.kind = 1; fb.len = C; }
    for (size_t k = 0; k < n / fb.len; ++k) fb.p.push_back(p + k * fb.len);
    return 0;
}

// Grid position -> storage position of entry k of a block of this kind.
inline size_t storage_index(const ofx_ctx* c, int kind, size_t k) {
    if (kind == 2) return c->fperm.empty() ? k : (size_t)c->fperm[k];
    if (c->cperm.empty() || k >= (size_t)c->m.C) return k;  // ghosts keep their positions
    return (size_t)c->cperm[k];
}
}  // namespace

extern "C" {

int ofx_get_field(ofx_ctx* c, int field, double* host) {
    CHECK_READY(c);
    if (field == OFX_F_DIST) OFX_TRY(ensure_dist(c));
    if (!host) { set_error("null host pointer"); return OFX_EINVAL; }
    FieldBlocks fb;
    if (field_blocks(c, field, fb)) { set_error("unknown field id"); return OFX_EINVAL; }
    if (field == OFX_F_DQDX || field == OFX_F_DTDX || field == OFX_F_TURBDQDX) OFX_TRY(fill_ghost_grad(c));
    if (field == OFX_F_TIMESTEP || field == OFX_F_IMPSR) OFX_TRY(materialize_timestep(c));
    const bool perm = fb.kind == 2 ? !c->fperm.empty() : !c->cperm.empty();
    std::vector<double> tmp(perm ? fb.len : 0);
    for (size_t b = 0; b < fb.p.size(); ++b) {
        double* dst = host + b * fb.len;
        OFX_HIP(hipMemcpyAsync(perm ? tmp.data() : dst, fb.p[b], fb.len * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
        if (!perm) continue;
        OFX_HIP(hipStreamSynchronize(c->stream));
        for (size_t k = 0; k < fb.len; ++k) dst[k] = tmp[storage_index(c, fb.kind, k)];
    }
    OFX_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int ofx_set_field(ofx_ctx* c, int field, const double* host) {
    CHECK_READY(c);
    if (!host) { set_error("null host pointer"); return OFX_EINVAL; }
    double* p;
    size_t n;
    if (field == OFX_F_DIST || field == OFX_F_VOL || field_view(c, field, &p, &n)) {
        set_error("field cannot be set"); return OFX_EINVAL;
    }
    FieldBlocks fb;
    if (field_blocks(c, field, fb)) { set_error("unknown field id"); return OFX_EINVAL; }
    const bool perm = !c->cperm.empty();
    std::vector<double> tmp(perm ? fb.len : 0);
    for (size_t b = 0; b < fb.p.size(); ++b) {
        const double* src = host + b * fb.len;
        if (perm) {
            for (size_t k = 0; k < fb.len; ++k) tmp[storage_index(c, fb.kind, k)] = src[k];
            src = tmp.data();
        }
        OFX_HIP(hipMemcpyAsync(fb.p[b], src, fb.len * sizeof(double), hipMemcpyHostToDevice, c->stream));
        if (perm) OFX_HIP(hipStreamSynchronize(c->stream));
    }
    if (field == OFX_F_Q) ++c->qgen;
    OFX_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int ofx_set_storage_order(ofx_ctx* c, int mode) {
    CHECK_CTX(c);
    if (c->has_mesh) { set_error("set_storage_order: call before ofx_mesh_upload"); return OFX_ESTATE; }
    if (mode != 0 && mode != 1) { set_error("set_storage_order: mode must be 0 or 1"); return OFX_EINVAL; }
    c->storage_order = mode;
    return 0;
}

int ofx_ns_boundary(ofx_ctx* c) { CHECK_READY(c); return ns_boundary(c); }
int ofx_sa_boundary(ofx_ctx* c) { CHECK_READY(c); return sa_boundary(c); }
int ofx_timestep(ofx_ctx* c, int iter) { CHECK_READY(c); return ns_timestep(c, iter); }
int ofx_residual_ns(ofx_ctx* c) { CHECK_READY(c); return ns_residual(c); }
int ofx_residual_sa(ofx_ctx* c) { CHECK_READY(c); return sa_residual(c); }
int ofx_lusgs_ns(ofx_ctx* c) { CHECK_READY(c); return ns_lusgs(c); }
int ofx_lusgs_sa(ofx_ctx* c) { CHECK_READY(c); return sa_lusgs(c); }
int ofx_update_ns(ofx_ctx* c) { CHECK_READY(c); return ns_update(c); }
int ofx_update_sa(ofx_ctx* c) { CHECK_READY(c); return sa_update(c); }
int ofx_lhs(ofx_ctx* c, int solver, double lhscoef) {
    CHECK_READY(c);
    if (solver != 0 && solver != 1) { set_error("lhs: solver must be 0 (NS) or 1 (SA)"); return OFX_EINVAL; }
    return lhs(c, solver, lhscoef);
}

}  // extern "C"

namespace ofx {
// OFX_ENAN if the current residual (NS or SA) holds a NaN.
int nan_stop(ofx_ctx* c) {
    ofx_norms nm;
    OFX_TRY(norms(c, &nm));
    bool bad = std::isnan(nm.sa);
    for (int e = 0; e < 5; ++e) bad = bad || std::isnan(nm.ns[e]);
    if (bad) { set_error("NotANumber in the residual"); return OFX_ENAN; }
    return 0;
}
}  // namespace ofx

extern "C" {

int ofx_norms_get(ofx_ctx* c, ofx_norms* out) {
    CHECK_READY(c);
    if (!out) { set_error("null out"); return OFX_EINVAL; }
    OFX_TRY(norms(c, out));
    for (int e = 0; e < 5; ++e)
        if (std::isnan(out->ns[e])) { set_error("NotANumber in NS residual"); return OFX_ENAN; }
    if (std::isnan(out->sa)) { set_error("NotANumber in SA residual"); return OFX_ENAN; }
    return 0;
}

// One StrongIter per step: NS then SA, each the TimeIntegral::Lusgs task list
// ZERO_DQ_FIELD, CALC_TIME_STEP, LOAD_RESIDUALS, UPDATE_RESIDUALS, LU-SGS
// sweeps, UPDATE_FLOWFIELD_LUSGS, CALC_BOUNDARY.
}  // extern "C"

namespace ofx {
// One pseudo-time step, kernels only (no host synchronisation, no memset), so
// it can also be recorded into a HIP graph.
int step_launch(ofx_ctx* c, int iter) {
    if (c->cfg.time_integral == 1) {
        // TimeIntegral::RungeKutta of NS, then of SA (codes/solver/src/TimeIntegral.cpp:81-110);
        // the SA solver registers no CALC_TIME_STEP and uses the NS time step
        OFX_TRY(ns_timestep(c, iter));
        for (int s = 0; s < c->cfg.rk_stage; ++s) {
            OFX_TRY(ns_residual(c, true));  // stage 0 reuses the time step's extrema
            OFX_TRY(lhs(c, 0, 1.0 / (s + 1)));
            OFX_TRY(ns_update_fused(c));
        }
        if (c->k.vismodel == 3) {
            for (int s = 0; s < c->cfg.rk_stage; ++s) {
                OFX_TRY(sa_residual(c));
                OFX_TRY(lhs(c, 1, 1.0 / (s + 1)));
                OFX_TRY(sa_update_fused(c));
            }
        }
        return 0;
    }
    OFX_TRY(ns_timestep(c, iter));
    // the first lower-sweep level is computed by the residual gathers when the
    // schedule allows it (fusable_l0), and the sweeps start at level 1
    const int L0 = fusable_l0(c);
    const int ns_l0 = c->gather_variant == 1 ? L0 : 0;
    c->fuse_l0 = ns_l0;
    int rc = ns_residual(c, true);
    c->fuse_l0 = 0;
    OFX_TRY(rc);
    OFX_TRY(ns_lusgs(c, ns_l0 > 0 ? 1 : 0));
    OFX_TRY(ns_update_fused(c));
    if (c->k.vismodel == 3) {
        c->fuse_l0 = L0;
        rc = sa_residual(c);
        c->fuse_l0 = 0;
        OFX_TRY(rc);
        OFX_TRY(sa_lusgs(c, L0 > 0 ? 1 : 0));
        OFX_TRY(sa_update_fused(c));
    }
    return 0;
}
}  // namespace ofx

extern "C" {

int ofx_iterate(ofx_ctx* c, int first_iter, int n, ofx_norms* norms_out) {
    CHECK_READY(c);
    OFX_TRY(ensure_dist(c));
    if (n < 0) { set_error("n must be >= 0"); return OFX_EINVAL; }
    for (int s = 0; s < n; ++s) {
        OFX_TRY(step_launch(c, first_iter + s));
        if (norms_out) OFX_TRY(ofx_norms_get(c, norms_out + s));
    }
    // Without per-step norms the NaN stop of DUMP_RESIDUAL
    // (codes/residual/src/ResidualTask.cpp:95-100) is checked once, on the
    // residual of the last step: a NaN state keeps producing NaN residuals.
    if (!norms_out && n > 0) OFX_TRY(nan_stop(c));
    OFX_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int ofx_wall_count(ofx_ctx* c, int* n) {
    CHECK_CTX(c);
    if (!c->has_mesh) { set_error("no mesh"); return OFX_ESTATE; }
    if (!n) { set_error("null out"); return OFX_EINVAL; }
    *n = c->m.nWall;
    return 0;
}

int ofx_wall_data(ofx_ctx* c, double* out) {
    CHECK_READY(c);
    if (!out) { set_error("null out"); return OFX_EINVAL; }
    return wall_data(c, out);
}

int ofx_aero(ofx_ctx* c, double* out4) {
    CHECK_READY(c);
    if (!out4) { set_error("null out"); return OFX_EINVAL; }
    return aero(c, out4);
}

int ofx_aero_row(ofx_ctx* c, double* out12) {
    CHECK_READY(c);
    if (!out12) { set_error("null out"); return OFX_EINVAL; }
    return aero_row_zone(c, 0, out12);
}

int ofx_time_kernel(ofx_ctx* c, int which, int iters, double* ms_out) {
    CHECK_READY(c);
    if (!ms_out || iters < 1 || which < 0 || which > 12) {
        set_error("time_kernel: which must be 0..12 and iters >= 1");
        return OFX_EINVAL;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    OFX_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); set_error("time_kernel: hipEventCreate"); return OFX_EHIP; }
    int rc = 0;
    auto once = [&]() -> int {
        if (which <= 3) return ns_residual_parts(c, 1u << which);
        if (which == 4) return ns_residual(c);
        if (which == 5) return sa_residual(c);
        if (which == 9) { OFX_TRY(ns_residual(c)); return sa_residual(c); }
        if (which == 10) return step_launch(c, c->cfg.ncfl + 1);
        if (which == 11) return ns_timestep(c, c->cfg.ncfl + 1);
        if (which == 12) return ns_lusgs(c, 0);
        return sa_residual_parts(c, 1u << (which - 6));
    };
    rc = once();  // warm-up launch
    if (!rc && hipEventRecord(e0, c->stream) != hipSuccess) rc = OFX_EHIP;
    for (int it = 0; it < iters && !rc; ++it) rc = once();
    if (!rc && hipEventRecord(e1, c->stream) != hipSuccess) rc = OFX_EHIP;
    float ms = 0;
    if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = OFX_EHIP;
    if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = OFX_EHIP;
    if (!rc) *ms_out = ms / iters;
    else if (rc == OFX_EHIP) set_error("time_kernel: HIP event error");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int ofx_set_face_tile(ofx_ctx* c, int ns_tile, int sa_tile) {
    CHECK_CTX(c);
    auto ok = [](int t) { return t == 64 || t == 128 || t == 256 || t == 512; };
    if (!ok(ns_tile) || !ok(sa_tile)) { set_error("face tile must be 64, 128, 256 or 512"); return OFX_EINVAL; }
    c->face_tb = ns_tile;
    c->saface_tb = sa_tile;
    ++c->gen;
    return 0;
}

}  // extern "C"

namespace {
// The launch-variant switch of kernel id 0..6 (0 and 1 both name the NS
// gradient kernel, which evaluates the limiter in the same pass; 6 the NS
// time step and LU-SGS sweeps).
int& variant_ref(ofx_ctx* c, int kernel) {
    switch (kernel) {
        case 2: return c->gather_variant;
        case 3: return c->face_split;
        case 4: return c->sa_geom_variant;
        case 5: return c->sagrad_variant;
        case 6: return c->normal_variant;
        default: return c->grad_variant;
    }
}
}  // namespace

extern "C" {

int ofx_set_kernel_variant(ofx_ctx* c, int kernel, int variant) {
    CHECK_CTX(c);
    if (kernel < 0 || kernel > 6 || variant < 0 || variant > 1) {
        set_error("kernel variant: kernel 0 or 1 (NS gradient with limiter), 2 (gather), 3 (NS face split), "
                  "4 (SA face geometry), 5 (SA gradient) or 6 (NS time step and sweeps), variant 0 or 1");
        return OFX_EINVAL;
    }
    variant_ref(c, kernel) = variant;
    ++c->gen;
    return 0;
}

int ofx_get_kernel_variant(ofx_ctx* c, int kernel, int* variant) {
    CHECK_CTX(c);
    if (!variant || kernel < 0 || kernel > 6) { set_error("kernel variant: kernel 0..6 and non-null out"); return OFX_EINVAL; }
    *variant = variant_ref(c, kernel);
    return 0;
}

}  // extern "C"

namespace {
// Median over 5 repetitions of the average launch time of one residual part
// (`iters` back-to-back launches each): a single short timing is at the mercy
// of clocks and co-running work (a profiler, for one).
int median_ms(ofx_ctx* c, int which, int iters, double* ms) {
    double t[5];
    for (double& x : t) OFX_TRY(ofx_time_kernel(c, which, iters, &x));
    std::sort(t, t + 5);
    *ms = t[2];
    return 0;
}
}  // namespace

extern "C" {

int ofx_autotune(ofx_ctx* c, int iters, int* chosen2) {
    CHECK_READY(c);
    if (iters < 1) { set_error("autotune: iters >= 1"); return OFX_EINVAL; }
    static const int cand[4] = {64, 128, 256, 512};
    double best = 1e300, ms = 0;
    int pick = c->face_tb;
    c->face_split = 0;
    for (int t : cand) {
        c->face_tb = t;
        OFX_TRY(median_ms(c, 2, iters, &ms));
        if (ms < best) { best = ms; pick = t; }
    }
    c->face_tb = pick;
    // the split face kernel (viscous FP64 runs), against the best tile
    if (c->k.vismodel != 0 && !c->k.flux_fp32) {
        c->face_split = 1;
        OFX_TRY(median_ms(c, 2, iters, &ms));
        c->face_split = ms < best ? 1 : 0;
    }
    // the residual gather's two forms (kernel part 3)
    best = 1e300;
    pick = c->gather_variant;
    for (int v = 0; v < 2; ++v) {
        c->gather_variant = v;
        OFX_TRY(median_ms(c, 3, iters, &ms));
        if (ms < best) { best = ms; pick = v; }
    }
    c->gather_variant = pick;
    // the NS gradient's face terms: by face id or the per-entry copies (part 0)
    best = 1e300;
    pick = c->grad_variant;
    for (int v = 0; v < 2; ++v) {
        c->grad_variant = v;
        OFX_TRY(median_ms(c, 0, iters, &ms));
        if (ms < best) { best = ms; pick = v; }
    }
    c->grad_variant = pick;
    // the NS time step's and sweeps' face normals (parts 11 and 12, on the
    // time step and residual of the current state)
    if (c->cfg.time_integral == 0) {
        best = 1e300;
        pick = c->normal_variant;
        for (int v = 0; v < 2; ++v) {
            c->normal_variant = v;
            double ms2 = 0;
            OFX_TRY(median_ms(c, 11, iters, &ms));
            OFX_TRY(median_ms(c, 12, iters, &ms2));
            if (ms + ms2 < best) { best = ms + ms2; pick = v; }
        }
        c->normal_variant = pick;
    }
    if (c->k.vismodel == 3) {
        // the same choice for the SA gradient (part 6)
        best = 1e300;
        pick = c->sagrad_variant;
        for (int v = 0; v < 2; ++v) {
            c->sagrad_variant = v;
            OFX_TRY(median_ms(c, 6, iters, &ms));
            if (ms < best) { best = ms; pick = v; }
        }
        c->sagrad_variant = pick;
        best = 1e300;
        pick = c->saface_tb;
        for (int t : cand) {
            c->saface_tb = t;
            OFX_TRY(median_ms(c, 5, iters, &ms));
            if (ms < best) { best = ms; pick = t; }
        }
        c->saface_tb = pick;
        // the TEST diffusion's face geometry: stored records or evaluated (part 7, the SA face terms)
        if (c->k.iturb_visflux == 2) {
            best = 1e300;
            pick = c->sa_geom_variant;
            for (int v = 0; v < 2; ++v) {
                c->sa_geom_variant = v;
                OFX_TRY(median_ms(c, 7, iters, &ms));
                if (ms < best) { best = ms; pick = v; }
            }
            c->sa_geom_variant = pick;
        }
    }
    ++c->gen;
    // leave the residual fields as one evaluation of the current state leaves them
    OFX_TRY(ns_residual(c));
    if (c->k.vismodel == 3) OFX_TRY(sa_residual(c));
    OFX_HIP(hipStreamSynchronize(c->stream));
    if (chosen2) { chosen2[0] = c->face_tb; chosen2[1] = c->saface_tb; }
    return 0;
}

int ofx_zones_set(ofx_ctx* c, int n, const int* cellStart, const int* bfaceStart, const ofx_config* cfgs) {
    CHECK_READY(c);
    DevMesh& m = c->m;
    const int C = m.C, B = m.B, F = m.F, T = m.T;
    if (n < 1 || !cellStart || !bfaceStart || !cfgs) { set_error("zones_set: bad arguments"); return OFX_EINVAL; }
    if (cellStart[0] != 0 || cellStart[n] != C || bfaceStart[0] != 0 || bfaceStart[n] != B) {
        set_error("zones_set: zone starts must begin at 0 and end at nCells / nBFaces");
        return OFX_EINVAL;
    }
    for (int z = 0; z < n; ++z)
        if (cellStart[z + 1] <= cellStart[z] || bfaceStart[z + 1] < bfaceStart[z]) {
            set_error("zones_set: every zone needs cells; starts must not decrease");
            return OFX_EINVAL;
        }
    for (int z = 0; z < n; ++z) {  // only the freestream (Mach, angles, Reynolds) may differ
        ofx_config a = cfgs[z], b = c->cfg;
        a.mach_ref = b.mach_ref; a.aoa_degree = b.aoa_degree; a.sideslip_degree = b.sideslip_degree;
        a.reynolds = b.reynolds;
        if (std::memcmp(&a, &b, sizeof(a)) != 0) {
            set_error("zones_set: zone configurations may differ only in mach_ref, aoa_degree, sideslip_degree, reynolds");
            return OFX_EINVAL;
        }
        if (!(cfgs[z].mach_ref > 0) || !(cfgs[z].reynolds > 0)) { set_error("zones_set: bad freestream"); return OFX_EINVAL; }
    }
    OFX_HIP(hipSetDevice(c->device));
    std::vector<int> lc(F), rc(F), wall(std::max(1, m.nWall));
    OFX_HIP(hipMemcpy(lc.data(), m.lc, sizeof(int) * F, hipMemcpyDeviceToHost));
    OFX_HIP(hipMemcpy(rc.data(), m.rc, sizeof(int) * F, hipMemcpyDeviceToHost));
    if (m.nWall) OFX_HIP(hipMemcpy(wall.data(), m.wallFaces, sizeof(int) * m.nWall, hipMemcpyDeviceToHost));
    std::vector<int> zcell(T), zface(F), zwall(n + 1, 0);
    for (int z = 0; z < n; ++z) {
        for (int i = cellStart[z]; i < cellStart[z + 1]; ++i) zcell[i] = z;
        for (int f = bfaceStart[z]; f < bfaceStart[z + 1]; ++f) { zface[f] = z; zcell[C + f] = z; }
    }
    for (int f = 0; f < F; ++f) {
        if (f >= B) zface[f] = zcell[lc[f]];
        if (zcell[lc[f]] != zface[f] || zcell[rc[f]] != zface[f]) {
            set_error("zones_set: a face joins two zones (zones must be disconnected, boundary faces zone by zone)");
            return OFX_EINVAL;
        }
    }
    for (int z = 0, w = 0; z <= n; ++z) {  // wall faces are in face order, so zone by zone
        int fz = (z < n) ? bfaceStart[z] : B;
        while (w < m.nWall && wall[w] < fz) ++w;
        zwall[z] = w;
    }
    c->zkh.resize(n);
    for (int z = 0; z < n; ++z) c->zkh[z] = zone_consts(cfgs[z]);
    auto A = [&](void** p, size_t bytes) { return alloc_bytes(c, p, bytes); };
    OFX_TRY(A((void**)&m.zcell, sizeof(int) * T));
    OFX_TRY(A((void**)&m.zface, sizeof(int) * F));
    OFX_TRY(A((void**)&m.zwall, sizeof(int) * (n + 1)));
    ZoneK* zk = nullptr;
    OFX_TRY(A((void**)&zk, sizeof(ZoneK) * n));
    OFX_TRY(upload(c, m.zcell, zcell.data(), sizeof(int) * T));
    OFX_TRY(upload(c, m.zface, zface.data(), sizeof(int) * F));
    OFX_TRY(upload(c, m.zwall, zwall.data(), sizeof(int) * (n + 1)));
    OFX_TRY(upload(c, zk, c->zkh.data(), sizeof(ZoneK) * n));
    m.zk = zk;
    // per-zone reduction slots (both parities; the min quantity starts at LARGE) and results
    DevFields& d = c->d;
    OFX_TRY(alloc_bytes(c, (void**)&d.slots, sizeof(unsigned long long) * 2 * (size_t)n * RQ_N * NSLOT));
    for (int p = 0; p < 2; ++p)
        for (int z = 0; z < n; ++z)
            k_fill_field<<<1, 64, 0, c->stream>>>(reinterpret_cast<double*>(d.slots) +
                                                      (((size_t)p * n + z) * RQ_N + RQ_DTMIN) * NSLOT, NSLOT, LARGE);
    OFX_HIP(hipGetLastError());
    OFX_TRY(A((void**)&d.zred, sizeof(double) * (size_t)n * RQ_N));
    c->par_dt = c->par_lim = c->par_fix = 0;
    c->qext_gen = -1;
    m.nz = n;
    c->zc.assign(cellStart, cellStart + n + 1);
    c->zb.assign(bfaceStart, bfaceStart + n + 1);
    c->zcfg.assign(cfgs, cfgs + n);
    c->dist_valid = false;  // per-zone walls from now on, computed on first use
    OFX_HIP(hipStreamSynchronize(c->stream));
    ++c->gen;
    return 0;
}

int ofx_zones_norms(ofx_ctx* c, ofx_norms* out) {
    CHECK_READY(c);
    if (!out) { set_error("null out"); return OFX_EINVAL; }
    if (c->m.nz == 1) return norms(c, out);
    for (int z = 0; z < c->m.nz; ++z) OFX_TRY(norms_range(c, c->zc[z], c->zc[z + 1], out + z));
    return 0;
}

int ofx_zones_aero(ofx_ctx* c, double* out) {
    CHECK_READY(c);
    if (!out) { set_error("null out"); return OFX_EINVAL; }
    for (int z = 0; z < c->m.nz; ++z) OFX_TRY(aero_zone(c, z, out + 4 * z));
    return 0;
}

// Bandwidth reference: one 16-byte load and store per thread.
__global__ void __launch_bounds__(256) k_copy16(const double2* __restrict__ a, double2* __restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

int ofx_copy_bench(ofx_ctx* c, size_t bytes, int iters, double* ms_out) {
    CHECK_CTX(c);
    if (!ms_out || iters < 1 || bytes < 16 || bytes % 16 != 0 || bytes / 16 > ((size_t)1 << 36)) {
        set_error("copy_bench: bytes must be a positive multiple of 16, iters >= 1");
        return OFX_EINVAL;
    }
    OFX_HIP(hipSetDevice(c->device));
    size_t n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    OFX_HIP(hipMalloc(&a, bytes));
    if (hipMalloc(&b, bytes) != hipSuccess) { (void)hipFree(a); set_error("copy_bench: out of device memory"); return OFX_ENOMEM; }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    do {
        if (hipMemsetAsync(a, 0, bytes, c->stream) != hipSuccess) { rc = OFX_EHIP; break; }
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) { rc = OFX_EHIP; break; }
        unsigned nb = (unsigned)((n + 255) / 256);
        k_copy16<<<nb, 256, 0, c->stream>>>((const double2*)a, (double2*)b, n);  // warm-up
        (void)hipEventRecord(e0, c->stream);
        for (int it = 0; it < iters; ++it) k_copy16<<<nb, 256, 0, c->stream>>>((const double2*)a, (double2*)b, n);
        (void)hipEventRecord(e1, c->stream);
        if (hipEventSynchronize(e1) != hipSuccess) { rc = OFX_EHIP; break; }
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = OFX_EHIP; break; }
        *ms_out = ms / iters;
    } while (0);
    if (rc) set_error("copy_bench: HIP error");
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

int ofx_sync(ofx_ctx* c) {
    CHECK_CTX(c);
    OFX_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"
