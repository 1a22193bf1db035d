// ofx_common.h -- context, device buffers and shared helpers of libofx.
//
// All floating-point code in this library is compiled with
// -ffp-contract=off: every a*b+c is two IEEE roundings, the same as the
// reference's x86-64 build, so the FP64 path reproduces the reference
// arithmetic operation for operation.  Division and sqrt in FP64 are the
// correctly rounded HIP implementations.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ofx.h"

namespace ofx {

// codes/basic/include/HXTypeBasic.h:33-34, codes/basic/include/Constant.h
constexpr double SMALL = 1.0e-40;
constexpr double LARGE = 1.0e40;
constexpr double PI = 3.141592653589793;

// codes/geometry/src/Boundary.cpp:35-46
constexpr int BC_EXTRAPOLATION = 1;
constexpr int BC_SOLID_SURFACE = 2;
constexpr int BC_SYMMETRY = 3;
constexpr int BC_FARFIELD = 4;
constexpr int BC_INFLOW = 5;
constexpr int BC_OUTFLOW = 6;

void set_error(const std::string& msg);

// Device-resident mesh of one zone.  Face arrays are SoA of length F, cell
// arrays of length T = C + B (ghost of boundary face f is cell C + f).
struct DevMesh {
    int dim = 2, nNodes = 0, C = 0, F = 0, B = 0, T = 0;
    int maxFacesPerCell = 0;
    int nLevLo = 0, nLevUp = 0;
    int nWall = 0;
    // geometry
    double *xfn = nullptr, *yfn = nullptr, *zfn = nullptr, *area = nullptr;
    double *xfc = nullptr, *yfc = nullptr, *zfc = nullptr;
    double *xcc = nullptr, *ycc = nullptr, *zcc = nullptr, *vol = nullptr;
    double *dist = nullptr;
    int *lc = nullptr, *rc = nullptr, *bct = nullptr;
    // cell -> face CSR in ascending face id; entry = f if the cell is lc[f], ~f if rc[f]
    int *c2fOff = nullptr, *c2f = nullptr;
    // LU-SGS level schedule: cells grouped by dependency depth
    int *loOrder = nullptr, *loOff = nullptr, *upOrder = nullptr, *upOff = nullptr;
    // nodes (mesh build + wall distance only)
    double *xn = nullptr, *yn = nullptr, *zn = nullptr;
    int *fnOff = nullptr, *fnNodes = nullptr, *wallFaces = nullptr;
};

// Host copies of the level schedule sizes for launch decisions.
struct HostSched {
    std::vector<int> loOff, upOff;
};

struct Consts {
    // NsCom / Chemical (codes/ns/src/NsCom.cpp:44-122, codes/chemical/src/Chemical.cpp:280-338)
    double gama, statecoef, oreynolds, reynolds, const_cp, oprl, oprt, twall, ts_suth;
    double inflow[5];
    double cfl;
    // SA (codes/turb/src/TurbCom.cpp:110-345)
    double sa_inflow, karm2, okarm2, cb1, cb2, cv13, cw2, cw36, cw1, osigma, sac2, sac3, cb2s, cw1k;
    double max_vis_ratio, rprod, turb_cfl_ratio;
    double vencat_coef, centropy1, centropy2, max_time_ratio;
    int ischeme, ilim, ieigenfix, ivischeme, vismodel, visSRModel, visTimeStepModel, isowallbc;
    int iturb_visflux, iprod_sa, flux_fp32;
};

struct DevFields {
    double *q = nullptr;        // [5][T] primitive
    double *tempr = nullptr;    // [T]
    double *visl = nullptr;     // [T]
    double *vist = nullptr;     // [T]
    double *turbq = nullptr;    // [T]
    double *res = nullptr;      // [5][T]
    double *turbres = nullptr;  // [T]
    double *dq = nullptr;       // [5][T]
    double *turbdq = nullptr;   // [T]
    double *grad = nullptr;     // [3][6][T]: x,y,z of rho,u,v,w,p,T
    double *tgrad = nullptr;    // [3][T]: SA gradient
    double *minmax = nullptr;   // [2][5][C] Venkat neighbour min / max differences
    double *limiter = nullptr;  // [5][T]
    double *timestep = nullptr; // [T]
    double *invsr = nullptr;    // [T]
    double *vissr = nullptr;    // [T]
    double *impsr = nullptr;    // [T]
    double *turbimpsr = nullptr;// [T]
    double *invflux = nullptr;  // [5][F]
    double *visflux = nullptr;  // [5][F]
    double *saface = nullptr;   // [5][F]: radius, matrix_l, matrix_r, conv flux, diff flux
    double *red = nullptr;      // reduction slots (uint64 bit patterns of non-negative doubles)
    double *norms = nullptr;    // [6] residual norm accumulators
};

}  // namespace ofx

struct ofx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool has_mesh = false, has_cfg = false;
    ofx_config cfg{};
    ofx::Consts k{};
    ofx::DevMesh m;
    ofx::HostSched sched;
    ofx::DevFields d;
    std::vector<void*> allocs;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double last_ms[2] = {0, 0};
};

#define OFX_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ofx::set_error(std::string(#call) + ": " + hipGetErrorString(e_));     \
            return OFX_EHIP;                                                       \
        }                                                                          \
    } while (0)

#define OFX_TRY(call)               \
    do {                            \
        int r_ = (call);            \
        if (r_ != 0) return r_;     \
    } while (0)

namespace ofx {

// Blocks of a 1-D launch are dispatched round-robin over the 8 XCDs; remap so
// that each XCD walks one contiguous range of cells/faces (shared L2 hits for
// neighbour gathers).  Bijective on [0, nblk): XCD x receives
// q + (x < r) consecutive logical blocks, q = nblk / 8, r = nblk % 8.
__host__ __device__ __forceinline__ int xcd_remap(int b, int nblk) {
    int q = nblk >> 3, r = nblk & 7;
    int x = b & 7, slot = b >> 3;
    return (x < r) ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}
__device__ __forceinline__ int xcd_block(int nblk) { return xcd_remap((int)blockIdx.x, nblk); }

__device__ __forceinline__ double dmin(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return (a > b) ? a : b; }
__device__ __forceinline__ double dabs(double a) { return (a < 0) ? -a : a; }
__device__ __forceinline__ double sqr3(double a, double b, double c) { return a * a + b * b + c * c; }

int alloc_bytes(ofx_ctx* c, void** p, size_t bytes);
int upload(ofx_ctx* c, void* dst, const void* src, size_t bytes);

// Kernel launch wrappers (defined in the .hip files)
int mesh_build(ofx_ctx* c, const double* x, const double* y, const double* z, const int* fnc,
               const int* fnn, const int* lc, const int* rc, const int* bct);
int ns_boundary(ofx_ctx* c);
int sa_boundary(ofx_ctx* c);
int ns_timestep(ofx_ctx* c, int iter);
int ns_residual(ofx_ctx* c);
int sa_residual(ofx_ctx* c);
int ns_lusgs(ofx_ctx* c);
int sa_lusgs(ofx_ctx* c);
int ns_update(ofx_ctx* c);
int sa_update(ofx_ctx* c);
int norms(ofx_ctx* c, ofx_norms* out);
int fill_ghost_grad(ofx_ctx* c);
int aero(ofx_ctx* c, double* out4);
int init_freestream(ofx_ctx* c);

}  // namespace ofx
