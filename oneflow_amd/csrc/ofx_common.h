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

#define OFX_HD __host__ __device__ __forceinline__
#include "ofx_math.h"

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
constexpr int MAXSEG = 9;  // cell segments <= 8 (colours); faces: boundary + 8

// Four doubles as a native vector: loads and stores of one are single
// 128-bit memory operations pairs kept in registers (HIP's double4 struct is
// copied through private memory when the compiler cannot split it).
typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef double dvec2 __attribute__((ext_vector_type(2)));

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
    // Static per-face terms, computed once from the metrics above by the same
    // expressions the kernels would evaluate (so bit for bit those values):
    //   fgg  = {cl, cr, nx A, ny A}: Green-Gauss inverse-distance weights of
    //          the left / right cell and the area-weighted normal;
    //   frec = {xf - xl, yf - yl, xf - xr, yf - yr}: face centre relative to
    //          the left and right cell centres (reconstruction, limiter);
    //   fz   = {nz A, zf - zl, zf - zr, 0} (3-D only);
    //   fzr  = {zf - zl, zf - zr} again (3-D only): the face flux's
    //          reconstruction reads these two and not the 32-byte fz.
    dvec4 *fgg = nullptr, *frec = nullptr, *fz = nullptr;
    dvec2* fzr = nullptr;
    //   fv0  = {dx, dy, dz, 1/|d|}: unit centre-to-centre vector of the STD
    //          viscous face gradient (VisGeom, method 0);
    //   fv2  = 5 per face: {d1, d2, angle1, angle2}, {dxnl, dynl, dznl, dxnr},
    //          {dynr, dznr, t1x, t1y}, {t1z, t2x, t2y, t2z}, {fw1, fw2, 0, 0}
    //          (method 2, TEST).
    dvec4 *fv0 = nullptr, *fv2 = nullptr;
    int *lc = nullptr, *rc = nullptr, *bct = nullptr;
    // The c2f / c2n lists again, cell-major with a fixed width ellw =
    // maxFacesPerCell: entry j of cell i at j * C + i (0 / i past the cell's
    // count).  A thread's loads of its cell's entries then depend on i only
    // and a wave's loads of entry j are contiguous.
    int ellw = 0;
    int *ellf = nullptr, *elln = nullptr;
    // Static face terms of the cell kernels, once per (cell, entry) in the
    // same cell-major layout: three planes of ellw * C records, plane p at
    // p * ellw * C, entry j of cell i at j * C + i (zeros past the cell's
    // count).  The gradients' terms, as seen from the cell:
    //   plane 0 {w_self, w_nb, nx A, ny A}: the Green-Gauss weights of the
    //     cell and of its neighbour (cl, cr when the cell is the face's left
    //     cell; cr, cl when it is the right one) and the area-weighted normal
    //     pointing out of the cell (negated for the right cell);
    //   plane 1 {nz A, xf - xc, yf - yc, zf - zc}: the normal's z part (out of
    //     the cell) and the face centre relative to the cell centre (z parts
    //     zero in 2-D).
    // The gradients' sums are the same bits as from fgg / frec / fz (a sum is
    // commutative, and acc - x == acc + (-x) in IEEE arithmetic).  For the
    // time step and the LU-SGS sweeps, as the face stores them:
    //   plane 2 {xfn, yfn, zfn, area}: unit normal (left to right) and area.
    // A cell kernel reads them with contiguous loads that do not wait for its
    // face ids.
    dvec4* ellg = nullptr;
    // Three more per-entry planes of doubles (same layout): 0 the distance
    // |n_out . (x_nb - x_c)| of the neighbour's centre from the cell's along
    // the normal out of the cell (the viscous spectral radius of models 2
    // and 3, evaluated as the sweeps evaluate it; equal bit for bit to the
    // time step's left-to-right form), 1 the neighbour's volume (model 1),
    // 2 nz A out of the cell again (plane 1's first value), so that the SA
    // gradient reads 40 bytes per entry rather than two records.
    double* elldv = nullptr;
    // cell -> face CSR in ascending face id; entry = f if the cell is lc[f], ~f if rc[f]
    int *c2fOff = nullptr, *c2f = nullptr;
    // per cell: bit j set when the neighbour of entry j of its face list has a
    // lower cell id in the grid's numbering (the LU-SGS lower / upper split;
    // storage positions need not follow that numbering)
    int *lomask = nullptr;
    int *c2n = nullptr;  // neighbour cell of each c2f entry (ghost C+f on boundary faces)
    // LU-SGS level schedule: cells grouped by dependency depth
    int *loOrder = nullptr, *loOff = nullptr, *upOrder = nullptr, *upOff = nullptr;
    // nodes (mesh build + wall distance only)
    double *xn = nullptr, *yn = nullptr, *zn = nullptr;
    int *fnOff = nullptr, *fnNodes = nullptr, *wallFaces = nullptr;
    // several independent zones (cases) in one context: zone of every cell
    // (ghosts included) and face, per-zone constants, per-zone wall-face
    // ranges [zwall[z], zwall[z+1]) of wallFaces.  nz == 1: all unused.
    int nz = 1;
    int *zcell = nullptr, *zface = nullptr, *zwall = nullptr;
    const struct ZoneK* zk = nullptr;
    // spatially interleaved segments of the cell and face numbering (the
    // colours of a multicolour numbering; one segment otherwise): item starts,
    // segc[nsegc] = C, segf[nsegf] = F.  Used to place one spatial region of
    // every segment on the same XCD (xcd_block_seg).
    int nsegc = 1, nsegf = 1;
    int segc[MAXSEG + 1] = {0}, segf[MAXSEG + 1] = {0};
};

// The case-dependent constants (Mach, AoA, sideslip, Reynolds) of one zone.
struct ZoneK {
    double statecoef, reynolds, oreynolds, const_cp;
    double inflow[5];
    double cosa, sina;  // of the angle of attack (wall friction sign)
};

// Host copies of the level schedule sizes for launch decisions.
struct HostSched {
    std::vector<int> loOff, upOff;
    // per level: first cell when the level's cells are one ascending run of
    // consecutive cells (a colouring stored colour by colour), else -1
    std::vector<int> loBase, upBase;
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
    double *limiter = nullptr;  // [5][T]
    double *timestep = nullptr; // [T]
    double *invsr = nullptr;    // [T]
    double *vissr = nullptr;    // [T]
    double *impsr = nullptr;    // [T]
    double *turbimpsr = nullptr;// [T]
    double *invflux = nullptr;  // [5][F]
    double *visflux = nullptr;  // [5][F]
    double *saface = nullptr;   // [5][F]: radius, matrix_l, matrix_r, conv flux, diff flux
    double *red = nullptr;      // reduction results (see RED_* in ofx_ns.hip)
    double *norms = nullptr;    // [6] residual norm accumulators
    unsigned long long *slots = nullptr;  // [2 parities][nz][RQ_N][NSLOT] grid-reduction slots
    double *zred = nullptr;     // [nz][RQ_N] per-zone reduced values (several zones only)
};

}  // namespace ofx

struct ofx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool has_mesh = false, has_cfg = false;
    // Storage order (chosen before the mesh upload): 0 (default) the grid's
    // own order, 1 spatial -- cells of each connected part in Morton order of their
    // centroids, interior faces by their lower cell, boundary faces and ghosts
    // unchanged.  cperm / fperm map a grid id to its storage position (empty:
    // identity); host copies of fields are permuted on the way in and out.
    int storage_order = 0;
    std::vector<int> cperm, fperm;
    ofx_config cfg{};
    ofx::Consts k{};
    ofx::DevMesh m;
    ofx::HostSched sched;
    ofx::DevFields d;
    std::vector<void*> allocs;
    int par_dt = 0, par_lim = 0, par_fix = 0;  // parity of the reduction buffers in use
    // qgen counts host calls that may change the primitive state; qext_gen is
    // its value when the limiter's grid extrema were last reduced (by the
    // time step or k_ns_qext), so a residual of the same state reuses them
    int qgen = 0, qext_gen = -1;
    int face_tb = 256, saface_tb = 256;        // face-kernel tile (threads per block)
    int gather_variant = 1;  // residual gather: 0 lanes per cell, 1 one thread per cell
    int sa_geom_variant = 0;  // SA diffusion, TEST method: 0 stored face records, 1 evaluated from the centres
    int face_split = 0;      // viscous NS face kernel: 1 inviscid and viscous parts on separate waves
    // cell gradients (NS gradient with limiter; SA gradient): 0 face terms
    // read by face id (fgg / frec / fz), 1 the per-entry copies (DevMesh::ellg)
    int grad_variant = 0, sagrad_variant = 0;
    // NS time step and LU-SGS sweeps: 0 face normals and areas read by face
    // id, 1 the per-entry copies (DevMesh::ellg plane 2)
    int normal_variant = 0;
    int gen = 0;  // bumped whenever launch parameters change (mesh, config, tiles)
    bool dist_valid = false;  // wall distance computed for the current mesh / zones
    int fuse_l0 = 0;  // > 0 while a full step fuses the first lower-sweep level into the gathers
    // several zones: cell / boundary-face starts (nz + 1 each), host copies of
    // the per-zone constants and configurations
    std::vector<int> zc, zb;
    std::vector<ofx::ZoneK> zkh;
    std::vector<ofx_config> zcfg;
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

// The same for an item range cut into nseg segments whose items are
// spatially interleaved (a multicolour numbering stores each colour as one
// segment): XCD x walks the x-th eighth of every segment, so one XCD's L2
// holds one spatial region of every colour.  Segment s owns logical blocks
// [ceil(start[s] / per), ceil(start[s+1] / per)); per segment each XCD gets
// q_s blocks and the r_s remainder blocks go to the next XCDs round robin
// (continuing across segments), so every XCD receives exactly the blocks
// the hardware dispatches to it (physical block b on XCD b % 8).  Bijective on
// [0, nblk) when nblk = ceil(start[nseg] / per); nseg = 1 is xcd_remap.
// CEIL(v) = ceil(v / per) for the item counts v >= 0 (a division, or a shift
// when per is a power of two known only at run time).
template <class CEIL>
__host__ __device__ __forceinline__ int xcd_remap_seg_by(int b, int nblk, CEIL ceil_div, const int* start, int nseg) {
    const int x = b & 7;
    int slot = b >> 3, rr = 0;
#pragma unroll
    for (int s = 0; s < MAXSEG; ++s) {
        if (s < nseg) {
            int b0 = ceil_div(start[s]);
            int b1 = (s + 1 == nseg) ? nblk : ceil_div(start[s + 1]);
            int n = b1 - b0, q = n >> 3, r = n & 7;
            // XCDs that receive one of the r remainder blocks: the cyclic
            // window [rr, rr + r) of 0..7, as a bit mask
            unsigned w = ((1u << r) - 1u) << rr;
            unsigned mask = (w | (w >> 8)) & 0xFFu;
            int cnt = q + (int)((mask >> x) & 1u);
            if (slot < cnt) {
                int before = x * q + __builtin_popcount(mask & ((1u << x) - 1u));
                return b0 + before + slot;
            }
            slot -= cnt;
            rr = (rr + r) & 7;
        }
    }
    return b;  // not reached for a consistent launch
}
__host__ __device__ __forceinline__ int xcd_remap_seg(int b, int nblk, int per, const int* start, int nseg) {
    return xcd_remap_seg_by(b, nblk, [per](int v) { return (v + per - 1) / per; }, start, nseg);
}
// The same for per = 1 << lg.
__host__ __device__ __forceinline__ int xcd_remap_seg_lg(int b, int nblk, int lg, const int* start, int nseg) {
    return xcd_remap_seg_by(
        b, nblk, [lg](int v) { return (int)(((unsigned)v + (1u << lg) - 1u) >> lg); }, start, nseg);
}
// Entries 0..MAXF-1 of cell i's face list (cell-major copy; entries past the
// list hold 0 / i and must be masked with the cell's count).
template <int MAXF>
__device__ __forceinline__ void cell_faces(const DevMesh& m, int i, int* ent, int* nb) {
#pragma unroll
    for (int j = 0; j < MAXF; ++j) {
        bool in = j < m.ellw;
        ent[j] = in ? m.ellf[(size_t)j * m.C + i] : 0;
        nb[j] = in ? m.elln[(size_t)j * m.C + i] : i;
    }
}

// The neighbour ids alone (the kernels that take their face terms from
// DevMesh::ellg need no face id).
template <int MAXF>
__device__ __forceinline__ void cell_nbs(const DevMesh& m, int i, int* nb) {
#pragma unroll
    for (int j = 0; j < MAXF; ++j) nb[j] = (j < m.ellw) ? m.elln[(size_t)j * m.C + i] : i;
}

// Whether cell i has a boundary face.  A cell's faces are listed in
// ascending face id and boundary faces (ids below B, whose left cell the
// cell is) come first, so its first entry tells: one contiguous load that
// does not depend on the cell's list offsets.
__device__ __forceinline__ bool has_bface(const DevMesh& m, int i) {
    const int e = m.ellf[i];
    return e >= 0 && e < m.B;
}

// Logical block of a launch over the cells (faces), `per` items per block.
__device__ __forceinline__ int cell_block(const DevMesh& m, int per) {
    return xcd_remap_seg((int)blockIdx.x, (int)gridDim.x, per, m.segc, m.nsegc);
}
// per: a power of two (the face tiles are 64 .. 512)
__device__ __forceinline__ int face_block(const DevMesh& m, int per) {
    return xcd_remap_seg_lg((int)blockIdx.x, (int)gridDim.x, __builtin_ctz((unsigned)per), m.segf, m.nsegf);
}

// x for a live entry, -0.0 otherwise: sum + live_or_nz(live, x) is bitwise
// `live ? sum + x : sum` (x + -0.0 == x for every x, signed zeros included),
// but keeps the load of x unconditional, so the compiler cannot sink it into
// a divergent branch and wait on it there.
__device__ __forceinline__ double live_or_nz(bool live, double x) { return live ? x : -0.0; }
// The same for a subtraction: acc - live_or_pz(live, x) is `live ? acc - x : acc`.
__device__ __forceinline__ double live_or_pz(bool live, double x) { return live ? x : 0.0; }

__device__ __forceinline__ double dmin(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return (a > b) ? a : b; }
__device__ __forceinline__ double dabs(double a) { return (a < 0) ? -a : a; }
__device__ __forceinline__ double sqr3(double a, double b, double c) { return a * a + b * b + c * c; }

// Grid-wide min/max without fences or extra launches.  Values are non-negative
// doubles (maxima) or positive doubles (minima), whose IEEE bit patterns order
// like the values, so integer atomics on the bits are exact.  Each block
// reduces in LDS and issues one atomic per quantity into one of NSLOT slots
// (spreading contention); the consumer reduces the NSLOT slots.  Two parity
// buffers alternate between calls: a reducing kernel writes buffer `par` and
// re-arms buffer 1-par (identity values) for the next call, whose consumer of
// this call has completed by stream order.
constexpr int NSLOT = 64;

// a where the mask bits are 0, b where they are 1 (m is all zeros or all ones).
__device__ __forceinline__ double bit_blend(double a, double b, long long m) {
    long long x = __double_as_longlong(a), y = __double_as_longlong(b);
    return __longlong_as_double((x & ~m) | (y & m));
}

// Wave reduction of NQ values per lane as a reduce-scatter butterfly: the
// values (padded to P, a power of two) halve at each exchange -- a lane keeps
// the half its offset bit selects and receives its partner's copy of that
// half -- so after log2(P) exchanges each lane holds one quantity, reduced
// over the lanes differing in those bits; plain exchanges over the remaining
// bits finish it: P - 1 + log2(64 / P) shuffles instead of 6 NQ.  Returns
// the wave's reduction of quantity q (set per lane); `holder` marks one lane
// per quantity q < NQ.  The halves are picked by a bit blend under an opaque
// mask (a select here is turned into a dynamic index into the value array,
// a long select chain).  Every lane of the wave must be active.
template <int NQ, bool IS_MAX>
__device__ __forceinline__ double wave_reduce_scatter(const double* v, int lane, int& q, bool& holder) {
    constexpr int P = NQ <= 1 ? 1 : NQ <= 2 ? 2 : NQ <= 4 ? 4 : NQ <= 8 ? 8 : 16;
    static_assert(NQ <= 16, "at most 16 quantities");
    const double ident = IS_MAX ? 0.0 : LARGE;
    double x[P];
#pragma unroll
    for (int t = 0; t < P; ++t) x[t] = t < NQ ? v[t] : ident;
    int qi = 0;
#pragma unroll
    for (int half = P / 2, o = 32; half >= 1; half >>= 1, o >>= 1) {
        long long m = (lane & o) ? -1ll : 0ll;
        asm volatile("" : "+v"(m));
#pragma unroll
        for (int k = 0; k < half; ++k) {
            double keep = bit_blend(x[k], x[k + half], m);
            double give = bit_blend(x[k + half], x[k], m);
            double got = __shfl_xor(give, o);
            x[k] = IS_MAX ? ((keep > got) ? keep : got) : ((keep < got) ? keep : got);
        }
        if (lane & o) qi += half;
    }
#pragma unroll
    for (int o = 32 / P; o > 0; o >>= 1) {
        double got = __shfl_xor(x[0], o);
        x[0] = IS_MAX ? ((x[0] > got) ? x[0] : got) : ((x[0] < got) ? x[0] : got);
    }
    q = qi;
    holder = (lane & (64 / P - 1)) == 0 && qi < NQ;
    return x[0];
}

template <int NQ, bool IS_MAX>
__device__ __forceinline__ void slot_reduce(const double* v, unsigned long long* buf, int par, int nq_total, int q0) {
    __shared__ double sm[NQ][16];
    const int nw = (int)(blockDim.x >> 6), w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    int qi;
    bool holder;
    const double xw = wave_reduce_scatter<NQ, IS_MAX>(v, lane, qi, holder);
    if (holder) sm[qi][w] = xw;
    __syncthreads();
    int slot = (int)(blockIdx.x % NSLOT);
    unsigned long long* mine = buf + (size_t)par * nq_total * NSLOT;
    if (threadIdx.x < NQ) {
        double x = sm[threadIdx.x][0];
        for (int j = 1; j < nw; ++j) {
            double y = sm[threadIdx.x][j];
            x = IS_MAX ? ((x > y) ? x : y) : ((x < y) ? x : y);
        }
        unsigned long long bits = (unsigned long long)__double_as_longlong(x);
        unsigned long long* a = mine + (q0 + threadIdx.x) * NSLOT + slot;
        if (IS_MAX) atomicMax(a, bits); else atomicMin(a, bits);
    }
    if (blockIdx.x == 0) {
        unsigned long long* other = buf + (size_t)(1 - par) * nq_total * NSLOT;
        unsigned long long ident = IS_MAX ? 0ull : (unsigned long long)__double_as_longlong(LARGE);
        for (int t = threadIdx.x; t < NQ * NSLOT; t += blockDim.x) other[q0 * NSLOT + t] = ident;
    }
}

// Consumer side: reduce the NSLOT slots of quantity q (whole block, result to all threads).
template <bool IS_MAX>
__device__ __forceinline__ double slot_result(const unsigned long long* buf, int par, int nq_total, int q) {
    __shared__ double r[16];
    const unsigned long long* mine = buf + (size_t)par * nq_total * NSLOT + q * NSLOT;
    double x = IS_MAX ? 0.0 : LARGE;
    if (threadIdx.x < NSLOT) x = __longlong_as_double((long long)mine[threadIdx.x]);
    if (threadIdx.x < 64) {
        for (int o = 32; o > 0; o >>= 1) {
            double y = __shfl_xor(x, o);
            x = IS_MAX ? ((x > y) ? x : y) : ((x < y) ? x : y);
        }
        if (threadIdx.x == 0) r[0] = x;
    }
    __syncthreads();
    double out = r[0];
    __syncthreads();
    return out;
}

// A double from lane l of the wave (l wave-uniform), whatever the exec mask.
__device__ __forceinline__ double readlane_d(double x, int l) {
    long long b = __double_as_longlong(x);
    int lo = __builtin_amdgcn_readlane((int)b, l);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Consumer side without LDS or barriers: every wave reduces the NSLOT slots of
// quantities q0 .. q0+NQ-1 itself, so that no block-wide barrier holds the
// wave's other loads back.  G lanes share a quantity (lane l takes quantity
// min(l / G, NQ - 1) and NSLOT / G consecutive slots of it), G - 1 shuffles
// combine them and out[t] is read from lane t*G.  Exact (a min or max, in any
// order): the value slot_result returns.  Every lane of the wave must be
// active, so the callers keep all lanes to this point (lanes past the last
// item work on the last item and store nothing).
template <int NQ, bool IS_MAX>
__device__ __forceinline__ void wave_slot_result(const unsigned long long* buf, int par, int nq_total, int q0,
                                                 double* out) {
    constexpr int G = NQ == 1 ? 64 : NQ == 2 ? 32 : NQ <= 4 ? 16 : NQ <= 8 ? 8 : NQ <= 16 ? 4 : 2;
    constexpr int PER = NSLOT / G;
    static_assert(NQ <= 32 && NSLOT == 64, "slot layout");
    const int lane = (int)(threadIdx.x & 63);
    const int q = (lane / G < NQ) ? lane / G : NQ - 1, part = lane % G;
    const unsigned long long* src = buf + ((size_t)par * nq_total + q0 + q) * NSLOT + part * PER;
    double x = __longlong_as_double((long long)src[0]);
#pragma unroll
    for (int t = 1; t < PER; ++t) {
        double y = __longlong_as_double((long long)src[t]);
        x = IS_MAX ? ((x > y) ? x : y) : ((x < y) ? x : y);
    }
#pragma unroll
    for (int o = 1; o < G; o <<= 1) {
        double y = __shfl_xor(x, o);
        x = IS_MAX ? ((x > y) ? x : y) : ((x < y) ? x : y);
    }
#pragma unroll
    for (int t = 0; t < NQ; ++t) out[t] = readlane_d(x, t * G);
}

// Lanes-per-cell helpers: a cell's faces are spread over LPC consecutive lanes
// of a wave (lane j holds face j of the cell's ascending list).  ordered_sum
// rebuilds the serial accumulation sum = ((0 + c0) + c1) + ... in face order,
// bit for bit, with in-wave shuffles.
template <int LPC>
__device__ __forceinline__ double ordered_sum(double c, int n) {
    int base = (int)(threadIdx.x & 63) / LPC * LPC;
    double sum = 0.0;
#pragma unroll
    for (int jj = 0; jj < LPC; ++jj) {
        double y = __shfl(c, base + jj, 64);
        if (jj < n) sum = sum + y;
    }
    return sum;
}
template <int LPC>
__device__ __forceinline__ double group_min(double x) {
#pragma unroll
    for (int o = LPC / 2; o > 0; o >>= 1) { double y = __shfl_xor(x, o, 64); x = (x < y) ? x : y; }
    return x;
}
template <int LPC>
__device__ __forceinline__ double group_max(double x) {
#pragma unroll
    for (int o = LPC / 2; o > 0; o >>= 1) { double y = __shfl_xor(x, o, 64); x = (x > y) ? x : y; }
    return x;
}

// Grid-reduction quantities (slot_reduce / slot_result):
//   0      : min time step (min)
//   1..5   : -qmin per primitive (Venkat, max)
//   6..10  : qmax per primitive (Venkat, max)
//   11     : 1 if some cell needs SolutionFix (max; one flag per context)
constexpr int RQ_DTMIN = 0, RQ_QMIN = 1, RQ_QMAX = 6, RQ_FIX = 11, RQ_N = 12;

// Zone of a cell / face (contexts holding several independent cases).
__device__ __forceinline__ int cell_zone(const DevMesh& m, int i) { return m.nz > 1 ? m.zcell[i] : 0; }
__device__ __forceinline__ int face_zone(const DevMesh& m, int f) { return m.nz > 1 ? m.zface[f] : 0; }

// Minimum time step of cell i's zone: the block-reduced value of a
// single-zone context, or the zone's extracted value.
__device__ __forceinline__ double zone_mindt(const DevMesh& m, const DevFields& d, int i, double mindt1) {
    return m.nz > 1 ? d.zred[(size_t)m.zcell[i] * RQ_N + RQ_DTMIN] : mindt1;
}

// Constants of the zone a cell or face belongs to (single-zone contexts: k).
__device__ __forceinline__ Consts zoned(const Consts& k, const DevMesh& m, int zone) {
    Consts r = k;
    if (m.nz > 1 && zone >= 0) {
        const ZoneK& z = m.zk[zone];
        r.statecoef = z.statecoef;
        r.reynolds = z.reynolds;
        r.oreynolds = z.oreynolds;
        r.const_cp = z.const_cp;
        for (int e = 0; e < 5; ++e) r.inflow[e] = z.inflow[e];
    }
    return r;
}

// Wave-level grid reductions (several zones, or kernels whose waves reduce
// different quantities).  Buffer layout [2 parities][nz][nq_total][NSLOT].
// Every lane of the wave calls; lanes without a cell pass zone = -1.  Each
// wave reduces the lanes of each zone it holds (zones are contiguous cell
// ranges, so usually one) and issues one atomic per zone and quantity.  The
// other parity's entries of these quantities are re-armed by the calling
// waves of all blocks together (waves of one block that reduce the same
// quantities write the same identity values).
template <int NQ, bool IS_MAX>
__device__ __forceinline__ void zslot_reduce(const double* v, int zone, unsigned long long* buf, int par,
                                             int nq_total, int q0, int nz) {
    const int lane = (int)(threadIdx.x & 63);
    const int slot = (int)(blockIdx.x % NSLOT);
    unsigned long long* mine = buf + (size_t)par * nz * nq_total * NSLOT;
    unsigned long long todo = __ballot(zone >= 0);
    while (todo) {
        int first = __ffsll((long long)todo) - 1;
        int z = __shfl(zone, first);
        bool in = (zone == z);
        double xin[NQ];
        for (int q = 0; q < NQ; ++q) xin[q] = in ? v[q] : (IS_MAX ? 0.0 : LARGE);
        int q;
        bool holder;
        const double x = wave_reduce_scatter<NQ, IS_MAX>(xin, lane, q, holder);
        if (holder) {
            unsigned long long bits = (unsigned long long)__double_as_longlong(x);
            unsigned long long* a = mine + ((size_t)z * nq_total + q0 + q) * NSLOT + slot;
            if (IS_MAX) atomicMax(a, bits); else atomicMin(a, bits);
        }
        todo &= ~__ballot(in);
    }
    // re-arm by the calling wave's lanes in every block (a block may hold
    // waves that reduce different quantities), grid-stride over the entries
    unsigned long long* other = buf + (size_t)(1 - par) * nz * nq_total * NSLOT;
    const unsigned long long ident = IS_MAX ? 0ull : (unsigned long long)__double_as_longlong(LARGE);
    size_t n = (size_t)nz * NQ * NSLOT;
    for (size_t t = (size_t)blockIdx.x * 64 + lane; t < n; t += (size_t)gridDim.x * 64) {
        size_t z = t / (NQ * NSLOT), r = t % (NQ * NSLOT);
        other[(z * nq_total + q0) * NSLOT + r] = ident;
    }
}

// Consumer side for several zones: one block per zone reduces the NSLOT slots
// of quantities q0..q0+nq-1 into zred[zone][nq_total] (read by the consumers).
template <bool IS_MAX>
__global__ void k_zone_extract(const unsigned long long* buf, int par, int nq_total, int q0, int nq, int nz,
                               double* zred) {
    const int z = blockIdx.x, lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const unsigned long long* mine = buf + (size_t)par * nz * nq_total * NSLOT;
    for (int q = w; q < nq; q += (int)(blockDim.x >> 6)) {
        double x = __longlong_as_double((long long)mine[((size_t)z * nq_total + q0 + q) * NSLOT + lane]);
        for (int o = 32; o > 0; o >>= 1) {
            double y = __shfl_xor(x, o);
            x = IS_MAX ? ((x > y) ? x : y) : ((x < y) ? x : y);
        }
        if (lane == 0) zred[(size_t)z * nq_total + q0 + q] = x;
    }
}

int alloc_bytes(ofx_ctx* c, void** p, size_t bytes);
int upload(ofx_ctx* c, void* dst, const void* src, size_t bytes);

// Kernel launch wrappers (defined in the .hip files)
int mesh_build(ofx_ctx* c, const double* x, const double* y, const double* z, const int* fnc,
               const int* fnn, const int* lc, const int* rc, const int* bct);
int ns_boundary(ofx_ctx* c);
int sa_boundary(ofx_ctx* c);
int ns_timestep(ofx_ctx* c, int iter);
int ns_residual(ofx_ctx* c, bool reuse_ext = false);
int ns_residual_parts(ofx_ctx* c, unsigned parts, bool reuse_ext = true);
int step_launch(ofx_ctx* c, int iter);
int sa_residual(ofx_ctx* c);
int sa_residual_parts(ofx_ctx* c, unsigned parts);  // bits: gradient, face terms, cell terms
int ns_lusgs(ofx_ctx* c, int lo_first = 0);
int fusable_l0(const ofx_ctx* c);
int sa_lusgs(ofx_ctx* c, int lo_first = 0);
int ns_update(ofx_ctx* c);
int sa_update(ofx_ctx* c);
int norms(ofx_ctx* c, ofx_norms* out);
int nan_stop(ofx_ctx* c);
int wall_distance(ofx_ctx* c);
inline int ensure_dist(ofx_ctx* c) { return c->dist_valid ? 0 : wall_distance(c); }
int norms_range(ofx_ctx* c, int c0, int c1, ofx_norms* out);
int aero_zone(ofx_ctx* c, int z, double* out4);
int aero_row_zone(ofx_ctx* c, int z, double* out12);
int wall_data(ofx_ctx* c, double* host_out);
int fill_ghost_grad(ofx_ctx* c);
int materialize_timestep(ofx_ctx* c);
int ns_update_fused(ofx_ctx* c);
int sa_update_fused(ofx_ctx* c);
int lhs(ofx_ctx* c, int solver, double lhscoef);

// The fields the update kernels read: under the multi-stage scheme the
// increment is the residual itself (UNsField::Init and UTurbField::Init bind
// dq to res when time_integral is MULTI_STAGE, codes/uns/src/UNsCom.cpp:85-88,
// codes/uturb/src/UTurbCom.cpp:79-82).
inline DevFields update_fields(const ofx_ctx* c) {
    DevFields d = c->d;
    if (c->cfg.time_integral == 1) { d.dq = d.res; d.turbdq = d.turbres; }
    return d;
}
int aero(ofx_ctx* c, double* out4);
int init_freestream(ofx_ctx* c);

}  // namespace ofx
