/*
 * ofx.h -- C ABI of the MI355X unstructured finite-volume hot path.
 *
 * The library (oneflow_amd/libofx.so) evaluates OneFLOW's unstructured-grid
 * compressible Navier-Stokes + Spalart-Allmaras residual and its implicit
 * LU-SGS update on one AMD Instinct MI355X (gfx950) with hand-written HIP
 * kernels.  Every entry point below replaces one CPU task of the reference
 * solver; the citation on each names the reference function it stands in for
 * (paths relative to the OneFLOW source tree).
 *
 * Conventions (all entry points):
 *   - return 0 on success, a negative OFX_E* code on error; the message of
 *     the last error on this host thread is returned by ofx_last_error();
 *   - nothing throws across the boundary;
 *   - device buffers are allocated and owned by the context; the host passes
 *     plain host pointers for upload / download;
 *   - one context = one HIP stream; calls on one context are serialised.
 *
 * Field layout (host and device): structure-of-arrays, one contiguous array
 * per equation, cells 0..nCells-1 followed by one ghost cell per boundary
 * face (ghost of boundary face f is cell nCells+f), faces ordered
 * boundary-first exactly as in the grid file (codes/geometry/src/UnsGrid.cpp:123-129).
 * Primitive variables are rho,u,v,w,p (codes/ns/include/NsIdx.h:33-44).
 */
#ifndef OFX_H
#define OFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------- */
#define OFX_OK            0
#define OFX_EINVAL       -1   /* bad argument / unsupported option        */
#define OFX_ESTATE       -2   /* call out of order (e.g. no mesh yet)     */
#define OFX_EHIP         -3   /* HIP runtime error                        */
#define OFX_ENOMEM       -4   /* device allocation failed                 */
#define OFX_ENAN         -5   /* NaN in a residual (reference: Stop())    */

/* ---- case configuration --------------------------------------------- */
/* The subset of the reference's script keys read on this path
 * (test/<case>/script/{cfd,inflow,turb,vismodel}.txt; readers
 * codes/ns/src/NsCom.cpp:44-122, codes/turb/src/TurbCom.cpp:43-380,
 * codes/global/src/Interation.cpp:52-64).  Plain old data, no pointers. */
typedef struct ofx_config {
    int32_t dim;               /* 2 or 3                                      */
    int32_t ischeme;           /* 1 = Roe, 12 = SLAU2                         */
    int32_t ilim;              /* -1 zero order, 0 none, 1 Barth, 2 Venkat    */
    int32_t ieigenfix;         /* Roe entropy fix on/off                      */
    int32_t ivischeme;         /* NS face gradient: 0 STD, 1 AVER, 2 TEST     */
    int32_t vismodel;          /* 0 Euler, 1 laminar, 3 SA                    */
    int32_t timestepModel;     /* 0 local time step                           */
    int32_t visSRModel;        /* viscous spectral radius model (1,2,3)       */
    int32_t visTimeStepModel;  /* >0: viscous time-step limit                 */
    int32_t ncfl;              /* CFL ramp length                             */
    int32_t isowallbc;         /* 0 adiabatic wall, 1 isothermal              */
    int32_t iturb_visflux;     /* SA face gradient method (0,1,2)             */
    int32_t iprod_sa;          /* SA production form                          */
    int32_t turb_ilim;         /* SA reconstruction (only -1 supported)       */
    int32_t tns_ilim;          /* NS state used by SA convection (only -1)    */
    int32_t flux_fp32;         /* 1: FP32 face fluxes, FP64 accumulation      */
    int32_t time_integral;     /* 1 multi-stage (Runge-Kutta), 2 LU-SGS       */
    int32_t rk_stage;          /* stages of the multi-stage scheme (>= 1)     */
    double  vencat_coef;
    double  centropy1, centropy2;
    double  cflst, cfled, max_time_ratio;
    double  mach_ref, aoa_degree, sideslip_degree, reynolds;
    double  gama_ref, prl, prt, tref_dim, twall_dim;
    double  ref_sa, max_vis_ratio, rprod, turb_cfl_ratio;
    double  aref;              /* reference area of the force coefficients    */
} ofx_config;

/* Field identifiers for ofx_get_field / ofx_set_field.  Sizes in doubles. */
enum ofx_field {
    OFX_F_Q = 0,        /* primitive state        [5][nTCell]               */
    OFX_F_TURBQ,        /* SA nu_tilde            [nTCell]                  */
    OFX_F_RES,          /* NS residual            [5][nTCell]               */
    OFX_F_TURBRES,      /* SA residual            [nTCell]                  */
    OFX_F_DQDX,         /* primitive gradients    [5][nTCell] x3 (x,y,z)   */
    OFX_F_DTDX,         /* temperature gradient   [nTCell] x3               */
    OFX_F_LIMITER,      /* NS limiter             [5][nTCell]               */
    OFX_F_TIMESTEP,     /* local time step        [nTCell]                  */
    OFX_F_INVSR,        /* inviscid spectral rad. [nTCell]                  */
    OFX_F_VISSR,        /* viscous spectral rad.  [nTCell]                  */
    OFX_F_DQ,           /* NS LU-SGS increment    [5][nTCell]               */
    OFX_F_TURBDQ,       /* SA LU-SGS increment    [nTCell]                  */
    OFX_F_TEMPR,        /* temperature            [nTCell]                  */
    OFX_F_VISL,         /* laminar viscosity      [nTCell]                  */
    OFX_F_VIST,         /* eddy viscosity         [nTCell]                  */
    OFX_F_DIST,         /* wall distance          [nCells]                  */
    OFX_F_VOL,          /* cell volume            [nTCell]                  */
    OFX_F_XCC,          /* cell centre            [nTCell] x3               */
    OFX_F_FACE_NORMAL,  /* unit normal + area     [F] x4 (nx,ny,nz,area)    */
    OFX_F_FACE_CENTER,  /* face centre            [F] x3                    */
    OFX_F_TURBDQDX,     /* SA gradient            [nTCell] x3               */
    OFX_F_TURBIMPSR,    /* SA implicit diagonal   [nTCell]                  */
    OFX_F_IMPSR,        /* NS implicit diagonal   [nTCell]                  */
    OFX_F_COUNT
};

/* Residual norms of one pseudo-time step, as written to res.dat/turbres.dat
 * (codes/residual/src/Residual.cpp:52-72): sqrt(sum_c res^2 / nCells). */
typedef struct ofx_norms {
    double ns[5];
    double sa;
} ofx_norms;

typedef struct ofx_ctx ofx_ctx;

/* Last error message of the calling thread ("" if none). */
const char* ofx_last_error(void);

/* Create / destroy a context bound to one HIP device and one stream. */
int ofx_ctx_create(ofx_ctx** out, int device);
int ofx_ctx_destroy(ofx_ctx* ctx);

/* Upload one zone's grid exactly as stored in a .ofl file and build the
 * geometry on the device: face normals/areas/centres and cell centres and
 * volumes (codes/geometry/src/UnsGrid.cpp:469-760, FaceMesh.cpp:112-287),
 * ghost-cell mirrors, the cell->face lists (codes/geometry/src/CellTopo.cpp:95-119),
 * the LU-SGS level schedule and the wall distance
 * (codes/geometry/src/WallDist.cpp:136-358).  rc[f] < 0 marks a boundary
 * face; boundary faces must come first.  faceNodeCount[nFaces],
 * faceNodes[sum(faceNodeCount)], bcType[nBFaces] (codes/geometry/src/Boundary.cpp:35-46). */
int ofx_mesh_upload(ofx_ctx* ctx, int dim, int nNodes,
                    const double* x, const double* y, const double* z,
                    int nFaces, int nCells, int nBFaces,
                    const int* faceNodeCount, const int* faceNodes,
                    const int* lc, const int* rc, const int* bcType);

/* Storage order of the device arrays (no counterpart in the reference; a
 * layout choice, results do not depend on it): 0 (default) the grid's own
 * cell and face order, 1 spatial -- the cells of each connected part of the
 * grid in Morton order of their centroids and the interior faces sorted by
 * their cells' positions, so that neighbours share cache lines whatever the
 * numbering.  On a grid numbered by colour (for the LU-SGS sweeps) the
 * spatial order scatters each sweep level over the whole array, so the
 * grid's own order is the default.  Call before
 * ofx_mesh_upload.  Every entry point keeps the grid's numbering: fields are
 * read and written in grid order, each cell sums its faces in ascending face
 * id and the LU-SGS sweeps split neighbours by cell id. */
int ofx_set_storage_order(ofx_ctx* ctx, int mode);

/* Mesh sizes: out[0..4] = nCells, nFaces, nBFaces, nTCell, nLevels (LU-SGS). */
int ofx_mesh_info(ofx_ctx* ctx, int* out5);

/* Case constants (codes/ns/src/NsCom.cpp:44-122, codes/chemical/src/Chemical.cpp:280-338,
 * codes/turb/src/TurbCom.cpp:110-380). */
int ofx_config_set(ofx_ctx* ctx, const ofx_config* cfg);

/* Freestream start: INIT_FLOWFIELD + INIT_FINAL of the NS and SA solvers
 * (codes/uns/src/UNsRestart.cpp:48-62, codes/ns/src/NsSolverImp.cpp:58-77,
 *  codes/uturb/src/UTurbRestart.cpp:48-60, codes/turb/src/TurbSolverImp.cpp:49-54). */
int ofx_init_freestream(ofx_ctx* ctx);

/* Host <-> device copies of any field (enum ofx_field, sizes above). */
int ofx_get_field(ofx_ctx* ctx, int field, double* host);
int ofx_set_field(ofx_ctx* ctx, int field, const double* host);

/* NsCalcBoundary: gamma/T, Sutherland viscosity, ghost-cell BCs
 * (codes/ns/src/NsSolverImp.cpp:72-79, codes/uns/src/UNsBcSolver.cpp:55-168). */
int ofx_ns_boundary(ofx_ctx* ctx);
/* TurbCalcBoundary: eddy viscosity + SA ghost cells
 * (codes/turb/src/TurbSolverImp.cpp:55-60). */
int ofx_sa_boundary(ofx_ctx* ctx);

/* CALC_TIME_STEP: UTimeStep::CalcTimeStep at pseudo-step `iter`
 * (codes/uns/src/UTimeStep.cpp:190-416, codes/ns/src/TimeStep.cpp:43-133). */
int ofx_timestep(ofx_ctx* ctx, int iter);

/* UPDATE_RESIDUALS of the NS solver: NsCalcRHS = UNsInvFlux::CalcFlux +
 * UNsVisFlux::CalcFlux with the ordered, atomic-free residual gather
 * (codes/ns/src/NsRhs.cpp:106-128, codes/uns/src/UNsInvFlux.cpp:123-192,
 *  codes/uns/src/UNsVisFlux.cpp:90-127, codes/usolver/src/UCom.cpp:144-171).
 * res is overwritten (ZERO_RESIDUALS + LOAD_RESIDUALS happen inside). */
int ofx_residual_ns(ofx_ctx* ctx);

/* UPDATE_RESIDUALS of the SA solver: TurbCalcRHS
 * (codes/turb/src/TurbRhs.cpp:57-71). */
int ofx_residual_sa(ofx_ctx* ctx);

/* INIT_LUSGS + LUSGS_LOWER_SWEEP + LUSGS_UPPER_SWEEP, level scheduled
 * (codes/uns/src/UNsLusgs.cpp:75-138, codes/ns/src/NsLusgs.cpp:147-330;
 *  codes/uturb/src/UTurbLusgs.cpp:90-304, codes/turb/src/TurbLusgs.cpp:123-254). */
int ofx_lusgs_ns(ofx_ctx* ctx);
int ofx_lusgs_sa(ofx_ctx* ctx);

/* UPDATE_FLOWFIELD_LUSGS (codes/uns/src/UNsUpdate.cpp:46-146,
 * codes/uturb/src/UTurbUpdate.cpp:85-120), with CALC_BOUNDARY of the same
 * solver folded in.  With time_integral = 1 (multi-stage) this is
 * UPDATE_FLOWFIELD: the increment is the residual field itself, as the
 * reference binds dq to res under MULTI_STAGE (codes/uns/src/UNsCom.cpp:85-88,
 * codes/uturb/src/UTurbCom.cpp:79-82), i.e. res / turbres after ofx_lhs. */
int ofx_update_ns(ofx_ctx* ctx);
int ofx_update_sa(ofx_ctx* ctx);

/* CALC_LHS of the multi-stage scheme: ULhs::CalcLHS
 * (codes/usolver/src/ULhs.cpp:47-62, registered at
 * codes/lhs/src/LhsTaskReg.cpp:38-44).  For every inner cell c and equation
 * e of the solver (0 NS: res[5], 1 SA: turbres[1]):
 * r[e][c] *= (dt[c] / vol[c]) * lhscoef, dt the time step of the last
 * ofx_timestep (after ModifyTimeStep).  Ghost entries are left alone. */
int ofx_lhs(ofx_ctx* ctx, int solver, double lhscoef);

/* Residual norms of the current res / turbres fields (DUMP_RESIDUAL,
 * codes/residual/src/ResidualTask.cpp:75-126). */
int ofx_norms_get(ofx_ctx* ctx, ofx_norms* out);

/* n whole pseudo-time steps starting at step `first_iter` (1-based outer
 * step, drives the CFL ramp), the task sequence of TimeIntegral::Lusgs for
 * NS then SA (codes/solver/src/TimeIntegral.cpp:113-130,
 * codes/multigrid/src/Multigrid.cpp:339-375).  With time_integral = 1 each
 * solver instead runs TimeIntegral::RungeKutta
 * (codes/solver/src/TimeIntegral.cpp:81-110): CALC_TIME_STEP, then per stage
 * s = 0..rk_stage-1 with lhscoef = 1/(s+1) (codes/global/src/Ctrl.cpp:74-79)
 * UPDATE_RESIDUALS, CALC_LHS, UPDATE_FLOWFIELD and CALC_BOUNDARY (LOAD_Q and
 * LOAD_RESIDUALS only move multigrid storage and leave a one-level run's
 * fields as they are).  norms_out (may be NULL) receives n entries, the norms
 * of res / turbres as the step leaves them (DUMP_RESIDUAL runs after it). */
int ofx_iterate(ofx_ctx* ctx, int first_iter, int n, ofx_norms* norms_out);

/* Wall force coefficients of the current state: out[0..3] = CL, CD, CDp, CDf
 * (codes/force/src/AeroForceTask.cpp:186-380). */
int ofx_aero(ofx_ctx* ctx, double* out4);

/* The value columns of one aero.dat row (AerodynamicForceTask::Dump,
 * codes/force/src/AeroForceTask.cpp:96-150, with AeroForceInfo::CalcCoef,
 * codes/force/src/Force.cpp:259-276): out[0..11] = CL, CD, CD_PR, CD_SF,
 * CDL2 (= CD - CL^2 / (pi 9.5)), Xcp, Fx, Fy, Fz, Cmx, Cmy, Cmz, where
 * F = total wall force / aref.  The reference's AeroForce::AddForce
 * (codes/force/src/Force.cpp:202-208) sums no moments, so its moment
 * coefficients are zero and Xcp is +-0 by the sign of Fy; these columns
 * reproduce that. */
int ofx_aero_row(ofx_ctx* ctx, double* out12);

/* DUMP_PRESSURE_COEFF (the idump_pres rows of CalcAeroForce,
 * codes/force/src/AeroForceTask.cpp:242-262,426-440, written to wallaeroFile):
 * the number of wall faces, and per wall face in face order x, y, z of the
 * face centre, -cp and cf -> out[5 k .. 5 k + 4]. */
int ofx_wall_count(ofx_ctx* ctx, int* n);
int ofx_wall_data(ofx_ctx* ctx, double* out);

/* Kernel timing (no counterpart in the reference; measurement only):
 * launches one part of the residual `iters` times back to back on the
 * context's stream between two HIP events, after one warm-up launch, and
 * returns the average ms per launch.  which: 0 NS gradient and limiter (one
 * kernel), 1 the grid extrema of the Venkatakrishnan epsilon (a pass of its
 * own only when no time step has reduced them on this state), 2 NS face
 * fluxes (MUSCL + Roe/SLAU2 + viscous), 3 NS ordered gather, 4 whole NS
 * residual, 5 whole SA residual, 6 SA gradient (+ ghost values), 7 SA face
 * terms, 8 SA cell terms (source + ordered gather), 9 NS + SA residual (one
 * cell-residual evaluation), 10 one whole pseudo-time step past the CFL ramp
 * (as ofx_iterate; this one advances the state), 11 the NS time step past
 * the CFL ramp, 12 the NS LU-SGS sweeps (lower and upper, every level) on the
 * current time step and residual.  Parts 0-9 are idempotent on the current
 * state, so the fields are left as one ofx_residual_ns / ofx_residual_sa
 * call leaves them; 11 and 12 rewrite the time step and dq of the state. */
int ofx_time_kernel(ofx_ctx* ctx, int which, int iters, double* ms_out);

/* Bandwidth reference (no counterpart in the reference): average duration
 * (ms) of a plain device-to-device copy kernel moving `bytes` bytes
 * (a multiple of 16), over `iters` launches after one warm-up launch.
 * Allocates and frees its own two buffers. */
int ofx_copy_bench(ofx_ctx* ctx, size_t bytes, int iters, double* ms_out);

/* Several independent cases in one context (no counterpart in the
 * reference, which solves zones one after another, codes/task/src/SimpleTask.cpp):
 * upload the cases' grids concatenated zone after zone -- cells, then all
 * zones' boundary faces, then all zones' interior faces, each part in zone
 * order with the zone's own order kept -- with ofx_mesh_upload, set the shared
 * options with ofx_config_set, then declare the zones: cellStart[n + 1] and
 * bfaceStart[n + 1] (zone z owns cells [cellStart[z], cellStart[z+1]) and
 * boundary faces [bfaceStart[z], bfaceStart[z+1])) and one configuration
 * per zone, which may differ from the shared one only in mach_ref,
 * aoa_degree, sideslip_degree and reynolds.  Every kernel then runs over all
 * zones at once with per-zone freestream constants, per-zone time-step and
 * limiter reductions and per-zone wall distance; each zone's results are
 * bit-identical to solving it alone.  Call again after ofx_config_set. */
int ofx_zones_set(ofx_ctx* ctx, int n, const int* cellStart, const int* bfaceStart, const ofx_config* cfgs);
/* Residual norms of each zone's current residual (out[n]). */
int ofx_zones_norms(ofx_ctx* ctx, ofx_norms* out);
/* Force coefficients of each zone (out[4 n]: CL, CD, CDp, CDf per zone). */
int ofx_zones_aero(ofx_ctx* ctx, double* out);

/* Face-kernel tile, the threads per block of the NS and SA face-flux kernels:
 * 64, 128, 256 (default) or 512.  Results do not depend on it (no
 * counterpart in the reference; a launch parameter). */
int ofx_set_face_tile(ofx_ctx* ctx, int ns_tile, int sa_tile);

/* Kernel variant (no counterpart in the reference; results do not depend on
 * it).  Kernels 0 and 1 both name the NS gradient kernel (it evaluates the
 * limiter in the same pass), kernel 5 the SA gradient: variant 0 (default)
 * reads the face terms (weights, normal, face centre offsets) by face id,
 * variant 1 reads the per-(cell, face) copies stored cell-major at mesh
 * upload (64 bytes per cell face), contiguous per wave.  kernel 6, the NS
 * time step and LU-SGS sweeps: variant 0 (default) reads the faces' normals
 * and areas by face id and evaluates the viscous radius's centre distance,
 * variant 1 reads per-(cell, face) copies of the normal, area, distance and
 * neighbour volume (48 bytes per cell face).  kernel 2, the residual
 * gather: variant 0 one lane per face of a cell, variant 1 (default) one
 * thread per cell.  kernel 3, the viscous FP64 NS face kernel: variant 0
 * (default) one thread per face, variant 1 two waves per 64 faces (inviscid
 * / viscous part).  kernel 4, the SA face kernel with the TEST diffusion
 * (iturb_visflux 2): variant 0 (default) reads the face's stored geometry
 * record, variant 1 evaluates it from the cell and face centres. */
int ofx_set_kernel_variant(ofx_ctx* ctx, int kernel, int variant);
int ofx_get_kernel_variant(ofx_ctx* ctx, int kernel, int* variant);

/* Autotuner (no counterpart in the reference): times the NS face kernel for
 * each face tile and split, the gather, both gradients and (LU-SGS runs)
 * the NS time step with the sweeps for each variant, and the SA residual for
 * each SA face tile and geometry variant on the current state (each the
 * median of five timings of `iters` launches), keeps the fastest, and
 * reports the face tiles in chosen2[0..1] (may be NULL).
 * Leaves res / turbres as one evaluation of the current state. */
int ofx_autotune(ofx_ctx* ctx, int iters, int* chosen2);

/* Batched cases (the reference solves zones one after another,
 * codes/task/src/SimpleTask.cpp; here many independent cases advance
 * together).  A batch references existing contexts (mesh, config and state
 * set, all on one device; the batch does not own them).  Pseudo-time steps
 * past every case's CFL ramp are recorded once into one HIP graph holding
 * steps_per_graph steps (rounded up to even) of every case, one independent
 * branch per case, and replayed; steps inside the ramp are launched directly.
 * Per case, the results are bit-identical to ofx_iterate on that context. */
typedef struct ofx_batch ofx_batch;
int ofx_batch_create(ofx_batch** out, ofx_ctx* const* ctxs, int n, int steps_per_graph);
int ofx_batch_destroy(ofx_batch* b);
/* Advance every case n steps from step first_iter (TimeIntegral::Lusgs for NS
 * then SA, as ofx_iterate); graph_launches (may be NULL) receives the number
 * of graph replays used. */
int ofx_batch_iterate(ofx_batch* b, int first_iter, int n, int* graph_launches);
/* Residual norms of each case's current residual, out[n]. */
int ofx_batch_norms(ofx_batch* b, ofx_norms* out);
/* ofx_autotune on every case; chosen (may be NULL) receives 2 tiles per case
 * (a recorded graph is recorded again at the next ofx_batch_iterate). */
int ofx_batch_autotune(ofx_batch* b, int iters, int* chosen);

/* ---- warm-start predictor (no counterpart in the reference) ----------
 * A BF16 multilayer perceptron, 9 inputs -> 5 hidden layers of 256 (SiLU)
 * -> 5 outputs, evaluated and trained by the library's own MFMA kernels
 * (FP32 accumulation, FP32 Adam master weights).  Inputs per sample:
 * [x, y, wall-distance feature, m, p, t, log10 Re, Mach, AoA] (scaled by the
 * caller); outputs: the normalised rho, u, v, p, nu_tilde targets the caller
 * trained on.  The parameter vector is, per layer, W[kin][kout] then b[kout]
 * with padded shapes kin = 16, 256, 256, 256, 256, 256 and
 * kout = 256, 256, 256, 256, 256, 64 (padding entries are zero). */
typedef struct ofx_mlp ofx_mlp;
int ofx_mlp_create(ofx_mlp** out, int device, unsigned long long seed);
int ofx_mlp_destroy(ofx_mlp* net);
int ofx_mlp_param_count(ofx_mlp* net, int* count);
int ofx_mlp_get_params(ofx_mlp* net, float* host);
int ofx_mlp_set_params(ofx_mlp* net, const float* host);
/* y[n][5] = net(x[n][9]), n <= 2^20 per call. */
int ofx_mlp_forward(ofx_mlp* net, const float* x, int n, float* y);
/* How ofx_mlp_forward and ofx_mlp_forward_noalloc evaluate the network (no
 * counterpart in the reference; results do not depend on it): variant 1
 * (default) all layers in one kernel, 128 samples per workgroup with the
 * hidden activations kept in LDS; variant 0 one GEMM kernel per layer, the
 * activations through device memory (the training pass's form). */
int ofx_mlp_set_forward_variant(ofx_mlp* net, int variant);
/* ofx_mlp_forward without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops).  x_dev [n][9] and y_dev [n][5] are DEVICE pointers
 * owned by the caller; the work is launched on `stream` (a hipStream_t; NULL:
 * the network's own stream) with no host synchronisation, copy, memset or
 * allocation.  n <= 8192: the work arrays are one fixed 10.25 MiB device
 * array per GPU, held until the process ends and shared by every call on
 * that GPU, so calls that overlap on one GPU race.  Larger n is rejected
 * (OFX_EINVAL).  Results equal ofx_mlp_forward's bit for bit. */
int ofx_mlp_forward_noalloc(ofx_mlp* net, const float* x_dev, int n, float* y_dev, void* stream);
/* Mean-squared-error loss over the 5 outputs and its gradient with respect to
 * every parameter (same layout as the parameter vector); no update. */
int ofx_mlp_gradients(ofx_mlp* net, const float* x, const float* t, int n, float* loss, float* grads);
/* `steps` Adam steps (learning rate lr) on mini-batches of `batch` samples
 * drawn without replacement from (x, t), reshuffled each epoch (seeded);
 * loss_out (may be NULL) receives each step's mini-batch loss. */
int ofx_mlp_train(ofx_mlp* net, const float* x, const float* t, int n, int batch, int steps, float lr,
                  unsigned long long seed, float* loss_out);

/* Synchronise the context's stream. */
int ofx_sync(ofx_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* OFX_H */
