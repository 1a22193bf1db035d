/*
 * oracle.h -- CPU oracle of the unstructured NS + SA hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a single-threaded C++ restatement of the
 * reference solver's CPU loops (OneFLOW codes/uns, codes/ns, codes/uturb,
 * codes/turb, codes/usolver, codes/grad), kept in the reference's face order,
 * cell order and accumulation order and compiled with -ffp-contract=off.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only to check results or as the timed CPU baseline.  The product
 * (oneflow_amd/libofx.so) never calls it.
 *
 * The entry points mirror include/ofx.h one to one with an ofo_ prefix.
 */
#ifndef OFX_ORACLE_H
#define OFX_ORACLE_H

#include "../include/ofx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ofo_case ofo_case;

const char* ofo_last_error(void);
ofo_case* ofo_create(int dim, int nNodes, const double* x, const double* y, const double* z,
                     int nFaces, int nCells, int nBFaces,
                     const int* faceNodeCount, const int* faceNodes,
                     const int* lc, const int* rc, const int* bcType);
void ofo_destroy(ofo_case* c);
int  ofo_mesh_info(ofo_case* c, int* out5);
int  ofo_config_set(ofo_case* c, const ofx_config* cfg);
int  ofo_init_freestream(ofo_case* c);
int  ofo_get_field(ofo_case* c, int field, double* host);
int  ofo_set_field(ofo_case* c, int field, const double* host);
int  ofo_ns_boundary(ofo_case* c);
int  ofo_sa_boundary(ofo_case* c);
int  ofo_timestep(ofo_case* c, int iter);
int  ofo_residual_ns(ofo_case* c);
int  ofo_residual_sa(ofo_case* c);
int  ofo_lusgs_ns(ofo_case* c);
int  ofo_lusgs_sa(ofo_case* c);
int  ofo_update_ns(ofo_case* c);
int  ofo_update_sa(ofo_case* c);
int  ofo_lhs(ofo_case* c, int solver, double lhscoef);
int  ofo_norms_get(ofo_case* c, ofx_norms* out);
int  ofo_iterate(ofo_case* c, int first_iter, int n, ofx_norms* norms_out);
int  ofo_aero(ofo_case* c, double* out4);
int  ofo_aero_row(ofo_case* c, double* out12);
/* Wall faces (face order): count, and x, y, z of the face centre, -cp, cf
 * per face -> out[5 * k .. 5 * k + 4]. */
int  ofo_wall_count(ofo_case* c);
int  ofo_wall_data(ofo_case* c, double* out);
/* Face flux arrays of the last ofo_residual_ns: inv[5][F], vis[5][F]. */
int  ofo_face_flux(ofo_case* c, double* inv, double* vis);
/* The shared x^y routine (ofx_math.h) for n pairs: out[i] = x[i]^y[i]. */
int  ofo_pow(const double* x, const double* y, int n, double* out);

#ifdef __cplusplus
}
#endif

#endif
