/* Mixed-precision variant of the split-explicit large step.
 *
 * libmpas_dycore_hip_mixed.so is built from the same kernel sources as
 * libmpas_dycore_hip.so with two changes:
 *
 * - storage: the edge- and vertex-located diagnostics, tendencies and
 *   fluxes are binary32 (v, pv_edge, rho_edge (h_edge) and vorticity in the
 *   state; the edge tendencies, cqu, the theta and tracer transport fluxes
 *   and the edge/vertex scratch in the workspace).  The normal velocity u
 *   and the mass fluxes ru, ru_p, ruAvg stay binary64: their divergence and
 *   curl nearly cancel, and binary32 rounding of the terms would carry into
 *   w, the divergence and the vorticity above 1e-5;
 * - arithmetic: the PV flux and kinetic-energy gradient of the u tendency,
 *   the theta advection fluxes and their divergence, the kinetic energy,
 *   tangential velocity and APVM potential vorticity of the diagnostic
 *   solve, and the tracer transport fluxes and their divergence compute in
 *   binary32; their slot sums carry a Neumaier compensation term (TwoSum).
 *
 * Binary64 throughout: the vertical-implicit coefficients and the column
 * tridiagonal solve, rho, theta_m, rho_theta', the exner/pressure recovery
 * from rho_theta, w, rw, the scalars, the vertical fluxes and the hydrostatic
 * base state in the mesh.
 *
 * The reference has no such variant (it runs float64 on the CPU only); this
 * is the reduced-precision form of mpas_hip_srk3_step
 * (include/mpas_hip.h).  The library also exports mpas_hip_tuning_set/get,
 * mpas_hip_tuning_set_order/get_order, mpas_hip_timer_arm/read and
 * mpas_hip_last_error with the meanings given in
 * mpas_hip.h, acting on its own kernels. */
#ifndef MPAS_HIP_MIXED_H
#define MPAS_HIP_MIXED_H

#include "mpas_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* mpas_hip_state with v, vorticity, pv_edge and rho_edge in binary32
 * (same field order). */
typedef struct mpas_hip_mixed_state {
    double *u;
    double *w, *theta_m, *rho_zz, *scalars, *rt_diabatic_tend;
    double *rho_p, *rtheta_p, *exner, *pressure_p;
    double *ru;
    double *rw;
    float *v;
    double *divergence;
    float *vorticity;
    double *ke;
    float *pv_edge, *rho_edge;
    double *ur_cell, *vr_cell;
    double *uReconstructX, *uReconstructY, *uReconstructZ;
    double *uReconstructZonal, *uReconstructMeridional;
    double *ruAvg;
    double *wwAvg;
} mpas_hip_mixed_state;

/* Doubles of workspace for one mixed-precision large step. */
int64_t mpas_hip_mixed_workspace_doubles(int32_t nCells, int32_t nEdges, int32_t nVertices,
                                         int32_t nLevels, int32_t nScalars);

/* One mixed-precision large step in place (the orchestration of
 * mpas_hip_srk3_step, time_integration.py:4897-4956). */
int mpas_hip_mixed_srk3_step(const mpas_hip_mesh *mesh, mpas_hip_mixed_state *state,
                             const mpas_hip_workspace *work, const mpas_hip_config *cfg, void *stream);

/* sizeof(mpas_hip_mixed_state), for binding checks. */
int32_t mpas_hip_mixed_state_size(void);

#ifdef __cplusplus
}
#endif

#endif /* MPAS_HIP_MIXED_H */
