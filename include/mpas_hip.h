/*
 * C ABI of the MI355X (gfx950) split-explicit dynamics step.
 *
 * Every pointer is a device pointer owned by the caller (PyTorch allocates
 * them); no entry point allocates or frees device memory, every entry point
 * runs asynchronously on the stream it is given and returns 0 on success or a
 * nonzero status with text available from mpas_hip_last_error().
 *
 * Layout matches the reference runtime tensors (mpas_pytorch docs/API.md
 * "State and controls"): row-major (rows, levels) with levels innermost and a
 * trailing padding row per entity.  Connectivity is 0-based int32 with -1 for
 * "no neighbour" (the host converts the reference's 1-based values).
 */
#ifndef MPAS_HIP_H
#define MPAS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Static mesh, vertical metrics and base state (read only).  Field names
 * follow the reference mesh dict (mpas_pytorch/ic_loader.py:1243-1440). */
typedef struct mpas_hip_mesh {
    int32_t nCells, nEdges, nVertices, nLevels;
    int32_t maxEdges, maxEdges2, vertexDegree, maxAdv;
    int32_t flatTerrain; /* 1: zb_cell/zb3_cell/zxu are all zero */
    int32_t reserved0;
    const int32_t *cellsOnEdge, *verticesOnEdge, *nEdgesOnCell, *edgesOnCell;
    const int32_t *cellsOnCell, *verticesOnCell, *kiteForCell, *edgesOnVertex;
    const int32_t *nEdgesOnEdge, *edgesOnEdge, *nAdvCellsForEdge, *advCellsForEdge;
    const double *edgesOnCell_sign, *edgesOnVertex_sign, *weightsOnEdge, *kiteAreasOnVertex;
    const double *dcEdge, *dvEdge, *invDcEdge, *invDvEdge, *invAreaCell, *invAreaTriangle;
    const double *fEdge, *fVertex, *meshScalingDel2, *meshScalingDel4;
    const double *adv_coefs, *adv_coefs_3rd, *defc_a, *defc_b, *coeffs_reconstruct;
    const double *latCell, *lonCell;
    const double *zz, *zxu, *dss, *zb_cell, *zb3_cell;
    const double *rdzw, *rdzu, *fzm, *fzp, *rdnw, *fnm, *fnp;
    double cf1, cf2, cf3, reserved1;
    const double *rho_base, *rtheta_base, *exner_base;
    /* Coriolis factors of the curvature terms (dyn_tend.py:2494-2502, :4258-4265), read only
     * with curvature on: (2 omega cos angleEdge) cos latEdge per edge and 2 omega cos latCell
     * per cell, omega = 7.29212e-5 */
    const double *curvEdge, *curvCell;
    /* vertical eddy viscosities (dyn_tend.py:2731-2781, :4633-4647, :5060-5097), read only with
     * them on: interface heights zgrid (cells x interfaces); without config_mix_full the
     * reference profiles t_init (cells x levels), u_init and v_init (levels) and cos/sin of
     * angleEdge (edges) */
    const double *zgrid, *t_init, *u_init, *v_init, *cosAngleEdge, *sinAngleEdge;
} mpas_hip_mesh;

/* Prognostic state and the diagnostics carried between large steps
 * (atmosphere_model.py:157-201).  Updated in place by mpas_hip_srk3_step. */
typedef struct mpas_hip_state {
    double *u, *w, *theta_m, *rho_zz, *scalars, *rt_diabatic_tend;
    double *rho_p, *rtheta_p, *exner, *pressure_p, *ru, *rw;
    double *v, *divergence, *vorticity, *ke, *pv_edge, *rho_edge;
    double *ur_cell, *vr_cell;
    double *uReconstructX, *uReconstructY, *uReconstructZ;
    double *uReconstructZonal, *uReconstructMeridional;
    double *ruAvg, *wwAvg; /* split-averaged transport fluxes of the last step */
} mpas_hip_state;

/* Caller-owned scratch for one large step; sizes are given by
 * mpas_hip_workspace_doubles() for the mesh dimensions. */
typedef struct mpas_hip_workspace {
    double *base;        /* one contiguous block carved by the library */
    int64_t n_doubles;   /* its length */
} mpas_hip_workspace;

/* Step configuration (the reference config keys in brackets). */
typedef struct mpas_hip_config {
    double dt;           /* large time step */
    double len_disp;     /* [config_len_disp] */
    double epssm;        /* [config_epssm] 0.1 */
    double smdiv;        /* [config_smdiv] 0.1 */
    double apvm;         /* [config_apvm_upwinding] 0.5 */
    double coef3;        /* [config_coef_3rd_order] 0.25 */
    double c_s;          /* [config_smagorinsky_coef] 0.125 */
    double visc4_2dsmag; /* [config_visc4_2dsmag] 0.05 */
    double del4u_div_factor; /* [config_del4u_div_factor] 10 */
    int32_t ns;          /* [config_number_of_sub_steps] 2 */
    int32_t split;       /* [config_dynamics_split_steps] 3 */
    int32_t monotonic;   /* [config_monotonic] or [config_positive_definite]: FCT in the last transport stage */
    int32_t nScalars;
    int32_t moist_start, moist_end; /* 1-based inclusive species range */
    int32_t diagnostics_current; /* 1: skip the first-stage diagnostic refresh */
    int32_t order;       /* [config_time_integration_order] 2 or 3 (0 is read as 3) */
    int32_t rayleigh_levels;      /* [config_number_rayleigh_damp_u_levels] when [config_rayleigh_damp_u], else 0 */
    double rayleigh_coef_inverse; /* 1 / (rayleigh_levels * [config_rayleigh_damp_u_timescale_days] * 86400) */
    /* [config_horiz_mixing]: 0 "2d_smagorinsky", 1 "2d_fixed" (dyn_tend.py:3367-3430); the three
     * fixed coefficients are read only in "2d_fixed" */
    int32_t mix_fixed;
    double h_theta_eddy_visc2;    /* [config_h_theta_eddy_visc2]: kdiff of every cell and level */
    double h_mom_eddy_visc4;      /* [config_h_mom_eddy_visc4] */
    double h_theta_eddy_visc4;    /* [config_h_theta_eddy_visc4] */
    /* CAM damping (dyn_tend.py:3434-3458): on when [config_mpas_cam_coef] > 0 and
     * [config_number_cam_damping_levels] > 0, else cam_levels = 0 */
    int32_t cam_levels;           /* [config_number_cam_damping_levels] */
    double cam_visc2;             /* 4 * 2.0833 * len_disp * [config_mpas_cam_coef] */
    int32_t curvature;            /* [config_curvature_u]: curvature terms of u and w */
    double v_mom_eddy_visc2;      /* [config_v_mom_eddy_visc2]: vertical mixing of u and w */
    double v_theta_eddy_visc2;    /* [config_v_theta_eddy_visc2]: vertical mixing of theta */
    int32_t mix_full;             /* [config_mix_full] 1; 0 mixes departures from u_init/v_init/t_init */
    int32_t scalar_advection;     /* [config_scalar_advection] 1; 0 leaves the scalars untransported */
} mpas_hip_config;

/* Number of doubles of workspace needed for one large step. */
int64_t mpas_hip_workspace_doubles(int32_t nCells, int32_t nEdges, int32_t nVertices,
                                   int32_t nLevels, int32_t nScalars);

/* One split-explicit large step in place (three RK stages with the RK3 or
 * RK2 weights of compute_rk_weights, time_integration.py:2787-2856).
 * Replaces the reference's atm_srk3 (time_integration.py:4897-4956). */
int mpas_hip_srk3_step(const mpas_hip_mesh *mesh, mpas_hip_state *state,
                       const mpas_hip_workspace *work, const mpas_hip_config *cfg,
                       void *stream);

/* The large step of mpas_hip_srk3_step without a workspace argument and
 * without any allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The workspace is a device array inside
 * the library (mpas_hip_srk3_step_noalloc_capacity() doubles, 128 MiB, held
 * by each GPU the library runs on until the process ends); a mesh whose
 * workspace (mpas_hip_workspace_doubles) exceeds it is rejected with an
 * error.  The call clears the part it uses with a fill kernel first.  All
 * calls on one GPU share the array: calls that overlap on one GPU race. */
int mpas_hip_srk3_step_noalloc(const mpas_hip_mesh *mesh, mpas_hip_state *state,
                               const mpas_hip_config *cfg, void *stream);
int64_t mpas_hip_srk3_step_noalloc_capacity(void);

/* Sizes of the ABI structs, for binding checks. */
int mpas_hip_struct_sizes(int32_t *mesh, int32_t *state, int32_t *workspace, int32_t *config);

/* ---------------------------------------------------------------------
 * Per-kernel entry points.  Each replaces one reference kernel function
 * (cited below) with the same argument meaning; the parity tests and the
 * per-kernel timings use them.  Every array has the reference shape with
 * its padding row; "scratch" arguments are caller-owned device buffers of
 * at least the size returned by mpas_hip_kernel_scratch_doubles().
 * ------------------------------------------------------------------- */

/* Scalars of atm_compute_dyn_tend_work evaluated on the host
 * (dyn_tend.py:3367-3400, :9616-9626). */
typedef struct mpas_hip_dyn_params {
    int32_t rk_step;
    double dt;
    double coef_3rd_order;
    double kdiff_coef;   /* (c_s * len_disp)**2 */
    double kdiff_cap;    /* (0.01 * len_disp**2) * (1/dt) */
    double h_mom_eddy_visc4;
    double h_theta_eddy_visc4;
    double del4u_div_factor;
    double prandtl_inv;
    /* Rayleigh damping of u on the top levels (dyn_tend.py:2942-2996, coefficient :10066-10072):
     * tend_u -= rho_edge u ((k + 1) - k0) rayleigh_coef_inverse for k >= k0 = nLevels - rayleigh_levels */
    int32_t rayleigh_levels;
    double rayleigh_coef_inverse;
    /* kdiff of the first RK stage (dyn_tend.py:3367-3458): the Smagorinsky value
     * min(kdiff_coef def, kdiff_cap), or kdiff_fixed everywhere when mix_fixed; then on the
     * top cam_levels levels max(kdiff, cam_visc2 (1 - (nLevels - (k + 1)) / cam_levels)) */
    int32_t mix_fixed;
    double kdiff_fixed;
    int32_t cam_levels;
    double cam_visc2;
    /* curvature terms (dyn_tend.py:2471-2505, :4241-4267): tend_u -= curvEdge rho_edge w_avg +
     * u w_avg rho_edge / r_earth; the w advective sum gains rho (ur^2 + vr^2) / r_earth +
     * curvCell ur rho at interfaces (cell winds of the step's input state), r_earth = 6371229 m */
    int32_t curvature;
    /* vertical eddy viscosities of the first RK stage, added to the Euler tendencies of levels
     * (u, theta: 1..L-2) and interfaces (w: 1..L-1); mix_full = 0 mixes u - u_init cos(angleEdge)
     * - v_init sin(angleEdge) and theta_m - t_init */
    double v_mom_eddy_visc2, v_theta_eddy_visc2;
    int32_t mix_full;
} mpas_hip_dyn_params;

/* Scalars of atm_advance_acoustic_step_work (acoustic_step.py:1297-1340). */
typedef struct mpas_hip_acoustic_params {
    int32_t small_step;
    double dts, epssm, resm, c2;
} mpas_hip_acoustic_params;

/* Scratch doubles needed by the kernels that take a scratch buffer. */
int64_t mpas_hip_kernel_scratch_doubles(int32_t nCells, int32_t nEdges, int32_t nVertices,
                                        int32_t nLevels);

/* atm_compute_moist_coefficients, moist_coefficients.py:124 */
int mpas_hip_moist_coefficients(const mpas_hip_mesh *m, const double *scalars, int32_t nScalars,
                                int32_t moist_start, int32_t moist_end, double *cqw, double *cqu,
                                double *qtot, void *stream);
/* atm_compute_vert_imp_coefs_work, vert_imp_coefs.py:188 */
int mpas_hip_vert_imp_coefs(const mpas_hip_mesh *m, double dts, double epssm, const double *cqw,
                            const double *p, const double *t, const double *rt, const double *qtot,
                            double *cofrz, double *cofwr, double *cofwz, double *coftz, double *cofwt,
                            double *a_tri, double *alpha_tri, double *gamma_tri, void *stream);
/* atm_compute_solve_diagnostics_work, solve_diagnostics.py:1052;
 * rk_step <= 0 stands for the reference's rk_step=None */
int mpas_hip_solve_diagnostics(const mpas_hip_mesh *m, double dt, double apvm, int32_t rk_step,
                               const double *u, const double *h, const double *v_in, double *h_edge,
                               double *v, double *divergence, double *vorticity, double *ke,
                               double *pv_edge, double *scratch, void *stream);
/* atm_compute_dyn_tend_work, dyn_tend.py:10836.  inputs: u, v, w, ru, rw,
 * ru_save, rw_save, theta_m, theta_m_save, rho_zz, rho_edge, divergence,
 * vorticity, ke, pv_edge, pp, rr_save, cqw, cqu, rt_diabatic_tend, qtot, ur_cell,
 * vr_cell (the last two read only with p->curvature).
 * carried (written at rk_step 1, read later): tend_rho, tend_u_euler,
 * tend_w_euler, tend_theta_euler, kdiff. */
int mpas_hip_dyn_tend(const mpas_hip_mesh *m, const mpas_hip_dyn_params *p,
                      const double *u, const double *v, const double *w, const double *ru,
                      const double *rw, const double *ru_save, const double *rw_save,
                      const double *theta_m, const double *theta_m_save, const double *rho_zz,
                      const double *rho_edge, const double *divergence, const double *vorticity,
                      const double *ke, const double *pv_edge, const double *pp,
                      const double *rr_save, const double *cqw, const double *cqu,
                      const double *rt_diabatic_tend, const double *qtot,
                      const double *ur_cell, const double *vr_cell,
                      double *tend_rho, double *tend_u_euler, double *tend_w_euler,
                      double *tend_theta_euler, double *kdiff,
                      double *tend_u, double *tend_w, double *tend_theta, double *h_divergence,
                      double *rthdynten, double *scratch, void *stream);
/* atm_set_smlstep_pert_variables_work, smlstep_pert.py:373 (w_tend in place) */
int mpas_hip_smlstep_pert(const mpas_hip_mesh *m, const double *u_tend, double *w_tend, void *stream);
/* atm_advance_acoustic_step_work, acoustic_step.py:1297.  theta_m is the
 * time-level-1 theta_m (theta_m_init). */
int mpas_hip_acoustic_step(const mpas_hip_mesh *m, const mpas_hip_acoustic_params *p,
                           const double *cofrz, const double *cofwr, const double *cofwz,
                           const double *coftz, const double *cofwt, const double *a_tri,
                           const double *alpha_tri, const double *gamma_tri, const double *rho_zz,
                           const double *theta_m, const double *exner, const double *cqu,
                           const double *tend_ru, const double *tend_rho, const double *tend_rt,
                           const double *tend_rw, const double *w, const double *rw,
                           const double *rw_save, double *ru_p, double *ruAvg, double *rho_pp,
                           double *rtheta_pp, double *rtheta_pp_old, double *rw_p, double *wwAvg,
                           void *stream);
/* atm_divergence_damping_3d, divergence_damping.py:299 (ru_p in place) */
int mpas_hip_divergence_damping(const mpas_hip_mesh *m, double dts, double smdiv, double len_disp,
                                const double *theta_m, const double *rtheta_pp,
                                const double *rtheta_pp_old, double *ru_p, void *stream);
/* atm_recover_large_step_variables_work, recover_large_step.py:1090 */
int mpas_hip_recover_large_step(const mpas_hip_mesh *m, int32_t rk_step, double dt, int32_t ns,
                                const double *rho_p_save, const double *rho_pp,
                                const double *rtheta_p_save, const double *rtheta_pp,
                                const double *rt_diabatic_tend, const double *rw_save,
                                const double *rw_p, const double *ru_save, const double *ru_p,
                                double *wwAvg, double *ruAvg, double *rho_p, double *rho_zz,
                                double *rtheta_p, double *theta_m, double *exner, double *pressure_p,
                                double *rw, double *w, double *ru, double *u, void *stream);
/* atm_advance_scalars_work, advance_scalars.py:1272 (weights per RK stage) */
int mpas_hip_advance_scalars(const mpas_hip_mesh *m, int32_t nScalars, double dt, double coef3,
                             double weight_old, double weight_new, const double *scalar_old,
                             const double *scalar_new_in, const double *rho_zz_old,
                             const double *rho_zz_new, const double *uhAvg, const double *wwAvg,
                             double *scalar_new_out, double *horiz_flux, void *stream);
/* atm_advance_scalars_mono_work, advance_scalars_mono.py:2878 */
int mpas_hip_advance_scalars_mono(const mpas_hip_mesh *m, int32_t nScalars, double dt, double coef3,
                                  const double *scalars_old, const double *scalars_new_in,
                                  const double *rho_zz_old, const double *uhAvg,
                                  const double *wwAvg, double *scalars_new_out,
                                  double *rho_zz_int, double *scratch, void *stream);
/* atm_init_coupled_diagnostics, init_coupled_diagnostics.py:147 (cold start
 * from raw theta, rho, qv, u, w; uses the mesh's rho_base, zz, fzm, fzp and
 * terrain terms; writes padded outputs with rho_zz's padding row = 1).  With
 * theta, rho and qv all null it takes the derived path of :256-262: theta_m
 * and rho_zz are read from the output arrays, which the caller fills. */
int mpas_hip_init_coupled_diagnostics(const mpas_hip_mesh *m, const double *theta, const double *rho,
                                      const double *qv, const double *u, const double *w,
                                      const double *theta_base, double *theta_m, double *rho_zz,
                                      double *ru, double *rw, double *rho_p, double *rtheta_base,
                                      double *rtheta_p, double *exner, double *exner_base,
                                      double *pressure_p, double *pressure_base, void *stream);
/* mpas_reconstruct_2d, vector_reconstruction.py:311 */
int mpas_hip_reconstruct(const mpas_hip_mesh *m, const double *u, double *ux, double *uy,
                         double *uz, double *uzonal, double *umerid, void *stream);

/* ------------------------------------------------------------------------
 * Adjoints (vector-Jacobian products).  The reference obtains them from
 * torch autograd through each atm_*_work function (with the hand-written
 * backward rules of TridiagonalFactorize, tridiagonal_factor.py:183-230, and
 * TridiagonalSolve, tridiagonal_solve.py:232-291), driven by
 * run_simulation.vjp (atmosphere_model.py:2101-2155).  Each entry point below
 * is the adjoint of the forward entry point of the same name:
 *   fwd[]  device pointers: the forward call's inputs, then its outputs, in
 *          the slot order stated per function (in/out arrays appear twice:
 *          before the call among the inputs, after it among the outputs);
 *   bar[]  adjoints of the same slots: input adjoints are accumulated into
 *          (+=), output adjoints are read and may be overwritten.  A null
 *          input slot is skipped; output slots named per function must be
 *          non-null.
 * Both arrays are host memory.  No kernel uses atomics: each element is
 * written by one thread per launch and every cross-entity sum is gathered
 * in slot order, so the adjoints are the same bit for bit on every run.
 * ---------------------------------------------------------------------- */
/* Inverse connectivity for the adjoint gathers: the counterpart of the
 * reference's ordered output transposes (build_ordered_output_transpose,
 * static_payloads.py:44-118).  CSR tables ordered by (source row, slot):
 *   advStart[c] .. advStart[c+1]-1  index the (advEdge, advSlot) pairs with
 *                                   advCellsForEdge[advEdge][advSlot] == c;
 *   eoeStart[e] .. eoeStart[e+1]-1  index the (eoeEdge, eoeSlot) pairs with
 *                                   edgesOnEdge[eoeEdge][eoeSlot] == e;
 *   cocStart[c] .. cocStart[c+1]-1  index the (cocCell, cocSlot) pairs with
 *                                   cellsOnCell[cocCell][cocSlot] == c.
 * Required by every *_vjp entry point that takes it (a null pointer is an error). */
typedef struct mpas_hip_inverse_tables {
    const int32_t *advStart, *advEdge, *advSlot; /* nCells + 1, then one entry per (edge, slot) */
    const int32_t *eoeStart, *eoeEdge, *eoeSlot; /* nEdges + 1, then one entry per (edge, slot) */
    const int32_t *cocStart, *cocCell, *cocSlot; /* nCells + 1, then one entry per (cell, slot) */
} mpas_hip_inverse_tables;

/* doubles of device scratch any *_vjp entry point below needs */
int64_t mpas_hip_vjp_scratch_doubles(int32_t nCells, int32_t nEdges, int32_t nVertices, int32_t nLevels,
                                     int32_t nScalars);
/* fwd: scalars | cqw, cqu, qtot */
int mpas_hip_moist_coefficients_vjp(const mpas_hip_mesh *m, int32_t nScalars, int32_t moist_start,
                                    int32_t moist_end, const double *const *fwd, double *const *bar,
                                    void *stream);
/* fwd: cqw, p, t, rt, qtot | cofrz, cofwr, cofwz, coftz, cofwt, a_tri, alpha_tri, gamma_tri
 * (cofrz and cofwr depend on the mesh only; cofwz/coftz/cofwt adjoints required) */
int mpas_hip_vert_imp_coefs_vjp(const mpas_hip_mesh *m, double dts, double epssm, const double *const *fwd,
                                double *const *bar, void *stream);
/* fwd: u, h, v_in | h_edge, v, divergence, vorticity, ke, pv_edge */
int mpas_hip_solve_diagnostics_vjp(const mpas_hip_mesh *m, double dt, double apvm, int32_t rk_step,
                                   const double *const *fwd, double *const *bar, double *scratch,
                                   const mpas_hip_inverse_tables *inv, void *stream);
/* fwd: the 21 inputs of mpas_hip_dyn_tend, tend_rho, tend_u_euler, tend_w_euler, tend_theta_euler, kdiff
 * (before) | the same five (after), tend_u, tend_w, tend_theta, h_divergence, rthdynten
 * (h_divergence and kdiff adjoints required) */
int mpas_hip_dyn_tend_vjp(const mpas_hip_mesh *m, const mpas_hip_dyn_params *p, const double *const *fwd,
                          double *const *bar, double *scratch, const mpas_hip_inverse_tables *inv,
                          void *stream);
/* fwd: u_tend, w_tend (before) | w_tend (after) */
int mpas_hip_smlstep_pert_vjp(const mpas_hip_mesh *m, const double *const *fwd, double *const *bar, void *stream);
/* fwd: the 19 inputs of mpas_hip_acoustic_step, ru_p, ruAvg, rho_pp, rtheta_pp, rtheta_pp_old, rw_p,
 * wwAvg (before) | the same seven (after); the ru_p adjoint is required */
int mpas_hip_acoustic_step_vjp(const mpas_hip_mesh *m, const mpas_hip_acoustic_params *p,
                               const double *const *fwd, double *const *bar, double *scratch, void *stream);
/* fwd: theta_m, rtheta_pp, rtheta_pp_old, ru_p (before) | ru_p (after) */
int mpas_hip_divergence_damping_vjp(const mpas_hip_mesh *m, double dts, double smdiv, double len_disp,
                                    const double *const *fwd, double *const *bar, void *stream);
/* fwd: rho_p_save, rho_pp, rtheta_p_save, rtheta_pp, rt_diabatic_tend, rw_save, rw_p, ru_save, ru_p,
 * wwAvg, ruAvg, exner, pressure_p (before) | wwAvg, ruAvg, rho_p, rho_zz, rtheta_p, theta_m, exner,
 * pressure_p, rw, w, ru, u (after); rho_zz and ru adjoints required */
int mpas_hip_recover_large_step_vjp(const mpas_hip_mesh *m, int32_t rk_step, double dt, int32_t ns,
                                    const double *const *fwd, double *const *bar, void *stream);
/* fwd: scalar_old, scalar_new_in, rho_zz_old, rho_zz_new, uhAvg, wwAvg | scalar_new_out, horiz_flux */
int mpas_hip_advance_scalars_vjp(const mpas_hip_mesh *m, int32_t nScalars, double dt, double coef3,
                                 double weight_old, double weight_new, const double *const *fwd,
                                 double *const *bar, double *scratch, const mpas_hip_inverse_tables *inv,
                                 void *stream);
/* fwd: scalars_old, scalars_new_in, rho_zz_old, uhAvg, wwAvg | scalars_new_out, rho_zz_int
 * (rho_zz_int adjoint required) */
int mpas_hip_advance_scalars_mono_vjp(const mpas_hip_mesh *m, int32_t nScalars, double dt, double coef3,
                                      const double *const *fwd, double *const *bar, double *scratch,
                                      const mpas_hip_inverse_tables *inv, void *stream);
/* fwd: u | uReconstructX, Y, Z, Zonal, Meridional */
int mpas_hip_reconstruct_vjp(const mpas_hip_mesh *m, const double *const *fwd, double *const *bar, void *stream);

/* ------------------------------------------------------------------------
 * Tangents (Jacobian-vector products).  The reference obtains them with torch
 * forward-mode AD (dual numbers) through each atm_*_work function and the
 * jvp rules of TridiagonalFactorize / TridiagonalSolve
 * (tridiagonal_factor.py:233, tridiagonal_solve.py:294), driven by
 * run_simulation.jvp (atmosphere_model.py:2041-2098).  Each entry point below
 * is the tangent of the forward entry point of the same name, with the slot
 * order of its *_vjp counterpart above:
 *   fwd[]  device pointers: the forward call's inputs, then its outputs;
 *   dot[]  tangents of the same slots: input tangents are read (a null slot
 *          is a zero tangent), output tangents are written (all must be
 *          non-null).
 * Derivatives follow the adjoints' conventions (branches frozen, torch's
 * one-sided values at switching points), so each is the exact transpose of
 * its *_vjp.  Every thread gathers; results are deterministic.  Entry points
 * taking scratch use the same buffer size as the *_vjp ones. */
int mpas_hip_moist_coefficients_jvp(const mpas_hip_mesh *m, int32_t nScalars, int32_t moist_start,
                                    int32_t moist_end, const double *const *fwd, double *const *dot,
                                    void *stream);
int mpas_hip_vert_imp_coefs_jvp(const mpas_hip_mesh *m, double dts, double epssm, const double *const *fwd,
                                double *const *dot, void *stream);
int mpas_hip_solve_diagnostics_jvp(const mpas_hip_mesh *m, double dt, double apvm, int32_t rk_step,
                                   const double *const *fwd, double *const *dot, double *scratch, void *stream);
int mpas_hip_dyn_tend_jvp(const mpas_hip_mesh *m, const mpas_hip_dyn_params *p, const double *const *fwd,
                          double *const *dot, double *scratch, void *stream);
int mpas_hip_smlstep_pert_jvp(const mpas_hip_mesh *m, const double *const *fwd, double *const *dot, void *stream);
int mpas_hip_acoustic_step_jvp(const mpas_hip_mesh *m, const mpas_hip_acoustic_params *p,
                               const double *const *fwd, double *const *dot, double *scratch, void *stream);
int mpas_hip_divergence_damping_jvp(const mpas_hip_mesh *m, double dts, double smdiv, double len_disp,
                                    const double *const *fwd, double *const *dot, void *stream);
int mpas_hip_recover_large_step_jvp(const mpas_hip_mesh *m, int32_t rk_step, double dt, int32_t ns,
                                    const double *const *fwd, double *const *dot, void *stream);
int mpas_hip_advance_scalars_jvp(const mpas_hip_mesh *m, int32_t nScalars, double dt, double coef3,
                                 double weight_old, double weight_new, const double *const *fwd,
                                 double *const *dot, void *stream);
int mpas_hip_advance_scalars_mono_jvp(const mpas_hip_mesh *m, int32_t nScalars, double dt, double coef3,
                                      const double *const *fwd, double *const *dot, double *scratch,
                                      void *stream);
int mpas_hip_reconstruct_jvp(const mpas_hip_mesh *m, const double *const *fwd, double *const *dot, void *stream);

/* History diagnostics from the device state (output_writer.py:264-282): on
 * the nCells x nLevels active rows of the padded inputs, theta = theta_m /
 * (1 + rvord qv) with qv = scalars[..., 0] (0 when scalars is null), rho =
 * rho_zz zz and pressure = pressure_p + pressure_base.  A null output is
 * skipped. */
int mpas_hip_history_fields(int32_t nCells, int32_t nLevels, int32_t nScalars, const double *theta_m,
                            const double *scalars, const double *rho_zz, const double *zz,
                            const double *pressure_p, const double *pressure_base, double *theta,
                            double *rho, double *pressure, void *stream);

/* Derivatives of the cold start (atm_init_coupled_diagnostics,
 * init_coupled_diagnostics.py:147-380) for differentiable runs from raw
 * fields (run_simulation -> _prepare_timestep_state, atmosphere_model.py:782).
 * Slots: theta, rho, qv, theta_m_in, rho_zz_in, u, w | theta_m, rho_zz, ru,
 * rw, rho_p, rtheta_p, exner, pressure_p.  derived = 0 reads theta, rho, qv;
 * derived = 1 reads theta_m_in, rho_zz_in (the unused slots may be null).  The
 * base-state outputs carry no derivative.  The VJP accumulates input
 * adjoints (gathers only, deterministic) and needs
 * mpas_hip_init_coupled_vjp_scratch_doubles() doubles of scratch; the JVP
 * writes every output tangent (all output slots non-null). */
int64_t mpas_hip_init_coupled_vjp_scratch_doubles(int32_t nEdges, int32_t nLevels);
int mpas_hip_init_coupled_diagnostics_vjp(const mpas_hip_mesh *m, int32_t derived, const double *theta_base,
                                          const double *const *fwd, double *const *bar, double *scratch,
                                          void *stream);
int mpas_hip_init_coupled_diagnostics_jvp(const mpas_hip_mesh *m, int32_t derived, const double *theta_base,
                                          const double *const *fwd, double *const *dot, void *stream);

/* ------------------------------------------------------------------------
 * Stand-alone batched tridiagonal operators (the large step fuses its own
 * copies of these recurrences).  a, b, c, alpha, gamma: (nColumns, nLevels)
 * row-major; rhs and the solution: (nColumns, nLevels + 1), the last level is
 * the boundary value.  Device pointers; every output is written in full
 * (level-0 entries the recurrence leaves undefined are written as zero); a
 * null tangent / cotangent input reads as zero.
 * ---------------------------------------------------------------------- */
/* TridiagonalFactorize.forward (autodiff/tridiagonal_factor.py:164-181):
 * alpha(k) = 1 / (b(k) - a(k) gamma(k-1)), gamma(k) = c(k) alpha(k), k = 1..nLevels-1 */
int mpas_hip_tridiagonal_factorize(int64_t nColumns, int32_t nLevels, const double *a, const double *b,
                                   const double *c, double *alpha, double *gamma, void *stream);
/* TridiagonalFactorize.backward (tridiagonal_factor.py:183-230) */
int mpas_hip_tridiagonal_factorize_vjp(int64_t nColumns, int32_t nLevels, const double *a, const double *c,
                                       const double *alpha, const double *gamma, const double *grad_alpha,
                                       const double *grad_gamma, double *grad_a, double *grad_b, double *grad_c,
                                       void *stream);
/* TridiagonalFactorize.jvp (tridiagonal_factor.py:232-277) */
int mpas_hip_tridiagonal_factorize_jvp(int64_t nColumns, int32_t nLevels, const double *a, const double *c,
                                       const double *alpha, const double *gamma, const double *da, const double *db,
                                       const double *dc, double *d_alpha, double *d_gamma, void *stream);
/* TridiagonalSolve.forward (autodiff/tridiagonal_solve.py:194-231):
 * u(k) = (rhs(k) - a(k) u(k-1)) alpha(k) for k = 1..nLevels-1, then
 * out(k) = u(k) - gamma(k) out(k+1) for k = nLevels-1..0, out(nLevels) = rhs(nLevels) */
int mpas_hip_tridiagonal_solve(int64_t nColumns, int32_t nLevels, const double *a, const double *alpha,
                               const double *gamma, const double *rhs, double *out, void *stream);
/* TridiagonalSolve.backward (tridiagonal_solve.py:233-291); out is the forward solution */
int mpas_hip_tridiagonal_solve_vjp(int64_t nColumns, int32_t nLevels, const double *a, const double *alpha,
                                   const double *gamma, const double *rhs, const double *out, const double *grad_out,
                                   double *grad_a, double *grad_alpha, double *grad_gamma, double *grad_rhs,
                                   void *stream);
/* TridiagonalSolve.jvp (tridiagonal_solve.py:293-352) */
int mpas_hip_tridiagonal_solve_jvp(int64_t nColumns, int32_t nLevels, const double *a, const double *alpha,
                                   const double *gamma, const double *rhs, const double *out, const double *da,
                                   const double *d_alpha, const double *d_gamma, const double *d_rhs, double *d_out,
                                   void *stream);

/* Launch-shape table, no counterpart in the reference (its kernels are torch
 * operators): wavefronts per block (1, 2, 4, 8 or 16; 0 restores the default 4) of a
 * wavefront-per-entity kernel, by HIP kernel name (e.g. "k_dyn_edges_a").
 * Written by the autotuner (autotune.py) and read at every launch. */
int mpas_hip_tuning_set(const char *kernel_name, int32_t waves_per_block);
int32_t mpas_hip_tuning_get(const char *kernel_name);
/* Block order of a wavefront-per-entity kernel: 0 (default) deals contiguous chunks of entities
 * to the 8 XCDs, so neighbouring entities share an L2; 1 keeps the hardware's round-robin
 * order.  No counterpart in the reference (launch tuning). */
int mpas_hip_tuning_set_order(const char *kernel_name, int32_t order);
int32_t mpas_hip_tuning_get_order(const char *kernel_name);

const char *mpas_hip_last_error(void);

/* Measurement hook, no counterpart in the reference (its timings come from
 * SectionTimer on the CPU): arm an event timer for the HIP kernel named
 * kernel_name (e.g. "k_acoustic_cells"); each of its next `capacity`
 * launches, from any entry point, is bracketed by HIP events on its stream.
 * mpas_hip_timer_read waits for the recorded events and returns the summed
 * device time and the launch count, then disarms. */
int mpas_hip_timer_arm(const char *kernel_name, int32_t capacity);
int mpas_hip_timer_read(double *total_ms, int32_t *launches);
/* The same, per launch in launch order (ms_each holds capacity entries). */
int mpas_hip_timer_read_each(double *ms_each, int32_t capacity, int32_t *launches);

/* Measurement hook, no counterpart in the reference: the column-stream probe
 * behind the level pitch of the device state.  One wavefront per column (lane =
 * level) over nColumns columns of nLevels live levels stored at row pitch
 * `pitch` doubles (nLevels <= pitch <= 64; at pitch nLevels + 1 the spare lane
 * writes the padding element).  variant 0 reads three arrays and writes two;
 * variant 1 reads one array at the column and at six neighbour columns (+-97,
 * +-194, +-291 in storage order) and writes one.  Allocates its own five
 * buffers of nColumns x pitch doubles, runs 3 untimed and `reps` timed launches
 * on the null stream and returns the mean launch time in *ms. */
int mpas_hip_layout_probe(int64_t nColumns, int32_t nLevels, int32_t pitch, int32_t variant, int32_t reps,
                          double *ms);

#ifdef __cplusplus
}
#endif

#endif /* MPAS_HIP_H */
