/*
 * CPU oracle for the split-explicit dynamics step (atm_srk3) - kernels.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the checker for the HIP kernels and
 * the timed CPU baseline.  It restates, loop by loop, the reference
 * mpas_pytorch v0.1.7 kernels named at each function, keeping the reference's
 * association of every floating-point expression and its left-to-right slot
 * accumulation (mpas_pytorch/dynamics/stencil_ops.py:145-183).  Build with
 * -ffp-contract=off so no multiply-add is fused.
 *
 * All connectivity is 0-based with -1 for "none"; all field arrays are
 * row-major (rows, levels) with the reference's trailing padding row.
 */
#include "mpas_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GRAVITY 9.80616
#define RGAS 287.0
#define CP 1004.5
#define P0 100000.0
#define R_EARTH 6371229.0 /* constants.py:84, the curvature terms' radius */

int orc_mesh_sizeof(void) { return (int)sizeof(orc_mesh); }

static inline double dmin(double a, double b) { return (b < a) ? b : a; }
static inline double dmax(double a, double b) { return (b > a) ? b : a; }
static inline double sign_pm(double x) { return (x >= 0.0) ? 1.0 : -1.0; }
static inline double max0(double x) { return (x < 0.0) ? 0.0 : x; }

/* flux4/flux3 of dyn_tend.py:839-890 and advance_scalars.py:128-162 */
static inline double flux4(double qm2, double qm1, double q0, double qp1, double ua) {
    return ua * (7.0 * (q0 + qm1) - (qp1 + qm2)) / 12.0;
}
static inline double flux3(double qm2, double qm1, double q0, double qp1, double ua, double coef3) {
    return flux4(qm2, qm1, q0, qp1, ua) + coef3 * fabs(ua) * ((qp1 - qm2) - 3.0 * (q0 - qm1)) / 12.0;
}

/* ------------------------------------------------------------------------
 * atm_compute_moist_coefficients, moist_coefficients.py:124-291
 * ---------------------------------------------------------------------- */
void orc_moist_coefficients(const orc_mesh *m, const double *scalars, int nScalars,
                            int moist_start, int moist_end,
                            double *cqw, double *cqu, double *qtot) {
    const int L = m->nLevels, nC = m->nCells, nE = m->nEdges, S = nScalars;
    const int q0 = moist_start - 1, q1 = moist_end; /* [q0, q1) */
    memset(qtot, 0, sizeof(double) * (size_t)(nC + 1) * L);
    memset(cqw, 0, sizeof(double) * (size_t)(nC + 1) * (L + 1));
    memset(cqu, 0, sizeof(double) * (size_t)(nE + 1) * L);
    for (int c = 0; c < nC; ++c)
        for (int k = 0; k < L; ++k) {
            double t = 0.0;
            if (q1 > q0) {
                t = scalars[((size_t)c * L + k) * S + q0];
                for (int q = q0 + 1; q < q1; ++q) t = t + scalars[((size_t)c * L + k) * S + q];
            }
            qtot[(size_t)c * L + k] = t;
        }
    for (int c = 0; c < nC; ++c)
        for (int k = 1; k < L; ++k) {
            double qw = 0.5 * (qtot[(size_t)c * L + k] + qtot[(size_t)c * L + k - 1]);
            cqw[(size_t)c * (L + 1) + k] = 1.0 / (1.0 + qw);
        }
    for (int e = 0; e < nE; ++e) {
        int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
        for (int k = 0; k < L; ++k) {
            double qu = 0.0;
            if (q1 > q0) {
                qu = 0.5 * (scalars[((size_t)c1 * L + k) * S + q0] + scalars[((size_t)c2 * L + k) * S + q0]);
                for (int q = q0 + 1; q < q1; ++q)
                    qu = qu + 0.5 * (scalars[((size_t)c1 * L + k) * S + q] + scalars[((size_t)c2 * L + k) * S + q]);
            }
            cqu[(size_t)e * L + k] = 1.0 / (1.0 + qu);
        }
    }
}

/* ------------------------------------------------------------------------
 * atm_compute_vert_imp_coefs_work, vert_imp_coefs.py:188-474
 * (coefficients :292-445, Thomas factor :100-121)
 * ---------------------------------------------------------------------- */
void orc_vert_imp_coefs(const orc_mesh *m, double dts, double epssm,
                        const double *cqw, const double *p, const double *t,
                        const double *rt, const double *qtot,
                        double *cofrz, double *cofwr, double *cofwz, double *coftz,
                        double *cofwt, double *a_tri, double *alpha_tri, double *gamma_tri) {
    const int L = m->nLevels, nC = m->nCells;
    const double dtseps = 0.5 * dts * (1.0 + epssm);
    const double rcv = RGAS / (CP - RGAS);
    const double c2 = CP * rcv;
    const double *zz = m->zz, *rb = m->rho_base, *rtb = m->rtheta_base, *pb = m->exner_base;
    const double *rdzw = m->rdzw, *fzm = m->fzm, *fzp = m->fzp, *rdzu = m->rdzu;
    for (int k = 0; k < L; ++k) cofrz[k] = dtseps * rdzw[k];
    memset(cofwr, 0, sizeof(double) * (size_t)(nC + 1) * L);
    memset(cofwz, 0, sizeof(double) * (size_t)(nC + 1) * L);
    memset(coftz, 0, sizeof(double) * (size_t)(nC + 1) * (L + 1));
    memset(cofwt, 0, sizeof(double) * (size_t)(nC + 1) * L);
    memset(a_tri, 0, sizeof(double) * (size_t)(nC + 1) * L);
    memset(alpha_tri, 0, sizeof(double) * (size_t)(nC + 1) * L);
    memset(gamma_tri, 0, sizeof(double) * (size_t)(nC + 1) * L);
    for (int c = 0; c < nC; ++c) {
        const size_t o = (size_t)c * L, ow = (size_t)c * (L + 1);
        for (int k = 1; k < L; ++k) {
            double zi = fzm[k] * zz[o + k] + fzp[k] * zz[o + k - 1];
            double pi = fzm[k] * p[o + k] + fzp[k] * p[o + k - 1];
            double ti = fzm[k] * t[o + k] + fzp[k] * t[o + k - 1];
            cofwr[o + k] = 0.5 * dtseps * GRAVITY * zi;
            cofwz[o + k] = dtseps * c2 * zi * rdzu[k] * cqw[ow + k] * pi;
            coftz[ow + k] = dtseps * ti;
        }
        for (int k = 0; k < L; ++k)
            cofwt[o + k] = 0.5 * dtseps * rcv * zz[o + k] * GRAVITY * rb[o + k] / (1.0 + qtot[o + k]) *
                           p[o + k] / ((rtb[o + k] + rt[o + k]) * pb[o + k]);
        double gprev = 0.0;
        for (int k = 1; k < L; ++k) {
            double a = -cofwz[o + k] * coftz[ow + k - 1] * rdzw[k - 1] * zz[o + k - 1] +
                       cofwr[o + k] * cofrz[k - 1] - cofwt[o + k - 1] * coftz[ow + k - 1] * rdzw[k - 1];
            double b = 1.0 +
                       cofwz[o + k] * (coftz[ow + k] * rdzw[k] * zz[o + k] + coftz[ow + k] * rdzw[k - 1] * zz[o + k - 1]) -
                       coftz[ow + k] * (cofwt[o + k] * rdzw[k] - cofwt[o + k - 1] * rdzw[k - 1]) +
                       cofwr[o + k] * (cofrz[k] - cofrz[k - 1]);
            double cc = -cofwz[o + k] * coftz[ow + k + 1] * rdzw[k] * zz[o + k] - cofwr[o + k] * cofrz[k] +
                        cofwt[o + k] * coftz[ow + k + 1] * rdzw[k];
            a_tri[o + k] = a;
            double al = 1.0 / (b - a * gprev);
            double ga = cc * al;
            alpha_tri[o + k] = al;
            gamma_tri[o + k] = ga;
            gprev = ga;
        }
    }
}

/* ------------------------------------------------------------------------
 * atm_compute_solve_diagnostics_work, solve_diagnostics.py:531-756
 * rk_step <= 0 stands for the reference's rk_step=None (v reconstructed).
 * ---------------------------------------------------------------------- */
void orc_solve_diagnostics(const orc_mesh *m, double dt, double apvm, int rk_step,
                           const double *u, const double *h, const double *v_in,
                           double *h_edge, double *v, double *divergence, double *vorticity,
                           double *ke, double *pv_edge) {
    const int L = m->nLevels, nC = m->nCells, nE = m->nEdges, nV = m->nVertices, ME = m->maxEdges;
    double *ke_edge = (double *)malloc(sizeof(double) * (size_t)nE * L);
    double *ke_vertex = (double *)malloc(sizeof(double) * (size_t)nV * L);
    double *pv_vertex = (double *)malloc(sizeof(double) * (size_t)nV * L);
    double *pv_cell = (double *)malloc(sizeof(double) * (size_t)nC * L);
    for (int e = 0; e < nE; ++e) {
        int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
        double efac = m->dcEdge[e] * m->dvEdge[e];
        for (int k = 0; k < L; ++k) {
            h_edge[(size_t)e * L + k] = 0.5 * (h[(size_t)c1 * L + k] + h[(size_t)c2 * L + k]);
            double uu = u[(size_t)e * L + k];
            ke_edge[(size_t)e * L + k] = efac * (uu * uu);
        }
    }
    for (int vx = 0; vx < nV; ++vx)
        for (int k = 0; k < L; ++k) {
            double s = 0.0, sk = 0.0;
            for (int i = 0; i < 3; ++i) {
                int e = m->edgesOnVertex[3 * vx + i];
                double w = m->edgesOnVertex_sign[3 * vx + i] * m->dcEdge[e];
                double c = w * u[(size_t)e * L + k];
                s = (i == 0) ? c : s + c;
                sk = (i == 0) ? ke_edge[(size_t)e * L + k] : sk + ke_edge[(size_t)e * L + k];
            }
            vorticity[(size_t)vx * L + k] = s * m->invAreaTriangle[vx];
            ke_vertex[(size_t)vx * L + k] = 0.25 * m->invAreaTriangle[vx] * sk;
            pv_vertex[(size_t)vx * L + k] = m->fVertex[vx] + vorticity[(size_t)vx * L + k];
        }
    const double ke_fact = 1.0 - 0.375;
    for (int c = 0; c < nC; ++c) {
        int n = m->nEdgesOnCell[c];
        for (int k = 0; k < L; ++k) {
            double d = 0.0, kc = 0.0, pc = 0.0;
            for (int i = 0; i < n; ++i) {
                int e = m->edgesOnCell[c * ME + i];
                double w = m->edgesOnCell_sign[c * ME + i] * m->dvEdge[e];
                double cd = w * u[(size_t)e * L + k];
                double ck = 0.25 * ke_edge[(size_t)e * L + k];
                d = (i == 0) ? cd : d + cd;
                kc = (i == 0) ? ck : kc + ck;
            }
            divergence[(size_t)c * L + k] = d * m->invAreaCell[c];
            double kk = kc * m->invAreaCell[c];
            double tot = ke_fact * kk;
            for (int i = 0; i < n; ++i) {
                int vx = m->verticesOnCell[c * ME + i];
                double ka = m->kiteAreasOnVertex[3 * vx + m->kiteForCell[c * ME + i]];
                double contrib = (1.0 - ke_fact) * ka;
                contrib = contrib * ke_vertex[(size_t)vx * L + k];
                contrib = contrib * m->invAreaCell[c];
                tot = tot + contrib;
                double cp = ka * pv_vertex[(size_t)vx * L + k] * m->invAreaCell[c];
                pc = (i == 0) ? cp : pc + cp;
            }
            ke[(size_t)c * L + k] = tot;
            pv_cell[(size_t)c * L + k] = pc;
        }
    }
    const int reconstruct_v = (rk_step <= 0 || rk_step == 3);
    for (int e = 0; e < nE; ++e) {
        int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
        int v1 = m->verticesOnEdge[2 * e], v2 = m->verticesOnEdge[2 * e + 1];
        int ne = m->nEdgesOnEdge[e];
        for (int k = 0; k < L; ++k) {
            double vv;
            if (reconstruct_v) {
                vv = 0.0;
                for (int j = 0; j < ne; ++j) {
                    int ee = m->edgesOnEdge[e * m->maxEdges2 + j];
                    double c = m->weightsOnEdge[e * m->maxEdges2 + j] * u[(size_t)ee * L + k];
                    vv = (j == 0) ? c : vv + c;
                }
            } else {
                vv = v_in[(size_t)e * L + k];
            }
            v[(size_t)e * L + k] = vv;
            double p1 = pv_vertex[(size_t)v1 * L + k], p2 = pv_vertex[(size_t)v2 * L + k];
            double pe = 0.5 * (p1 + p2);
            if (apvm > 0.0) {
                double grad_pvt = (p2 - p1) * m->invDvEdge[e];
                double grad_pvn = (pv_cell[(size_t)c2 * L + k] - pv_cell[(size_t)c1 * L + k]) * m->invDcEdge[e];
                double delta = (apvm * dt) * (vv * grad_pvt + u[(size_t)e * L + k] * grad_pvn);
                pe = pe - delta;
            }
            pv_edge[(size_t)e * L + k] = pe;
        }
    }
    free(ke_edge);
    free(ke_vertex);
    free(pv_vertex);
    free(pv_cell);
}

/* ------------------------------------------------------------------------
 * atm_compute_dyn_tend_work, dyn_tend.py:9390-10833 with sections
 * :1045 (density), :1097/:1242 (PV), :1455/:1698/:2178 (u), :2508/:2559
 * (u mixing), :2956 (u final), :3299/:3367 (Smagorinsky), :3627/:3678
 * (edge advection), :4174/:4312 (w/theta horizontal advection), :4415-4657
 * (w vertical), :4803-5057 (theta vertical), :5315-5843 (w/theta mixing),
 * :5901 (theta perturbation flux), :6485 (ordered del4 subtraction).
 * Active configuration: 2d_smagorinsky or 2d_fixed horizontal mixing
 * (:3367-3430), optional CAM damping of kdiff (:3434-3458), optional vertical
 * eddy viscosities of u, w and theta with config_mix_full either way (:2731-2781,
 * :4633-4647, :5060-5097), optional Rayleigh damping of u on the top levels (:2942-2996,
 * coefficient :10066-10072), optional curvature terms of u (:2471-2505) and w
 * (:4241-4267; config_curvature_u), no physics tendencies.
 * ---------------------------------------------------------------------- */
typedef struct orc_dyn_params {
    int rk_step;
    double dt;
    double coef_3rd_order;
    double kdiff_coef;   /* (c_s * len_disp)**2 */
    double kdiff_cap;    /* (0.01 * len_disp**2) * (1/dt) */
    double h_mom_eddy_visc4;
    double h_theta_eddy_visc4;
    double del4u_div_factor;
    double prandtl_inv;
    int rayleigh_levels;          /* damped top levels, 0 without Rayleigh damping of u */
    double rayleigh_coef_inverse; /* 1 / (levels * timescale_days * 86400) */
    int mix_fixed;                /* 1: "2d_fixed", kdiff = kdiff_fixed (config_h_theta_eddy_visc2) */
    double kdiff_fixed;
    int cam_levels;               /* CAM-damped top levels, 0 without CAM damping */
    double cam_visc2;             /* 4 * 2.0833 * len_disp * config_mpas_cam_coef */
    int curvature;                /* config_curvature_u */
    double v_mom_eddy_visc2;      /* vertical mixing of u and w, first RK stage */
    double v_theta_eddy_visc2;    /* vertical mixing of theta, first RK stage */
    int mix_full;                 /* 0: mix u and theta as departures from u_init/v_init/t_init */
} orc_dyn_params;

int orc_dyn_params_sizeof(void) { return (int)sizeof(orc_dyn_params); }

void orc_dyn_tend(const orc_mesh *m, const orc_dyn_params *P,
                  const double *u, const double *v, const double *w, const double *ru,
                  const double *rw, const double *ru_save, const double *rw_save,
                  const double *theta_m, const double *theta_m_save, const double *rho_zz,
                  const double *rho_edge, const double *divergence, const double *vorticity,
                  const double *ke, const double *pv_edge, const double *pp,
                  const double *rr_save, const double *cqw, const double *cqu,
                  const double *rt_diabatic_tend, const double *qtot,
                  /* curvature: reconstructed cell winds of the step's input state, and
                   * (2 omega cos angleEdge) cos latEdge per edge, 2 omega cos latCell per cell */
                  const double *ur_cell, const double *vr_cell, const double *curv_edge,
                  const double *curv_cell,
                  /* vertical mixing: zgrid (cells x interfaces), t_init (cells x levels),
                   * u_init, v_init (levels), cos and sin of angleEdge (edges) */
                  const double *zgrid, const double *t_init, const double *u_init,
                  const double *v_init, const double *cos_edge, const double *sin_edge,
                  /* in (rk>1) / out (rk==1) */
                  double *tend_rho, double *tend_u_euler, double *tend_w_euler,
                  double *tend_theta_euler, double *kdiff,
                  /* out */
                  double *tend_u, double *tend_w, double *tend_theta, double *h_divergence,
                  double *rthdynten) {
    const int L = m->nLevels, nC = m->nCells, nE = m->nEdges, nV = m->nVertices;
    const int ME = m->maxEdges, ME2 = m->maxEdges2, MA = m->maxAdv;
    const int rk1 = (P->rk_step == 1);
    const double *rdzw = m->rdzw, *rdzu = m->rdzu, *fzm = m->fzm, *fzp = m->fzp;
    const size_t CL = (size_t)(nC + 1) * L, CW = (size_t)(nC + 1) * (L + 1), EL = (size_t)(nE + 1) * L;
    double *dpdz = (double *)calloc(CL, sizeof(double));
    double *delsq_u = (double *)calloc(EL, sizeof(double));
    double *delsq_div = (double *)calloc(CL, sizeof(double));
    double *delsq_vort = (double *)calloc((size_t)(nV + 1) * L, sizeof(double));
    double *delsq_theta = (double *)calloc(CL, sizeof(double));
    double *delsq_w = (double *)calloc(CW, sizeof(double));
    double *colbuf = (double *)malloc(sizeof(double) * (size_t)(L + 1) * 4);

    memset(tend_u, 0, sizeof(double) * EL);
    memset(tend_w, 0, sizeof(double) * CW);
    memset(tend_theta, 0, sizeof(double) * CL);
    memset(h_divergence, 0, sizeof(double) * CL);
    memset(rthdynten, 0, sizeof(double) * CL);

    /* mixing initialisation (dyn_tend.py:3367-3465): Smagorinsky kdiff, or the
     * fixed h_theta_eddy_visc2; then max(kdiff, visc2cam(k)) on the CAM levels with
     * visc2cam(k) = base * (1 - (L - (k + 1)) / levels) */
    if (rk1) {
        memset(kdiff, 0, sizeof(double) * CL);
        memset(tend_u_euler, 0, sizeof(double) * EL);
        memset(tend_w_euler, 0, sizeof(double) * CW);
        memset(tend_theta_euler, 0, sizeof(double) * CL);
        memset(tend_rho, 0, sizeof(double) * CL);
        for (int c = 0; c < nC; ++c) {
            int n = m->nEdgesOnCell[c];
            for (int k = 0; k < L; ++k) {
                double kd;
                if (P->mix_fixed) {
                    kd = P->kdiff_fixed;
                } else {
                    double dd = 0.0, dof = 0.0;
                    for (int i = 0; i < n; ++i) {
                        int e = m->edgesOnCell[c * ME + i];
                        double a = m->defc_a[c * ME + i], b = m->defc_b[c * ME + i];
                        double us = u[(size_t)e * L + k], vs = v[(size_t)e * L + k];
                        dd = (dd + a * us) - b * vs;
                        dof = (dof + b * us) + a * vs;
                    }
                    double def = sqrt(dd * dd + dof * dof);
                    kd = dmin(P->kdiff_coef * def, P->kdiff_cap);
                }
                if (P->cam_levels > 0 && k >= L - P->cam_levels) {
                    double factor = 1.0 - (double)(L - (k + 1)) / (double)P->cam_levels;
                    kd = dmax(kd, P->cam_visc2 * factor);
                }
                kdiff[(size_t)c * L + k] = kd;
            }
        }
    }

    /* density tendency and horizontal mass-flux divergence (:1045-1094) */
    for (int c = 0; c < nC; ++c) {
        int n = m->nEdgesOnCell[c];
        for (int k = 0; k < L; ++k) {
            double s = 0.0;
            for (int i = 0; i < n; ++i) {
                int e = m->edgesOnCell[c * ME + i];
                double wgt = m->edgesOnCell_sign[c * ME + i] * m->dvEdge[e];
                double cc = wgt * ru[(size_t)e * L + k];
                s = (i == 0) ? cc : s + cc;
            }
            h_divergence[(size_t)c * L + k] = s * m->invAreaCell[c];
        }
        if (rk1)
            for (int k = 0; k < L; ++k) {
                const size_t o = (size_t)c * L + k;
                tend_rho[o] = -h_divergence[o] - rdzw[k] * (rw[(size_t)c * (L + 1) + k + 1] - rw[(size_t)c * (L + 1) + k]);
                dpdz[o] = -GRAVITY * (m->rho_base[o] * qtot[o] + rr_save[o] * (1.0 + qtot[o]));
            }
    }

    /* normal-velocity tendency (:1455-2470) */
    for (int e = 0; e < nE; ++e) {
        const int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
        const double *ue = u + (size_t)e * L;
        double *wduz = colbuf;
        /* interface fluxes, index = interface k (1..L-1) */
        for (int k = 1; k < L; ++k) {
            double erw = 0.5 * (rw[(size_t)c1 * (L + 1) + k] + rw[(size_t)c2 * (L + 1) + k]);
            if (k == 1 || k == L - 1)
                wduz[k] = erw * (fzm[k] * ue[k] + fzp[k] * ue[k - 1]);
            else
                wduz[k] = flux3(ue[k - 2], ue[k - 1], ue[k], ue[k + 1], erw, 1.0);
        }
        for (int k = 0; k < L; ++k) {
            double delta;
            if (L == 1) delta = 0.0;
            else if (k == 0) delta = wduz[1];
            else if (k == L - 1) delta = -wduz[L - 1];
            else delta = wduz[k + 1] - wduz[k];
            double vert = -rdzw[k] * delta;
            /* q: Ringler et al. 2010 PV flux (:1097-1131) */
            double q = 0.0;
            int ne = m->nEdgesOnEdge[e];
            for (int j = 0; j < ne; ++j) {
                int ee = m->edgesOnEdge[e * ME2 + j];
                double workpv = 0.5 * (pv_edge[(size_t)e * L + k] + pv_edge[(size_t)ee * L + k]);
                double cc = m->weightsOnEdge[e * ME2 + j] * u[(size_t)ee * L + k] * workpv;
                q = (j == 0) ? cc : q + cc;
            }
            double ke_delta = ke[(size_t)c2 * L + k] - ke[(size_t)c1 * L + k];
            double hsum = h_divergence[(size_t)c1 * L + k] + h_divergence[(size_t)c2 * L + k];
            double base = rho_edge[(size_t)e * L + k] * (q - ke_delta * m->invDcEdge[e]) - ue[k] * 0.5 * hsum;
            tend_u[(size_t)e * L + k] = vert + base;
            if (P->curvature) {
                /* (:2471-2505) tend_u -= curv_edge rho_edge w_avg + u w_avg rho_edge / r_earth */
                const double *w1 = w + (size_t)c1 * (L + 1), *w2 = w + (size_t)c2 * (L + 1);
                const double w_avg = 0.25 * (((w1[k] + w1[k + 1]) + w2[k]) + w2[k + 1]);
                const double re = rho_edge[(size_t)e * L + k];
                const double curv = curv_edge[e] * re * w_avg + ue[k] * w_avg * re * (1.0 / R_EARTH);
                tend_u[(size_t)e * L + k] = tend_u[(size_t)e * L + k] - curv;
            }
            if (rk1) {
                double zzs = m->zz[(size_t)c2 * L + k] + m->zz[(size_t)c1 * L + k];
                double ppd = pp[(size_t)c2 * L + k] - pp[(size_t)c1 * L + k];
                double dps = dpdz[(size_t)c1 * L + k] + dpdz[(size_t)c2 * L + k];
                tend_u_euler[(size_t)e * L + k] =
                    -cqu[(size_t)e * L + k] * (ppd * m->invDcEdge[e] / (0.5 * zzs) - 0.5 * m->zxu[(size_t)e * L + k] * dps);
            }
        }
    }

    /* horizontal mixing for u, first RK stage only (:2508-2730) */
    if (rk1) {
        for (int e = 0; e < nE; ++e) {
            const int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
            const int v1 = m->verticesOnEdge[2 * e], v2 = m->verticesOnEdge[2 * e + 1];
            double r_dc = m->invDcEdge[e];
            double r_dv = dmin(m->invDvEdge[e], 4.0 * m->invDcEdge[e]);
            for (int k = 0; k < L; ++k) {
                double ud = (divergence[(size_t)c2 * L + k] - divergence[(size_t)c1 * L + k]) * r_dc -
                            (vorticity[(size_t)v2 * L + k] - vorticity[(size_t)v1 * L + k]) * r_dv;
                delsq_u[(size_t)e * L + k] = ud;
                double kdiffu = 0.5 * (kdiff[(size_t)c1 * L + k] + kdiff[(size_t)c2 * L + k]);
                tend_u_euler[(size_t)e * L + k] = tend_u_euler[(size_t)e * L + k] +
                                                  rho_edge[(size_t)e * L + k] * kdiffu * ud * m->meshScalingDel2[e];
            }
        }
        if (P->h_mom_eddy_visc4 > 0.0) {
            for (int vx = 0; vx < nV; ++vx)
                for (int k = 0; k < L; ++k) {
                    double s = 0.0;
                    for (int i = 0; i < 3; ++i) {
                        int e = m->edgesOnVertex[3 * vx + i];
                        double wgt = m->invAreaTriangle[vx] * m->dcEdge[e] * m->edgesOnVertex_sign[3 * vx + i];
                        double cc = wgt * delsq_u[(size_t)e * L + k];
                        s = (i == 0) ? cc : s + cc;
                    }
                    delsq_vort[(size_t)vx * L + k] = s;
                }
            for (int c = 0; c < nC; ++c) {
                int n = m->nEdgesOnCell[c];
                for (int k = 0; k < L; ++k) {
                    double s = 0.0;
                    for (int i = 0; i < n; ++i) {
                        int e = m->edgesOnCell[c * ME + i];
                        double wgt = m->invAreaCell[c] * m->dvEdge[e] * m->edgesOnCell_sign[c * ME + i];
                        double cc = wgt * delsq_u[(size_t)e * L + k];
                        s = (i == 0) ? cc : s + cc;
                    }
                    delsq_div[(size_t)c * L + k] = s;
                }
            }
            for (int e = 0; e < nE; ++e) {
                const int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
                const int v1 = m->verticesOnEdge[2 * e], v2 = m->verticesOnEdge[2 * e + 1];
                double ums = m->meshScalingDel4[e] * P->h_mom_eddy_visc4;
                double r_dc = ums * P->del4u_div_factor * m->invDcEdge[e];
                double r_dv = ums * dmin(m->invDvEdge[e], 4.0 * m->invDcEdge[e]);
                for (int k = 0; k < L; ++k) {
                    double ud = rho_edge[(size_t)e * L + k] *
                                ((delsq_div[(size_t)c2 * L + k] - delsq_div[(size_t)c1 * L + k]) * r_dc -
                                 (delsq_vort[(size_t)v2 * L + k] - delsq_vort[(size_t)v1 * L + k]) * r_dv);
                    tend_u_euler[(size_t)e * L + k] = tend_u_euler[(size_t)e * L + k] - ud;
                }
            }
        }
        /* vertical mixing of u (:2731-2781), levels 1..L-2, on the full u or on its
         * departure u - u_init cos(angleEdge) - v_init sin(angleEdge); heights are
         * the means of the edge's two cell columns */
        if (P->v_mom_eddy_visc2 > 0.0 && L > 2) {
            double *um = colbuf;
            for (int e = 0; e < nE; ++e) {
                const int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
                const double *g1 = zgrid + (size_t)c1 * (L + 1), *g2 = zgrid + (size_t)c2 * (L + 1);
                for (int k = 0; k < L; ++k) {
                    const double uk = u[(size_t)e * L + k];
                    um[k] = P->mix_full ? uk : (uk - u_init[k] * cos_edge[e]) - v_init[k] * sin_edge[e];
                }
                for (int k = 1; k + 1 < L; ++k) {
                    const double z1 = 0.5 * (g1[k - 1] + g2[k - 1]), z2 = 0.5 * (g1[k] + g2[k]);
                    const double z3 = 0.5 * (g1[k + 1] + g2[k + 1]), z4 = 0.5 * (g1[k + 2] + g2[k + 2]);
                    const double zm = 0.5 * (z1 + z2), z0 = 0.5 * (z2 + z3), zp = 0.5 * (z3 + z4);
                    const double vm = rho_edge[(size_t)e * L + k] * P->v_mom_eddy_visc2 *
                                      ((um[k + 1] - um[k]) / (zp - z0) - (um[k] - um[k - 1]) / (z0 - zm)) /
                                      (0.5 * (zp - zm));
                    tend_u_euler[(size_t)e * L + k] = tend_u_euler[(size_t)e * L + k] + vm;
                }
            }
        }
    }
    /* Rayleigh damping of u on the top levels, before the Euler part is added
     * (:2969-2994): tend_u -= rho_edge u damp(k), damp(k) = ((k + 1) - k0) / (levels tau) */
    if (P->rayleigh_levels > 0) {
        const int k0 = L - P->rayleigh_levels;
        for (int e = 0; e < nE; ++e)
            for (int k = k0 < 0 ? 0 : k0; k < L; ++k) {
                const size_t i = (size_t)e * L + k;
                const double damp = (((double)k + 1.0) - (double)k0) * P->rayleigh_coef_inverse;
                tend_u[i] = tend_u[i] - rho_edge[i] * u[i] * damp;
            }
    }
    for (size_t i = 0; i < (size_t)nE * L; ++i) tend_u[i] = tend_u[i] + tend_u_euler[i];

    /* w: horizontal advection (:4174-4210, :3627-3747) */
    double *hadv = (double *)calloc(CW, sizeof(double));
    for (int c = 0; c < nC; ++c) {
        int n = m->nEdgesOnCell[c];
        for (int i = 0; i < n; ++i) {
            int e = m->edgesOnCell[c * ME + i];
            double sgn = m->edgesOnCell_sign[c * ME + i];
            int na = m->nAdvCellsForEdge[e];
            for (int k = 1; k < L; ++k) {
                double rue = fzm[k] * ru[(size_t)e * L + k] + fzp[k] * ru[(size_t)e * L + k - 1];
                double sr = sign_pm(rue);
                double fs = 0.0;
                for (int j = 0; j < na; ++j) {
                    int ac = m->advCellsForEdge[e * MA + j];
                    double wgt = m->adv_coefs[e * MA + j] + sr * m->adv_coefs_3rd[e * MA + j];
                    double cc = wgt * w[(size_t)ac * (L + 1) + k];
                    fs = (j == 0) ? cc : fs + cc;
                }
                double prod = rue * fs;
                double cc = -sgn * prod;
                double *acc = &hadv[(size_t)c * (L + 1) + k];
                *acc = (i == 0) ? cc : *acc + cc;
            }
        }
        if (P->curvature)
            /* (:4241-4267) added to the advective sum before its 1/area scaling */
            for (int k = 1; k < L; ++k) {
                const size_t o = (size_t)c * L + k;
                const double ur = fzm[k] * ur_cell[o] + fzp[k] * ur_cell[o - 1];
                const double vr = fzm[k] * vr_cell[o] + fzp[k] * vr_cell[o - 1];
                const double rh = rho_zz[o] * fzm[k] + rho_zz[o - 1] * fzp[k];
                double *acc = &hadv[(size_t)c * (L + 1) + k];
                *acc = *acc + rh * (ur * ur + vr * vr) / R_EARTH;
                *acc = *acc + curv_cell[c] * ur * rh;
            }
    }
    /* w: horizontal mixing, first RK stage (:5315-5842, :6485) */
    if (rk1) {
        double *wturb = (double *)calloc(CW, sizeof(double));
        for (int c = 0; c < nC; ++c) {
            int n = m->nEdgesOnCell[c];
            for (int i = 0; i < n; ++i) {
                int e = m->edgesOnCell[c * ME + i];
                int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
                double es = 0.5 * m->invAreaCell[c] * m->edgesOnCell_sign[c * ME + i] * m->dvEdge[e] * m->invDcEdge[e];
                for (int k = 1; k < L; ++k) {
                    double bf = es * (rho_edge[(size_t)e * L + k] + rho_edge[(size_t)e * L + k - 1]) *
                                (w[(size_t)c2 * (L + 1) + k] - w[(size_t)c1 * (L + 1) + k]);
                    double sf = bf * m->meshScalingDel2[e] * 0.25 *
                                (kdiff[(size_t)c1 * L + k] + kdiff[(size_t)c2 * L + k] + kdiff[(size_t)c1 * L + k - 1] +
                                 kdiff[(size_t)c2 * L + k - 1]);
                    size_t o = (size_t)c * (L + 1) + k;
                    delsq_w[o] = (i == 0) ? bf : delsq_w[o] + bf;
                    wturb[o] = (i == 0) ? sf : wturb[o] + sf;
                }
            }
        }
        if (P->h_mom_eddy_visc4 > 0.0)
            for (int c = 0; c < nC; ++c) {
                int n = m->nEdgesOnCell[c];
                double r_area = P->h_mom_eddy_visc4 * m->invAreaCell[c];
                for (int i = 0; i < n; ++i) {
                    int e = m->edgesOnCell[c * ME + i];
                    int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
                    double es = m->meshScalingDel4[e] * r_area * m->dvEdge[e] * m->edgesOnCell_sign[c * ME + i] * m->invDcEdge[e];
                    for (int k = 1; k < L; ++k) {
                        size_t o = (size_t)c * (L + 1) + k;
                        wturb[o] = wturb[o] - es * (delsq_w[(size_t)c2 * (L + 1) + k] - delsq_w[(size_t)c1 * (L + 1) + k]);
                    }
                }
            }
        for (int c = 0; c < nC; ++c)
            for (int k = 1; k < L; ++k) {
                size_t o = (size_t)c * (L + 1) + k, oc = (size_t)c * L;
                double e0 = wturb[o];
                e0 = e0 - cqw[o] * (rdzu[k] * (pp[oc + k] - pp[oc + k - 1]) - (fzm[k] * dpdz[oc + k] + fzp[k] * dpdz[oc + k - 1]));
                if (P->v_mom_eddy_visc2 > 0.0) {
                    /* vertical mixing of w (:4633-4647) */
                    const double *wc = w + (size_t)c * (L + 1);
                    const double rho_half = 0.5 * (rho_zz[oc + k] + rho_zz[oc + k - 1]);
                    const double w_up = (wc[k + 1] - wc[k]) * rdzw[k], w_down = (wc[k] - wc[k - 1]) * rdzw[k - 1];
                    e0 = e0 + P->v_mom_eddy_visc2 * rho_half * (w_up - w_down) * rdzu[k];
                }
                tend_w_euler[o] = e0;
            }
        free(wturb);
    }
    /* w: vertical transport and final w tendency (:4415-4657) */
    for (int c = 0; c < nC; ++c) {
        const double *wc = w + (size_t)c * (L + 1), *rwc = rw + (size_t)c * (L + 1);
        double *wdwz = colbuf; /* mass-level fluxes 0..L-2 */
        for (int j = 0; j + 1 < L; ++j) {
            if (j == 0 || j == L - 2)
                wdwz[j] = 0.25 * (rwc[j + 1] + rwc[j]) * (wc[j + 1] + wc[j]);
            else
                wdwz[j] = flux3(wc[j - 1], wc[j], wc[j + 1], wc[j + 2], 0.5 * (rwc[j + 1] + rwc[j]), 1.0);
        }
        for (int k = 1; k < L; ++k) {
            double delta = (k == L - 1) ? -wdwz[L - 2] : wdwz[k] - wdwz[k - 1];
            size_t o = (size_t)c * (L + 1) + k;
            double inner = hadv[o] * m->invAreaCell[c] - rdzu[k] * delta;
            tend_w[o] = inner + tend_w_euler[o];
        }
    }
    free(hadv);

    /* theta: horizontal advection (:4312-4380, :5901-6014) */
    double *thadv = (double *)calloc(CL, sizeof(double));
    for (int c = 0; c < nC; ++c) {
        int n = m->nEdgesOnCell[c];
        for (int i = 0; i < n; ++i) {
            int e = m->edgesOnCell[c * ME + i];
            double sgn = m->edgesOnCell_sign[c * ME + i];
            int na = m->nAdvCellsForEdge[e];
            for (int k = 0; k < L; ++k) {
                double rue = ru[(size_t)e * L + k];
                double sr = sign_pm(rue);
                double fs = 0.0;
                for (int j = 0; j < na; ++j) {
                    int ac = m->advCellsForEdge[e * MA + j];
                    double wgt = m->adv_coefs[e * MA + j] + sr * m->adv_coefs_3rd[e * MA + j];
                    double cc = wgt * theta_m[(size_t)ac * L + k];
                    fs = (j == 0) ? cc : fs + cc;
                }
                double cc = -sgn * (rue * fs);
                double *acc = &thadv[(size_t)c * L + k];
                *acc = (i == 0) ? cc : *acc + cc;
            }
        }
        if (!rk1)
            for (int i = 0; i < n; ++i) {
                int e = m->edgesOnCell[c * ME + i];
                int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
                double sdv = m->edgesOnCell_sign[c * ME + i] * m->dvEdge[e];
                for (int k = 0; k < L; ++k) {
                    double face = 0.5 * (theta_m_save[(size_t)c2 * L + k] + theta_m_save[(size_t)c1 * L + k]);
                    double cc = sdv * (ru_save[(size_t)e * L + k] - ru[(size_t)e * L + k]) * face;
                    thadv[(size_t)c * L + k] = thadv[(size_t)c * L + k] - cc;
                }
            }
    }
    /* theta: horizontal mixing, first RK stage (:5563-5900) */
    if (rk1) {
        for (int c = 0; c < nC; ++c) {
            int n = m->nEdgesOnCell[c];
            for (int i = 0; i < n; ++i) {
                int e = m->edgesOnCell[c * ME + i];
                int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
                double es = m->invAreaCell[c] * m->edgesOnCell_sign[c * ME + i] * m->dvEdge[e] * m->invDcEdge[e];
                double prs = P->prandtl_inv * m->meshScalingDel2[e];
                for (int k = 0; k < L; ++k) {
                    double bf = es * (theta_m[(size_t)c2 * L + k] - theta_m[(size_t)c1 * L + k]) * rho_edge[(size_t)e * L + k];
                    double sf = bf * 0.5 * (kdiff[(size_t)c1 * L + k] + kdiff[(size_t)c2 * L + k]) * prs;
                    size_t o = (size_t)c * L + k;
                    delsq_theta[o] = (i == 0) ? bf : delsq_theta[o] + bf;
                    tend_theta_euler[o] = (i == 0) ? sf : tend_theta_euler[o] + sf;
                }
            }
        }
        if (P->h_theta_eddy_visc4 > 0.0)
            for (int c = 0; c < nC; ++c) {
                int n = m->nEdgesOnCell[c];
                double r_area = P->h_theta_eddy_visc4 * P->prandtl_inv * m->invAreaCell[c];
                for (int i = 0; i < n; ++i) {
                    int e = m->edgesOnCell[c * ME + i];
                    int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
                    double es = m->meshScalingDel4[e] * r_area * m->dvEdge[e] * m->edgesOnCell_sign[c * ME + i] * m->invDcEdge[e];
                    for (int k = 0; k < L; ++k) {
                        size_t o = (size_t)c * L + k;
                        tend_theta_euler[o] = tend_theta_euler[o] - es * (delsq_theta[(size_t)c2 * L + k] - delsq_theta[(size_t)c1 * L + k]);
                    }
                }
            }
        /* vertical mixing of theta (:5060-5097), levels 1..L-2, on theta_m or on
         * theta_m - t_init */
        if (P->v_theta_eddy_visc2 > 0.0 && L > 2) {
            double *tw = colbuf;
            for (int c = 0; c < nC; ++c) {
                const size_t oc = (size_t)c * L;
                const double *zg = zgrid + (size_t)c * (L + 1);
                for (int k = 0; k < L; ++k) tw[k] = P->mix_full ? theta_m[oc + k] : theta_m[oc + k] - t_init[oc + k];
                for (int k = 1; k + 1 < L; ++k) {
                    const double zm = 0.5 * (zg[k - 1] + zg[k]), z0 = 0.5 * (zg[k] + zg[k + 1]);
                    const double zp = 0.5 * (zg[k + 1] + zg[k + 2]);
                    const double diff = (tw[k + 1] - tw[k]) / (zp - z0) - (tw[k] - tw[k - 1]) / (z0 - zm);
                    tend_theta_euler[oc + k] = tend_theta_euler[oc + k] + P->v_theta_eddy_visc2 * P->prandtl_inv *
                                                                              rho_zz[oc + k] * diff / (0.5 * (zp - zm));
                }
            }
        }
    }
    /* theta: vertical transport, diabatic forcing and final (:4803-5057, :5097) */
    for (int c = 0; c < nC; ++c) {
        const size_t oc = (size_t)c * L, ow = (size_t)c * (L + 1);
        const double *th = theta_m + oc, *ths = theta_m_save + oc, *rwc = rw + ow, *rws = rw_save + ow;
        double *wdtz = colbuf; /* interface k: 1..L-1 */
        for (int k = 1; k < L; ++k) {
            if (k == 1)
                wdtz[k] = rwc[k] * (fzm[k] * th[k] + fzp[k] * th[k - 1]) +
                          (rws[k] - rwc[k]) * (fzm[k] * ths[k] + fzp[k] * ths[k - 1]);
            else if (k == L - 1)
                wdtz[k] = rws[k] * (fzm[k] * th[k] + fzp[k] * th[k - 1]);
            else
                wdtz[k] = flux3(th[k - 2], th[k - 1], th[k], th[k + 1], rwc[k], P->coef_3rd_order) +
                          (rws[k] - rwc[k]) * (fzm[k] * ths[k] + fzp[k] * ths[k - 1]);
        }
        for (int k = 0; k < L; ++k) {
            double delta;
            if (L == 1) delta = 0.0;
            else if (k == 0) delta = wdtz[1];
            else if (k == L - 1) delta = -wdtz[L - 1];
            else delta = wdtz[k + 1] - wdtz[k];
            double base = thadv[oc + k] * m->invAreaCell[c] - rdzw[k] * delta;
            rthdynten[oc + k] = (base - tend_rho[oc + k] * th[k]) / rho_zz[oc + k];
            double tt = base + rho_zz[oc + k] * rt_diabatic_tend[oc + k];
            tend_theta[oc + k] = tt + tend_theta_euler[oc + k];
        }
    }
    free(thadv);

    free(dpdz);
    free(delsq_u);
    free(delsq_div);
    free(delsq_vort);
    free(delsq_theta);
    free(delsq_w);
    free(colbuf);
}

/* ------------------------------------------------------------------------
 * atm_set_smlstep_pert_variables_work, smlstep_pert.py:197-249
 * ---------------------------------------------------------------------- */
void orc_smlstep_pert(const orc_mesh *m, const double *u_tend, double *w_tend) {
    const int L = m->nLevels, nC = m->nCells, ME = m->maxEdges;
    for (int c = 0; c < nC; ++c) {
        int n = m->nEdgesOnCell[c];
        for (int k = 1; k < L; ++k) {
            double wv = w_tend[(size_t)c * (L + 1) + k];
            for (int i = 0; i < n; ++i) {
                int e = m->edgesOnCell[c * ME + i];
                double uk = u_tend[(size_t)e * L + k];
                double ui = m->fzm[k] * uk + m->fzp[k] * u_tend[(size_t)e * L + k - 1];
                double flux = m->edgesOnCell_sign[c * ME + i] * ui;
                size_t zo = ((size_t)c * ME + i) * (L + 1) + k;
                double coef = m->zb_cell[zo] + sign_pm(uk) * m->zb3_cell[zo];
                wv = wv - coef * flux;
            }
            double zi = m->fzm[k] * m->zz[(size_t)c * L + k] + m->fzp[k] * m->zz[(size_t)c * L + k - 1];
            w_tend[(size_t)c * (L + 1) + k] = zi * wv;
        }
    }
}

/* ------------------------------------------------------------------------
 * atm_advance_acoustic_step_work, acoustic_step.py:456-889, 1297-1951
 * theta_m is the time-level-1 potential temperature (theta_m_init).
 * ---------------------------------------------------------------------- */
typedef struct orc_acoustic_params {
    int small_step;
    double dts, epssm, resm, c2;
} orc_acoustic_params;

int orc_acoustic_params_sizeof(void) { return (int)sizeof(orc_acoustic_params); }

void orc_acoustic_step(const orc_mesh *m, const orc_acoustic_params *P,
                       const double *cofrz, const double *cofwr, const double *cofwz,
                       const double *coftz, const double *cofwt, const double *a_tri,
                       const double *alpha_tri, const double *gamma_tri,
                       const double *rho_zz, const double *theta_m, const double *exner,
                       const double *cqu, const double *tend_ru, const double *tend_rho,
                       const double *tend_rt, const double *tend_rw, const double *w,
                       const double *rw, const double *rw_save,
                       double *ru_p, double *ruAvg, double *rho_pp, double *rtheta_pp,
                       double *rtheta_pp_old, double *rw_p, double *wwAvg) {
    const int L = m->nLevels, nC = m->nCells, nE = m->nEdges, ME = m->maxEdges;
    const double dts = P->dts, epssm = P->epssm, resm = P->resm, c2 = P->c2;
    const double *fzm = m->fzm, *fzp = m->fzp, *rdzw = m->rdzw, *zz = m->zz;
    if (P->small_step != 1) {
        for (int e = 0; e < nE; ++e) {
            int c1 = m->cellsOnEdge[2 * e], c2c = m->cellsOnEdge[2 * e + 1];
            for (int k = 0; k < L; ++k) {
                size_t o = (size_t)e * L + k;
                double pg = ((rtheta_pp[(size_t)c2c * L + k] - rtheta_pp[(size_t)c1 * L + k]) * m->invDcEdge[e]) /
                            (0.5 * (zz[(size_t)c2c * L + k] + zz[(size_t)c1 * L + k]));
                pg = cqu[o] * 0.5 * c2 * (exner[(size_t)c1 * L + k] + exner[(size_t)c2c * L + k]) * pg;
                pg = pg + 0.5 * m->zxu[o] * GRAVITY * (rho_pp[(size_t)c1 * L + k] + rho_pp[(size_t)c2c * L + k]);
                ru_p[o] = ru_p[o] + dts * (tend_ru[o] - pg);
                ruAvg[o] = ruAvg[o] + ru_p[o];
            }
        }
    } else {
        for (int e = 0; e < nE; ++e)
            for (int k = 0; k < L; ++k) {
                size_t o = (size_t)e * L + k;
                ru_p[o] = dts * tend_ru[o];
                ruAvg[o] = ru_p[o];
            }
    }
    for (int c = 0; c < nC; ++c)
        for (int k = 0; k < L; ++k)
            rtheta_pp_old[(size_t)c * L + k] = (P->small_step == 1) ? 0.0 : rtheta_pp[(size_t)c * L + k];

    double *rs = (double *)malloc(sizeof(double) * (size_t)L * 2);
    double *ts = rs + L;
    for (int c = 0; c < nC; ++c) {
        const size_t oc = (size_t)c * L, ow = (size_t)c * (L + 1);
        if (P->small_step == 1) {
            for (int k = 0; k <= L; ++k) {
                wwAvg[ow + k] = 0.0;
                rw_p[ow + k] = 0.0;
            }
            for (int k = 0; k < L; ++k) {
                rho_pp[oc + k] = 0.0;
                rtheta_pp[oc + k] = 0.0;
            }
        }
        int n = m->nEdgesOnCell[c];
        for (int k = 0; k < L; ++k) {
            double rsum = 0.0, tsum = 0.0;
            for (int i = 0; i < n; ++i) {
                int e = m->edgesOnCell[c * ME + i];
                int c1 = m->cellsOnEdge[2 * e], c2c = m->cellsOnEdge[2 * e + 1];
                double metric = m->edgesOnCell_sign[c * ME + i] * m->dvEdge[e];
                double flux = dts * metric * ru_p[(size_t)e * L + k] * m->invAreaCell[c];
                double thedge = 0.5 * (theta_m[(size_t)c2c * L + k] + theta_m[(size_t)c1 * L + k]);
                double rcon = -flux, tcon = -(flux * thedge);
                rsum = (i == 0) ? rcon : rsum + rcon;
                tsum = (i == 0) ? tcon : tsum + tcon;
            }
            rs[k] = rho_pp[oc + k] + dts * tend_rho[oc + k] + rsum - cofrz[k] * resm * (rw_p[ow + k + 1] - rw_p[ow + k]);
            ts[k] = rtheta_pp[oc + k] + dts * tend_rt[oc + k] + tsum -
                    resm * rdzw[k] * (coftz[ow + k + 1] * rw_p[ow + k + 1] - coftz[ow + k] * rw_p[ow + k]);
        }
        for (int k = 1; k < L; ++k)
            wwAvg[ow + k] = wwAvg[ow + k] + 0.5 * (1.0 - epssm) * rw_p[ow + k];
        /* right-hand side of the implicit rw_p equation (:719-755) */
        for (int k = 1; k < L; ++k) {
            double r = rw_p[ow + k] + dts * tend_rw[ow + k] -
                       cofwz[oc + k] * ((zz[oc + k] * ts[k] - zz[oc + k - 1] * ts[k - 1]) +
                                        resm * (zz[oc + k] * rtheta_pp[oc + k] - zz[oc + k - 1] * rtheta_pp[oc + k - 1])) -
                       cofwr[oc + k] * ((rs[k] + rs[k - 1]) + resm * (rho_pp[oc + k] + rho_pp[oc + k - 1])) +
                       cofwt[oc + k] * (ts[k] + resm * rtheta_pp[oc + k]) +
                       cofwt[oc + k - 1] * (ts[k - 1] + resm * rtheta_pp[oc + k - 1]);
            rw_p[ow + k] = r; /* level k's right-hand side reads only its own rw_p */
        }
        /* Thomas sweeps (tridiagonal_solve.py:50-63) */
        for (int k = 1; k < L; ++k)
            rw_p[ow + k] = (rw_p[ow + k] - a_tri[oc + k] * rw_p[ow + k - 1]) * alpha_tri[oc + k];
        for (int k = L - 1; k >= 0; --k)
            rw_p[ow + k] = rw_p[ow + k] - gamma_tri[oc + k] * rw_p[ow + k + 1];
        /* implicit Rayleigh damping of w and wwAvg (:838-870) */
        for (int k = 1; k < L; ++k) {
            double zi = fzm[k] * zz[oc + k] + fzp[k] * zz[oc + k - 1];
            double ri = fzm[k] * rho_zz[oc + k] + fzp[k] * rho_zz[oc + k - 1];
            double rwd = rw_save[ow + k] - rw[ow + k];
            double rr = (rw_p[ow + k] + rwd - dts * m->dss[oc + k] * zi * ri * w[ow + k]) / (1.0 + dts * m->dss[oc + k]) - rwd;
            rw_p[ow + k] = rr;
            wwAvg[ow + k] = wwAvg[ow + k] + 0.5 * (1.0 + epssm) * rr;
        }
        for (int k = 0; k < L; ++k) {
            rho_pp[oc + k] = rs[k] - cofrz[k] * (rw_p[ow + k + 1] - rw_p[ow + k]);
            rtheta_pp[oc + k] = ts[k] - rdzw[k] * (coftz[ow + k + 1] * rw_p[ow + k + 1] - coftz[ow + k] * rw_p[ow + k]);
        }
    }
    free(rs);
}

/* ------------------------------------------------------------------------
 * atm_divergence_damping_3d, divergence_damping.py:187-209, coefficient :267
 * theta_m is the time-level-1 potential temperature (theta_m_init).
 * ---------------------------------------------------------------------- */
void orc_divergence_damping(const orc_mesh *m, double dts, double smdiv, double len_disp,
                            const double *theta_m, const double *rtheta_pp,
                            const double *rtheta_pp_old, double *ru_p) {
    const int L = m->nLevels, nE = m->nEdges;
    const double coef = 2.0 * smdiv * len_disp / dts;
    for (int e = 0; e < nE; ++e) {
        int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
        for (int k = 0; k < L; ++k) {
            double d1 = -(rtheta_pp[(size_t)c1 * L + k] - rtheta_pp_old[(size_t)c1 * L + k]);
            double d2 = -(rtheta_pp[(size_t)c2 * L + k] - rtheta_pp_old[(size_t)c2 * L + k]);
            double ts = theta_m[(size_t)c1 * L + k] + theta_m[(size_t)c2 * L + k];
            double damp = (fabs(ts) >= 1.0e-30) ? coef * (d2 - d1) / ts : 0.0;
            ru_p[(size_t)e * L + k] = ru_p[(size_t)e * L + k] + damp;
        }
    }
}

/* ------------------------------------------------------------------------
 * atm_recover_large_step_variables_work, recover_large_step.py:311-711
 * ---------------------------------------------------------------------- */
void orc_recover_large_step(const orc_mesh *m, int rk_step, double dt, int ns,
                            const double *rho_p_save, const double *rho_pp,
                            const double *rtheta_p_save, const double *rtheta_pp,
                            const double *rt_diabatic_tend, const double *rw_save,
                            const double *rw_p, const double *ru_save, const double *ru_p,
                            double *wwAvg, double *ruAvg,
                            double *rho_p, double *rho_zz, double *rtheta_p, double *theta_m,
                            double *exner, double *pressure_p, double *rw, double *w,
                            double *ru, double *u) {
    const int L = m->nLevels, nC = m->nCells, nE = m->nEdges, ME = m->maxEdges;
    const double invNs = 1.0 / (double)ns;
    const double rcv = RGAS / (CP - RGAS);
    const double *fzm = m->fzm, *fzp = m->fzp, *zz = m->zz;
    for (int c = 0; c < nC; ++c)
        for (int k = 0; k < L; ++k) {
            size_t o = (size_t)c * L + k;
            double rp = rho_p_save[o] + rho_pp[o];
            double rz = rp + m->rho_base[o];
            double rtp;
            if (rk_step == 3)
                rtp = rtheta_p_save[o] + rtheta_pp[o] - dt * rz * rt_diabatic_tend[o];
            else
                rtp = rtheta_p_save[o] + rtheta_pp[o];
            rho_p[o] = rp;
            rho_zz[o] = rz;
            rtheta_p[o] = rtp;
            theta_m[o] = (rtp + m->rtheta_base[o]) / rz;
            if (rk_step == 3) {
                double ex = pow(zz[o] * (RGAS / P0) * (rtp + m->rtheta_base[o]), rcv);
                exner[o] = ex;
                pressure_p[o] = zz[o] * RGAS * (ex * rtp + m->rtheta_base[o] * (ex - m->exner_base[o]));
            }
        }
    /* padding-row sentinel (recover_large_step.py:985-998) */
    for (int k = 0; k < L; ++k) rho_zz[(size_t)nC * L + k] = 1.0;
    for (int e = 0; e < nE; ++e) {
        int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
        for (int k = 0; k < L; ++k) {
            size_t o = (size_t)e * L + k;
            ruAvg[o] = ru_save[o] + (ruAvg[o] * invNs);
            double rn = ru_save[o] + ru_p[o];
            ru[o] = rn;
            u[o] = 2.0 * rn / (rho_zz[(size_t)c1 * L + k] + rho_zz[(size_t)c2 * L + k]);
        }
    }
    for (int c = 0; c < nC; ++c) {
        const size_t oc = (size_t)c * L, ow = (size_t)c * (L + 1);
        int n = m->nEdgesOnCell[c];
        for (int k = 1; k < L; ++k) {
            wwAvg[ow + k] = rw_save[ow + k] + (wwAvg[ow + k] * invNs);
            rw[ow + k] = rw_save[ow + k] + rw_p[ow + k];
        }
        /* surface */
        double sacc = 0.0;
        for (int i = 0; i < n; ++i) {
            int e = m->edgesOnCell[c * ME + i];
            const double *re = ru + (size_t)e * L;
            double fs = m->cf1 * re[0] + m->cf2 * re[1] + m->cf3 * re[2];
            size_t zo = ((size_t)c * ME + i) * (L + 1);
            double coef = m->zb_cell[zo] + sign_pm(fs) * m->zb3_cell[zo];
            double con = m->edgesOnCell_sign[c * ME + i] * coef * fs;
            sacc = (i == 0) ? con : sacc + con;
        }
        w[ow + 0] = sacc / (m->cf1 * rho_zz[oc + 0] + m->cf2 * rho_zz[oc + 1] + m->cf3 * rho_zz[oc + 2]);
        for (int k = 1; k < L; ++k) {
            double zi = fzm[k] * zz[oc + k] + fzp[k] * zz[oc + k - 1];
            double acc = rw[ow + k] / zi;
            for (int i = 0; i < n; ++i) {
                int e = m->edgesOnCell[c * ME + i];
                const double *re = ru + (size_t)e * L;
                double fi = fzm[k] * re[k] + fzp[k] * re[k - 1];
                size_t zo = ((size_t)c * ME + i) * (L + 1) + k;
                double coef = m->zb_cell[zo] + sign_pm(fi) * m->zb3_cell[zo];
                acc = acc + m->edgesOnCell_sign[c * ME + i] * coef * fi;
            }
            w[ow + k] = acc / (fzm[k] * rho_zz[oc + k] + fzp[k] * rho_zz[oc + k - 1]);
        }
    }
}

/* ------------------------------------------------------------------------
 * Pattern-C high-order edge value, pattern_c.py:582-733 (slot order)
 * ---------------------------------------------------------------------- */
static double patternc_value(const orc_mesh *m, int e, double sgn, const double *q, int L, int S, int k, int s) {
    int na = m->nAdvCellsForEdge[e];
    double f = 0.0;
    for (int j = 0; j < na; ++j) {
        int ac = m->advCellsForEdge[e * m->maxAdv + j];
        double wgt = m->adv_coefs[e * m->maxAdv + j] + sgn * m->adv_coefs_3rd[e * m->maxAdv + j];
        double cc = wgt * q[((size_t)ac * L + k) * S + s];
        f = (j == 0) ? cc : f + cc;
    }
    return f;
}

/* ------------------------------------------------------------------------
 * atm_advance_scalars_work (explicit RK stages), advance_scalars.py:128-700
 * ---------------------------------------------------------------------- */
void orc_advance_scalars(const orc_mesh *m, int nScalars, double dt, double coef3,
                         double weight_old, double weight_new,
                         const double *scalar_old, const double *scalar_new_in,
                         const double *rho_zz_old, const double *rho_zz_new,
                         const double *uhAvg, const double *wwAvg,
                         double *scalar_new_out, double *horiz_flux) {
    const int L = m->nLevels, nC = m->nCells, nE = m->nEdges, ME = m->maxEdges, S = nScalars;
    for (int e = 0; e < nE; ++e)
        for (int k = 0; k < L; ++k) {
            double sgn = sign_pm(uhAvg[(size_t)e * L + k]);
            for (int s = 0; s < S; ++s)
                horiz_flux[((size_t)e * L + k) * S + s] = patternc_value(m, e, sgn, scalar_new_in, L, S, k, s);
        }
    double *wdtn = (double *)malloc(sizeof(double) * (size_t)(L + 1));
    for (int c = 0; c < nC; ++c) {
        const size_t oc = (size_t)c * L, ow = (size_t)c * (L + 1);
        int n = m->nEdgesOnCell[c];
        for (int s = 0; s < S; ++s) {
            const double *sn = scalar_new_in;
#define SN(kk) sn[((size_t)c * L + (kk)) * S + s]
            for (int k = 0; k <= L; ++k) wdtn[k] = 0.0;
            if (L >= 2) wdtn[1] = wwAvg[ow + 1] * (m->fnm[1] * SN(1) + m->fnp[1] * SN(0));
            for (int k = 2; k <= L - 2; ++k)
                wdtn[k] = flux3(SN(k - 2), SN(k - 1), SN(k), SN(k + 1), wwAvg[ow + k], coef3);
            if (L > 2) wdtn[L - 1] = wwAvg[ow + L - 1] * (m->fnm[L - 1] * SN(L - 1) + m->fnp[L - 1] * SN(L - 2));
#undef SN
            for (int k = 0; k < L; ++k) {
                double hd = 0.0;
                for (int i = 0; i < n; ++i) {
                    int e = m->edgesOnCell[c * ME + i];
                    double cc = -m->edgesOnCell_sign[c * ME + i] * uhAvg[(size_t)e * L + k] *
                                horiz_flux[((size_t)e * L + k) * S + s];
                    hd = (i == 0) ? cc : hd + cc;
                }
                double tend_col = hd * m->invAreaCell[c] + 0.0;
                double vdiv;
                if (k == 0) vdiv = m->rdnw[0] * wdtn[1];
                else if (k == L - 1) vdiv = m->rdnw[L - 1] * (-wdtn[L - 1]);
                else vdiv = m->rdnw[k] * (wdtn[k + 1] - wdtn[k]);
                double rho_i = weight_old * rho_zz_old[oc + k] + weight_new * rho_zz_new[oc + k];
                double inv = 1.0 / rho_i;
                scalar_new_out[(oc + k) * S + s] =
                    (scalar_old[(oc + k) * S + s] * rho_zz_old[oc + k] + dt * (tend_col - vdiv)) * inv;
            }
        }
    }
    free(wdtn);
}

/* ------------------------------------------------------------------------
 * atm_advance_scalars_mono_work (monotonic FCT), advance_scalars_mono.py:
 * 3332-4037 and the tile numeric :1476-1818 with helpers :145-422, :583,
 * :1402-1473.  scalars_old is advanced by dt*scalar_tend/rho_zz_old (zero
 * tendency here), eps = 1e-20.
 * ---------------------------------------------------------------------- */
void orc_advance_scalars_mono(const orc_mesh *m, int nScalars, double dt, double coef3,
                              const double *scalars_old, const double *scalars_new_in,
                              const double *rho_zz_old, const double *uhAvg,
                              const double *wwAvg, double *scalars_new_out,
                              double *rho_zz_int) {
    const int L = m->nLevels, nC = m->nCells, nE = m->nEdges, ME = m->maxEdges, S = nScalars;
    const double eps = 1.0e-20;
    const size_t CL = (size_t)(nC + 1) * L, EL = (size_t)(nE + 1) * L;
    /* density: rho_zz_int = rho_old + dt*(horiz_div - vert_div) */
    for (int c = 0; c < nC; ++c) {
        int n = m->nEdgesOnCell[c];
        const size_t oc = (size_t)c * L, ow = (size_t)c * (L + 1);
        for (int k = 0; k < L; ++k) {
            double hd = 0.0;
            for (int i = 0; i < n; ++i) {
                int e = m->edgesOnCell[c * ME + i];
                double metric = -m->edgesOnCell_sign[c * ME + i] * m->dvEdge[e];
                double base = uhAvg[(size_t)e * L + k] * metric;
                double cc = base * m->invAreaCell[c];
                hd = (i == 0) ? cc : hd + cc;
            }
            double vd = m->rdnw[k] * (wwAvg[ow + k + 1] - wwAvg[ow + k]);
            rho_zz_int[oc + k] = rho_zz_old[oc + k] + dt * (hd - vd);
        }
    }
    double *so = (double *)malloc(sizeof(double) * CL);
    double *smax = (double *)malloc(sizeof(double) * CL);
    double *smin = (double *)malloc(sizeof(double) * CL);
    double *wdtn = (double *)calloc((size_t)(nC + 1) * (L + 1), sizeof(double));
    double *wmid = (double *)calloc((size_t)(nC + 1) * (L + 1), sizeof(double));
    double *scell = (double *)malloc(sizeof(double) * CL);
    double *sin_ = (double *)malloc(sizeof(double) * CL);
    double *sout = (double *)malloc(sizeof(double) * CL);
    double *fedge = (double *)malloc(sizeof(double) * EL);
    double *fup = (double *)malloc(sizeof(double) * EL);
    double *ftmp = (double *)malloc(sizeof(double) * EL);
    for (int s = 0; s < S; ++s) {
        for (int c = 0; c < nC; ++c)
            for (int k = 0; k < L; ++k) {
                size_t o = (size_t)c * L + k;
                so[o] = scalars_old[o * S + s] + dt * 0.0 / rho_zz_old[o];
            }
#define SNV(cc, kk) scalars_new_in[((size_t)(cc) * L + (kk)) * S + s]
        /* local vertical bounds and high-order vertical flux */
        for (int c = 0; c < nC; ++c) {
            const size_t oc = (size_t)c * L, ow = (size_t)c * (L + 1);
            const double *sc = so + oc;
            smax[oc + 0] = dmax(sc[0], sc[1]);
            smin[oc + 0] = dmin(sc[0], sc[1]);
            smax[oc + 1] = dmax(dmax(sc[0], sc[1]), sc[2]);
            smin[oc + 1] = dmin(dmin(sc[0], sc[1]), sc[2]);
            for (int k = 0; k <= L; ++k) wdtn[ow + k] = 0.0;
            wdtn[ow + 1] = wwAvg[ow + 1] * (m->fnm[1] * SNV(c, 1) + m->fnp[1] * SNV(c, 0));
            for (int k = 2; k <= L - 2; ++k) {
                wdtn[ow + k] = flux3(SNV(c, k - 2), SNV(c, k - 1), SNV(c, k), SNV(c, k + 1), wwAvg[ow + k], coef3);
                smax[oc + k] = dmax(dmax(sc[k - 1], sc[k]), sc[k + 1]);
                smin[oc + k] = dmin(dmin(sc[k - 1], sc[k]), sc[k + 1]);
            }
            int kl = L - 1;
            wdtn[ow + kl] = wwAvg[ow + kl] * (m->fnm[kl] * SNV(c, kl) + m->fnp[kl] * SNV(c, kl - 1));
            smax[oc + kl] = dmax(sc[kl], sc[kl - 1]);
            smin[oc + kl] = dmin(sc[kl], sc[kl - 1]);
            int n = m->nEdgesOnCell[c];
            for (int k = 0; k < L; ++k) {
                double mx = smax[oc + k], mn = smin[oc + k];
                for (int i = 0; i < n; ++i) {
                    int nb = m->cellsOnCell[c * ME + i];
                    mx = dmax(mx, so[(size_t)nb * L + k]);
                    mn = dmin(mn, so[(size_t)nb * L + k]);
                }
                smax[oc + k] = mx;
                smin[oc + k] = mn;
            }
        }
        /* horizontal fluxes: high order, upwind, antidiffusive */
        for (int e = 0; e < nE; ++e) {
            int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
            for (int k = 0; k < L; ++k) {
                size_t o = (size_t)e * L + k;
                double uh = uhAvg[o];
                double hv = patternc_value(m, e, sign_pm(uh), scalars_new_in, L, S, k, s);
                fedge[o] = uh * hv;
                double up = (uh >= 0.0) ? so[(size_t)c1 * L + k] : so[(size_t)c2 * L + k];
                fup[o] = m->dvEdge[e] * dt * (uh * up);
                ftmp[o] = dt * fedge[o] - fup[o];
            }
        }
        /* vertical upwind update and limiter scale accumulators */
        for (int c = 0; c < nC; ++c) {
            const size_t oc = (size_t)c * L, ow = (size_t)c * (L + 1);
            double *sr = scell + oc;
            for (int k = 0; k < L; ++k) sr[k] = so[oc + k] * rho_zz_old[oc + k];
            /* flux_upwind_mid at interfaces 1..L-1 stored in wmid */
            for (int k = 1; k < L; ++k) {
                double vel = wwAvg[ow + k];
                double val = (vel >= 0.0) ? so[oc + k - 1] : so[oc + k];
                wmid[ow + k] = dt * (vel * val);
            }
            for (int k = 0; k < L - 1; ++k) sr[k] = sr[k] - wmid[ow + k + 1] * m->rdnw[k];
            for (int k = 1; k < L; ++k) sr[k] = sr[k] + wmid[ow + k] * m->rdnw[k];
            /* antidiffusive vertical flux: wdtn_mid = dt*wdtn - flux_upwind */
            for (int k = 1; k < L; ++k) wmid[ow + k] = dt * wdtn[ow + k] - wmid[ow + k];
            for (int k = 0; k < L; ++k) {
                double prev = (k == 0) ? wdtn[ow + 0] : wmid[ow + k];
                double next = (k == L - 1) ? wdtn[ow + L] : wmid[ow + k + 1];
                sin_[oc + k] = -m->rdnw[k] * ((-max0(-next)) - max0(prev));
                sout[oc + k] = -m->rdnw[k] * (max0(next) - (-max0(-prev)));
            }
            int n = m->nEdgesOnCell[c];
            for (int i = 0; i < n; ++i) {
                int e = m->edgesOnCell[c * ME + i];
                double sgn = m->edgesOnCell_sign[c * ME + i];
                double w_area = m->invAreaCell[c];
                double signed_inv = sgn * w_area;
                for (int k = 0; k < L; ++k) {
                    size_t o = (size_t)e * L + k;
                    double sc_ = fup[o] * signed_inv;
                    double sf = sgn * ftmp[o];
                    double outc = max0(sf) * w_area;
                    double inc = (-max0(-sf)) * w_area;
                    sr[k] = sr[k] - sc_;
                    sout[oc + k] = sout[oc + k] - outc;
                    sin_[oc + k] = sin_[oc + k] - inc;
                }
            }
            for (int k = 0; k < L; ++k) {
                double fi = (smax[oc + k] * rho_zz_int[oc + k] - sr[k]) / (sin_[oc + k] + eps);
                double fo = (smin[oc + k] * rho_zz_int[oc + k] - sr[k]) / (sout[oc + k] - eps);
                sin_[oc + k] = dmin(dmax(fi, 0.0), 1.0);
                sout[oc + k] = dmin(dmax(fo, 0.0), 1.0);
            }
        }
        /* limited horizontal fluxes */
        for (int e = 0; e < nE; ++e) {
            int c1 = m->cellsOnEdge[2 * e], c2 = m->cellsOnEdge[2 * e + 1];
            for (int k = 0; k < L; ++k) {
                size_t o = (size_t)e * L + k;
                double f = ftmp[o];
                double lim = (f >= 0.0) ? dmin(sout[(size_t)c1 * L + k], sin_[(size_t)c2 * L + k])
                                        : dmin(sin_[(size_t)c1 * L + k], sout[(size_t)c2 * L + k]);
                fedge[o] = f * lim;
            }
        }
        /* limited vertical fluxes, final update and positivity */
        for (int c = 0; c < nC; ++c) {
            const size_t oc = (size_t)c * L, ow = (size_t)c * (L + 1);
            double *sr = scell + oc;
            for (int k = 1; k < L; ++k) {
                double f = wmid[ow + k];
                double lim = (f >= 0.0) ? dmin(sout[oc + k - 1], sin_[oc + k]) : dmin(sout[oc + k], sin_[oc + k - 1]);
                wmid[ow + k] = f * lim;
            }
            int n = m->nEdgesOnCell[c];
            for (int i = 0; i < n; ++i) {
                int e = m->edgesOnCell[c * ME + i];
                double signed_inv = m->edgesOnCell_sign[c * ME + i] * m->invAreaCell[c];
                for (int k = 0; k < L; ++k) sr[k] = sr[k] - fedge[(size_t)e * L + k] * signed_inv;
            }
            for (int k = 0; k < L; ++k) {
                double lo = (k == 0) ? wdtn[ow + 0] : wmid[ow + k];
                double hi = (k == L - 1) ? wdtn[ow + L] : wmid[ow + k + 1];
                double delta = hi - lo;
                double st2 = (sr[k] - delta * m->rdnw[k]) / rho_zz_int[oc + k];
                scalars_new_out[(oc + k) * S + s] = max0(st2);
            }
        }
#undef SNV
    }
    free(so);
    free(smax);
    free(smin);
    free(wdtn);
    free(wmid);
    free(scell);
    free(sin_);
    free(sout);
    free(fedge);
    free(fup);
    free(ftmp);
}

/* ------------------------------------------------------------------------
 * mpas_reconstruct_2d, vector_reconstruction.py:311-441 (on a sphere)
 * ---------------------------------------------------------------------- */
void orc_reconstruct(const orc_mesh *m, const double *u, double *ux, double *uy, double *uz,
                     double *uzonal, double *umerid) {
    const int L = m->nLevels, nC = m->nCells, ME = m->maxEdges;
    for (int c = 0; c < nC; ++c) {
        int n = m->nEdgesOnCell[c];
        double clat = cos(m->latCell[c]), slat = sin(m->latCell[c]);
        double clon = cos(m->lonCell[c]), slon = sin(m->lonCell[c]);
        for (int k = 0; k < L; ++k) {
            double x = 0.0, y = 0.0, z = 0.0;
            for (int i = 0; i < n; ++i) {
                int e = m->edgesOnCell[c * ME + i];
                const double *cf = m->coeffs_reconstruct + ((size_t)c * ME + i) * 3;
                double ue = u[(size_t)e * L + k];
                double cx = cf[0] * ue, cy = cf[1] * ue, cz = cf[2] * ue;
                x = (i == 0) ? cx : x + cx;
                y = (i == 0) ? cy : y + cy;
                z = (i == 0) ? cz : z + cz;
            }
            size_t o = (size_t)c * L + k;
            ux[o] = x;
            uy[o] = y;
            uz[o] = z;
            uzonal[o] = -x * slon + y * clon;
            umerid[o] = -(x * clon + y * slon) * slat + z * clat;
        }
    }
}
