// Adjoint (vector-Jacobian) bodies of the split-explicit step's kernels.
//
// Each forward kernel F(inputs) -> outputs of the large step has an adjoint
// that reads the forward inputs and outputs of one call (fwd[]), reads the
// adjoints of its outputs and accumulates the adjoints of its inputs
// (bar[], same slot order as fwd[]).  Derivatives are those of the hard
// primal the forward evaluates, with the conventions of torch's autograd that
// the reference differentiates (mpas_pytorch/autodiff/soft_limiter.py:72-186):
// d|x|/dx = sign(x) with 0 at 0, min/max send the adjoint to the selected
// operand and split it evenly on a tie, sign selectors and branch masks carry
// no derivative.
//
// Output adjoints may be used as accumulators for contributions a kernel adds
// to its own outputs (an output read again inside the same call); the tape
// consumes each output adjoint exactly once, so this is safe.
//
// The including file defines acc(p, v) (add v to *p, atomically on the GPU)
// before including this header.  A null bar[] slot means "no adjoint wanted"
// and is skipped.
#pragma once

#include "../../include/mpas_hip.h"
#include "scalar_math.hpp"

namespace mpas {
namespace adj {

using Fwd = const double *const *;
using Bar = double *const *;

constexpr int kMaxEdgesAdj = 8;  // cell slot cap (meshes beyond it are rejected on entry)

__device__ __forceinline__ void add(Bar B, int slot, long long i, double v) {
    if (B[slot]) acc(B[slot] + i, v);
}
// For an element only this thread touches during the launch.
__device__ __forceinline__ void add_own(Bar B, int slot, long long i, double v) {
    if (B[slot]) B[slot][i] += v;
}
__device__ __forceinline__ double bar_of(Bar B, int slot, long long i) { return B[slot] ? B[slot][i] : 0.0; }

// sign of edge e in the edge list of cell c
__device__ __forceinline__ double edge_sign_in_cell(const mpas_hip_mesh &m, int c, int e) {
    const int n = m.nEdgesOnCell[c];
    for (int j = 0; j < n; ++j)
        if (m.edgesOnCell[c * m.maxEdges + j] == e) return m.edgesOnCell_sign[c * m.maxEdges + j];
    return 0.0;
}

// d|x|/dx as torch.abs differentiates it
__device__ __forceinline__ double sgn0(double x) { return (x > 0.0) ? 1.0 : ((x < 0.0) ? -1.0 : 0.0); }

// Partial derivatives of flux3(qm2, qm1, q0, qp1, ua, c3): d[0..3] for the q
// arguments, d[4] for ua.
__device__ __forceinline__ void flux3_partials(double qm2, double qm1, double q0, double qp1, double ua,
                                               double c3, double d[5]) {
    const double au = fabs(ua);
    d[0] = (-ua - c3 * au) / 12.0;
    d[1] = (7.0 * ua + 3.0 * c3 * au) / 12.0;
    d[2] = (7.0 * ua - 3.0 * c3 * au) / 12.0;
    d[3] = (-ua + c3 * au) / 12.0;
    d[4] = (7.0 * (q0 + qm1) - (qp1 + qm2)) / 12.0 + c3 * sgn0(ua) * ((qp1 - qm2) - 3.0 * (q0 - qm1)) / 12.0;
}

// min(a, b) / max(a, b): the adjoint goes to the selected operand, half each on a tie
__device__ __forceinline__ void min_partials(double a, double b, double &da, double &db) {
    if (a < b) { da = 1.0; db = 0.0; }
    else if (b < a) { da = 0.0; db = 1.0; }
    else { da = 0.5; db = 0.5; }
}
__device__ __forceinline__ void max_partials(double a, double b, double &da, double &db) {
    if (a > b) { da = 1.0; db = 0.0; }
    else if (b > a) { da = 0.0; db = 1.0; }
    else { da = 0.5; db = 0.5; }
}

// ===========================================================================
// atm_set_smlstep_pert_variables_work (smlstep_pert.py:197-249)
// fwd: u_tend, w_tend (before), w_tend (after)
// ===========================================================================
namespace smlstep {
enum { u_tend, w_tend_in, w_tend_out, N };
}

// Gather form (no atomics): the cell pass keeps the w_tend part (thread (cell
// c, interface k), c in [0, nCells], k in [0, L]); the edge pass sums what the
// edge's two cells send its level k from interfaces k and k+1.
__device__ __forceinline__ void smlstep_cells_vjp(const mpas_hip_mesh &m, Bar B, int c, int k) {
    using namespace smlstep;
    const int L = m.nLevels, Lw = L + 1;
    const long long i = (long long)c * Lw + k;
    const double g = bar_of(B, w_tend_out, i);
    if (c >= m.nCells || k < 1 || k >= L) {
        add_own(B, w_tend_in, i, g);
        return;
    }
    add_own(B, w_tend_in, i, (m.fzm[k] * m.zz[c * L + k] + m.fzp[k] * m.zz[c * L + k - 1]) * g);
}
// gf of cell c, slot j (edge e) at interface kk in 1..L-1
__device__ __forceinline__ double smlstep_gf(const mpas_hip_mesh &m, Fwd F, Bar B, int c, int j, int e, int kk) {
    using namespace smlstep;
    const int L = m.nLevels, ME = m.maxEdges, Lw = L + 1;
    const double zi = m.fzm[kk] * m.zz[c * L + kk] + m.fzp[kk] * m.zz[c * L + kk - 1];
    const double gw = zi * bar_of(B, w_tend_out, (long long)c * Lw + kk);
    const double uk = F[u_tend][(long long)e * L + kk];
    const long long zo = ((long long)c * ME + j) * Lw + kk;
    const double coef = m.zb_cell[zo] + sign_pm(uk) * m.zb3_cell[zo];
    return -coef * m.edgesOnCell_sign[c * ME + j] * gw;
}
__device__ __forceinline__ void smlstep_edges_vjp(const mpas_hip_mesh &m, Fwd F, Bar B, int e, int k) {
    using namespace smlstep;
    const int L = m.nLevels, ME = m.maxEdges;
    if (e >= m.nEdges || k >= L) return;
    double gu = 0.0;
    for (int q = 0; q < 2; ++q) {
        const int c = m.cellsOnEdge[2 * e + q];
        if (c >= m.nCells) continue;
        const int n = m.nEdgesOnCell[c];
        int j = 0;
        while (j < n && m.edgesOnCell[c * ME + j] != e) ++j;
        if (j == n) continue;
        if (k >= 1) gu += m.fzm[k] * smlstep_gf(m, F, B, c, j, e, k);
        if (k + 1 < L) gu += m.fzp[k + 1] * smlstep_gf(m, F, B, c, j, e, k + 1);
    }
    add_own(B, u_tend, (long long)e * L + k, gu);
}

// ===========================================================================
// atm_divergence_damping_3d (divergence_damping.py:187-209, coefficient :267)
// fwd: theta_m, rtheta_pp, rtheta_pp_old, ru_p (before), ru_p (after)
// ===========================================================================
namespace divdamp {
enum { theta_m, rtheta_pp, rtheta_pp_old, ru_p_in, ru_p_out, N };
}

// thread (edge e, level k), e in [0, nEdges]
// Gather form of the cell adjoints below, thread (cell c, level k): the sum
// over the cell's edges of what divdamp_vjp scatters to that cell (used with
// divdamp_vjp(..., cells = false), which then keeps the edge part only).
__device__ __forceinline__ void divdamp_cells_vjp(const mpas_hip_mesh &m, double coef, Fwd F, Bar B, int c, int k) {
    using namespace divdamp;
    const int L = m.nLevels, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const int n = m.nEdgesOnCell[c];
    double grt = 0.0, gold = 0.0, gth = 0.0;
    for (int j = 0; j < n; ++j) {
        const int e = m.edgesOnCell[c * ME + j];
        const double g = bar_of(B, ru_p_out, (long long)e * L + k);
        if (g == 0.0) continue;
        const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
        const long long o1 = (long long)c1 * L + k, o2 = (long long)c2 * L + k;
        const double d1 = -(F[rtheta_pp][o1] - F[rtheta_pp_old][o1]);
        const double d2 = -(F[rtheta_pp][o2] - F[rtheta_pp_old][o2]);
        const double ts = F[theta_m][o1] + F[theta_m][o2];
        if (!(fabs(ts) >= 1.0e-30)) continue;
        const double gd = coef * g / ts;
        const double gts = -(coef * (d2 - d1) / ts) / ts * g;
        const double sg = (c == c2) ? -1.0 : 1.0;
        grt += sg * gd;
        gold -= sg * gd;
        gth += gts;
    }
    const long long o = (long long)c * L + k;
    add_own(B, rtheta_pp, o, grt);
    add_own(B, rtheta_pp_old, o, gold);
    add_own(B, theta_m, o, gth);
}

__device__ __forceinline__ void divdamp_vjp(const mpas_hip_mesh &m, double coef, Fwd F, Bar B, int e, int k,
                                            bool cells = true) {
    using namespace divdamp;
    const int L = m.nLevels;
    const long long i = (long long)e * L + k;
    const double g = bar_of(B, ru_p_out, i);
    add_own(B, ru_p_in, i, g);
    if (!cells || e >= m.nEdges || g == 0.0) return;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const long long o1 = (long long)c1 * L + k, o2 = (long long)c2 * L + k;
    const double d1 = -(F[rtheta_pp][o1] - F[rtheta_pp_old][o1]);
    const double d2 = -(F[rtheta_pp][o2] - F[rtheta_pp_old][o2]);
    const double ts = F[theta_m][o1] + F[theta_m][o2];
    if (!(fabs(ts) >= 1.0e-30)) return;
    const double gd = coef * g / ts;  // adjoint of (d2 - d1)
    const double gts = -(coef * (d2 - d1) / ts) / ts * g;
    add(B, rtheta_pp, o2, -gd);
    add(B, rtheta_pp_old, o2, gd);
    add(B, rtheta_pp, o1, gd);
    add(B, rtheta_pp_old, o1, -gd);
    add(B, theta_m, o1, gts);
    add(B, theta_m, o2, gts);
}

// ===========================================================================
// atm_compute_moist_coefficients (moist_coefficients.py:124-291)
// fwd: scalars, cqw, cqu, qtot; moist species [ms-1, me)
// ===========================================================================
namespace moist {
enum { scalars, cqw, cqu, qtot, N };
}

// thread (cell c, level k): qtot and cqw adjoints into the scalars
__device__ __forceinline__ void moist_cell_vjp(const mpas_hip_mesh &m, int S, int q0, int q1, Fwd F, Bar B, int c,
                                               int k) {
    using namespace moist;
    const int L = m.nLevels, Lw = L + 1;
    if (c >= m.nCells || k >= L || q1 <= q0) return;
    double g = bar_of(B, qtot, (long long)c * L + k);
    // cqw(k') = 1 / (1 + 0.5 (qtot(k') + qtot(k'-1))) for k' = k and k + 1
    for (int kk = k; kk <= k + 1; ++kk) {
        if (kk < 1 || kk >= L) continue;
        const double cw = F[cqw][(long long)c * Lw + kk];
        g += 0.5 * (-(cw * cw) * bar_of(B, cqw, (long long)c * Lw + kk));
    }
    for (int q = q0; q < q1; ++q) add(B, scalars, ((long long)c * L + k) * S + q, g);
}

// thread (edge e, level k): cqu adjoint into the scalars of both edge cells
__device__ __forceinline__ void moist_edge_vjp(const mpas_hip_mesh &m, int S, int q0, int q1, Fwd F, Bar B, int e,
                                               int k) {
    using namespace moist;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L || q1 <= q0) return;
    const long long i = (long long)e * L + k;
    const double cu = F[cqu][i];
    const double g = -(cu * cu) * bar_of(B, cqu, i);
    if (g == 0.0) return;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    for (int q = q0; q < q1; ++q) {
        add(B, scalars, ((long long)c1 * L + k) * S + q, 0.5 * g);
        add(B, scalars, ((long long)c2 * L + k) * S + q, 0.5 * g);
    }
}

// ===========================================================================
// mpas_reconstruct_2d (vector_reconstruction.py:311-441)
// fwd: u, uReconstructX, Y, Z, Zonal, Meridional
// ===========================================================================
namespace recon {
enum { u, ux, uy, uz, uzonal, umerid, N };
}

// thread (cell c, level k)
__device__ __forceinline__ void reconstruct_vjp(const mpas_hip_mesh &m, Fwd F, Bar B, int c, int k) {
    using namespace recon;
    (void)F;
    const int L = m.nLevels, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const long long o = (long long)c * L + k;
    const double clat = cos(m.latCell[c]), slat = sin(m.latCell[c]);
    const double clon = cos(m.lonCell[c]), slon = sin(m.lonCell[c]);
    const double gz = bar_of(B, uzonal, o), gm = bar_of(B, umerid, o);
    const double gx = bar_of(B, ux, o) - slon * gz - clon * slat * gm;
    const double gy = bar_of(B, uy, o) + clon * gz - slon * slat * gm;
    const double gzz = bar_of(B, uz, o) + clat * gm;
    const int n = m.nEdgesOnCell[c];
    for (int j = 0; j < n; ++j) {
        const int e = m.edgesOnCell[c * ME + j];
        const double *cf = m.coeffs_reconstruct + ((long long)c * ME + j) * 3;
        add(B, u, (long long)e * L + k, cf[0] * gx + cf[1] * gy + cf[2] * gzz);
    }
}

// ===========================================================================
// atm_recover_large_step_variables_work (recover_large_step.py:311-711)
// ===========================================================================
namespace recover {
enum {
    rho_p_save, rho_pp, rtheta_p_save, rtheta_pp, rt_diabatic_tend, rw_save, rw_p, ru_save, ru_p,
    wwAvg_in, ruAvg_in, exner_in, pressure_p_in,
    wwAvg, ruAvg, rho_p, rho_zz, rtheta_p, theta_m, exner, pressure_p, rw, w, ru, u, N
};
}

// pass A, thread (cell c, interface k): w and rw/wwAvg; adds into the rho_zz
// and ru output adjoints
__device__ __forceinline__ void recover_w_vjp(const mpas_hip_mesh &m, double invNs, Fwd F, Bar B, int c, int k) {
    using namespace recover;
    const int L = m.nLevels, ME = m.maxEdges, Lw = L + 1;
    const long long ow = (long long)c * Lw;
    if (c >= m.nCells || k > L) return;
    const long long oc = (long long)c * L;
    const double *rz = F[rho_zz] + oc;
    // the edge loops carry the terrain terms only (nothing on a flat mesh)
    const int n = m.flatTerrain ? 0 : m.nEdgesOnCell[c];
    const double gw = bar_of(B, w, ow + k);
    if (k == 0) {
        const double den = m.cf1 * rz[0] + m.cf2 * rz[1] + m.cf3 * rz[2];
        const double w0 = F[w][ow];
        const double gs = gw / den, gden = -(w0 / den) * gw;
        add(B, rho_zz, oc + 0, m.cf1 * gden);
        add(B, rho_zz, oc + 1, m.cf2 * gden);
        add(B, rho_zz, oc + 2, m.cf3 * gden);
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[c * ME + j];
            const double *re = F[ru] + (long long)e * L;
            const double fs = m.cf1 * re[0] + m.cf2 * re[1] + m.cf3 * re[2];
            const long long zo = ((long long)c * ME + j) * Lw;
            const double coef = m.zb_cell[zo] + sign_pm(fs) * m.zb3_cell[zo];
            const double gf = m.edgesOnCell_sign[c * ME + j] * coef * gs;
            if (gf == 0.0) continue;
            add(B, ru, (long long)e * L + 0, m.cf1 * gf);
            add(B, ru, (long long)e * L + 1, m.cf2 * gf);
            add(B, ru, (long long)e * L + 2, m.cf3 * gf);
        }
        add_own(B, wwAvg_in, ow + k, bar_of(B, wwAvg, ow + k));
        return;
    }
    if (k == L) {
        add_own(B, wwAvg_in, ow + k, bar_of(B, wwAvg, ow + k));
        return;
    }
    const double zi = m.fzm[k] * m.zz[oc + k] + m.fzp[k] * m.zz[oc + k - 1];
    const double den = m.fzm[k] * rz[k] + m.fzp[k] * rz[k - 1];
    const double wk = F[w][ow + k];
    const double gnum = gw / den, gden = -(wk / den) * gw;
    add(B, rho_zz, oc + k, m.fzm[k] * gden);
    add(B, rho_zz, oc + k - 1, m.fzp[k] * gden);
    for (int j = 0; j < n; ++j) {
        const int e = m.edgesOnCell[c * ME + j];
        const double *re = F[ru] + (long long)e * L;
        const double fi = m.fzm[k] * re[k] + m.fzp[k] * re[k - 1];
        const long long zo = ((long long)c * ME + j) * Lw + k;
        const double coef = m.zb_cell[zo] + sign_pm(fi) * m.zb3_cell[zo];
        const double gf = m.edgesOnCell_sign[c * ME + j] * coef * gnum;
        if (gf == 0.0) continue;
        add(B, ru, (long long)e * L + k, m.fzm[k] * gf);
        add(B, ru, (long long)e * L + k - 1, m.fzp[k] * gf);
    }
    const double grw = bar_of(B, rw, ow + k) + gnum / zi;
    const double gww = bar_of(B, wwAvg, ow + k);
    add_own(B, rw_save, ow + k, grw + gww);
    add_own(B, rw_p, ow + k, grw);
    add_own(B, wwAvg_in, ow + k, invNs * gww);
}

// pass B, thread (edge e, level k): ruAvg, ru, u; adds into the rho_zz output adjoint
__device__ __forceinline__ void recover_edge_vjp(const mpas_hip_mesh &m, double invNs, Fwd F, Bar B, int e, int k) {
    using namespace recover;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const double s = F[rho_zz][(long long)c1 * L + k] + F[rho_zz][(long long)c2 * L + k];
    const double gu = bar_of(B, u, i);
    const double gru = bar_of(B, ru, i) + 2.0 * gu / s;
    const double gs = -(F[u][i] / s) * gu;
    add(B, rho_zz, (long long)c1 * L + k, gs);
    add(B, rho_zz, (long long)c2 * L + k, gs);
    const double gavg = bar_of(B, ruAvg, i);
    add_own(B, ru_save, i, gru + gavg);
    add_own(B, ru_p, i, gru);
    add_own(B, ruAvg_in, i, invNs * gavg);
}

// pass C, thread (cell c, level k): densities, theta_m, exner, pressure
__device__ __forceinline__ void recover_cell_vjp(const mpas_hip_mesh &m, int rk_step, double dt, Fwd F, Bar B, int c,
                                                 int k) {
    using namespace recover;
    const int L = m.nLevels;
    if (c >= m.nCells || k >= L) return;
    const long long o = (long long)c * L + k;
    const double rcv = MPAS_RGAS / (MPAS_CP - MPAS_RGAS);
    const double rz = F[rho_zz][o];
    const double rtp = F[rtheta_p][o];
    const double rtb = m.rtheta_base[o];
    double grtp = bar_of(B, rtheta_p, o);
    double grz = bar_of(B, rho_zz, o);
    if (rk_step == 3) {
        const double ex = F[exner][o];
        const double gp = bar_of(B, pressure_p, o);
        const double zr = m.zz[o] * MPAS_RGAS;
        const double gex = bar_of(B, exner, o) + zr * (rtp + rtb) * gp;
        grtp += zr * ex * gp;
        grtp += gex * (rcv * ex / (rtp + rtb));
    } else {
        add_own(B, exner_in, o, bar_of(B, exner, o));
        add_own(B, pressure_p_in, o, bar_of(B, pressure_p, o));
    }
    const double gth = bar_of(B, theta_m, o);
    grtp += gth / rz;
    grz += -(F[theta_m][o] / rz) * gth;
    add_own(B, rtheta_p_save, o, grtp);
    add_own(B, rtheta_pp, o, grtp);
    if (rk_step == 3) {
        grz += -(dt * F[rt_diabatic_tend][o]) * grtp;
        add_own(B, rt_diabatic_tend, o, -(dt * rz) * grtp);
    }
    const double grp = bar_of(B, rho_p, o) + grz;
    add_own(B, rho_p_save, o, grp);
    add_own(B, rho_pp, o, grp);
}

// ===========================================================================
// atm_compute_solve_diagnostics_work (solve_diagnostics.py:531-756)
// fwd: u, h, v_in, h_edge, v, divergence, vorticity, ke, pv_edge
// scratch adjoints: pv_vertex (V), ke_vertex (V), pv_cell (C), ke_edge (E),
// all (N+1) x L and zeroed before pass 1
// ===========================================================================
namespace diag {
enum { u, h, v_in, h_edge, v, divergence, vorticity, ke, pv_edge, N };
}
struct DiagScratch {
    double *g_pvv, *g_kev, *g_pvc, *g_kee;
    double *pvc;  // gather form: pv_cell of every cell (C), formed once by the cell pass
};

__device__ __forceinline__ double diag_pv_cell(const mpas_hip_mesh &m, Fwd F, int c, int k) {
    using namespace diag;
    const int L = m.nLevels, ME = m.maxEdges;
    const int n = m.nEdgesOnCell[c];
    double pc = 0.0;
    for (int i = 0; i < n; ++i) {
        const int vx = m.verticesOnCell[c * ME + i];
        const double ka = m.kiteAreasOnVertex[3 * vx + m.kiteForCell[c * ME + i]];
        const double pvv = m.fVertex[vx] + F[vorticity][(long long)vx * L + k];
        const double cp = ka * pvv * m.invAreaCell[c];
        pc = (i == 0) ? cp : pc + cp;
    }
    return pc;
}

// pass 1, thread (edge e, level k): pv_edge and v
__device__ __forceinline__ void diag_edge2_vjp(const mpas_hip_mesh &m, double dt, double apvm, int rk_step, Fwd F,
                                               Bar B, const DiagScratch &S, int e, int k) {
    using namespace diag;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const int v1 = m.verticesOnEdge[2 * e], v2 = m.verticesOnEdge[2 * e + 1];
    const double gpe = bar_of(B, pv_edge, i);
    double gv = bar_of(B, v, i);
    double gp1 = 0.5 * gpe, gp2 = 0.5 * gpe;
    if (apvm > 0.0 && gpe != 0.0) {
        const double p1 = m.fVertex[v1] + F[vorticity][(long long)v1 * L + k];
        const double p2 = m.fVertex[v2] + F[vorticity][(long long)v2 * L + k];
        const double gpt = (p2 - p1) * m.invDvEdge[e];
        const double pc1 = diag_pv_cell(m, F, c1, k), pc2 = diag_pv_cell(m, F, c2, k);
        const double gpn = (pc2 - pc1) * m.invDcEdge[e];
        const double t = (apvm * dt) * (-gpe);
        gv += t * gpt;
        add(B, u, i, t * gpn);
        const double ggt = t * F[v][i] * m.invDvEdge[e];
        const double ggn = t * F[u][i] * m.invDcEdge[e];
        gp2 += ggt;
        gp1 -= ggt;
        acc(S.g_pvc + (long long)c2 * L + k, ggn);
        acc(S.g_pvc + (long long)c1 * L + k, -ggn);
    }
    acc(S.g_pvv + (long long)v1 * L + k, gp1);
    acc(S.g_pvv + (long long)v2 * L + k, gp2);
    if (rk_step <= 0 || rk_step == 3) {
        const int ne = m.nEdgesOnEdge[e];
        for (int j = 0; j < ne; ++j) {
            const int ee = m.edgesOnEdge[e * m.maxEdges2 + j];
            add(B, u, (long long)ee * L + k, m.weightsOnEdge[e * m.maxEdges2 + j] * gv);
        }
    } else {
        add(B, v_in, i, gv);
    }
}

// pass 2, thread (cell c, level k): divergence, ke, pv_cell
__device__ __forceinline__ void diag_cell_vjp(const mpas_hip_mesh &m, Fwd F, Bar B, const DiagScratch &S, int c,
                                              int k) {
    using namespace diag;
    (void)F;
    const int L = m.nLevels, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const long long o = (long long)c * L + k;
    const double ia = m.invAreaCell[c];
    const double gd = bar_of(B, divergence, o), gk = bar_of(B, ke, o), gpc = S.g_pvc[o];
    const double ke_fact = 1.0 - 0.375;
    const int n = m.nEdgesOnCell[c];
    for (int i = 0; i < n; ++i) {
        const int e = m.edgesOnCell[c * ME + i];
        add(B, u, (long long)e * L + k, ia * m.edgesOnCell_sign[c * ME + i] * m.dvEdge[e] * gd);
        acc(S.g_kee + (long long)e * L + k, 0.25 * ia * ke_fact * gk);
        const int vx = m.verticesOnCell[c * ME + i];
        const double ka = m.kiteAreasOnVertex[3 * vx + m.kiteForCell[c * ME + i]];
        acc(S.g_kev + (long long)vx * L + k, (1.0 - ke_fact) * ka * ia * gk);
        acc(S.g_pvv + (long long)vx * L + k, ka * ia * gpc);
    }
}

// pass 3, thread (vertex vx, level k): vorticity and vertex KE
__device__ __forceinline__ void diag_vertex_vjp(const mpas_hip_mesh &m, Fwd F, Bar B, const DiagScratch &S, int vx,
                                                int k) {
    using namespace diag;
    (void)F;
    const int L = m.nLevels;
    if (vx >= m.nVertices || k >= L) return;
    const long long o = (long long)vx * L + k;
    const double gz = bar_of(B, vorticity, o) + S.g_pvv[o];
    const double gkv = S.g_kev[o];
    const double iat = m.invAreaTriangle[vx];
    for (int j = 0; j < 3; ++j) {
        const int e = m.edgesOnVertex[3 * vx + j];
        add(B, u, (long long)e * L + k, iat * m.edgesOnVertex_sign[3 * vx + j] * m.dcEdge[e] * gz);
        acc(S.g_kee + (long long)e * L + k, 0.25 * iat * gkv);
    }
}

// pass 4, thread (edge e, level k): h_edge and edge KE
__device__ __forceinline__ void diag_edge1_vjp(const mpas_hip_mesh &m, Fwd F, Bar B, const DiagScratch &S, int e,
                                               int k) {
    using namespace diag;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const double gh = 0.5 * bar_of(B, h_edge, i);
    add(B, h, (long long)c1 * L + k, gh);
    add(B, h, (long long)c2 * L + k, gh);
    add(B, u, i, (m.dcEdge[e] * m.dvEdge[e]) * 2.0 * F[u][i] * S.g_kee[i]);
}

// ---------------------------------------------------------------------------
// Gather form of the diagnostics adjoint (no atomics), given the inverse
// edgesOnEdge table.  Each target sums its sources; the per-edge apvm terms
// are recomputed where they are consumed, from the output adjoints, u, v and
// vorticity.
//   G1 edges:    adjoint of v (own + apvm tangential-gradient term), kept in
//                g_kee, or added to v_in when v was not reconstructed
//   G2 cells:    the pv_cell adjoint (kept in g_pvc) and h
//   G3 vertices: vorticity / pv_vertex adjoint (kept in g_pvv) and the
//                vertex-KE adjoint (kept in g_kev)
//   G4 edges:    u from its own apvm term, the v transpose over the inverse
//                table, the cells' divergence / KE and the vertices' terms
// ---------------------------------------------------------------------------
// t = apvm dt (-g_pv_edge) of the edge row i (0 without apvm)
__device__ __forceinline__ double diag_t(double apvm_dt, bool apvm_on, Bar B, long long i) {
    using namespace diag;
    const double gpe = bar_of(B, pv_edge, i);
    return (apvm_on && gpe != 0.0) ? apvm_dt * (-gpe) : 0.0;
}
__device__ __forceinline__ void diag_g1(const mpas_hip_mesh &m, double dt, double apvm, int rk_step, Fwd F, Bar B,
                                        const DiagScratch &S, int e, int k) {
    using namespace diag;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    double gv = bar_of(B, v, i);
    const double t = diag_t(apvm * dt, apvm > 0.0, B, i);
    if (t != 0.0) {
        const int v1 = m.verticesOnEdge[2 * e], v2 = m.verticesOnEdge[2 * e + 1];
        const double p1 = m.fVertex[v1] + F[vorticity][(long long)v1 * L + k];
        const double p2 = m.fVertex[v2] + F[vorticity][(long long)v2 * L + k];
        gv += t * ((p2 - p1) * m.invDvEdge[e]);
    }
    if (rk_step <= 0 || rk_step == 3) S.g_kee[i] = gv;
    else add_own(B, v_in, i, gv);
}
__device__ __forceinline__ void diag_g2(const mpas_hip_mesh &m, double dt, double apvm, Fwd F, Bar B,
                                        const DiagScratch &S, int c, int k) {
    using namespace diag;
    const int L = m.nLevels, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const double apvm_dt = apvm * dt;
    const bool on = apvm > 0.0;
    const int n = m.nEdgesOnCell[c];
    // pv_cell of this cell, read by the edge pass for both cells of each edge
    if (on) S.pvc[(long long)c * L + k] = diag_pv_cell(m, F, c, k);
    double gpc = 0.0, gh = 0.0;
    for (int j = 0; j < n; ++j) {
        const int e = m.edgesOnCell[c * ME + j];
        const long long ie = (long long)e * L + k;
        const double t = diag_t(apvm_dt, on, B, ie);
        if (t != 0.0) {
            const double ggn = t * F[u][ie] * m.invDcEdge[e];
            gpc += (m.cellsOnEdge[2 * e + 1] == c) ? ggn : -ggn;
        }
        gh += 0.5 * bar_of(B, h_edge, ie);
    }
    const long long o = (long long)c * L + k;
    S.g_pvc[o] = gpc;
    add_own(B, h, o, gh);
}
__device__ __forceinline__ void diag_g3(const mpas_hip_mesh &m, double dt, double apvm, Fwd F, Bar B,
                                        const DiagScratch &S, int vx, int k) {
    using namespace diag;
    const int L = m.nLevels, ME = m.maxEdges;
    if (vx >= m.nVertices || k >= L) return;
    const double apvm_dt = apvm * dt;
    const bool on = apvm > 0.0;
    const long long o = (long long)vx * L + k;
    const double ke_fact = 1.0 - 0.375;
    double gz = bar_of(B, vorticity, o), gkv = 0.0;
    int seen0 = -1, seen1 = -1, seen2 = -1;
    for (int j = 0; j < 3; ++j) {
        const int e = m.edgesOnVertex[3 * vx + j];
        const long long ie = (long long)e * L + k;
        gz += 0.5 * bar_of(B, pv_edge, ie);
        const double t = diag_t(apvm_dt, on, B, ie);
        if (t != 0.0) {
            const double ggt = t * F[v][ie] * m.invDvEdge[e];
            gz += (m.verticesOnEdge[2 * e + 1] == vx) ? ggt : -ggt;
        }
        // the kite terms of the cells around the vertex, each cell once (the
        // three edges name each of the three cells twice)
        for (int q = 0; q < 2; ++q) {
            const int c = m.cellsOnEdge[2 * e + q];
            if (c == seen0 || c == seen1 || c == seen2) continue;
            if (seen0 < 0) seen0 = c;
            else if (seen1 < 0) seen1 = c;
            else seen2 = c;
            const int n = m.nEdgesOnCell[c];
            for (int i = 0; i < n; ++i) {
                if (m.verticesOnCell[c * ME + i] != vx) continue;
                const double ka = m.kiteAreasOnVertex[3 * vx + m.kiteForCell[c * ME + i]];
                const long long oc = (long long)c * L + k;
                const double ia = m.invAreaCell[c];
                gz += ka * ia * S.g_pvc[oc];
                gkv += (1.0 - ke_fact) * ka * ia * bar_of(B, ke, oc);
                break;
            }
        }
    }
    S.g_pvv[o] = gz;
    S.g_kev[o] = gkv;
}
__device__ __forceinline__ void diag_g4(const mpas_hip_mesh &m, double dt, double apvm, int rk_step, Fwd F, Bar B,
                                        const DiagScratch &S, const mpas_hip_inverse_tables &inv, int e, int k) {
    using namespace diag;
    const int L = m.nLevels, ME2 = m.maxEdges2;
    if (e >= m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    const double ke_fact = 1.0 - 0.375;
    double gu = 0.0, gkee = 0.0;
    const double t = diag_t(apvm * dt, apvm > 0.0, B, i);
    if (t != 0.0) {
        const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
        const double pc1 = S.pvc[(long long)c1 * L + k], pc2 = S.pvc[(long long)c2 * L + k];
        gu += t * ((pc2 - pc1) * m.invDcEdge[e]);
    }
    if (rk_step <= 0 || rk_step == 3) {
        const int j1 = inv.eoeStart[e + 1];
        for (int j = inv.eoeStart[e]; j < j1; ++j) {
            const int eo = inv.eoeEdge[j];
            gu += m.weightsOnEdge[eo * ME2 + inv.eoeSlot[j]] * S.g_kee[(long long)eo * L + k];
        }
    }
    for (int q = 0; q < 2; ++q) {
        const int c = m.cellsOnEdge[2 * e + q];
        const long long oc = (long long)c * L + k;
        const double ia = m.invAreaCell[c];
        gu += ia * edge_sign_in_cell(m, c, e) * m.dvEdge[e] * bar_of(B, divergence, oc);
        gkee += 0.25 * ia * ke_fact * bar_of(B, ke, oc);
        const int vx = m.verticesOnEdge[2 * e + q];
        const long long ov = (long long)vx * L + k;
        const double iat = m.invAreaTriangle[vx];
        double sv = 0.0;
        for (int j = 0; j < 3; ++j)
            if (m.edgesOnVertex[3 * vx + j] == e) sv = m.edgesOnVertex_sign[3 * vx + j];
        gu += iat * sv * m.dcEdge[e] * S.g_pvv[ov];
        gkee += 0.25 * iat * S.g_kev[ov];
    }
    gu += (m.dcEdge[e] * m.dvEdge[e]) * 2.0 * F[u][i] * gkee;
    add_own(B, u, i, gu);
}

// ===========================================================================
// atm_compute_vert_imp_coefs_work (vert_imp_coefs.py:292-451, Thomas factor
// :100-121; the factor's adjoint is that of tridiagonal_factor.py:183-230)
// fwd: cqw, p (exner), t (theta_m), rt (rtheta_p), qtot, then the outputs
// ===========================================================================
namespace vimp {
enum { cqw, p, t, rt, qtot, cofrz, cofwr, cofwz, coftz, cofwt, a_tri, alpha_tri, gamma_tri, N };
}

// thread per column c; the cofwz/coftz/cofwt output adjoints of the column
// are used as accumulators
__device__ __forceinline__ void vimp_col_vjp(const mpas_hip_mesh &m, double dts, double epssm, Fwd F, Bar B, int c) {
    using namespace vimp;
    const int L = m.nLevels, Lw = L + 1;
    if (c >= m.nCells) return;
    const double dtseps = 0.5 * dts * (1.0 + epssm);
    const double rcv = MPAS_RGAS / (MPAS_CP - MPAS_RGAS);
    const double c2 = MPAS_CP * rcv;
    const long long o = (long long)c * L, ow = (long long)c * Lw;
    const double *zz = m.zz + o, *rdzw = m.rdzw;
    const double *cofwz_ = F[cofwz] + o, *coftz_ = F[coftz] + ow, *cofwt_ = F[cofwt] + o;
    const double *cofwr_ = F[cofwr] + o, *cofrz_ = F[cofrz];
    const double *a_ = F[a_tri] + o, *al_ = F[alpha_tri] + o, *ga_ = F[gamma_tri] + o;
    double *gwz = B[cofwz] ? B[cofwz] + o : nullptr;
    double *gtz = B[coftz] ? B[coftz] + ow : nullptr;
    double *gwt = B[cofwt] ? B[cofwt] + o : nullptr;
    if (!gwz || !gtz || !gwt) return;  // the orchestration always provides them
    // reverse Thomas factor
    double gg_carry = 0.0;  // adjoint of gamma(k) from level k + 1
    for (int kk = L - 1; kk >= 1; --kk) {
        const double al = al_[kk], a = a_[kk];
        const double gprev = (kk > 1) ? ga_[kk - 1] : 0.0;
        const double cc = -cofwz_[kk] * coftz_[kk + 1] * rdzw[kk] * zz[kk] - cofwr_[kk] * cofrz_[kk] +
                          cofwt_[kk] * coftz_[kk + 1] * rdzw[kk];
        const double g_ga = bar_of(B, gamma_tri, o + kk) + gg_carry;
        const double g_c = al * g_ga;
        const double g_al = bar_of(B, alpha_tri, o + kk) + cc * g_ga;
        const double g_den = -(al * al) * g_al;
        const double g_b = g_den;
        const double g_a = bar_of(B, a_tri, o + kk) - gprev * g_den;
        gg_carry = -a * g_den;
        // a = -cofwz(k) coftz(k-1) rdzw(k-1) zz(k-1) + cofwr(k) cofrz(k-1) - cofwt(k-1) coftz(k-1) rdzw(k-1)
        gwz[kk] += -coftz_[kk - 1] * rdzw[kk - 1] * zz[kk - 1] * g_a;
        gtz[kk - 1] += (-cofwz_[kk] * rdzw[kk - 1] * zz[kk - 1] - cofwt_[kk - 1] * rdzw[kk - 1]) * g_a;
        gwt[kk - 1] += -coftz_[kk - 1] * rdzw[kk - 1] * g_a;
        // b = 1 + cofwz(k) (coftz(k) rdzw(k) zz(k) + coftz(k) rdzw(k-1) zz(k-1))
        //       - coftz(k) (cofwt(k) rdzw(k) - cofwt(k-1) rdzw(k-1)) + cofwr(k) (cofrz(k) - cofrz(k-1))
        gwz[kk] += (coftz_[kk] * rdzw[kk] * zz[kk] + coftz_[kk] * rdzw[kk - 1] * zz[kk - 1]) * g_b;
        gtz[kk] += (cofwz_[kk] * (rdzw[kk] * zz[kk] + rdzw[kk - 1] * zz[kk - 1]) -
                    (cofwt_[kk] * rdzw[kk] - cofwt_[kk - 1] * rdzw[kk - 1])) * g_b;
        gwt[kk] += -coftz_[kk] * rdzw[kk] * g_b;
        gwt[kk - 1] += coftz_[kk] * rdzw[kk - 1] * g_b;
        // c = -cofwz(k) coftz(k+1) rdzw(k) zz(k) - cofwr(k) cofrz(k) + cofwt(k) coftz(k+1) rdzw(k)
        gwz[kk] += -coftz_[kk + 1] * rdzw[kk] * zz[kk] * g_c;
        gtz[kk + 1] += (-cofwz_[kk] * rdzw[kk] * zz[kk] + cofwt_[kk] * rdzw[kk]) * g_c;
        gwt[kk] += coftz_[kk + 1] * rdzw[kk] * g_c;
    }
    // coefficients -> state
    const double *P = F[p] + o, *T = F[t] + o, *RT = F[rt] + o, *Q = F[qtot] + o;
    const double *cq = F[cqw] + ow;
    for (int kk = 1; kk < L; ++kk) {
        const double zi = m.fzm[kk] * zz[kk] + m.fzp[kk] * zz[kk - 1];
        const double pi = m.fzm[kk] * P[kk] + m.fzp[kk] * P[kk - 1];
        const double base = dtseps * c2 * zi * m.rdzu[kk];
        add_own(B, cqw, ow + kk, base * pi * gwz[kk]);
        const double gpi = base * cq[kk] * gwz[kk];
        add_own(B, p, o + kk, m.fzm[kk] * gpi);
        add_own(B, p, o + kk - 1, m.fzp[kk] * gpi);
        const double gti = dtseps * gtz[kk];
        add_own(B, t, o + kk, m.fzm[kk] * gti);
        add_own(B, t, o + kk - 1, m.fzp[kk] * gti);
    }
    for (int kk = 0; kk < L; ++kk) {
        const double g = gwt[kk];
        if (g == 0.0) continue;
        const double A0 = 0.5 * dtseps * rcv * zz[kk] * MPAS_GRAVITY * m.rho_base[o + kk];
        const double q1 = 1.0 + Q[kk], rr = m.rtheta_base[o + kk] + RT[kk], pb = m.exner_base[o + kk];
        const double cw = cofwt_[kk];
        add_own(B, p, o + kk, A0 / q1 / (rr * pb) * g);
        add_own(B, qtot, o + kk, -(cw / q1) * g);
        add_own(B, rt, o + kk, -(cw / rr) * g);
    }
}

// ===========================================================================
// atm_advance_acoustic_step_work (acoustic_step.py:456-889, 1297-1951);
// the column solve's adjoint is the transposed Thomas solve
// (tridiagonal_solve.py:232-291)
// ===========================================================================
namespace acoustic {
enum {
    cofrz, cofwr, cofwz, coftz, cofwt, a_tri, alpha_tri, gamma_tri, rho_zz, theta_m, exner, cqu, tend_ru,
    tend_rho, tend_rt, tend_rw, w, rw, rw_save,
    ru_p_in, ruAvg_in, rho_pp_in, rtheta_pp_in, rtheta_pp_old_in, rw_p_in, wwAvg_in,
    ru_p, ruAvg, rho_pp, rtheta_pp, rtheta_pp_old, rw_p, wwAvg, N
};
constexpr int kScratchArrays = 8;  // per column, (L + 1) each: rs, ts, y, x and their adjoints
}

// Column pass, thread per column c.  Column scratch arrays are interleaved
// across columns (element (a, k) of column c at ((a (L+1) + k) ncol + c)).
// Adds into the ru_p output adjoint (the cell pass reads the updated ru_p),
// so it runs before the edge pass.
//
// Forward, per column (inputs read as zero on the first substep):
//   rs, ts                      explicit parts of rho'' and (rho theta)''
//   wwAvg += (1 - epssm)/2 rw_p                       (interior interfaces)
//   r(k)                        right-hand side of the implicit rw_p equation
//   y(0) = rw_p(0); y(k) = (r(k) - a(k) y(k-1)) alpha(k)          (k = 1..L-1)
//   x(L) = rw_p(L); x(k) = y(k) - gamma(k) x(k+1)                 (k = L-1..0)
//   rw_p'(k) = rr(k), implicit Rayleigh damping of x(k)           (k = 1..L-1)
//   wwAvg += (1 + epssm)/2 rr
//   rho_pp' = rs - cofrz d(rw_p'), rtheta_pp' = ts - rdzw d(coftz rw_p')
__device__ __forceinline__ void acoustic_col_vjp(const mpas_hip_mesh &m, const mpas_hip_acoustic_params &P, Fwd F,
                                                 Bar B, double *scr, int c) {
    using namespace acoustic;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges, nC = m.nCells;
    if (c >= nC) return;
    const long long ncol = nC;
    auto S = [&](int a, int k) -> double & { return scr[((long long)a * Lw + k) * ncol + c]; };
    enum { RS, TS, Y, X, GRS, GTS, GY, GX };
    const double dts = P.dts, epssm = P.epssm, resm = P.resm;
    const bool first = (P.small_step == 1);
    const long long oc = (long long)c * L, ow = (long long)c * Lw;
    const double *zz = m.zz + oc, *rdzw = m.rdzw, *fzm = m.fzm, *fzp = m.fzp;
    const double *cofrz_ = F[cofrz], *cofwz_ = F[cofwz] + oc, *cofwr_ = F[cofwr] + oc, *cofwt_ = F[cofwt] + oc;
    const double *coftz_ = F[coftz] + ow, *a_ = F[a_tri] + oc, *al_ = F[alpha_tri] + oc, *ga_ = F[gamma_tri] + oc;
    auto RPP = [&](int k) { return first ? 0.0 : F[rho_pp_in][oc + k]; };
    auto RTPP = [&](int k) { return first ? 0.0 : F[rtheta_pp_in][oc + k]; };
    auto RWP = [&](int k) { return first ? 0.0 : F[rw_p_in][ow + k]; };
    auto RWF = [&](int k) { return F[rw_p][ow + k]; };
    const int n = m.nEdgesOnCell[c];
    const double ia = m.invAreaCell[c];
    // ---------------- forward recompute
    for (int k = 0; k < L; ++k) {
        double rsum = 0.0, tsum = 0.0;
        for (int i = 0; i < n; ++i) {
            const int e = m.edgesOnCell[c * ME + i];
            const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
            const double metric = m.edgesOnCell_sign[c * ME + i] * m.dvEdge[e];
            const double flux = dts * metric * F[ru_p][(long long)e * L + k] * ia;
            const double thedge = 0.5 * (F[theta_m][(long long)c2 * L + k] + F[theta_m][(long long)c1 * L + k]);
            const double rcon = -flux, tcon = -(flux * thedge);
            rsum = (i == 0) ? rcon : rsum + rcon;
            tsum = (i == 0) ? tcon : tsum + tcon;
        }
        S(RS, k) = RPP(k) + dts * F[tend_rho][oc + k] + rsum - cofrz_[k] * resm * (RWP(k + 1) - RWP(k));
        S(TS, k) = RTPP(k) + dts * F[tend_rt][oc + k] + tsum -
                   resm * rdzw[k] * (coftz_[k + 1] * RWP(k + 1) - coftz_[k] * RWP(k));
    }
    auto rhs = [&](int k) -> double {
        return RWP(k) + dts * F[tend_rw][ow + k] -
               cofwz_[k] * ((zz[k] * S(TS, k) - zz[k - 1] * S(TS, k - 1)) +
                            resm * (zz[k] * RTPP(k) - zz[k - 1] * RTPP(k - 1))) -
               cofwr_[k] * ((S(RS, k) + S(RS, k - 1)) + resm * (RPP(k) + RPP(k - 1))) +
               cofwt_[k] * (S(TS, k) + resm * RTPP(k)) + cofwt_[k - 1] * (S(TS, k - 1) + resm * RTPP(k - 1));
    };
    S(Y, 0) = RWP(0);
    for (int k = 1; k < L; ++k) S(Y, k) = (rhs(k) - a_[k] * S(Y, k - 1)) * al_[k];
    S(X, L) = RWP(L);
    for (int k = L - 1; k >= 0; --k) S(X, k) = S(Y, k) - ga_[k] * S(X, k + 1);
    for (int k = 0; k <= L; ++k) {
        S(GRS, k) = 0.0;
        S(GTS, k) = 0.0;
        S(GY, k) = 0.0;
        S(GX, k) = 0.0;
    }
    // ---------------- adjoint
    // final rw_p' adjoint, then the final rho_pp' / rtheta_pp' update
    for (int k = 0; k <= L; ++k) S(GX, k) = bar_of(B, rw_p, ow + k);  // g_rwF for now
    for (int k = 0; k < L; ++k) {
        const double gr = bar_of(B, rho_pp, oc + k), gt = bar_of(B, rtheta_pp, oc + k);
        S(GRS, k) += gr;
        S(GTS, k) += gt;
        S(GX, k + 1) += -cofrz_[k] * gr - rdzw[k] * coftz_[k + 1] * gt;
        S(GX, k) += cofrz_[k] * gr + rdzw[k] * coftz_[k] * gt;
        add_own(B, coftz, ow + k + 1, -rdzw[k] * RWF(k + 1) * gt);
        add_own(B, coftz, ow + k, rdzw[k] * RWF(k) * gt);
    }
    // damping, interior interfaces: g_rwF(k) -> g_x(k)
    for (int k = 1; k < L; ++k) {
        const double gww = bar_of(B, wwAvg, ow + k);
        const double grr = S(GX, k) + 0.5 * (1.0 + epssm) * gww;
        const double D = 1.0 + dts * m.dss[oc + k];
        const double zi = fzm[k] * zz[k] + fzp[k] * zz[k - 1];
        const double ri = fzm[k] * F[rho_zz][oc + k] + fzp[k] * F[rho_zz][oc + k - 1];
        const double gq = grr / D;
        const double grwd = gq - grr;
        const double gri = -gq * dts * m.dss[oc + k] * zi * F[w][ow + k];
        add_own(B, w, ow + k, -gq * dts * m.dss[oc + k] * zi * ri);
        add(B, rho_zz, oc + k, fzm[k] * gri);
        add(B, rho_zz, oc + k - 1, fzp[k] * gri);
        add_own(B, rw_save, ow + k, grwd);
        add_own(B, rw, ow + k, -grwd);
        S(GX, k) = gq;
    }
    // back substitution, reversed (k = 0 .. L-1)
    for (int k = 0; k < L; ++k) {
        const double gx = S(GX, k);
        S(GY, k) += gx;
        add_own(B, gamma_tri, oc + k, -S(X, k + 1) * gx);
        S(GX, k + 1) += -ga_[k] * gx;
    }
    // elimination, reversed (k = L-1 .. 1), and the right-hand side
    for (int k = L - 1; k >= 1; --k) {
        const double gy = S(GY, k);
        const double gr = al_[k] * gy;
        add_own(B, alpha_tri, oc + k, (rhs(k) - a_[k] * S(Y, k - 1)) * gy);
        add_own(B, a_tri, oc + k, -S(Y, k - 1) * gr);
        S(GY, k - 1) += -a_[k] * gr;
        const double A = zz[k] * S(TS, k) - zz[k - 1] * S(TS, k - 1);
        const double Bz = zz[k] * RTPP(k) - zz[k - 1] * RTPP(k - 1);
        const double R2 = (S(RS, k) + S(RS, k - 1)) + resm * (RPP(k) + RPP(k - 1));
        add_own(B, tend_rw, ow + k, dts * gr);
        add_own(B, cofwz, oc + k, -(A + resm * Bz) * gr);
        add_own(B, cofwr, oc + k, -R2 * gr);
        add_own(B, cofwt, oc + k, (S(TS, k) + resm * RTPP(k)) * gr);
        add_own(B, cofwt, oc + k - 1, (S(TS, k - 1) + resm * RTPP(k - 1)) * gr);
        S(GTS, k) += (-cofwz_[k] * zz[k] + cofwt_[k]) * gr;
        S(GTS, k - 1) += (cofwz_[k] * zz[k - 1] + cofwt_[k - 1]) * gr;
        S(GRS, k) += -cofwr_[k] * gr;
        S(GRS, k - 1) += -cofwr_[k] * gr;
        if (!first) {
            add_own(B, rw_p_in, ow + k, gr);
            add_own(B, rtheta_pp_in, oc + k, (-cofwz_[k] * resm * zz[k] + cofwt_[k] * resm) * gr);
            add_own(B, rtheta_pp_in, oc + k - 1, (cofwz_[k] * resm * zz[k - 1] + cofwt_[k - 1] * resm) * gr);
            add_own(B, rho_pp_in, oc + k, -cofwr_[k] * resm * gr);
            add_own(B, rho_pp_in, oc + k - 1, -cofwr_[k] * resm * gr);
        }
    }
    // rs, ts -> operands
    for (int k = 0; k < L; ++k) {
        const double grs = S(GRS, k), gts = S(GTS, k);
        add_own(B, tend_rho, oc + k, dts * grs);
        add_own(B, tend_rt, oc + k, dts * gts);
        add_own(B, coftz, ow + k + 1, -resm * rdzw[k] * RWP(k + 1) * gts);
        add_own(B, coftz, ow + k, resm * rdzw[k] * RWP(k) * gts);
        for (int i = 0; i < n; ++i) {
            const int e = m.edgesOnCell[c * ME + i];
            const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
            const double metric = m.edgesOnCell_sign[c * ME + i] * m.dvEdge[e];
            const double thedge = 0.5 * (F[theta_m][(long long)c2 * L + k] + F[theta_m][(long long)c1 * L + k]);
            const double flux = dts * metric * F[ru_p][(long long)e * L + k] * ia;
            const double gflux = -grs - thedge * gts;
            const double gth = -flux * gts;
            add(B, ru_p, (long long)e * L + k, dts * metric * ia * gflux);
            add(B, theta_m, (long long)c2 * L + k, 0.5 * gth);
            add(B, theta_m, (long long)c1 * L + k, 0.5 * gth);
        }
        if (!first) {
            add_own(B, rho_pp_in, oc + k, grs);
            add_own(B, rtheta_pp_in, oc + k, gts + bar_of(B, rtheta_pp_old, oc + k));
            add_own(B, rw_p_in, ow + k + 1, -cofrz_[k] * resm * grs - resm * rdzw[k] * coftz_[k + 1] * gts);
            add_own(B, rw_p_in, ow + k, cofrz_[k] * resm * grs + resm * rdzw[k] * coftz_[k] * gts);
        }
    }
    if (!first) {
        add_own(B, rw_p_in, ow + 0, S(GY, 0));
        add_own(B, rw_p_in, ow + L, S(GX, L));
        for (int k = 0; k <= L; ++k) {
            const double gww = bar_of(B, wwAvg, ow + k);
            add_own(B, wwAvg_in, ow + k, gww);
            if (k >= 1 && k < L) add_own(B, rw_p_in, ow + k, 0.5 * (1.0 - epssm) * gww);
        }
    }
}

// Edge pass, thread (edge e, level k), after the column pass.
__device__ __forceinline__ void acoustic_edge_vjp(const mpas_hip_mesh &m, const mpas_hip_acoustic_params &P, Fwd F,
                                                  Bar B, int e, int k) {
    using namespace acoustic;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    const double dts = P.dts;
    const double gavg = bar_of(B, ruAvg, i);
    const double g = bar_of(B, ru_p, i) + gavg;
    if (P.small_step == 1) {
        add_own(B, tend_ru, i, dts * g);
        return;
    }
    add_own(B, ruAvg_in, i, gavg);
    add_own(B, ru_p_in, i, g);
    add_own(B, tend_ru, i, dts * g);
    const double gpg = -dts * g;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const long long o1 = (long long)c1 * L + k, o2 = (long long)c2 * L + k;
    const double zs = 0.5 * (m.zz[o2] + m.zz[o1]);
    const double pg1 = ((F[rtheta_pp_in][o2] - F[rtheta_pp_in][o1]) * m.invDcEdge[e]) / zs;
    const double exs = F[exner][o1] + F[exner][o2];
    const double cq = F[cqu][i];
    const double gr = 0.5 * m.zxu[i] * MPAS_GRAVITY * gpg;
    add(B, rho_pp_in, o1, gr);
    add(B, rho_pp_in, o2, gr);
    add_own(B, cqu, i, 0.5 * P.c2 * exs * pg1 * gpg);
    const double gex = cq * 0.5 * P.c2 * pg1 * gpg;
    add(B, exner, o1, gex);
    add(B, exner, o2, gex);
    const double gp1 = cq * 0.5 * P.c2 * exs * gpg;
    const double gd = m.invDcEdge[e] / zs * gp1;
    add(B, rtheta_pp_in, o2, gd);
    add(B, rtheta_pp_in, o1, -gd);
}

}  // namespace adj
}  // namespace mpas
