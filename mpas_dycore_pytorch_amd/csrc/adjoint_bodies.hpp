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
// No thread writes an element another thread of the same launch writes.  A
// contribution an entity owes its neighbours (the transpose of a forward
// gather) is formed by the neighbour's own thread, which walks its stencil in
// slot order and either recomputes the term from the forward values and the
// output adjoints or reads it from a per-entity array an earlier launch
// stored; contributions between the levels of one column cross lanes of the
// column's wavefront with shuffles.  There are no atomics, so every adjoint
// is deterministic bit for bit.  A null bar[] slot means "no adjoint wanted"
// and is skipped.
#pragma once

#include "../../include/mpas_hip.h"
#include "scalar_math.hpp"

namespace mpas {
namespace adj {

using Fwd = const double *const *;
using Bar = double *const *;

constexpr int kMaxEdgesAdj = 8;  // cell slot cap (meshes beyond it are rejected on entry)

// Add v to element i of an adjoint this thread owns during the launch.
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

// Value of v held by lane k + d of the wavefront (d > 0 above, d < 0 below),
// 0 where that lane lies outside 0..hi.  Every lane must run it.
__device__ __forceinline__ double from_lane(double v, int d, int k, int hi) {
    const double t = (d > 0) ? __shfl_down(v, (unsigned)d) : __shfl_up(v, (unsigned)(-d));
    const int src = k + d;
    return (src >= 0 && src <= hi) ? t : 0.0;
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

// thread (cell c, level k): the sum over the cell's edges of the damping
// term's adjoints for this cell, recomputed from the edge's ru_p adjoint
__device__ __forceinline__ void divdamp_cells_vjp(const mpas_hip_mesh &m, double coef, Fwd F, Bar B, int c, int k) {
    using namespace divdamp;
    const int L = m.nLevels, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const int n = m.nEdgesOnCell[c];
    // indices first, then every slot's loads (the cell's own columns once,
    // the other cell's per edge), then the sums in slot order
    int eo[kMaxEdges], nbr[kMaxEdges];
    bool own2[kMaxEdges];  // this cell is the edge's second cell
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        eo[j] = m.edgesOnCell[c * ME + clamp_slot(j, n)];
        const int c1 = m.cellsOnEdge[2 * eo[j]], c2 = m.cellsOnEdge[2 * eo[j] + 1];
        own2[j] = (c2 == c);
        nbr[j] = (c2 == c) ? c1 : c2;
    }
    const long long o = (long long)c * L + k;
    const double rtc = F[rtheta_pp][o], roc = F[rtheta_pp_old][o], thc = F[theta_m][o];
    double gv[kMaxEdges], rtn[kMaxEdges], ron[kMaxEdges], thn[kMaxEdges];
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        gv[j] = rtn[j] = ron[j] = thn[j] = 0.0;
        if (j < n) {  // wave-uniform
            const long long on = (long long)nbr[j] * L + k;
            gv[j] = bar_of(B, ru_p_out, (long long)eo[j] * L + k);
            rtn[j] = F[rtheta_pp][on];
            ron[j] = F[rtheta_pp_old][on];
            thn[j] = F[theta_m][on];
        }
    }
    double grt = 0.0, gold = 0.0, gth = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        if (j >= n) break;
        const double g = gv[j];
        if (g == 0.0) continue;
        const double rt1 = own2[j] ? rtn[j] : rtc, rt2 = own2[j] ? rtc : rtn[j];
        const double ro1 = own2[j] ? ron[j] : roc, ro2 = own2[j] ? roc : ron[j];
        const double th1 = own2[j] ? thn[j] : thc, th2 = own2[j] ? thc : thn[j];
        const double d1 = -(rt1 - ro1);
        const double d2 = -(rt2 - ro2);
        const double ts = th1 + th2;
        if (!(fabs(ts) >= 1.0e-30)) continue;
        const double gd = coef * g / ts;  // adjoint of (d2 - d1)
        const double gts = -(coef * (d2 - d1) / ts) / ts * g;
        const double sg = own2[j] ? -1.0 : 1.0;
        grt += sg * gd;
        gold -= sg * gd;
        gth += gts;
    }
    add_own(B, rtheta_pp, o, grt);
    add_own(B, rtheta_pp_old, o, gold);
    add_own(B, theta_m, o, gth);
}

// thread (edge e, level k), e in [0, nEdges]: the edge's own part (the update
// is ru_p += term, so ru_p passes its adjoint through)
__device__ __forceinline__ void divdamp_edges_vjp(const mpas_hip_mesh &m, Bar B, int e, int k) {
    using namespace divdamp;
    const long long i = (long long)e * m.nLevels + k;
    add_own(B, ru_p_in, i, bar_of(B, ru_p_out, i));
}

// ===========================================================================
// atm_compute_moist_coefficients (moist_coefficients.py:124-291)
// fwd: scalars, cqw, cqu, qtot; moist species [ms-1, me)
// ===========================================================================
namespace moist {
enum { scalars, cqw, cqu, qtot, N };
}

// thread (cell c, level k): the qtot and cqw adjoints of the column and the
// cqu adjoints of the cell's edges (cqu(e) = 1 / (1 + 0.5 (qtot(c1) + qtot(c2)))),
// in slot order, into the moist scalars of the cell
__device__ __forceinline__ void moist_cell_vjp(const mpas_hip_mesh &m, int S, int q0, int q1, Fwd F, Bar B, int c,
                                               int k) {
    using namespace moist;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c >= m.nCells || k >= L || q1 <= q0) return;
    double g = bar_of(B, qtot, (long long)c * L + k);
    // cqw(k') = 1 / (1 + 0.5 (qtot(k') + qtot(k'-1))) for k' = k and k + 1
    for (int kk = k; kk <= k + 1; ++kk) {
        if (kk < 1 || kk >= L) continue;
        const double cw = F[cqw][(long long)c * Lw + kk];
        g += 0.5 * (-(cw * cw) * bar_of(B, cqw, (long long)c * Lw + kk));
    }
    if (B[cqu]) {
        const int n = m.nEdgesOnCell[c];
        for (int j = 0; j < n; ++j) {
            const long long i = (long long)m.edgesOnCell[c * ME + j] * L + k;
            const double cu = F[cqu][i];
            g += 0.5 * (-(cu * cu) * B[cqu][i]);
        }
    }
    for (int q = q0; q < q1; ++q) add_own(B, scalars, ((long long)c * L + k) * S + q, g);
}

// ===========================================================================
// mpas_reconstruct_2d (vector_reconstruction.py:311-441)
// fwd: u, uReconstructX, Y, Z, Zonal, Meridional
// ===========================================================================
namespace recon {
enum { u, ux, uy, uz, uzonal, umerid, N };
}

// adjoint of the Cartesian reconstruction of cell c at level k (the zonal and
// meridional projections folded in)
__device__ __forceinline__ void recon_cell_bar(const mpas_hip_mesh &m, Bar B, int c, int k, double &gx, double &gy,
                                               double &gz) {
    using namespace recon;
    const long long o = (long long)c * m.nLevels + k;
    const double clat = cos(m.latCell[c]), slat = sin(m.latCell[c]);
    const double clon = cos(m.lonCell[c]), slon = sin(m.lonCell[c]);
    const double gzon = bar_of(B, uzonal, o), gmer = bar_of(B, umerid, o);
    gx = bar_of(B, ux, o) - slon * gzon - clon * slat * gmer;
    gy = bar_of(B, uy, o) + clon * gzon - slon * slat * gmer;
    gz = bar_of(B, uz, o) + clat * gmer;
}

// thread (edge e, level k): u(e) feeds the reconstruction of its two cells,
// each through the coefficients of the cell's slot for e
__device__ __forceinline__ void reconstruct_vjp(const mpas_hip_mesh &m, Fwd F, Bar B, int e, int k) {
    using namespace recon;
    (void)F;
    const int L = m.nLevels, ME = m.maxEdges;
    if (e >= m.nEdges || k >= L) return;
    double g = 0.0;
    for (int q = 0; q < 2; ++q) {
        const int c = m.cellsOnEdge[2 * e + q];
        if (c >= m.nCells) continue;
        const int n = m.nEdgesOnCell[c];
        double gx, gy, gz;
        recon_cell_bar(m, B, c, k, gx, gy, gz);
        for (int j = 0; j < n; ++j) {
            if (m.edgesOnCell[c * ME + j] != e) continue;
            const double *cf = m.coeffs_reconstruct + ((long long)c * ME + j) * 3;
            g += cf[0] * gx + cf[1] * gy + cf[2] * gz;
        }
    }
    add_own(B, u, (long long)e * L + k, g);
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

// The w recovery of interface kk (0..L-1) of cell c, w = num / den with
// den = cf1 rz(0) + cf2 rz(1) + cf3 rz(2) at the surface and
// fzm rz(kk) + fzp rz(kk-1) inside: the adjoints of num and of den.
__device__ __forceinline__ void recover_w_partials(const mpas_hip_mesh &m, Fwd F, Bar B, int c, int kk, double &gnum,
                                                   double &gden) {
    using namespace recover;
    const int L = m.nLevels, Lw = L + 1;
    const double *rz = F[rho_zz] + (long long)c * L;
    const long long iw = (long long)c * Lw + kk;
    const double gw = bar_of(B, w, iw);
    const double den = (kk == 0) ? m.cf1 * rz[0] + m.cf2 * rz[1] + m.cf3 * rz[2]
                                 : m.fzm[kk] * rz[kk] + m.fzp[kk] * rz[kk - 1];
    gnum = gw / den;
    gden = -(F[w][iw] / den) * gw;
}
__device__ __forceinline__ double recover_cf(const mpas_hip_mesh &m, int j) {
    return (j == 0) ? m.cf1 : ((j == 1) ? m.cf2 : m.cf3);
}

// pass A, thread (cell c, interface k): rw / wwAvg
__device__ __forceinline__ void recover_w_vjp(const mpas_hip_mesh &m, double invNs, Fwd F, Bar B, int c, int k) {
    using namespace recover;
    const int L = m.nLevels, Lw = L + 1;
    const long long ow = (long long)c * Lw;
    if (c >= m.nCells || k > L) return;
    if (k == 0 || k == L) {
        add_own(B, wwAvg_in, ow + k, bar_of(B, wwAvg, ow + k));
        return;
    }
    const long long oc = (long long)c * L;
    const double zi = m.fzm[k] * m.zz[oc + k] + m.fzp[k] * m.zz[oc + k - 1];
    double gnum, gden;
    recover_w_partials(m, F, B, c, k, gnum, gden);
    const double grw = bar_of(B, rw, ow + k) + gnum / zi;
    const double gww = bar_of(B, wwAvg, ow + k);
    add_own(B, rw_save, ow + k, grw + gww);
    add_own(B, rw_p, ow + k, grw);
    add_own(B, wwAvg_in, ow + k, invNs * gww);
}

// terrain part of the w recovery of the cells of edge e: the adjoint it sends
// ru(e, k), from the surface interface (k <= 2) and interfaces k and k + 1
__device__ __forceinline__ double recover_terrain_ru(const mpas_hip_mesh &m, Fwd F, Bar B, int e, int k) {
    using namespace recover;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    const double *re = F[ru] + (long long)e * L;
    double g = 0.0;
    for (int q = 0; q < 2; ++q) {
        const int c = m.cellsOnEdge[2 * e + q];
        if (c >= m.nCells) continue;
        const int n = m.nEdgesOnCell[c];
        int j = 0;
        while (j < n && m.edgesOnCell[c * ME + j] != e) ++j;
        if (j == n) continue;
        const double sg = m.edgesOnCell_sign[c * ME + j];
        const long long zo = ((long long)c * ME + j) * Lw;
        double gnum, gden;
        if (k <= 2) {
            recover_w_partials(m, F, B, c, 0, gnum, gden);
            const double fs = m.cf1 * re[0] + m.cf2 * re[1] + m.cf3 * re[2];
            g += recover_cf(m, k) * (sg * (m.zb_cell[zo] + sign_pm(fs) * m.zb3_cell[zo]) * gnum);
        }
        for (int kk = k; kk <= k + 1; ++kk) {
            if (kk < 1 || kk > L - 1) continue;
            recover_w_partials(m, F, B, c, kk, gnum, gden);
            const double fi = m.fzm[kk] * re[kk] + m.fzp[kk] * re[kk - 1];
            const double gf = sg * (m.zb_cell[zo + kk] + sign_pm(fi) * m.zb3_cell[zo + kk]) * gnum;
            g += ((kk == k) ? m.fzm[kk] : m.fzp[kk]) * gf;
        }
    }
    return g;
}

// pass B, thread (edge e, level k): ruAvg, ru, u
__device__ __forceinline__ void recover_edge_vjp(const mpas_hip_mesh &m, double invNs, Fwd F, Bar B, int e, int k) {
    using namespace recover;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const double s = F[rho_zz][(long long)c1 * L + k] + F[rho_zz][(long long)c2 * L + k];
    const double gu = bar_of(B, u, i);
    double gru = bar_of(B, ru, i) + 2.0 * gu / s;
    if (!m.flatTerrain) gru += recover_terrain_ru(m, F, B, e, k);
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
    {
        // rho_zz(k) in the w recovery of the surface interface (k <= 2) and
        // of interfaces k and k + 1, then in u = 2 ru / (rho_zz(c1) + rho_zz(c2))
        // of the cell's edges
        double gnum, gden;
        if (k <= 2) {
            recover_w_partials(m, F, B, c, 0, gnum, gden);
            grz += recover_cf(m, k) * gden;
        }
        if (k >= 1) {
            recover_w_partials(m, F, B, c, k, gnum, gden);
            grz += m.fzm[k] * gden;
        }
        if (k + 1 <= L - 1) {
            recover_w_partials(m, F, B, c, k + 1, gnum, gden);
            grz += m.fzp[k + 1] * gden;
        }
        const int n = m.nEdgesOnCell[c];
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[c * m.maxEdges + j];
            const long long i = (long long)e * L + k;
            const double s = F[rho_zz][(long long)m.cellsOnEdge[2 * e] * L + k] +
                             F[rho_zz][(long long)m.cellsOnEdge[2 * e + 1] * L + k];
            grz += -(F[u][i] / s) * bar_of(B, u, i);
        }
    }
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
// scratch: the adjoints of pv_vertex (V), ke_vertex (V), pv_cell (C), of v
// (E, kept in g_kee) and pv_cell itself (C), each (N+1) x L and written by
// the pass that forms it
// ===========================================================================
namespace diag {
enum { u, h, v_in, h_edge, v, divergence, vorticity, ke, pv_edge, N };
}
struct DiagScratch {
    double *g_pvv, *g_kev, *g_pvc, *g_kee;
    double *pvc;  // pv_cell of every cell (C), formed once by the cell pass
};

__device__ __forceinline__ double diag_pv_cell(const mpas_hip_mesh &m, Fwd F, int c, int k) {
    using namespace diag;
    const int L = m.nLevels, ME = m.maxEdges;
    const int n = m.nEdgesOnCell[c];
    // the vertices first, then every slot's vorticity, then the sum in slot order
    int vxs[kMaxEdges];
#pragma unroll
    for (int i = 0; i < kMaxEdges; ++i) vxs[i] = m.verticesOnCell[c * ME + clamp_slot(i, n)];
    double zv[kMaxEdges];
#pragma unroll
    for (int i = 0; i < kMaxEdges; ++i) {
        zv[i] = 0.0;
        if (i < n) zv[i] = F[vorticity][(long long)vxs[i] * L + k];  // wave-uniform
    }
    double pc = 0.0;
#pragma unroll
    for (int i = 0; i < kMaxEdges; ++i) {
        if (i >= n) break;
        const int vx = vxs[i];
        const double ka = m.kiteAreasOnVertex[3 * vx + m.kiteForCell[c * ME + i]];
        const double pvv = m.fVertex[vx] + zv[i];
        const double cp = ka * pvv * m.invAreaCell[c];
        pc = (i == 0) ? cp : pc + cp;
    }
    return pc;
}

// ---------------------------------------------------------------------------
// The diagnostics adjoint as four gather passes over the inverse edgesOnEdge
// table.  Each target sums its sources; the per-edge apvm terms
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
    // (formed here, stored with the other results at the end)
    const double pvc = on ? diag_pv_cell(m, F, c, k) : 0.0;
    // the edges first, then every slot's loads, then the sums in slot order
    int eo[kMaxEdges];
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) eo[j] = m.edgesOnCell[c * ME + clamp_slot(j, n)];
    double gpe[kMaxEdges], uv[kMaxEdges], ghe[kMaxEdges];
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        gpe[j] = uv[j] = ghe[j] = 0.0;
        if (j < n) {  // wave-uniform
            const long long ie = (long long)eo[j] * L + k;
            gpe[j] = bar_of(B, pv_edge, ie);
            uv[j] = F[u][ie];
            ghe[j] = bar_of(B, h_edge, ie);
        }
    }
    double gpc = 0.0, gh = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        if (j >= n) break;
        const int e = eo[j];
        // t = apvm dt (-g_pv_edge), as diag_t forms it
        const double t = (on && gpe[j] != 0.0) ? apvm_dt * (-gpe[j]) : 0.0;
        if (t != 0.0) {
            const double ggn = t * uv[j] * m.invDcEdge[e];
            gpc += (m.cellsOnEdge[2 * e + 1] == c) ? ggn : -ggn;
        }
        gh += 0.5 * ghe[j];
    }
    const long long o = (long long)c * L + k;
    if (on) S.pvc[o] = pvc;
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
    // The vertex's edges and the cells of each edge, each cell taken at its
    // first naming (the three edges name each of the three cells twice); then
    // every load (edge adjoints, each cell's kite area by its slot for the
    // vertex, its pv_cell and KE adjoints); then the sums in the order
    // edge 0, its new cells, edge 1, ...
    int ev[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) ev[j] = m.edgesOnVertex[3 * vx + j];
    int cj[3][2];
    bool pr[3][2];
    int seen0 = -1, seen1 = -1, seen2 = -1;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int c = m.cellsOnEdge[2 * ev[j] + q];
            cj[j][q] = c;
            pr[j][q] = !(c == seen0 || c == seen1 || c == seen2);
            if (pr[j][q]) {
                if (seen0 < 0) seen0 = c;
                else if (seen1 < 0) seen1 = c;
                else seen2 = c;
            }
        }
    }
    double gpe[3], vv[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const long long ie = (long long)ev[j] * L + k;
        gpe[j] = bar_of(B, pv_edge, ie);
        vv[j] = F[v][ie];
    }
    double kav[3][2], iav[3][2], gpc[3][2], gke[3][2];
    bool hit[3][2];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            kav[j][q] = iav[j][q] = gpc[j][q] = gke[j][q] = 0.0;
            hit[j][q] = false;
            if (!pr[j][q]) continue;  // wave-uniform
            const int c = cj[j][q];
            const int n = m.nEdgesOnCell[c];
            int h = -1;  // the cell's first slot naming the vertex
#pragma unroll
            for (int i = kMaxEdges - 1; i >= 0; --i) {
                const int vc = m.verticesOnCell[c * ME + clamp_slot(i, n)];
                if (i < n && vc == vx) h = i;
            }
            if (h < 0) continue;
            const double ka = m.kiteAreasOnVertex[3 * vx + m.kiteForCell[c * ME + h]];
            const long long oc = (long long)c * L + k;
            kav[j][q] = ka;
            iav[j][q] = m.invAreaCell[c];
            gpc[j][q] = S.g_pvc[oc];
            gke[j][q] = bar_of(B, ke, oc);
            hit[j][q] = true;
        }
    }
    double gz = bar_of(B, vorticity, o), gkv = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int e = ev[j];
        gz += 0.5 * gpe[j];
        // t = apvm dt (-g_pv_edge), as diag_t forms it
        const double t = (on && gpe[j] != 0.0) ? apvm_dt * (-gpe[j]) : 0.0;
        if (t != 0.0) {
            const double ggt = t * vv[j] * m.invDvEdge[e];
            gz += (m.verticesOnEdge[2 * e + 1] == vx) ? ggt : -ggt;
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (!hit[j][q]) continue;
            gz += kav[j][q] * iav[j][q] * gpc[j][q];
            gkv += (1.0 - ke_fact) * kav[j][q] * iav[j][q] * gke[j][q];
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
        // the (edge, slot) entries naming e, L at a time across the lanes
        // (lanes 0..L-1 are all present), taken back by readlane; four
        // entries' gathers are issued before the first is used
        const int j0 = inv.eoeStart[e], cnt = inv.eoeStart[e + 1] - j0;
        for (int b0 = 0; b0 < cnt; b0 += L) {
            const int t_l = b0 + k;
            int eo_l = 0;
            double w_l = 0.0;
            if (t_l < cnt) {
                eo_l = inv.eoeEdge[j0 + t_l];
                w_l = m.weightsOnEdge[eo_l * ME2 + inv.eoeSlot[j0 + t_l]];
            }
            const int nt = (cnt - b0 < L) ? cnt - b0 : L;
            for (int t0 = 0; t0 < nt; t0 += 4) {
                double gk[4];
#pragma unroll
                for (int u4 = 0; u4 < 4; ++u4) {
                    gk[u4] = 0.0;
                    if (t0 + u4 < nt) gk[u4] = S.g_kee[(long long)bcast(eo_l, t0 + u4) * L + k];  // wave-uniform
                }
#pragma unroll
                for (int u4 = 0; u4 < 4; ++u4) {
                    if (t0 + u4 >= nt) break;
                    gu += bcast(w_l, t0 + u4) * gk[u4];
                }
            }
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

// Edge pass, thread (edge e, level k), after the column pass: the edge
// update ru_p' = ru_p + dts (tend_ru - pgrad) and ruAvg' = ruAvg + ru_p'.
// The pressure-gradient term's adjoints for the two cells are gathered by
// acoustic_pg_cell.
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
    add_own(B, cqu, i, 0.5 * P.c2 * exs * pg1 * gpg);
}

// thread (cell c, level k), after the edge pass (later substeps only): the
// pressure-gradient terms of the cell's edges, in slot order, into the
// rho_pp, exner and rtheta_pp adjoints of the cell.  Reads the complete ru_p
// adjoint the edge pass saw.
__device__ __forceinline__ void acoustic_pg_cell(const mpas_hip_mesh &m, const mpas_hip_acoustic_params &P, Fwd F,
                                                 Bar B, int c, int k) {
    using namespace acoustic;
    const int L = m.nLevels, ME = m.maxEdges;
    if (c >= m.nCells || k >= L || P.small_step == 1) return;
    const int n = m.nEdgesOnCell[c];
    double grp = 0.0, gex = 0.0, grt = 0.0;
    for (int j = 0; j < n; ++j) {
        const int e = m.edgesOnCell[c * ME + j];
        const long long i = (long long)e * L + k;
        const double gpg = -P.dts * (bar_of(B, ru_p, i) + bar_of(B, ruAvg, i));
        const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
        const long long o1 = (long long)c1 * L + k, o2 = (long long)c2 * L + k;
        const double zs = 0.5 * (m.zz[o2] + m.zz[o1]);
        const double pg1 = ((F[rtheta_pp_in][o2] - F[rtheta_pp_in][o1]) * m.invDcEdge[e]) / zs;
        const double exs = F[exner][o1] + F[exner][o2];
        const double cq = F[cqu][i];
        grp += 0.5 * m.zxu[i] * MPAS_GRAVITY * gpg;
        gex += cq * 0.5 * P.c2 * pg1 * gpg;
        const double gd = m.invDcEdge[e] / zs * (cq * 0.5 * P.c2 * exs * gpg);
        grt += (c == c2) ? gd : -gd;
    }
    const long long o = (long long)c * L + k;
    add_own(B, rho_pp_in, o, grp);
    add_own(B, exner, o, gex);
    add_own(B, rtheta_pp_in, o, grt);
}

}  // namespace adj
}  // namespace mpas
