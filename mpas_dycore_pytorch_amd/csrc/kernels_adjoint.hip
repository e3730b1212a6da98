// Adjoint kernels of the large step: launch wrappers over the bodies in
// adjoint_bodies.hpp / adjoint_dyn.hpp / adjoint_transport.hpp and the
// extern "C" entry points mpas_hip_*_vjp.
//
// Every entry point takes fwd[] (the forward inputs, then the forward
// outputs, of one call) and bar[] (their adjoints, same slots; input
// adjoints are accumulated into, output adjoints are read and may be used as
// accumulators).  Both are host arrays of device pointers.  No kernel uses
// atomics: every element is written by one thread of a launch, contributions
// between entities are gathered in slot order, so the adjoints are the same
// bit for bit on every run.
#include "adjoint_launch.hpp"

namespace mpas {
namespace adj {


ADJ_WAVE_KERNEL(k_adj_smlstep_cells, smlstep_cells_vjp(m, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_smlstep_edges, smlstep_edges_vjp(m, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_divdamp, divdamp_edges_vjp(m, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_divdamp_cells, divdamp_cells_vjp(m, A.coef, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_moist_cell, moist_cell_vjp(m, A.S, A.q0, A.q1, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_reconstruct, reconstruct_vjp(m, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_recover_w, recover_w_vjp(m, A.coef, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_recover_edge, recover_edge_vjp(m, A.coef, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_recover_cell, recover_cell_vjp(m, A.rk, A.dt, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_diag_g1, diag_g1(m, A.dt, A.apvm, A.rk, ADJ_F, ADJ_B, A.ds, ent, k))
ADJ_WAVE_KERNEL(k_adj_diag_g2, diag_g2(m, A.dt, A.apvm, ADJ_F, ADJ_B, A.ds, ent, k))
ADJ_WAVE_KERNEL(k_adj_diag_g3, diag_g3(m, A.dt, A.apvm, ADJ_F, ADJ_B, A.ds, ent, k))
ADJ_WAVE_KERNEL(k_adj_diag_g4, diag_g4(m, A.dt, A.apvm, A.rk, ADJ_F, ADJ_B, A.ds, A.inv, ent, k))
ADJ_WAVE_KERNEL(k_adj_acoustic_edge, acoustic_edge_vjp(m, A.ap, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_fwd_cells, dyn_fwd_cells(m, ADJ_F, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_fwd_edges, dyn_fwd_edges(m, ADJ_F, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_fwd_del4, dyn_fwd_del4(m, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_a1_del4, dyn_a1_del4(m, A.dp, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_a2_edges, dyn_a2_edges(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_a2_cells, dyn_a2_cells(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_a3_cells, dyn_a3_cells(m, A.dp, ADJ_F, ADJ_B, A.dy, A.inv, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_a3_edges, dyn_a3_edges(m, A.dp, ADJ_F, ADJ_B, A.dy, A.inv, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_a6, dyn_a6(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_a7, dyn_a7(m, A.dp, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_scal_cell, scal_cell_vjp(m, A.S, A.dt, A.coef3, A.w_old, A.w_new, ADJ_F, ADJ_B, A.sc, ent, k))
ADJ_WAVE_KERNEL(k_adj_scal_edge, scal_edge_vjp(m, A.S, ADJ_F, ADJ_B, A.sc, ent, k))
ADJ_COLUMN_KERNEL(k_adj_scal_col,
                  scal_col_vjp(m, A.inv, A.S, A.dt, A.coef3, A.w_old, A.w_new, ADJ_F, ADJ_B, A.sc, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_f1, mono_f1(m, A.S, A.s, A.coef3, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_f2, mono_f2(m, A.S, A.s, A.dt, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_f3, mono_f3(m, A.S, A.s, A.dt, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_f4, mono_f4(m, A.S, A.s, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r1, mono_r1(m, A.S, A.s, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r2, mono_r2(m, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r4, mono_r4(m, A.S, A.s, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r5, mono_r5(m, A.S, A.s, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_COLUMN_KERNEL(k_adj_mono_r6, mono_r6(m, A.S, A.s, A.dt, A.coef3, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r7, mono_r7(m, A.S, A.s, A.dt, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r7g, mono_r7g(m, A.inv, A.S, A.s, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r8, mono_r8(m, A.dt, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r8e, mono_r8e(m, A.dt, ADJ_B, ent, k))

ADJ_COLUMN_KERNEL(k_adj_dyn_a1, dyn_a1<false>(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_COLUMN_KERNEL(k_adj_dyn_a1_opt, dyn_a1<true>(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))

// ---------------------------------------------------------------------------
// Dynamics-tendency edge adjoint (A3), one wavefront per edge (lane = level):
//  * contributions to the edge's own column (u, ru, rho_edge, pv_edge, ...)
//    are summed in registers, the ones a lane owes the levels next to it
//    (the flux3 stencil of the u vertical flux, the averaged ru of the w
//    flux) cross lanes with shuffles, and each lane writes its level once;
//  * what the edge owes its two cells and vertices is stored per edge (the
//    x[] arrays) and summed by the cell / vertex gather pass, as are the
//    advection-stencil and PV-flux transposes (g_tf, g_wf, g_q).
// All 64 lanes run through the shuffles; lanes past the column compute on a
// clamped level and contribute nothing.
// ---------------------------------------------------------------------------
// kOpt builds in the curvature and vertical-mixing terms
template <bool kOpt>
__device__ __forceinline__ void adj_dyn_a3w(long long nb, mpas_hip_mesh m, Slots S_, KArgs A) {
    using namespace dyn;
    const int e = wave_entity_base(nb);
    if (e >= m.nEdges) return;  // wave-uniform
    const double *const *F = S_.f;
    double *const *B = S_.b;
    const DynAdjScratch &S = A.dy;
    const mpas_hip_dyn_params &P = A.dp;
    const int L = m.nLevels, Lw = L + 1, MA = m.maxAdv, ME2 = m.maxEdges2;
    const int lane = wave_lane();
    const bool act = lane < L;
    const double on = act ? 1.0 : 0.0;
    const int k = act ? lane : L - 1;
    const bool rk1 = (P.rk_step == 1);
    const long long i = (long long)e * L + k;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const double s1 = edge_sign_in_cell(m, c1, e), s2 = edge_sign_in_cell(m, c2, e);
    const double *ru_ = F[ru] + (long long)e * L;
    const int na = m.nAdvCellsForEdge[e];
    const int ne = m.nEdgesOnEdge[e];
    double g_ru_own = 0.0, g_ru_dn = 0.0;  // to ru(k) and ru(k-1)
    double x_ths = 0.0;
    // Slot rows across the wavefront: lane j loads slot j of the edge's
    // advection and PV-flux rows (clamped to the live count) with one vector
    // load each, and every slot loop takes its slot's value from lane j by
    // readlane.  The stencil gathers of each slot loop are all issued before
    // the first is used (a loop that loads and accumulates slot by slot waits
    // one full memory round trip per slot).
    const int sa = lane < na ? lane : (na > 0 ? na - 1 : 0);
    const int acv_l = m.advCellsForEdge[e * MA + sa];
    const double c0_l = m.adv_coefs[e * MA + sa], c3_l = m.adv_coefs_3rd[e * MA + sa];
    const int so = lane < ne ? lane : (ne > 0 ? ne - 1 : 0);
    const int eev = m.edgesOnEdge[e * ME2 + so];
    const double wv_l = m.weightsOnEdge[e * ME2 + so];
    // this edge's column at the neighbouring levels, from the neighbouring
    // lanes (lanes past the column hold the top level and feed nothing used)
    const double rut = ru_[k];
    const double rut_m1 = __shfl_up(rut, 1);
    const double *ue = F[u] + (long long)e * L;
    const double u0 = ue[k];
    const double um1 = __shfl_up(u0, 1), um2 = __shfl_up(u0, 2), up1 = __shfl_down(u0, 1);
    const double gtu = bar_of(B, tend_u, i);
    const double gtu_m1 = __shfl_up(gtu, 1);
    // ---- theta edge flux
    {
        const double gf = -(s1 * S.g_thadv[(long long)c1 * L + k] + s2 * S.g_thadv[(long long)c2 * L + k]);
        const double srt = sign_pm(rut);
        double ft = 0.0;
#pragma unroll
        for (int a0 = 0; a0 < kMaxAdv; a0 += 8) {
            if (a0 >= na) break;
            double qv[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                qv[t] = 0.0;
                if (a0 + t < na) qv[t] = F[theta_m][(long long)bcast(acv_l, a0 + t) * L + k];
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int a = a0 + t;
                if (a >= na) break;
                const double wgt = bcast(c0_l, a) + srt * bcast(c3_l, a);
                const double cc = wgt * qv[t];
                ft = (a == 0) ? cc : ft + cc;
            }
        }
        if (act) S.g_tf[i] = gf;
        g_ru_own += ft * gf;
        if (!rk1) {
            const double face = 0.5 * (F[theta_m_save][(long long)c2 * L + k] + F[theta_m_save][(long long)c1 * L + k]);
            const double dv = m.dvEdge[e];
            if (act) add_own(B, ru_save, i, dv * face * gf);
            g_ru_own += -dv * face * gf;
            x_ths = 0.5 * (dv * (F[ru_save][i] - rut) * gf);
        }
    }
    // ---- w edge flux at interface k (1..L-1)
    if (k >= 1) {
        const double gf = -(s1 * S.g_hadv[(long long)c1 * Lw + k] + s2 * S.g_hadv[(long long)c2 * Lw + k]);
        const double ruw = m.fzm[k] * rut + m.fzp[k] * rut_m1;
        const double srw = sign_pm(ruw);
        double fw = 0.0;
#pragma unroll
        for (int a0 = 0; a0 < kMaxAdv; a0 += 8) {
            if (a0 >= na) break;
            double qv[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                qv[t] = 0.0;
                if (a0 + t < na) qv[t] = F[w][(long long)bcast(acv_l, a0 + t) * Lw + k];
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int a = a0 + t;
                if (a >= na) break;
                const double wgt = bcast(c0_l, a) + srw * bcast(c3_l, a);
                const double cc = wgt * qv[t];
                fw = (a == 0) ? cc : fw + cc;
            }
        }
        if (act) S.g_wf[i] = gf;
        g_ru_own += m.fzm[k] * fw * gf;
        g_ru_dn = m.fzp[k] * fw * gf;
    }
    // ---- u tendency
    const double re = F[rho_edge][i];
    double g_re = 0.0, g_u_own = 0.0, g_pv_own = 0.0;
    double x_ke, x_hd, x_cw = 0.0;
    {
        const double pvs = F[pv_edge][i];
        const double gq = re * gtu;
        double q = 0.0, gpe = 0.0;
#pragma unroll
        for (int j0 = 0; j0 < kMaxEdges2; j0 += 8) {
            if (j0 >= ne) break;
            double uv[8], pv[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                uv[t] = pv[t] = 0.0;
                if (j0 + t < ne) {
                    const long long io = (long long)bcast(eev, j0 + t) * L + k;
                    uv[t] = F[u][io];
                    pv[t] = F[pv_edge][io];
                }
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int j = j0 + t;
                if (j >= ne) break;
                const double wv = bcast(wv_l, j);
                const double cc = wv * uv[t] * (0.5 * (pvs + pv[t]));
                q = (j == 0) ? cc : q + cc;
                gpe += wv * uv[t] * 0.5 * gq;
            }
        }
        const double ked = F[ke][(long long)c2 * L + k] - F[ke][(long long)c1 * L + k];
        const double hsum = F[h_divergence][(long long)c1 * L + k] + F[h_divergence][(long long)c2 * L + k];
        g_re += (q - ked * m.invDcEdge[e]) * gtu;
        x_ke = re * m.invDcEdge[e] * gtu;
        g_u_own += -0.5 * hsum * gtu;
        // Rayleigh damping of u on the top levels: tend_u -= rho_edge u damp(k)
        const int k0 = L - P.rayleigh_levels;
        if (P.rayleigh_levels > 0 && k >= k0) {
            const double damp = (((double)k + 1.0) - (double)k0) * P.rayleigh_coef_inverse;
            g_re += -(u0 * damp) * gtu;
            g_u_own += -(re * damp) * gtu;
        }
        // curvature: tend_u -= curvEdge re w_avg + u w_avg re / r_earth
        if (kOpt && P.curvature) {
            const double *w1 = F[w] + (long long)c1 * Lw, *w2 = F[w] + (long long)c2 * Lw;
            const double wa = 0.25 * (((w1[k] + w1[k + 1]) + w2[k]) + w2[k + 1]);
            const double ce = m.curvEdge[e], ir = 1.0 / kEarthRadius;
            g_re += -(ce * wa + u0 * wa * ir) * gtu;
            g_u_own += -(wa * re * ir) * gtu;
            x_cw = -0.25 * (ce * re + u0 * re * ir) * gtu;
        }
        x_hd = -u0 * 0.5 * gtu;
        g_pv_own += gpe;
        if (act) S.g_q[i] = gq;
    }
    // ---- vertical flux of u at interface k: contributions to u(k-2..k+1)
    double c_m2 = 0.0, c_m1 = 0.0, c_p1 = 0.0, x_rw = 0.0;
    if (k >= 1) {
        const double g = -m.rdzw[k - 1] * gtu_m1 + m.rdzw[k] * gtu;
        const double erw = 0.5 * (F[rw][(long long)c1 * Lw + k] + F[rw][(long long)c2 * Lw + k]);
        double gerw;
        if (k == 1 || k == L - 1) {
            const double ui = m.fzm[k] * u0 + m.fzp[k] * um1;
            gerw = ui * g;
            g_u_own += erw * m.fzm[k] * g;
            c_m1 = erw * m.fzp[k] * g;
        } else {
            double d[5];
            flux3_partials(um2, um1, u0, up1, erw, 1.0, d);
            c_m2 = d[0] * g;
            c_m1 = d[1] * g;
            g_u_own += d[2] * g;
            c_p1 = d[3] * g;
            gerw = d[4] * g;
        }
        x_rw = 0.5 * gerw;
    }
    // ---- tend_u_euler
    const double gtue = bar_of(B, tend_u_euler, i) + gtu;
    double x_pp = 0.0, x_dp = 0.0, x_kd = 0.0, x_dd = 0.0, x_dv = 0.0, g_dsu = 0.0;
    if (!rk1) {
        if (act) add_own(B, tend_u_euler_in, i, gtue);
    } else {
        const double zzs = m.zz[(long long)c2 * L + k] + m.zz[(long long)c1 * L + k];
        const double ppd = F[pp][(long long)c2 * L + k] - F[pp][(long long)c1 * L + k];
        const double dps = S.dpdz[(long long)c1 * L + k] + S.dpdz[(long long)c2 * L + k];
        const double cq = F[cqu][i];
        if (act) add_own(B, cqu, i, -(ppd * m.invDcEdge[e] / (0.5 * zzs) - 0.5 * m.zxu[i] * dps) * gtue);
        x_pp = cq * m.invDcEdge[e] / (0.5 * zzs) * gtue;  // pp(c1) += x, pp(c2) -= x
        x_dp = cq * 0.5 * m.zxu[i] * gtue;
        const double ud = S.delsq_u[i];
        const double kdu = 0.5 * (F[kdiff][(long long)c1 * L + k] + F[kdiff][(long long)c2 * L + k]);
        const double ms2 = m.meshScalingDel2[e];
        g_re += kdu * ud * ms2 * gtue;
        x_kd = 0.5 * re * ud * ms2 * gtue;
        g_dsu = re * kdu * ms2 * gtue;
        if (P.h_mom_eddy_visc4 > 0.0) {
            const int v1 = m.verticesOnEdge[2 * e], v2 = m.verticesOnEdge[2 * e + 1];
            const double ums = m.meshScalingDel4[e] * P.h_mom_eddy_visc4;
            const double r_dc = ums * P.del4u_div_factor * m.invDcEdge[e];
            const double r_dv = ums * dmin(m.invDvEdge[e], 4.0 * m.invDcEdge[e]);
            const double dd = S.delsq_div[(long long)c2 * L + k] - S.delsq_div[(long long)c1 * L + k];
            const double dvv = S.delsq_vort[(long long)v2 * L + k] - S.delsq_vort[(long long)v1 * L + k];
            g_re += -(dd * r_dc - dvv * r_dv) * gtue;
            x_dd = re * r_dc * gtue;
            x_dv = re * r_dv * gtue;
        }
        if (kOpt && P.v_mom_eddy_visc2 > 0.0 && k >= 1 && k <= L - 2) {
            // vertical mixing of u: re visc ((c - b)/(zp - z0) - (b - a)/(z0 - zm)) / h,
            // a, b, c the mixed u of levels k-1, k, k+1
            const VmixZ z = vmix_z_edge(m.zgrid + c1 * Lw, m.zgrid + c2 * Lw, k);
            const double ip = 1.0 / (z.zp - z.z0), im = 1.0 / (z.z0 - z.zm);
            const double a = vmix_u(m, P, um1, e, k - 1), b = vmix_u(m, P, u0, e, k), cc = vmix_u(m, P, up1, e, k + 1);
            const double sv = P.v_mom_eddy_visc2 / (0.5 * (z.zp - z.zm));
            g_re += sv * ((cc - b) * ip - (b - a) * im) * gtue;
            const double gr = re * sv * gtue;
            c_p1 += gr * ip;
            g_u_own += -gr * (ip + im);
            c_m1 += gr * im;
        }
    }
    // ---- lanes exchange what they owe the neighbouring levels
    c_m2 *= on;
    c_m1 *= on;
    c_p1 *= on;
    g_ru_dn *= on;
    const double from_p1 = __shfl_down(c_m1, 1);   // lane k+1's share for level k
    const double from_p2 = __shfl_down(c_m2, 2);   // lane k+2's share for level k
    const double from_m1 = __shfl_up(c_p1, 1);     // lane k-1's share for level k
    const double ru_from_p1 = __shfl_down(g_ru_dn, 1);
    if (!act) return;
    double gu = g_u_own;
    if (k + 1 <= L - 1) gu += from_p1;
    if (k + 2 <= L - 1) gu += from_p2;
    if (k >= 2) gu += from_m1;
    add_own(B, u, i, gu);
    add_own(B, ru, i, g_ru_own + ((k + 1 <= L - 1) ? ru_from_p1 : 0.0));
    add_own(B, rho_edge, i, g_re);
    add_own(B, pv_edge, i, g_pv_own);
    // what the two cells (and vertices) of the edge receive, summed there by
    // k_adj_dyn_a3_cells
    S.x[0][i] = x_ke;
    S.x[1][i] = x_hd;
    S.x[2][i] = (k >= 1) ? x_rw : 0.0;
    if (kOpt && P.curvature) S.x[8][i] = x_cw;
    if (rk1) {
        S.x[3][i] = x_pp;
        S.x[4][i] = x_dp;
        S.x[5][i] = x_kd;
        S.g_delsq_u[i] = g_dsu;
        if (P.h_mom_eddy_visc4 > 0.0) {
            S.x[6][i] = x_dd;
            S.x[7][i] = x_dv;
        }
    } else {
        S.x[3][i] = x_ths;
    }
}
__global__ __launch_bounds__(256) void k_adj_dyn_a3w(long long nb, long long n_ent, int n_lev, mpas_hip_mesh m,
                                                     Slots S_, KArgs A) {
    adj_dyn_a3w<false>(nb, m, S_, A);
}
__global__ __launch_bounds__(256) void k_adj_dyn_a3w_opt(long long nb, long long n_ent, int n_lev, mpas_hip_mesh m,
                                                         Slots S_, KArgs A) {
    adj_dyn_a3w<true>(nb, m, S_, A);
}

// ---------------------------------------------------------------------------
// Column adjoints, one wavefront per column (lane = level or interface).  The
// level-parallel parts use all lanes; the Thomas recurrences run on lane 0
// over the column staged in LDS, as in the forward kernels.
// ---------------------------------------------------------------------------
constexpr int kColWaves = 4;

// Adjoint of the acoustic column update (acoustic_col_vjp, same algebra).
__global__ __launch_bounds__(64 * kColWaves) void k_adj_acoustic_colw(mpas_hip_mesh m, Slots S_, KArgs A) {
    using namespace acoustic;
    __shared__ double s_rs[kColWaves][65], s_ts[kColWaves][65], s_r[kColWaves][65], s_y[kColWaves][65];
    __shared__ double s_x[kColWaves][65], s_gx[kColWaves][65], s_gy[kColWaves][65], s_gri[kColWaves][65];
    __shared__ double s_gr[kColWaves][65], s_grs[kColWaves][65], s_gts[kColWaves][65];
    __shared__ double s_fa[kColWaves][65], s_fal[kColWaves][65], s_fga[kColWaves][65];  // staged factor
    const double *const *F = S_.f;
    double *const *B = S_.b;
    const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
    // the column index is wave-uniform, so the cell's connectivity and
    // metrics are scalar loads
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.y), k = threadIdx.x;
    const int c = blockIdx.x * kColWaves + wv;
    const bool live = c < nC;
    const long long oc = (long long)c * L, ow = (long long)c * Lw;
    const mpas_hip_acoustic_params &P = A.ap;
    const double dts = P.dts, epssm = P.epssm, resm = P.resm;
    const bool first = (P.small_step == 1);
    const double *zz = m.zz + oc, *rdzw = m.rdzw, *fzm = m.fzm, *fzp = m.fzp;
    const double *cofrz_ = F[cofrz];
    auto RPP = [&](int kk) { return first ? 0.0 : F[rho_pp_in][oc + kk]; };
    auto RTPP = [&](int kk) { return first ? 0.0 : F[rtheta_pp_in][oc + kk]; };
    auto RWP = [&](int kk) { return first ? 0.0 : F[rw_p_in][ow + kk]; };
    auto COFWZ = [&](int kk) { return F[cofwz][oc + kk]; };
    auto COFWR = [&](int kk) { return F[cofwr][oc + kk]; };
    auto COFWT = [&](int kk) { return F[cofwt][oc + kk]; };
    auto COFTZ = [&](int kk) { return F[coftz][ow + kk]; };
    const int n = live ? m.nEdgesOnCell[c] : 0;
    const double ia = live ? m.invAreaCell[c] : 0.0;
    // ---- phase 1: rs, ts; stage the factor for the serial sweeps
    if (live && k < L) {
        s_fa[wv][k] = F[a_tri][oc + k];
        s_fal[wv][k] = F[alpha_tri][oc + k];
        s_fga[wv][k] = F[gamma_tri][oc + k];
        // indices first, then every slot's loads (the cell's own theta once,
        // the other cell's per edge), then the sums in slot order
        int eo[kMaxEdges], nbr[kMaxEdges];
        bool own2[kMaxEdges];
#pragma unroll
        for (int i = 0; i < kMaxEdges; ++i) {
            eo[i] = m.edgesOnCell[c * ME + clamp_slot(i, n)];
            const int c1 = m.cellsOnEdge[2 * eo[i]], c2 = m.cellsOnEdge[2 * eo[i] + 1];
            own2[i] = (c2 == c);
            nbr[i] = (c2 == c) ? c1 : c2;
        }
        const double tc = F[theta_m][oc + k];
        double rp[kMaxEdges], tn[kMaxEdges];
#pragma unroll
        for (int i = 0; i < kMaxEdges; ++i) {
            rp[i] = tn[i] = 0.0;
            if (i < n) {  // wave-uniform
                rp[i] = F[ru_p][(long long)eo[i] * L + k];
                tn[i] = F[theta_m][(long long)nbr[i] * L + k];
            }
        }
        double rsum = 0.0, tsum = 0.0;
#pragma unroll
        for (int i = 0; i < kMaxEdges; ++i) {
            if (i >= n) break;
            const double metric = m.edgesOnCell_sign[c * ME + i] * m.dvEdge[eo[i]];
            const double flux = dts * metric * rp[i] * ia;
            const double t2 = own2[i] ? tc : tn[i], t1 = own2[i] ? tn[i] : tc;
            const double thedge = 0.5 * (t2 + t1);
            const double rcon = -flux, tcon = -(flux * thedge);
            rsum = (i == 0) ? rcon : rsum + rcon;
            tsum = (i == 0) ? tcon : tsum + tcon;
        }
        s_rs[wv][k] = RPP(k) + dts * F[tend_rho][oc + k] + rsum - cofrz_[k] * resm * (RWP(k + 1) - RWP(k));
        s_ts[wv][k] = RTPP(k) + dts * F[tend_rt][oc + k] + tsum -
                      resm * rdzw[k] * (COFTZ(k + 1) * RWP(k + 1) - COFTZ(k) * RWP(k));
    }
    __syncthreads();
    // ---- phase 2: right-hand side r(k), 1 <= k < L
    if (live && k >= 1 && k < L) {
        const double tsk = s_ts[wv][k], tskm = s_ts[wv][k - 1], rsk = s_rs[wv][k], rskm = s_rs[wv][k - 1];
        s_r[wv][k] = RWP(k) + dts * F[tend_rw][ow + k] -
                     COFWZ(k) * ((zz[k] * tsk - zz[k - 1] * tskm) + resm * (zz[k] * RTPP(k) - zz[k - 1] * RTPP(k - 1))) -
                     COFWR(k) * ((rsk + rskm) + resm * (RPP(k) + RPP(k - 1))) + COFWT(k) * (tsk + resm * RTPP(k)) +
                     COFWT(k - 1) * (tskm + resm * RTPP(k - 1));
    }
    __syncthreads();
    // ---- phase 3: forward sweeps (lane 0)
    if (live && k == 0) {
        double y = RWP(0);
        s_y[wv][0] = y;
        for (int kk = 1; kk < L; ++kk) {
            y = (s_r[wv][kk] - s_fa[wv][kk] * y) * s_fal[wv][kk];
            s_y[wv][kk] = y;
        }
        double x = RWP(L);
        s_x[wv][L] = x;
        for (int kk = L - 1; kk >= 0; --kk) {
            x = s_y[wv][kk] - s_fga[wv][kk] * x;
            s_x[wv][kk] = x;
        }
    }
    __syncthreads();
    // ---- phase 4: adjoint of the final update and of the damping
    const double gr_k = (live && k < L) ? bar_of(B, rho_pp, oc + k) : 0.0;
    const double gt_k = (live && k < L) ? bar_of(B, rtheta_pp, oc + k) : 0.0;
    double ctz_acc = 0.0;  // adjoint of coftz(k), accumulated over the phases
    if (live && k <= L) {
        double gx = bar_of(B, rw_p, ow + k);
        if (k < L) gx += cofrz_[k] * gr_k + rdzw[k] * COFTZ(k) * gt_k;
        if (k >= 1) {
            const double grm = bar_of(B, rho_pp, oc + k - 1), gtm = bar_of(B, rtheta_pp, oc + k - 1);
            gx += -cofrz_[k - 1] * grm - rdzw[k - 1] * COFTZ(k) * gtm;
            ctz_acc += -rdzw[k - 1] * F[rw_p][ow + k] * gtm;
        }
        if (k < L) ctz_acc += rdzw[k] * F[rw_p][ow + k] * gt_k;
        double gri = 0.0;
        if (k >= 1 && k < L) {
            const double gww = bar_of(B, wwAvg, ow + k);
            const double grr = gx + 0.5 * (1.0 + epssm) * gww;
            const double dss = m.dss[oc + k];
            const double D = 1.0 + dts * dss;
            const double zi = fzm[k] * zz[k] + fzp[k] * zz[k - 1];
            const double ri = fzm[k] * F[rho_zz][oc + k] + fzp[k] * F[rho_zz][oc + k - 1];
            const double gq = grr / D;
            const double grwd = gq - grr;
            gri = -gq * dts * dss * zi * F[w][ow + k];
            add_own(B, w, ow + k, -gq * dts * dss * zi * ri);
            add_own(B, rw_save, ow + k, grwd);
            add_own(B, rw, ow + k, -grwd);
            gx = gq;
        }
        s_gx[wv][k] = gx;
        s_gri[wv][k] = gri;
    }
    __syncthreads();
    // ---- phase 5: rho_zz (all lanes); reverse sweeps (lane 0)
    if (live && k < L) {
        double g = 0.0;
        if (k >= 1) g += fzm[k] * s_gri[wv][k];
        if (k + 1 <= L - 1) g += fzp[k + 1] * s_gri[wv][k + 1];
        if (g != 0.0) add_own(B, rho_zz, oc + k, g);
    }
    if (live && k == 0) {
        for (int kk = 0; kk < L; ++kk) s_gx[wv][kk + 1] += -s_fga[wv][kk] * s_gx[wv][kk];
        for (int kk = 0; kk < L; ++kk) s_gy[wv][kk] = s_gx[wv][kk];
        for (int kk = L - 1; kk >= 1; --kk) s_gy[wv][kk - 1] += -s_fa[wv][kk] * s_fal[wv][kk] * s_gy[wv][kk];
    }
    __syncthreads();
    // ---- phase 6: factor adjoints and the right-hand-side adjoint g_r
    if (live && k < L) {
        add_own(B, gamma_tri, oc + k, -s_x[wv][k + 1] * s_gx[wv][k]);
        double grk = 0.0;
        if (k >= 1) {
            const double al = F[alpha_tri][oc + k], a = F[a_tri][oc + k];
            const double gy = s_gy[wv][k];
            grk = al * gy;
            add_own(B, alpha_tri, oc + k, (s_r[wv][k] - a * s_y[wv][k - 1]) * gy);
            add_own(B, a_tri, oc + k, -s_y[wv][k - 1] * grk);
        }
        s_gr[wv][k] = grk;
    }
    if (live && k == L) s_gr[wv][k] = 0.0;
    __syncthreads();
    // ---- phase 7: right-hand side -> rs, ts and the coefficients
    double grs = 0.0, gts = 0.0;
    if (live && k < L) {
        const double g0 = s_gr[wv][k];                         // r(k), zero at k = 0
        const double g1 = (k + 1 <= L - 1) ? s_gr[wv][k + 1] : 0.0;  // r(k+1)
        grs = gr_k - COFWR(k) * g0 - ((k + 1 <= L - 1) ? COFWR(k + 1) * g1 : 0.0);
        gts = gt_k + (-COFWZ(k) * zz[k] + COFWT(k)) * g0 +
              ((k + 1 <= L - 1) ? (COFWZ(k + 1) * zz[k] + COFWT(k)) * g1 : 0.0);
        add_own(B, cofwt, oc + k, (s_ts[wv][k] + resm * RTPP(k)) * (g0 + g1));
        if (k >= 1) {
            const double Az = zz[k] * s_ts[wv][k] - zz[k - 1] * s_ts[wv][k - 1];
            const double Bz = zz[k] * RTPP(k) - zz[k - 1] * RTPP(k - 1);
            const double R2 = (s_rs[wv][k] + s_rs[wv][k - 1]) + resm * (RPP(k) + RPP(k - 1));
            add_own(B, tend_rw, ow + k, dts * g0);
            add_own(B, cofwz, oc + k, -(Az + resm * Bz) * g0);
            add_own(B, cofwr, oc + k, -R2 * g0);
        }
        if (!first) {
            double grt = (-COFWZ(k) * resm * zz[k] + COFWT(k) * resm) * g0;
            double grp = -COFWR(k) * resm * g0;
            if (k + 1 <= L - 1) {
                grt += (COFWZ(k + 1) * resm * zz[k] + COFWT(k) * resm) * g1;
                grp += -COFWR(k + 1) * resm * g1;
            }
            add_own(B, rtheta_pp_in, oc + k, grt + gts + bar_of(B, rtheta_pp_old, oc + k));
            add_own(B, rho_pp_in, oc + k, grp + grs);
        }
        add_own(B, tend_rho, oc + k, dts * grs);
        add_own(B, tend_rt, oc + k, dts * gts);
        s_grs[wv][k] = grs;
        s_gts[wv][k] = gts;
    }
    __syncthreads();
    // ---- phase 8: interface adjoints (coftz, rw_p, wwAvg) and the edge scatter
    if (live && k <= L) {
        const double gts_k = (k < L) ? s_gts[wv][k] : 0.0, grs_k = (k < L) ? s_grs[wv][k] : 0.0;
        const double gts_m = (k >= 1) ? s_gts[wv][k - 1] : 0.0, grs_m = (k >= 1) ? s_grs[wv][k - 1] : 0.0;
        ctz_acc += resm * ((k < L) ? rdzw[k] : 0.0) * RWP(k) * gts_k;
        if (k >= 1) ctz_acc += -resm * rdzw[k - 1] * RWP(k) * gts_m;
        add_own(B, coftz, ow + k, ctz_acc);
        if (!first) {
            double g = 0.0;
            if (k < L) g += cofrz_[k] * resm * grs_k + resm * rdzw[k] * COFTZ(k) * gts_k;
            if (k >= 1) g += -cofrz_[k - 1] * resm * grs_m - resm * rdzw[k - 1] * COFTZ(k) * gts_m;
            const double gww = bar_of(B, wwAvg, ow + k);
            if (k >= 1 && k < L) g += s_gr[wv][k] + 0.5 * (1.0 - epssm) * gww;
            if (k == 0) g += s_gy[wv][0];
            if (k == L) g += s_gx[wv][L];
            add_own(B, rw_p_in, ow + k, g);
            add_own(B, wwAvg_in, ow + k, gww);
        }
    }
    if (live && k < L) {
        // The horizontal flux terms reach the cell's edges and neighbours:
        // keep this cell's rs/ts adjoints for the gather passes
        // (k_adj_acoustic_edge_flux, k_adj_acoustic_theta) and add the part
        // of theta_m(c) that uses them here.
        const double grs_k = s_grs[wv][k], gts_k = s_gts[wv][k];
        A.col[oc + k] = grs_k;
        A.col[(long long)(nC + 1) * L + oc + k] = gts_k;
        int eo[kMaxEdges];
#pragma unroll
        for (int i = 0; i < kMaxEdges; ++i) eo[i] = m.edgesOnCell[c * ME + clamp_slot(i, n)];
        double rp[kMaxEdges];
#pragma unroll
        for (int i = 0; i < kMaxEdges; ++i) {
            rp[i] = 0.0;
            if (i < n) rp[i] = F[ru_p][(long long)eo[i] * L + k];
        }
        double gth = 0.0;
#pragma unroll
        for (int i = 0; i < kMaxEdges; ++i) {
            if (i >= n) break;
            const double metric = m.edgesOnCell_sign[c * ME + i] * m.dvEdge[eo[i]];
            const double flux = dts * metric * rp[i] * ia;
            gth += 0.5 * (-flux * gts_k);
        }
        add_own(B, theta_m, oc + k, gth);
    }
}

// thread (edge e, level k) after k_adj_acoustic_colw: the ru_p adjoint of the
// two cells' flux terms, gathered from their rs/ts adjoints, then the edge
// update's own adjoint.
__device__ __forceinline__ void acoustic_flux_gather(const mpas_hip_mesh &m, const mpas_hip_acoustic_params &P,
                                                     Fwd F, Bar B, const double *col, int e, int k) {
    using namespace acoustic;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const long long o1 = (long long)c1 * L + k, o2 = (long long)c2 * L + k;
    const long long gts0 = (long long)(m.nCells + 1) * L;
    const double thedge = 0.5 * (F[theta_m][o2] + F[theta_m][o1]);
    const double dv = m.dvEdge[e];
    const double s1 = edge_sign_in_cell(m, c1, e), s2 = edge_sign_in_cell(m, c2, e);
    const double g1 = -col[o1] - thedge * col[gts0 + o1];
    const double g2 = -col[o2] - thedge * col[gts0 + o2];
    add_own(B, ru_p, (long long)e * L + k,
            P.dts * (s1 * dv) * m.invAreaCell[c1] * g1 + P.dts * (s2 * dv) * m.invAreaCell[c2] * g2);
}
__global__ __launch_bounds__(256) void k_adj_acoustic_edge_flux(long long nb, long long n_ent, int n_lev,
                                                                mpas_hip_mesh m, Slots S_, KArgs A) {
    const int ent = wave_entity_base(nb);
    const int k = wave_lane();
    if (ent >= n_ent || k >= n_lev) return;
    acoustic_flux_gather(m, A.ap, S_.f, S_.b, A.col, ent, k);
    acoustic_edge_vjp(m, A.ap, S_.f, S_.b, ent, k);
}
// thread (cell c, level k), after k_adj_acoustic_edge_flux: theta_m(c)
// through the flux terms of its neighbours (the cell's own were added in
// k_adj_acoustic_colw), then the pressure-gradient terms of its edges
__global__ __launch_bounds__(256) void k_adj_acoustic_theta(long long nb, long long n_ent, int n_lev,
                                                            mpas_hip_mesh m, Slots S_, KArgs A) {
    using namespace acoustic;
    const int c = wave_entity_base(nb);
    const int k = wave_lane();
    if (c >= n_ent || k >= n_lev) return;
    const double *const *F = S_.f;
    const int L = m.nLevels, ME = m.maxEdges;
    const long long gts0 = (long long)(m.nCells + 1) * L;
    const int n = m.nEdgesOnCell[c];
    int eo[kMaxEdges], cnv[kMaxEdges];
#pragma unroll
    for (int i = 0; i < kMaxEdges; ++i) {
        eo[i] = m.edgesOnCell[c * ME + clamp_slot(i, n)];
        const int c1 = m.cellsOnEdge[2 * eo[i]], c2 = m.cellsOnEdge[2 * eo[i] + 1];
        cnv[i] = (c1 == c) ? c2 : c1;
    }
    double rp[kMaxEdges], gt[kMaxEdges];
#pragma unroll
    for (int i = 0; i < kMaxEdges; ++i) {
        rp[i] = gt[i] = 0.0;
        if (i < n) {
            rp[i] = F[ru_p][(long long)eo[i] * L + k];
            gt[i] = A.col[gts0 + (long long)cnv[i] * L + k];
        }
    }
    double gth = 0.0;
#pragma unroll
    for (int i = 0; i < kMaxEdges; ++i) {
        if (i >= n) break;
        const double metric_n = -m.edgesOnCell_sign[c * ME + i] * m.dvEdge[eo[i]];  // sign of e in cn
        const double flux = A.ap.dts * metric_n * rp[i] * m.invAreaCell[cnv[i]];
        gth += 0.5 * (-flux * gt[i]);
    }
    add_own(S_.b, theta_m, (long long)c * L + k, gth);
    acoustic_pg_cell(m, A.ap, F, S_.b, c, k);
}

// Adjoint of the vertical-implicit coefficients (vimp_col_vjp, same algebra).
__global__ __launch_bounds__(64 * kColWaves) void k_adj_vimp_colw(mpas_hip_mesh m, Slots S_, KArgs A) {
    using namespace vimp;
    __shared__ double s_ga[kColWaves][65], s_gb[kColWaves][65], s_gc[kColWaves][65];
    __shared__ double s_gpi[kColWaves][65], s_gti[kColWaves][65];
    // the column the serial sweep walks, staged by all lanes
    __shared__ double s_al[kColWaves][65], s_a[kColWaves][65], s_gam[kColWaves][65], s_cc[kColWaves][65];
    __shared__ double s_bg[kColWaves][65], s_bl[kColWaves][65], s_ba[kColWaves][65];
    const double *const *F = S_.f;
    double *const *B = S_.b;
    const int L = m.nLevels, nC = m.nCells, Lw = L + 1;
    const int wv = threadIdx.y, k = threadIdx.x;
    const int c = blockIdx.x * kColWaves + wv;
    const bool live = c < nC;
    const double dtseps = 0.5 * A.dt * (1.0 + A.epssm);
    const double rcv = MPAS_RGAS / (MPAS_CP - MPAS_RGAS);
    const double c2 = MPAS_CP * rcv;
    const long long o = (long long)c * L, ow = (long long)c * Lw;
    const double *zz = m.zz + o, *rdzw = m.rdzw;
    auto WZ = [&](int kk) { return F[cofwz][o + kk]; };
    auto TZ = [&](int kk) { return F[coftz][ow + kk]; };
    auto WT = [&](int kk) { return F[cofwt][o + kk]; };
    auto WR = [&](int kk) { return F[cofwr][o + kk]; };
    // ---- phase 0: stage the factor, its adjoints and the upper diagonal c (all lanes)
    if (live && k < L) {
        s_al[wv][k] = F[alpha_tri][o + k];
        s_a[wv][k] = F[a_tri][o + k];
        s_gam[wv][k] = F[gamma_tri][o + k];
        s_cc[wv][k] = (k >= 1) ? -WZ(k) * TZ(k + 1) * rdzw[k] * zz[k] - WR(k) * F[cofrz][k] +
                                     WT(k) * TZ(k + 1) * rdzw[k]
                               : 0.0;
        s_bg[wv][k] = bar_of(B, gamma_tri, o + k);
        s_bl[wv][k] = bar_of(B, alpha_tri, o + k);
        s_ba[wv][k] = bar_of(B, a_tri, o + k);
    }
    __syncthreads();
    // ---- phase 1: reverse Thomas factor (lane 0, over LDS) -> adjoints of a, b, c
    if (live && k == 0) {
        double gg_carry = 0.0;
        s_ga[wv][0] = s_gb[wv][0] = s_gc[wv][0] = 0.0;
        s_ga[wv][L] = s_gb[wv][L] = s_gc[wv][L] = 0.0;
        for (int kk = L - 1; kk >= 1; --kk) {
            const double al = s_al[wv][kk], a = s_a[wv][kk];
            const double gprev = (kk > 1) ? s_gam[wv][kk - 1] : 0.0;
            const double cc = s_cc[wv][kk];
            const double g_ga = s_bg[wv][kk] + gg_carry;
            const double g_al = s_bl[wv][kk] + cc * g_ga;
            const double g_den = -(al * al) * g_al;
            s_gc[wv][kk] = al * g_ga;
            s_gb[wv][kk] = g_den;
            s_ga[wv][kk] = s_ba[wv][kk] - gprev * g_den;
            gg_carry = -a * g_den;
        }
    }
    __syncthreads();
    // ---- phase 2: totals of the coefficient adjoints; interface partials
    double Gwt = 0.0;
    if (live && k <= L) {
        auto GA = [&](int kk) { return (kk >= 1 && kk <= L - 1) ? s_ga[wv][kk] : 0.0; };
        auto GB = [&](int kk) { return (kk >= 1 && kk <= L - 1) ? s_gb[wv][kk] : 0.0; };
        auto GC = [&](int kk) { return (kk >= 1 && kk <= L - 1) ? s_gc[wv][kk] : 0.0; };
        double Gwz = 0.0, Gtz = 0.0;
        if (k >= 1 && k <= L - 1) {
            Gwz = bar_of(B, cofwz, o + k) - TZ(k - 1) * rdzw[k - 1] * zz[k - 1] * GA(k) +
                  (TZ(k) * rdzw[k] * zz[k] + TZ(k) * rdzw[k - 1] * zz[k - 1]) * GB(k) -
                  TZ(k + 1) * rdzw[k] * zz[k] * GC(k);
        }
        Gtz = bar_of(B, coftz, ow + k);
        if (k + 1 <= L - 1) Gtz += (-WZ(k + 1) * rdzw[k] * zz[k] - WT(k) * rdzw[k]) * GA(k + 1);
        if (k >= 1 && k <= L - 1)
            Gtz += (WZ(k) * (rdzw[k] * zz[k] + rdzw[k - 1] * zz[k - 1]) - (WT(k) * rdzw[k] - WT(k - 1) * rdzw[k - 1])) *
                   GB(k);
        if (k - 1 >= 1) Gtz += (-WZ(k - 1) * rdzw[k - 1] * zz[k - 1] + WT(k - 1) * rdzw[k - 1]) * GC(k - 1);
        if (k < L) {
            Gwt = bar_of(B, cofwt, o + k) - TZ(k) * rdzw[k] * GA(k + 1) - TZ(k) * rdzw[k] * GB(k) +
                  TZ(k + 1) * rdzw[k] * GB(k + 1) + TZ(k + 1) * rdzw[k] * GC(k);
        }
        double gpi = 0.0, gti = 0.0;
        if (k >= 1 && k <= L - 1) {
            const double zi = m.fzm[k] * zz[k] + m.fzp[k] * zz[k - 1];
            const double pi = m.fzm[k] * F[p][o + k] + m.fzp[k] * F[p][o + k - 1];
            const double base = dtseps * c2 * zi * m.rdzu[k];
            add_own(B, cqw, ow + k, base * pi * Gwz);
            gpi = base * F[cqw][ow + k] * Gwz;
            gti = dtseps * Gtz;
        }
        s_gpi[wv][k] = gpi;
        s_gti[wv][k] = gti;
    }
    __syncthreads();
    // ---- phase 3: level adjoints of p, t, rt, qtot
    if (live && k < L) {
        double gp = 0.0, gt = 0.0;
        if (k >= 1) {
            gp += m.fzm[k] * s_gpi[wv][k];
            gt += m.fzm[k] * s_gti[wv][k];
        }
        if (k + 1 <= L - 1) {
            gp += m.fzp[k + 1] * s_gpi[wv][k + 1];
            gt += m.fzp[k + 1] * s_gti[wv][k + 1];
        }
        if (Gwt != 0.0) {
            const double A0 = 0.5 * dtseps * rcv * zz[k] * MPAS_GRAVITY * m.rho_base[o + k];
            const double q1 = 1.0 + F[qtot][o + k], rr = m.rtheta_base[o + k] + F[rt][o + k], pb = m.exner_base[o + k];
            const double cw = WT(k);
            gp += A0 / q1 / (rr * pb) * Gwt;
            add_own(B, qtot, o + k, -(cw / q1) * Gwt);
            add_own(B, rt, o + k, -(cw / rr) * Gwt);
        }
        add_own(B, p, o + k, gp);
        add_own(B, t, o + k, gt);
    }
}


size_t vjp_scratch_doubles(const mpas_hip_mesh &m, int S) {
    const size_t L = m.nLevels, C = (m.nCells + 1) * L, E = (m.nEdges + 1) * L, V = (m.nVertices + 1) * L;
    size_t n = dyn_adj_doubles(m);
    n = std::max(n, 2 * V + 2 * C + E);                                      // diagnostics
    n = std::max(n, (size_t)acoustic::kScratchArrays * (L + 1) * m.nCells);  // acoustic columns
    n = std::max(n, (C + E) * (size_t)(S > 0 ? S : 1));                      // explicit transport
    n = std::max(n, mono_adj_doubles(m));                                    // monotonic transport
    return n;
}

}  // namespace adj
}  // namespace mpas

using namespace mpas;
using namespace mpas::adj;


static int vjp_check(const mpas_hip_mesh *m, const char *what) {
    if (!m) {
        set_error(std::string(what) + ": null mesh");
        return 1;
    }
    if (m->nLevels < 3 || m->nLevels > 63 || m->maxEdges > kMaxEdges || m->maxEdges2 > kMaxEdges2 ||
        m->maxAdv > kMaxAdv) {
        set_error(std::string(what) + ": mesh outside the kernels' limits");
        return 1;
    }
    return 0;
}
static int need(double *const *bar, int slot, const char *what) {
    if (!bar[slot]) {
        set_error(std::string(what) + ": this output adjoint must be provided (it is used as an accumulator)");
        return 1;
    }
    return 0;
}

extern "C" {

int64_t mpas_hip_vjp_scratch_doubles(int32_t nCells, int32_t nEdges, int32_t nVertices, int32_t nLevels,
                                     int32_t nScalars) {
    mpas_hip_mesh m{};
    m.nCells = nCells;
    m.nEdges = nEdges;
    m.nVertices = nVertices;
    m.nLevels = nLevels;
    return (int64_t)vjp_scratch_doubles(m, nScalars);
}

int mpas_hip_smlstep_pert_vjp(const mpas_hip_mesh *mesh, const double *const *fwd, double *const *bar, void *stream) {
    if (vjp_check(mesh, "smlstep_pert_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(smlstep::N, fwd, bar);
    KArgs A{};
    // gather form: w_tend per (cell, interface), u_tend per (edge, level)
    ADJ_LAUNCH(k_adj_smlstep_cells, m.nCells + 1, m.nLevels + 1, STREAM(stream));
    // u_tend enters through the terrain coefficients only: nothing on a flat mesh
    if (!m.flatTerrain) ADJ_LAUNCH(k_adj_smlstep_edges, m.nEdges, m.nLevels, STREAM(stream));
    return check_launch("smlstep_pert_vjp");
}

int mpas_hip_divergence_damping_vjp(const mpas_hip_mesh *mesh, double dts, double smdiv, double len_disp,
                                    const double *const *fwd, double *const *bar, void *stream) {
    if (vjp_check(mesh, "divergence_damping_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(divdamp::N, fwd, bar);
    KArgs A{};
    A.coef = 2.0 * smdiv * len_disp / dts;
    ADJ_LAUNCH(k_adj_divdamp, m.nEdges + 1, m.nLevels, STREAM(stream));
    ADJ_LAUNCH(k_adj_divdamp_cells, m.nCells, m.nLevels, STREAM(stream));
    return check_launch("divergence_damping_vjp");
}

int mpas_hip_moist_coefficients_vjp(const mpas_hip_mesh *mesh, int32_t nScalars, int32_t moist_start,
                                    int32_t moist_end, const double *const *fwd, double *const *bar, void *stream) {
    if (vjp_check(mesh, "moist_coefficients_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(moist::N, fwd, bar);
    KArgs A{};
    A.S = nScalars;
    A.q0 = moist_start - 1;
    A.q1 = moist_end;
    ADJ_LAUNCH(k_adj_moist_cell, m.nCells, m.nLevels, STREAM(stream));
    return check_launch("moist_coefficients_vjp");
}

int mpas_hip_reconstruct_vjp(const mpas_hip_mesh *mesh, const double *const *fwd, double *const *bar, void *stream) {
    if (vjp_check(mesh, "reconstruct_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(recon::N, fwd, bar);
    KArgs A{};
    ADJ_LAUNCH(k_adj_reconstruct, m.nEdges, m.nLevels, STREAM(stream));
    return check_launch("reconstruct_vjp");
}

int mpas_hip_recover_large_step_vjp(const mpas_hip_mesh *mesh, int32_t rk_step, double dt, int32_t ns,
                                    const double *const *fwd, double *const *bar, void *stream) {
    if (vjp_check(mesh, "recover_large_step_vjp")) return 1;
    if (need(bar, recover::rho_zz, "recover_large_step_vjp") || need(bar, recover::ru, "recover_large_step_vjp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(recover::N, fwd, bar);
    KArgs A{};
    A.rk = rk_step;
    A.dt = dt;
    A.coef = 1.0 / (double)ns;
    ADJ_LAUNCH(k_adj_recover_w, m.nCells, m.nLevels + 1, STREAM(stream));
    ADJ_LAUNCH(k_adj_recover_edge, m.nEdges, m.nLevels, STREAM(stream));
    ADJ_LAUNCH(k_adj_recover_cell, m.nCells, m.nLevels, STREAM(stream));
    return check_launch("recover_large_step_vjp");
}

int mpas_hip_solve_diagnostics_vjp(const mpas_hip_mesh *mesh, double dt, double apvm, int32_t rk_step,
                                   const double *const *fwd, double *const *bar, double *scratch,
                                   const mpas_hip_inverse_tables *inv, void *stream) {
    if (vjp_check(mesh, "solve_diagnostics_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    const size_t L = m.nLevels, C = (m.nCells + 1) * L, E = (m.nEdges + 1) * L, V = (m.nVertices + 1) * L;
    Slots S_ = make_slots(diag::N, fwd, bar);
    KArgs A{};
    A.dt = dt;
    A.apvm = apvm;
    A.rk = rk_step;
    A.ds = DiagScratch{scratch, scratch + V, scratch + 2 * V, scratch + 2 * V + C, scratch + 2 * V + C + E};
    if (!inv) {
        set_error("solve_diagnostics_vjp: the inverse connectivity tables are required");
        return 1;
    }
    // every live row of the scratch is written before it is read
    A.inv = *inv;
    ADJ_LAUNCH(k_adj_diag_g1, m.nEdges, m.nLevels, s);
    ADJ_LAUNCH(k_adj_diag_g2, m.nCells, m.nLevels, s);
    ADJ_LAUNCH(k_adj_diag_g3, m.nVertices, m.nLevels, s);
    ADJ_LAUNCH(k_adj_diag_g4, m.nEdges, m.nLevels, s);
    return check_launch("solve_diagnostics_vjp");
}

int mpas_hip_vert_imp_coefs_vjp(const mpas_hip_mesh *mesh, double dts, double epssm, const double *const *fwd,
                                double *const *bar, void *stream) {
    if (vjp_check(mesh, "vert_imp_coefs_vjp")) return 1;
    if (need(bar, vimp::cofwz, "vert_imp_coefs_vjp") || need(bar, vimp::coftz, "vert_imp_coefs_vjp") ||
        need(bar, vimp::cofwt, "vert_imp_coefs_vjp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(vimp::N, fwd, bar);
    KArgs A{};
    A.dt = dts;
    A.epssm = epssm;
    MPAS_LAUNCH(k_adj_vimp_colw, dim3((m.nCells + kColWaves - 1) / kColWaves), dim3(64, kColWaves), 0,
                STREAM(stream), m, S_, A);
    return check_launch("vert_imp_coefs_vjp");
}

int mpas_hip_acoustic_step_vjp(const mpas_hip_mesh *mesh, const mpas_hip_acoustic_params *p, const double *const *fwd,
                               double *const *bar, double *scratch, void *stream) {
    if (vjp_check(mesh, "acoustic_step_vjp")) return 1;
    if (need(bar, acoustic::ru_p, "acoustic_step_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(acoustic::N, fwd, bar);
    KArgs A{};
    A.ap = *p;
    A.col = scratch;
    MPAS_LAUNCH(k_adj_acoustic_colw, dim3((m.nCells + kColWaves - 1) / kColWaves), dim3(64, kColWaves), 0, s, m,
                S_, A);
    ADJ_LAUNCH(k_adj_acoustic_edge_flux, m.nEdges, m.nLevels, s);
    ADJ_LAUNCH(k_adj_acoustic_theta, m.nCells, m.nLevels, s);
    return check_launch("acoustic_step_vjp");
}

int mpas_hip_dyn_tend_vjp(const mpas_hip_mesh *mesh, const mpas_hip_dyn_params *p, const double *const *fwd,
                          double *const *bar, double *scratch, const mpas_hip_inverse_tables *inv, void *stream) {
    if (vjp_check(mesh, "dyn_tend_vjp")) return 1;
    if (need(bar, dyn::h_divergence, "dyn_tend_vjp") || need(bar, dyn::kdiff, "dyn_tend_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(dyn::N, fwd, bar);
    KArgs A{};
    A.dp = *p;
    A.dy = carve_dyn(m, scratch);
    if (!inv) {
        set_error("dyn_tend_vjp: the inverse connectivity tables are required");
        return 1;
    }
    A.inv = *inv;
    const int L = m.nLevels, nC = m.nCells, nE = m.nEdges, nV = m.nVertices;
    const bool rk1 = (p->rk_step == 1);
    const bool del4 = rk1 && p->h_mom_eddy_visc4 > 0.0;
    const int ncv = nC > nV ? nC : nV;
    // every scratch element a pass reads was written by an earlier pass of this call
    if (rk1) {
        ADJ_LAUNCH(k_adj_dyn_fwd_cells, nC, L, s);
        ADJ_LAUNCH(k_adj_dyn_fwd_edges, nE, L, s);
        ADJ_LAUNCH(k_adj_dyn_fwd_del4, ncv, L, s);
    }
    if (dyn_options(*p))
        ADJ_LAUNCH(k_adj_dyn_a1_opt, nC, L + 1, s);
    else
        ADJ_LAUNCH(k_adj_dyn_a1, nC, L + 1, s);
    if (rk1) {
        ADJ_LAUNCH(k_adj_dyn_a1_del4, nC, L + 1, s);
        ADJ_LAUNCH(k_adj_dyn_a2_edges, nE, L, s);
        ADJ_LAUNCH(k_adj_dyn_a2_cells, nC, L, s);
    }
    if (dyn_options(*p))
        ADJ_LAUNCH(k_adj_dyn_a3w_opt, nE, L, s);
    else
        ADJ_LAUNCH(k_adj_dyn_a3w, nE, L, s);
    ADJ_LAUNCH(k_adj_dyn_a3_cells, del4 ? ncv : nC, L, s);
    ADJ_LAUNCH(k_adj_dyn_a3_edges, nE, L, s);
    ADJ_LAUNCH(k_adj_dyn_a6, rk1 ? ncv : nC, L, s);
    ADJ_LAUNCH(k_adj_dyn_a7, nE, L, s);
    return check_launch("dyn_tend_vjp");
}

int mpas_hip_advance_scalars_vjp(const mpas_hip_mesh *mesh, int32_t nScalars, double dt, double coef3,
                                 double weight_old, double weight_new, const double *const *fwd, double *const *bar,
                                 double *scratch, const mpas_hip_inverse_tables *inv, void *stream) {
    if (vjp_check(mesh, "advance_scalars_vjp")) return 1;
    if (!inv) {
        set_error("advance_scalars_vjp: the inverse connectivity tables are required");
        return 1;
    }
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(scal::N, fwd, bar);
    KArgs A{};
    A.S = nScalars;
    A.dt = dt;
    A.coef3 = coef3;
    A.w_old = weight_old;
    A.w_new = weight_new;
    A.inv = *inv;
    const size_t C = (size_t)(m.nCells + 1) * m.nLevels;
    A.sc = ScalAdjScratch{scratch, scratch + C * (size_t)(nScalars > 0 ? nScalars : 1)};
    ADJ_LAUNCH(k_adj_scal_cell, m.nCells, m.nLevels, s);
    ADJ_LAUNCH(k_adj_scal_edge, m.nEdges, m.nLevels, s);
    ADJ_LAUNCH(k_adj_scal_col, m.nCells, m.nLevels + 1, s);
    return check_launch("advance_scalars_vjp");
}

int mpas_hip_advance_scalars_mono_vjp(const mpas_hip_mesh *mesh, int32_t nScalars, double dt, double coef3,
                                      const double *const *fwd, double *const *bar, double *scratch,
                                      const mpas_hip_inverse_tables *inv, void *stream) {
    if (vjp_check(mesh, "advance_scalars_mono_vjp")) return 1;
    if (need(bar, mono::rho_zz_int, "advance_scalars_mono_vjp")) return 1;
    if (!inv) {
        set_error("advance_scalars_mono_vjp: the inverse connectivity tables are required");
        return 1;
    }
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(mono::N, fwd, bar);
    KArgs A{};
    A.S = nScalars;
    A.dt = dt;
    A.coef3 = coef3;
    A.inv = *inv;
    A.mo = carve_mono(m, scratch);
    const int L = m.nLevels, nC = m.nCells, nE = m.nEdges;
    // every scratch element a pass reads was written by an earlier pass for the same scalar
    for (int sc = 0; sc < nScalars; ++sc) {
        A.s = sc;
        ADJ_LAUNCH(k_adj_mono_f1, nC, L + 1, s);
        ADJ_LAUNCH(k_adj_mono_f2, nE, L, s);
        ADJ_LAUNCH(k_adj_mono_f3, nC, L + 1, s);
        ADJ_LAUNCH(k_adj_mono_f4, nC, L, s);
        ADJ_LAUNCH(k_adj_mono_r1, nC, L, s);
        ADJ_LAUNCH(k_adj_mono_r2, nC, L + 1, s);
        ADJ_LAUNCH(k_adj_mono_r4, nC, L, s);
        ADJ_LAUNCH(k_adj_mono_r5, nC, L, s);
        ADJ_LAUNCH(k_adj_mono_r6, nC, L + 1, s);
        ADJ_LAUNCH(k_adj_mono_r7, nE, L, s);
        ADJ_LAUNCH(k_adj_mono_r7g, nC, L, s);
    }
    ADJ_LAUNCH(k_adj_mono_r8, nC, L + 1, s);
    ADJ_LAUNCH(k_adj_mono_r8e, nE, L, s);
    return check_launch("advance_scalars_mono_vjp");
}

}  // extern "C"
