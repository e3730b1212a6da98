// Tangent-linear model of atm_compute_dyn_tend_work (dyn_tend.py:9390-10833;
// the same sections as the forward kernels).  Included after adjoint_dyn.hpp:
// it shares the slot enum and the scratch layout, whose forward arrays
// (dpdz, delsq_u, delsq_div, delsq_vort) are filled by the same recompute
// passes and whose g_* arrays hold tangents here.
//
// Passes (each a launch; every thread gathers, none scatters):
//   F0  forward recompute of dpdz, delsq_u, delsq_div, delsq_vort (rk 1)
//   J1  cells x interfaces: h_divergence, tend_rho, kdiff; dpdz, delsq_theta
//       and delsq_w tangents (rk 1)
//   J2  edges (rk 1): delsq_u tangent
//   J3  cells + vertices (rk 1, del4): delsq_div, delsq_vort tangents
//   J4  edges: tend_u, tend_u_euler, and the per-edge advective fluxes of
//       theta and w
//   J5  cells x interfaces: tend_w_euler, tend_w, tend_theta_euler,
//       tend_theta, rthdynten
#pragma once

namespace mpas {
namespace adj {

// ---------------------------------------------------------------- J1
__device__ __forceinline__ void dyn_j1(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Fwd F, Dot T,
                                       const DynAdjScratch &S, int c, int k) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c > m.nCells || k > L) return;
    const bool rk1 = (P.rk_step == 1);
    const long long o = (long long)c * L + k, ow = (long long)c * Lw;
    if (c == m.nCells) {
        if (k < L) {
            dot_put(T, h_divergence, o, 0.0);
            dot_put(T, tend_rho, o, rk1 ? 0.0 : dot_of(T, tend_rho_in, o));
            dot_put(T, kdiff, o, rk1 ? 0.0 : dot_of(T, kdiff_in, o));
            S.g_dpdz[o] = 0.0;
            S.g_delsq_theta[o] = 0.0;
        }
        S.g_delsq_w[ow + k] = 0.0;
        return;
    }
    const int n = m.nEdgesOnCell[c];
    const double ia = m.invAreaCell[c];
    if (k < L) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[c * ME + j];
            s += m.edgesOnCell_sign[c * ME + j] * m.dvEdge[e] * dot_of(T, ru, (long long)e * L + k);
        }
        const double dhd = s * ia;
        dot_put(T, h_divergence, o, dhd);
        if (rk1) {
            dot_put(T, tend_rho, o,
                    -dhd - m.rdzw[k] * (dot_of(T, rw, ow + k + 1) - dot_of(T, rw, ow + k)));
            const double q = F[qtot][o];
            S.g_dpdz[o] = -MPAS_GRAVITY * ((m.rho_base[o] + F[rr_save][o]) * dot_of(T, qtot, o) +
                                           (1.0 + q) * dot_of(T, rr_save, o));
            // kdiff = min(coef def, cap), def = sqrt(dd^2 + dof^2), floored on the
            // CAM levels by max(kdiff, cam); a fixed kdiff has no tangent
            double dk = 0.0;
            if (!P.mix_fixed) {
                double dd = 0.0, dof = 0.0, ddd = 0.0, ddof = 0.0;
                for (int j = 0; j < n; ++j) {
                    const int e = m.edgesOnCell[c * ME + j];
                    const double a = m.defc_a[c * ME + j], b = m.defc_b[c * ME + j];
                    const long long ie = (long long)e * L + k;
                    const double us = F[u][ie], vs = F[v][ie], dus = dot_of(T, u, ie), dvs = dot_of(T, v, ie);
                    dd = (dd + a * us) - b * vs;
                    dof = (dof + b * us) + a * vs;
                    ddd += a * dus - b * dvs;
                    ddof += b * dus + a * dvs;
                }
                const double def = sqrt(dd * dd + dof * dof);
                double dsel, dcap;
                min_partials(P.kdiff_coef * def, P.kdiff_cap, dsel, dcap);
                if (dsel != 0.0 && def != 0.0) dk = P.kdiff_coef * dsel * (dd * ddd + dof * ddof) / def;
                if (k >= L - P.cam_levels) {
                    double dsm, dcam;
                    max_partials(dmin(P.kdiff_coef * def, P.kdiff_cap), cam_kdiff(P, L, k), dsm, dcam);
                    dk *= dsm;
                }
            }
            dot_put(T, kdiff, o, dk);
            // delsq_theta = sum es (th2 - th1) rho_edge
            double st = 0.0;
            for (int j = 0; j < n; ++j) {
                const int e = m.edgesOnCell[c * ME + j];
                const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
                const double es = ia * m.edgesOnCell_sign[c * ME + j] * m.dvEdge[e] * m.invDcEdge[e];
                const long long o1 = (long long)c1 * L + k, o2 = (long long)c2 * L + k, ie = (long long)e * L + k;
                st += es * ((dot_of(T, theta_m, o2) - dot_of(T, theta_m, o1)) * F[rho_edge][ie] +
                            (F[theta_m][o2] - F[theta_m][o1]) * dot_of(T, rho_edge, ie));
            }
            S.g_delsq_theta[o] = st;
        } else {
            dot_put(T, tend_rho, o, dot_of(T, tend_rho_in, o));
            dot_put(T, kdiff, o, dot_of(T, kdiff_in, o));
        }
    }
    // delsq_w at interface k = sum es (re_k + re_{k-1}) (w2 - w1)
    double sw = 0.0;
    if (rk1 && k >= 1 && k < L) {
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[c * ME + j];
            const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
            const double es = 0.5 * ia * m.edgesOnCell_sign[c * ME + j] * m.dvEdge[e] * m.invDcEdge[e];
            const long long ie = (long long)e * L + k;
            const long long w1 = (long long)c1 * Lw + k, w2 = (long long)c2 * Lw + k;
            sw += es * ((dot_of(T, rho_edge, ie) + dot_of(T, rho_edge, ie - 1)) * (F[w][w2] - F[w][w1]) +
                        (F[rho_edge][ie] + F[rho_edge][ie - 1]) * (dot_of(T, w, w2) - dot_of(T, w, w1)));
        }
    }
    S.g_delsq_w[ow + k] = sw;
}

// ---------------------------------------------------------------- J2
__device__ __forceinline__ void dyn_j2(const mpas_hip_mesh &m, Dot T, const DynAdjScratch &S, int e, int k) {
    using namespace dyn;
    const int L = m.nLevels;
    if (e > m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    if (e == m.nEdges) {
        S.g_delsq_u[i] = 0.0;
        return;
    }
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const int v1 = m.verticesOnEdge[2 * e], v2 = m.verticesOnEdge[2 * e + 1];
    const double r_dc = m.invDcEdge[e], r_dv = dmin(m.invDvEdge[e], 4.0 * m.invDcEdge[e]);
    S.g_delsq_u[i] = (dot_of(T, divergence, (long long)c2 * L + k) - dot_of(T, divergence, (long long)c1 * L + k)) * r_dc -
                     (dot_of(T, vorticity, (long long)v2 * L + k) - dot_of(T, vorticity, (long long)v1 * L + k)) * r_dv;
}

// ---------------------------------------------------------------- J3
__device__ __forceinline__ void dyn_j3(const mpas_hip_mesh &m, const DynAdjScratch &S, int i, int k) {
    const int L = m.nLevels, ME = m.maxEdges;
    if (k >= L) return;
    if (i < m.nVertices) {
        double s = 0.0;
        for (int j = 0; j < 3; ++j) {
            const int e = m.edgesOnVertex[3 * i + j];
            s += m.invAreaTriangle[i] * m.dcEdge[e] * m.edgesOnVertex_sign[3 * i + j] * S.g_delsq_u[(long long)e * L + k];
        }
        S.g_delsq_vort[(long long)i * L + k] = s;
    }
    if (i < m.nCells) {
        const int n = m.nEdgesOnCell[i];
        double s = 0.0;
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[i * ME + j];
            s += m.invAreaCell[i] * m.dvEdge[e] * m.edgesOnCell_sign[i * ME + j] * S.g_delsq_u[(long long)e * L + k];
        }
        S.g_delsq_div[(long long)i * L + k] = s;
    }
}

// ---------------------------------------------------------------- J4
// tangent of the u vertical flux at interface kk (1..L-1) of edge e
__device__ __forceinline__ double dyn_dwduz(const mpas_hip_mesh &m, Fwd F, Dot T, int e, int c1, int c2, int kk) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1;
    const long long eu = (long long)e * L;
    const double erw = 0.5 * (F[rw][(long long)c1 * Lw + kk] + F[rw][(long long)c2 * Lw + kk]);
    const double derw = 0.5 * (dot_of(T, rw, (long long)c1 * Lw + kk) + dot_of(T, rw, (long long)c2 * Lw + kk));
    if (kk == 1 || kk == L - 1) {
        return derw * (m.fzm[kk] * F[u][eu + kk] + m.fzp[kk] * F[u][eu + kk - 1]) +
               erw * (m.fzm[kk] * dot_of(T, u, eu + kk) + m.fzp[kk] * dot_of(T, u, eu + kk - 1));
    }
    return flux3_dot(F[u][eu + kk - 2], F[u][eu + kk - 1], F[u][eu + kk], F[u][eu + kk + 1], erw, 1.0,
                     dot_of(T, u, eu + kk - 2), dot_of(T, u, eu + kk - 1), dot_of(T, u, eu + kk),
                     dot_of(T, u, eu + kk + 1), derw);
}

// kOpt builds in the curvature and vertical-mixing terms (J4 and J5)
template <bool kOpt>
__device__ __forceinline__ void dyn_j4(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Fwd F, Dot T,
                                       const DynAdjScratch &S, int e, int k) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1, ME2 = m.maxEdges2, MA = m.maxAdv;
    if (e > m.nEdges || k >= L) return;
    const bool rk1 = (P.rk_step == 1);
    const long long i = (long long)e * L + k;
    if (e == m.nEdges) {
        dot_put(T, tend_u, i, 0.0);
        dot_put(T, tend_u_euler, i, rk1 ? 0.0 : dot_of(T, tend_u_euler_in, i));
        S.g_tf[i] = 0.0;
        S.g_wf[i] = 0.0;
        return;
    }
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const long long o1 = (long long)c1 * L + k, o2 = (long long)c2 * L + k;
    // vertical transport of u
    double ddelta = 0.0;
    if (L > 1) {
        if (k == 0) ddelta = dyn_dwduz(m, F, T, e, c1, c2, 1);
        else if (k == L - 1) ddelta = -dyn_dwduz(m, F, T, e, c1, c2, L - 1);
        else ddelta = dyn_dwduz(m, F, T, e, c1, c2, k + 1) - dyn_dwduz(m, F, T, e, c1, c2, k);
    }
    const double dvert = -m.rdzw[k] * ddelta;
    // PV flux q; four slots' gathers are issued before the first is used
    double q = 0.0, dq = 0.0;
    const int ne = m.nEdgesOnEdge[e];
    const double pve = F[pv_edge][i], dpve = dot_of(T, pv_edge, i);
#pragma unroll
    for (int j0 = 0; j0 < kMaxEdges2; j0 += 4) {
        if (j0 >= ne) break;
        double uj[4], pj[4], duj[4], dpj[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            uj[t] = pj[t] = duj[t] = dpj[t] = 0.0;
            if (j0 + t < ne) {  // wave-uniform
                const long long ie = (long long)m.edgesOnEdge[e * ME2 + j0 + t] * L + k;
                uj[t] = F[u][ie];
                pj[t] = F[pv_edge][ie];
                duj[t] = dot_of(T, u, ie);
                dpj[t] = dot_of(T, pv_edge, ie);
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int j = j0 + t;
            if (j >= ne) break;
            const double wgt = m.weightsOnEdge[e * ME2 + j];
            const double workpv = 0.5 * (pve + pj[t]);
            const double cc = wgt * uj[t] * workpv;
            q = (j == 0) ? cc : q + cc;
            dq += wgt * (duj[t] * workpv + uj[t] * 0.5 * (dpve + dpj[t]));
        }
    }
    const double ked = F[ke][o2] - F[ke][o1], dked = dot_of(T, ke, o2) - dot_of(T, ke, o1);
    const double hsum = F[h_divergence][o1] + F[h_divergence][o2];
    const double dhsum = dot_of(T, h_divergence, o1) + dot_of(T, h_divergence, o2);
    const double re = F[rho_edge][i], dre = dot_of(T, rho_edge, i), ue = F[u][i], due = dot_of(T, u, i);
    const double dbase = dre * (q - ked * m.invDcEdge[e]) + re * (dq - dked * m.invDcEdge[e]) - due * 0.5 * hsum -
                         ue * 0.5 * dhsum;
    double dtue;
    if (rk1) {
        // pressure gradient and buoyancy
        const double zzs = m.zz[o2] + m.zz[o1];
        const double ppd = F[pp][o2] - F[pp][o1], dppd = dot_of(T, pp, o2) - dot_of(T, pp, o1);
        const double dps = S.dpdz[o1] + S.dpdz[o2], ddps = S.g_dpdz[o1] + S.g_dpdz[o2];
        const double X = ppd * m.invDcEdge[e] / (0.5 * zzs) - 0.5 * m.zxu[i] * dps;
        const double dX = dppd * m.invDcEdge[e] / (0.5 * zzs) - 0.5 * m.zxu[i] * ddps;
        dtue = -dot_of(T, cqu, i) * X - F[cqu][i] * dX;
        // del2 mixing
        const double ud = S.delsq_u[i], dud = S.g_delsq_u[i];
        const double kdu = 0.5 * (F[kdiff][o1] + F[kdiff][o2]);
        const double dkdu = 0.5 * (dot_of(T, kdiff, o1) + dot_of(T, kdiff, o2));
        dtue += (dre * kdu * ud + re * dkdu * ud + re * kdu * dud) * m.meshScalingDel2[e];
        // del4 mixing
        if (P.h_mom_eddy_visc4 > 0.0) {
            const int v1 = m.verticesOnEdge[2 * e], v2 = m.verticesOnEdge[2 * e + 1];
            const double ums = m.meshScalingDel4[e] * P.h_mom_eddy_visc4;
            const double r_dc = ums * P.del4u_div_factor * m.invDcEdge[e];
            const double r_dv = ums * dmin(m.invDvEdge[e], 4.0 * m.invDcEdge[e]);
            const long long w1 = (long long)v1 * L + k, w2 = (long long)v2 * L + k;
            const double Y = (S.delsq_div[o2] - S.delsq_div[o1]) * r_dc - (S.delsq_vort[w2] - S.delsq_vort[w1]) * r_dv;
            const double dY =
                (S.g_delsq_div[o2] - S.g_delsq_div[o1]) * r_dc - (S.g_delsq_vort[w2] - S.g_delsq_vort[w1]) * r_dv;
            dtue -= dre * Y + re * dY;
        }
        if (kOpt && P.v_mom_eddy_visc2 > 0.0 && k >= 1 && k <= L - 2) {
            // vertical mixing of u (the mixed u's departures are constant shifts)
            const VmixZ z = vmix_z_edge(m.zgrid + c1 * Lw, m.zgrid + c2 * Lw, k);
            const double ip = 1.0 / (z.zp - z.z0), im = 1.0 / (z.z0 - z.zm);
            const double a = vmix_u(m, P, F[u][i - 1], e, k - 1), b = vmix_u(m, P, ue, e, k);
            const double cc = vmix_u(m, P, F[u][i + 1], e, k + 1);
            const double da = dot_of(T, u, i - 1), db = due, dc = dot_of(T, u, i + 1);
            const double sv = P.v_mom_eddy_visc2 / (0.5 * (z.zp - z.zm));
            dtue += sv * (dre * ((cc - b) * ip - (b - a) * im) + re * ((dc - db) * ip - (db - da) * im));
        }
    } else {
        dtue = dot_of(T, tend_u_euler_in, i);
    }
    // curvature: tend_u -= curvEdge re w_avg + u w_avg re / r_earth
    double dcurv = 0.0;
    if (kOpt && P.curvature) {
        const long long w1 = (long long)c1 * Lw + k, w2 = (long long)c2 * Lw + k;
        const double wa = 0.25 * (((F[w][w1] + F[w][w1 + 1]) + F[w][w2]) + F[w][w2 + 1]);
        const double dwa =
            0.25 * (((dot_of(T, w, w1) + dot_of(T, w, w1 + 1)) + dot_of(T, w, w2)) + dot_of(T, w, w2 + 1));
        const double ce = m.curvEdge[e], ir = 1.0 / kEarthRadius;
        dcurv = ce * (dre * wa + re * dwa) + ((due * wa + ue * dwa) * re + ue * wa * dre) * ir;
    }
    // Rayleigh damping of u on the top levels: tend_u -= rho_edge u damp(k)
    double dray = 0.0;
    const int k0 = L - P.rayleigh_levels;
    if (P.rayleigh_levels > 0 && k >= k0) {
        const double damp = (((double)k + 1.0) - (double)k0) * P.rayleigh_coef_inverse;
        dray = (dre * ue + re * due) * damp;
    }
    // per-edge advective fluxes: theta at level k, w at interface k (the
    // stores of this thread all come last, so the stencil rows stay scalar
    // loads; four slots' gathers are issued before the first is used)
    const int na = m.nAdvCellsForEdge[e];
    const double ru0 = F[ru][i], dru0 = dot_of(T, ru, i);
    const double ru_m1 = __shfl_up(ru0, 1), dru_m1 = __shfl_up(dru0, 1);  // level k - 1, lane k - 1
    double g_tf, g_wf = 0.0;
    {
        const double sr = sign_pm(ru0);
        double fs = 0.0, dfs = 0.0;
#pragma unroll
        for (int a0 = 0; a0 < kMaxAdv; a0 += 4) {
            if (a0 >= na) break;
            double qv[4], dqv[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                qv[t] = dqv[t] = 0.0;
                if (a0 + t < na) {  // wave-uniform
                    const long long ia = (long long)m.advCellsForEdge[e * MA + a0 + t] * L + k;
                    qv[t] = F[theta_m][ia];
                    dqv[t] = dot_of(T, theta_m, ia);
                }
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int a = a0 + t;
                if (a >= na) break;
                const double wgt = m.adv_coefs[e * MA + a] + sr * m.adv_coefs_3rd[e * MA + a];
                fs += wgt * qv[t];
                dfs += wgt * dqv[t];
            }
        }
        g_tf = dru0 * fs + ru0 * dfs;
    }
    if (k >= 1) {
        const double rue = m.fzm[k] * ru0 + m.fzp[k] * ru_m1;
        const double drue = m.fzm[k] * dru0 + m.fzp[k] * dru_m1;
        const double sr = sign_pm(rue);
        double fs = 0.0, dfs = 0.0;
#pragma unroll
        for (int a0 = 0; a0 < kMaxAdv; a0 += 4) {
            if (a0 >= na) break;
            double qv[4], dqv[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                qv[t] = dqv[t] = 0.0;
                if (a0 + t < na) {  // wave-uniform
                    const long long ia = (long long)m.advCellsForEdge[e * MA + a0 + t] * Lw + k;
                    qv[t] = F[w][ia];
                    dqv[t] = dot_of(T, w, ia);
                }
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int a = a0 + t;
                if (a >= na) break;
                const double wgt = m.adv_coefs[e * MA + a] + sr * m.adv_coefs_3rd[e * MA + a];
                fs += wgt * qv[t];
                dfs += wgt * dqv[t];
            }
        }
        g_wf = drue * fs + rue * dfs;
    }
    dot_put(T, tend_u_euler, i, dtue);
    dot_put(T, tend_u, i, dvert + dbase - dcurv - dray + dtue);
    S.g_tf[i] = g_tf;
    S.g_wf[i] = g_wf;
}

// ---------------------------------------------------------------- J5
// tangent of the w vertical flux at mass level j (0..L-2) of cell c
__device__ __forceinline__ double dyn_dwdwz(const mpas_hip_mesh &m, Fwd F, Dot T, int c, int j) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1;
    const long long ow = (long long)c * Lw;
    const double *wc = F[w] + ow, *rwc = F[rw] + ow;
    auto dw = [&](int kk) { return dot_of(T, w, ow + kk); };
    auto drw = [&](int kk) { return dot_of(T, rw, ow + kk); };
    if (j == 0 || j == L - 2)
        return 0.25 * ((drw(j + 1) + drw(j)) * (wc[j + 1] + wc[j]) + (rwc[j + 1] + rwc[j]) * (dw(j + 1) + dw(j)));
    return flux3_dot(wc[j - 1], wc[j], wc[j + 1], wc[j + 2], 0.5 * (rwc[j + 1] + rwc[j]), 1.0, dw(j - 1), dw(j),
                     dw(j + 1), dw(j + 2), 0.5 * (drw(j + 1) + drw(j)));
}
// tangent of the theta vertical flux at interface kk (1..L-1) of cell c
__device__ __forceinline__ double dyn_dwdtz(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Fwd F, Dot T,
                                            int c, int kk) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1;
    const long long oc = (long long)c * L, ow = (long long)c * Lw;
    const double *th = F[theta_m] + oc, *ths = F[theta_m_save] + oc, *rwc = F[rw] + ow, *rws = F[rw_save] + ow;
    auto dth = [&](int j) { return dot_of(T, theta_m, oc + j); };
    auto dths = [&](int j) { return dot_of(T, theta_m_save, oc + j); };
    const double fm = m.fzm[kk], fp = m.fzp[kk];
    const double drwc = dot_of(T, rw, ow + kk), drws = dot_of(T, rw_save, ow + kk);
    const double ti = fm * th[kk] + fp * th[kk - 1], dti = fm * dth(kk) + fp * dth(kk - 1);
    const double tsi = fm * ths[kk] + fp * ths[kk - 1], dtsi = fm * dths(kk) + fp * dths(kk - 1);
    if (kk == 1)
        return drwc * ti + rwc[kk] * dti + (drws - drwc) * tsi + (rws[kk] - rwc[kk]) * dtsi;
    if (kk == L - 1) return drws * ti + rws[kk] * dti;
    return flux3_dot(th[kk - 2], th[kk - 1], th[kk], th[kk + 1], rwc[kk], P.coef_3rd_order, dth(kk - 2),
                     dth(kk - 1), dth(kk), dth(kk + 1), drwc) +
           (drws - drwc) * tsi + (rws[kk] - rwc[kk]) * dtsi;
}

template <bool kOpt>
__device__ __forceinline__ void dyn_j5(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Fwd F, Dot T,
                                       const DynAdjScratch &S, int c, int k) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c > m.nCells || k > L) return;
    const bool rk1 = (P.rk_step == 1);
    const long long oc = (long long)c * L, ow = (long long)c * Lw;
    if (c == m.nCells) {
        dot_put(T, tend_w, ow + k, 0.0);
        dot_put(T, tend_w_euler, ow + k, rk1 ? 0.0 : dot_of(T, tend_w_euler_in, ow + k));
        if (k < L) {
            dot_put(T, tend_theta, oc + k, 0.0);
            dot_put(T, rthdynten, oc + k, 0.0);
            dot_put(T, tend_theta_euler, oc + k, rk1 ? 0.0 : dot_of(T, tend_theta_euler_in, oc + k));
        }
        return;
    }
    const int n = m.nEdgesOnCell[c];
    const double ia = m.invAreaCell[c];
    // ------------------------------------------------ w, interface k
    if (k == 0 || k == L) {
        dot_put(T, tend_w, ow + k, 0.0);
        dot_put(T, tend_w_euler, ow + k, rk1 ? 0.0 : dot_of(T, tend_w_euler_in, ow + k));
    } else {
        double dtwe;
        if (rk1) {
            dtwe = 0.0;
            for (int j = 0; j < n; ++j) {
                const int e = m.edgesOnCell[c * ME + j];
                const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
                const double es = 0.5 * ia * m.edgesOnCell_sign[c * ME + j] * m.dvEdge[e] * m.invDcEdge[e];
                const long long ie = (long long)e * L + k;
                const long long w1 = (long long)c1 * Lw + k, w2 = (long long)c2 * Lw + k;
                const long long a1 = (long long)c1 * L + k, a2 = (long long)c2 * L + k;
                const double res = F[rho_edge][ie] + F[rho_edge][ie - 1];
                const double dres = dot_of(T, rho_edge, ie) + dot_of(T, rho_edge, ie - 1);
                const double wd = F[w][w2] - F[w][w1], dwd = dot_of(T, w, w2) - dot_of(T, w, w1);
                const double K = F[kdiff][a1] + F[kdiff][a2] + F[kdiff][a1 - 1] + F[kdiff][a2 - 1];
                const double dK = dot_of(T, kdiff, a1) + dot_of(T, kdiff, a2) + dot_of(T, kdiff, a1 - 1) +
                                  dot_of(T, kdiff, a2 - 1);
                const double bf = es * res * wd, dbf = es * (dres * wd + res * dwd);
                dtwe += (dbf * K + bf * dK) * m.meshScalingDel2[e] * 0.25;
                if (P.h_mom_eddy_visc4 > 0.0) {
                    const double es4 = m.meshScalingDel4[e] * (P.h_mom_eddy_visc4 * ia) * m.dvEdge[e] *
                                       m.edgesOnCell_sign[c * ME + j] * m.invDcEdge[e];
                    dtwe -= es4 * (S.g_delsq_w[w2] - S.g_delsq_w[w1]);
                }
            }
            const double *pp_ = F[pp] + oc;
            const double Y = m.rdzu[k] * (pp_[k] - pp_[k - 1]) - (m.fzm[k] * S.dpdz[oc + k] + m.fzp[k] * S.dpdz[oc + k - 1]);
            const double dY = m.rdzu[k] * (dot_of(T, pp, oc + k) - dot_of(T, pp, oc + k - 1)) -
                              (m.fzm[k] * S.g_dpdz[oc + k] + m.fzp[k] * S.g_dpdz[oc + k - 1]);
            dtwe -= dot_of(T, cqw, ow + k) * Y + F[cqw][ow + k] * dY;
            if (kOpt && P.v_mom_eddy_visc2 > 0.0) {
                // vertical mixing of w
                const double *wc = F[w] + ow, *rz = F[rho_zz] + oc;
                const double rho_half = 0.5 * (rz[k] + rz[k - 1]);
                const double drho_half = 0.5 * (dot_of(T, rho_zz, oc + k) + dot_of(T, rho_zz, oc + k - 1));
                const double wud = (wc[k + 1] - wc[k]) * m.rdzw[k] - (wc[k] - wc[k - 1]) * m.rdzw[k - 1];
                const double dwud = (dot_of(T, w, ow + k + 1) - dot_of(T, w, ow + k)) * m.rdzw[k] -
                                    (dot_of(T, w, ow + k) - dot_of(T, w, ow + k - 1)) * m.rdzw[k - 1];
                dtwe += P.v_mom_eddy_visc2 * (drho_half * wud + rho_half * dwud) * m.rdzu[k];
            }
        } else {
            dtwe = dot_of(T, tend_w_euler_in, ow + k);
        }
        dot_put(T, tend_w_euler, ow + k, dtwe);
        double dh = 0.0;
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[c * ME + j];
            dh -= m.edgesOnCell_sign[c * ME + j] * S.g_wf[(long long)e * L + k];
        }
        if (kOpt && P.curvature) {
            // curvature: rh (ur^2 + vr^2) / r_earth + curvCell ur rh in the advective sum
            const double fk = m.fzm[k], gk = m.fzp[k];
            const long long a0 = oc + k, a1 = oc + k - 1;
            const double ur = fk * F[ur_cell][a0] + gk * F[ur_cell][a1];
            const double vr = fk * F[vr_cell][a0] + gk * F[vr_cell][a1];
            const double rh = F[rho_zz][a0] * fk + F[rho_zz][a1] * gk;
            const double dur = fk * dot_of(T, ur_cell, a0) + gk * dot_of(T, ur_cell, a1);
            const double dvr = fk * dot_of(T, vr_cell, a0) + gk * dot_of(T, vr_cell, a1);
            const double drh = dot_of(T, rho_zz, a0) * fk + dot_of(T, rho_zz, a1) * gk;
            dh += (drh * (ur * ur + vr * vr) + rh * 2.0 * (ur * dur + vr * dvr)) / kEarthRadius +
                  m.curvCell[c] * (dur * rh + ur * drh);
        }
        const double ddelta =
            (k == L - 1) ? -dyn_dwdwz(m, F, T, c, L - 2) : dyn_dwdwz(m, F, T, c, k) - dyn_dwdwz(m, F, T, c, k - 1);
        dot_put(T, tend_w, ow + k, dh * ia - m.rdzu[k] * ddelta + dtwe);
    }
    if (k >= L) return;
    // ------------------------------------------------ theta, level k
    const long long o = oc + k;
    double dth_adv = 0.0;
    for (int j = 0; j < n; ++j) {
        const int e = m.edgesOnCell[c * ME + j];
        dth_adv -= m.edgesOnCell_sign[c * ME + j] * S.g_tf[(long long)e * L + k];
    }
    if (!rk1) {
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[c * ME + j];
            const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
            const double sdv = m.edgesOnCell_sign[c * ME + j] * m.dvEdge[e];
            const long long ie = (long long)e * L + k, a1 = (long long)c1 * L + k, a2 = (long long)c2 * L + k;
            const double face = 0.5 * (F[theta_m_save][a2] + F[theta_m_save][a1]);
            const double dface = 0.5 * (dot_of(T, theta_m_save, a2) + dot_of(T, theta_m_save, a1));
            const double rd = F[ru_save][ie] - F[ru][ie], drd = dot_of(T, ru_save, ie) - dot_of(T, ru, ie);
            dth_adv -= sdv * (drd * face + rd * dface);
        }
    }
    double dtte;
    if (rk1) {
        dtte = 0.0;
        const double r_area = P.h_theta_eddy_visc4 * P.prandtl_inv * ia;
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[c * ME + j];
            const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
            const double es = ia * m.edgesOnCell_sign[c * ME + j] * m.dvEdge[e] * m.invDcEdge[e];
            const double prs = P.prandtl_inv * m.meshScalingDel2[e];
            const long long ie = (long long)e * L + k, a1 = (long long)c1 * L + k, a2 = (long long)c2 * L + k;
            const double td = F[theta_m][a2] - F[theta_m][a1], dtd = dot_of(T, theta_m, a2) - dot_of(T, theta_m, a1);
            const double re = F[rho_edge][ie], dre = dot_of(T, rho_edge, ie);
            const double bf = es * td * re, dbf = es * (dtd * re + td * dre);
            const double K = F[kdiff][a1] + F[kdiff][a2], dK = dot_of(T, kdiff, a1) + dot_of(T, kdiff, a2);
            dtte += (dbf * K + bf * dK) * 0.5 * prs;
            if (P.h_theta_eddy_visc4 > 0.0) {
                const double es4 = m.meshScalingDel4[e] * r_area * m.dvEdge[e] * m.edgesOnCell_sign[c * ME + j] *
                                   m.invDcEdge[e];
                dtte -= es4 * (S.g_delsq_theta[a2] - S.g_delsq_theta[a1]);
            }
        }
        if (kOpt && P.v_theta_eddy_visc2 > 0.0 && k >= 1 && k <= L - 2) {
            // vertical mixing of theta_m (or of its departure from t_init, a constant shift)
            const VmixZ z = vmix_z_cell(m.zgrid + c * Lw, k);
            const double ip = 1.0 / (z.zp - z.z0), im = 1.0 / (z.z0 - z.zm);
            const double *thc = F[theta_m] + oc, *ti = m.t_init + oc;
            const double a = P.mix_full ? thc[k - 1] : thc[k - 1] - ti[k - 1];
            const double b = P.mix_full ? thc[k] : thc[k] - ti[k];
            const double cc = P.mix_full ? thc[k + 1] : thc[k + 1] - ti[k + 1];
            const double da = dot_of(T, theta_m, o - 1), db = dot_of(T, theta_m, o), dc = dot_of(T, theta_m, o + 1);
            const double sv = P.v_theta_eddy_visc2 * P.prandtl_inv / (0.5 * (z.zp - z.zm));
            dtte += sv * (dot_of(T, rho_zz, o) * ((cc - b) * ip - (b - a) * im) +
                          F[rho_zz][o] * ((dc - db) * ip - (db - da) * im));
        }
    } else {
        dtte = dot_of(T, tend_theta_euler_in, o);
    }
    dot_put(T, tend_theta_euler, o, dtte);
    double ddelta = 0.0;
    if (L > 1) {
        if (k == 0) ddelta = dyn_dwdtz(m, P, F, T, c, 1);
        else if (k == L - 1) ddelta = -dyn_dwdtz(m, P, F, T, c, L - 1);
        else ddelta = dyn_dwdtz(m, P, F, T, c, k + 1) - dyn_dwdtz(m, P, F, T, c, k);
    }
    const double dbase = dth_adv * ia - m.rdzw[k] * ddelta;
    const double rz = F[rho_zz][o], drz = dot_of(T, rho_zz, o), th = F[theta_m][o];
    const double dtrho = dot_of(T, tend_rho, o);
    dot_put(T, rthdynten, o,
            (dbase - dtrho * th - F[tend_rho][o] * dot_of(T, theta_m, o)) / rz - F[rthdynten][o] / rz * drz);
    const double dtt = dbase + drz * F[rt_diabatic_tend][o] + rz * dot_of(T, rt_diabatic_tend, o);
    dot_put(T, tend_theta, o, dtt + dtte);
}

}  // namespace adj
}  // namespace mpas
