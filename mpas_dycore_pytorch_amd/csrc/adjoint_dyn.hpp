// Adjoint of atm_compute_dyn_tend_work (dyn_tend.py:9390-10833; sections as
// in kernels_dyn.hip).  Included after adjoint_bodies.hpp.
//
// Passes (each a launch; later passes read adjoints completed by earlier ones):
//   F0  cells/edges/vertices (rk 1): forward recompute of dpdz, delsq_u,
//       delsq_div, delsq_vort
//   A1  cells x interfaces: theta and w final tendencies, their vertical
//       transport, rthdynten, tend_rho, the PGF/buoyancy part of
//       tend_w_euler and the del4 scatter of the w / theta Laplacians
//   A2  cells (rk 1): del2 w / theta mixing sums
//   A3  edges: edge advective fluxes of w and theta, the u tendency and the
//       non-Laplacian part of tend_u_euler
//   A4  cells + vertices (rk 1, del4): delsq_div / delsq_vort -> delsq_u
//   A5  edges (rk 1): delsq_u -> divergence, vorticity
//   A6  cells: tend_rho / h_divergence -> ru, rw; dpdz -> qtot, rr_save;
//       Smagorinsky kdiff -> u, v
#pragma once

namespace mpas {
namespace adj {

namespace dyn {
enum {
    u, v, w, ru, rw, ru_save, rw_save, theta_m, theta_m_save, rho_zz, rho_edge, divergence, vorticity, ke,
    pv_edge, pp, rr_save, cqw, cqu, rt_diabatic_tend, qtot,
    tend_rho_in, tend_u_euler_in, tend_w_euler_in, tend_theta_euler_in, kdiff_in,
    tend_rho, tend_u_euler, tend_w_euler, tend_theta_euler, kdiff, tend_u, tend_w, tend_theta, h_divergence,
    rthdynten, N
};
}

// Forward values and adjoint scratch of one dyn_tend adjoint (each (N+1)-row)
struct DynAdjScratch {
    double *dpdz, *delsq_u, *delsq_div, *delsq_vort;                         // forward (rk 1)
    double *g_thadv, *g_hadv, *g_dpdz, *g_delsq_w, *g_delsq_theta, *g_delsq_div, *g_delsq_vort, *g_delsq_u;
    // per-edge adjoints of the theta flux, w flux and PV-flux sum, consumed by
    // the gather passes (used when inverse tables are given)
    double *g_tf, *g_wf, *g_q;
};

// ---------------------------------------------------------------- F0
__device__ __forceinline__ void dyn_fwd_cells(const mpas_hip_mesh &m, Fwd F, const DynAdjScratch &S, int c, int k) {
    using namespace dyn;
    const int L = m.nLevels;
    if (c >= m.nCells || k >= L) return;
    const long long o = (long long)c * L + k;
    S.dpdz[o] = -MPAS_GRAVITY * (m.rho_base[o] * F[qtot][o] + F[rr_save][o] * (1.0 + F[qtot][o]));
}
__device__ __forceinline__ void dyn_fwd_edges(const mpas_hip_mesh &m, Fwd F, const DynAdjScratch &S, int e, int k) {
    using namespace dyn;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const int v1 = m.verticesOnEdge[2 * e], v2 = m.verticesOnEdge[2 * e + 1];
    const double r_dc = m.invDcEdge[e], r_dv = dmin(m.invDvEdge[e], 4.0 * m.invDcEdge[e]);
    S.delsq_u[(long long)e * L + k] =
        (F[divergence][(long long)c2 * L + k] - F[divergence][(long long)c1 * L + k]) * r_dc -
        (F[vorticity][(long long)v2 * L + k] - F[vorticity][(long long)v1 * L + k]) * r_dv;
}
__device__ __forceinline__ void dyn_fwd_del4(const mpas_hip_mesh &m, const DynAdjScratch &S, int i, int k) {
    // i < nVertices: delsq_vort; i < nCells: delsq_div
    const int L = m.nLevels, ME = m.maxEdges;
    if (k >= L) return;
    if (i < m.nVertices) {
        double s = 0.0;
        for (int j = 0; j < 3; ++j) {
            const int e = m.edgesOnVertex[3 * i + j];
            const double wgt = m.invAreaTriangle[i] * m.dcEdge[e] * m.edgesOnVertex_sign[3 * i + j];
            const double cc = wgt * S.delsq_u[(long long)e * L + k];
            s = (j == 0) ? cc : s + cc;
        }
        S.delsq_vort[(long long)i * L + k] = s;
    }
    if (i < m.nCells) {
        const int n = m.nEdgesOnCell[i];
        double s = 0.0;
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[i * ME + j];
            const double wgt = m.invAreaCell[i] * m.dvEdge[e] * m.edgesOnCell_sign[i * ME + j];
            const double cc = wgt * S.delsq_u[(long long)e * L + k];
            s = (j == 0) ? cc : s + cc;
        }
        S.delsq_div[(long long)i * L + k] = s;
    }
}

// ---------------------------------------------------------------- A1
// g_base(k) of the theta tendency: base feeds tend_theta and rthdynten
__device__ __forceinline__ double dyn_g_theta_base(Fwd F, Bar B, long long o) {
    using namespace dyn;
    return bar_of(B, tend_theta, o) + bar_of(B, rthdynten, o) / F[rho_zz][o];
}

__device__ __forceinline__ void dyn_a1(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Fwd F, Bar B,
                                       const DynAdjScratch &S, int c, int k) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c >= m.nCells || k > L) return;
    const bool rk1 = (P.rk_step == 1);
    const long long oc = (long long)c * L, ow = (long long)c * Lw;
    const double ia = m.invAreaCell[c];
    const int n = m.nEdgesOnCell[c];
    const double *fzm = m.fzm, *fzp = m.fzp;
    // ------------------------------------------------ theta, level k
    if (k < L) {
        const long long o = oc + k;
        const double rz = F[rho_zz][o], th = F[theta_m][o];
        const double gtt = bar_of(B, tend_theta, o), grth = bar_of(B, rthdynten, o);
        const double gbase = gtt + grth / rz;
        const double rth = F[rthdynten][o];
        // rthdynten = (base - tend_rho th) / rho_zz; tt = base + rho_zz rt_diab
        const double g_trho = bar_of(B, tend_rho, o) - (th / rz) * grth;
        add(B, theta_m, o, -(F[tend_rho][o] / rz) * grth);
        add_own(B, rho_zz, o, -(rth / rz) * grth + F[rt_diabatic_tend][o] * gtt);
        add_own(B, rt_diabatic_tend, o, rz * gtt);
        S.g_thadv[o] = ia * gbase;
        // tend_rho: computed here on the first stage, carried otherwise
        if (rk1) {
            add(B, rw, ow + k + 1, -m.rdzw[k] * g_trho);
            add(B, rw, ow + k, m.rdzw[k] * g_trho);
            acc(B[h_divergence] + o, -g_trho);
        } else {
            add_own(B, tend_rho_in, o, g_trho);
        }
        // tend_theta_euler
        const double gtte = bar_of(B, tend_theta_euler, o) + gtt;
        if (!rk1) {
            add_own(B, tend_theta_euler_in, o, gtte);
        } else if (P.h_theta_eddy_visc4 > 0.0) {
            // tte -= es (delsq_theta(c2) - delsq_theta(c1)), es per slot
            const double r_area = P.h_theta_eddy_visc4 * P.prandtl_inv * ia;
            for (int j = 0; j < n; ++j) {
                const int e = m.edgesOnCell[c * ME + j];
                const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
                const double es = m.meshScalingDel4[e] * r_area * m.dvEdge[e] * m.edgesOnCell_sign[c * ME + j] *
                                  m.invDcEdge[e];
                acc(S.g_delsq_theta + (long long)c2 * L + k, -es * gtte);
                acc(S.g_delsq_theta + (long long)c1 * L + k, es * gtte);
            }
        }
    }
    // theta vertical flux at interface k (1..L-1): g_wdtz(k) = g_delta(k-1) - g_delta(k)
    if (k >= 1 && k < L) {
        const double gd_km1 = -m.rdzw[k - 1] * dyn_g_theta_base(F, B, oc + k - 1);
        const double gd_k = -m.rdzw[k] * dyn_g_theta_base(F, B, oc + k);
        const double g = gd_km1 - gd_k;
        const double *th = F[theta_m] + oc, *ths = F[theta_m_save] + oc;
        const double *rwc = F[rw] + ow, *rws = F[rw_save] + ow;
        if (k == L - 1 && k != 1) {
            // rws (fzm th(k) + fzp th(k-1))
            add(B, rw_save, ow + k, (fzm[k] * th[k] + fzp[k] * th[k - 1]) * g);
            add(B, theta_m, oc + k, rws[k] * fzm[k] * g);
            add(B, theta_m, oc + k - 1, rws[k] * fzp[k] * g);
        } else {
            // (rws - rwc) (fzm ths(k) + fzp ths(k-1)) plus the advective part
            const double tsi = fzm[k] * ths[k] + fzp[k] * ths[k - 1];
            const double dr = rws[k] - rwc[k];
            add(B, rw_save, ow + k, tsi * g);
            add(B, rw, ow + k, -tsi * g);
            add(B, theta_m_save, oc + k, dr * fzm[k] * g);
            add(B, theta_m_save, oc + k - 1, dr * fzp[k] * g);
            if (k == 1) {
                add(B, rw, ow + k, (fzm[k] * th[k] + fzp[k] * th[k - 1]) * g);
                add(B, theta_m, oc + k, rwc[k] * fzm[k] * g);
                add(B, theta_m, oc + k - 1, rwc[k] * fzp[k] * g);
            } else {
                double d[5];
                flux3_partials(th[k - 2], th[k - 1], th[k], th[k + 1], rwc[k], P.coef_3rd_order, d);
                add(B, theta_m, oc + k - 2, d[0] * g);
                add(B, theta_m, oc + k - 1, d[1] * g);
                add(B, theta_m, oc + k, d[2] * g);
                add(B, theta_m, oc + k + 1, d[3] * g);
                add(B, rw, ow + k, d[4] * g);
            }
        }
    }
    // ------------------------------------------------ w, interface k (1..L-1)
    const long long i = ow + k;
    if (!rk1 && (k == 0 || k == L)) add_own(B, tend_w_euler_in, i, bar_of(B, tend_w_euler, i));
    if (k >= 1 && k < L) {
        const double gtw = bar_of(B, tend_w, i);
        S.g_hadv[i] = ia * gtw;
        const double gtwe = bar_of(B, tend_w_euler, i) + gtw;
        if (!rk1) {
            add_own(B, tend_w_euler_in, i, gtwe);
        } else {
            // e0 = wturb - cqw (rdzu (pp(k) - pp(k-1)) - (fzm dpdz(k) + fzp dpdz(k-1)))
            const double *pp_ = F[pp] + oc, *dp = S.dpdz + oc;
            const double bracket = m.rdzu[k] * (pp_[k] - pp_[k - 1]) - (fzm[k] * dp[k] + fzp[k] * dp[k - 1]);
            const double cq = F[cqw][i];
            add_own(B, cqw, i, -bracket * gtwe);
            add(B, pp, oc + k, -cq * m.rdzu[k] * gtwe);
            add(B, pp, oc + k - 1, cq * m.rdzu[k] * gtwe);
            acc(S.g_dpdz + oc + k, cq * fzm[k] * gtwe);
            acc(S.g_dpdz + oc + k - 1, cq * fzp[k] * gtwe);
            // wturb = sum_slots sf - sum_slots es4 (delsq_w(c2) - delsq_w(c1))
            if (P.h_mom_eddy_visc4 > 0.0) {
                const double r_area = P.h_mom_eddy_visc4 * ia;
                for (int j = 0; j < n; ++j) {
                    const int e = m.edgesOnCell[c * ME + j];
                    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
                    const double es = m.meshScalingDel4[e] * r_area * m.dvEdge[e] * m.edgesOnCell_sign[c * ME + j] *
                                      m.invDcEdge[e];
                    acc(S.g_delsq_w + (long long)c2 * Lw + k, -es * gtwe);
                    acc(S.g_delsq_w + (long long)c1 * Lw + k, es * gtwe);
                }
            }
        }
    }
    // w vertical flux at mass level j = k (0..L-2): g_wdwz(j) = g_delta(j) [j>=1] - g_delta(j+1)
    if (k <= L - 2) {
        const int j = k;
        auto gdel = [&](int kk) -> double { return -m.rdzu[kk] * bar_of(B, tend_w, ow + kk); };
        double g = -gdel(j + 1);
        if (j >= 1) g += gdel(j);
        const double *wc = F[w] + ow, *rwc = F[rw] + ow;
        if (j == 0 || j == L - 2) {
            const double rs = rwc[j + 1] + rwc[j], ws = wc[j + 1] + wc[j];
            add(B, rw, ow + j + 1, 0.25 * ws * g);
            add(B, rw, ow + j, 0.25 * ws * g);
            add(B, w, ow + j + 1, 0.25 * rs * g);
            add(B, w, ow + j, 0.25 * rs * g);
        } else {
            double d[5];
            flux3_partials(wc[j - 1], wc[j], wc[j + 1], wc[j + 2], 0.5 * (rwc[j + 1] + rwc[j]), 1.0, d);
            add(B, w, ow + j - 1, d[0] * g);
            add(B, w, ow + j, d[1] * g);
            add(B, w, ow + j + 1, d[2] * g);
            add(B, w, ow + j + 2, d[3] * g);
            add(B, rw, ow + j + 1, 0.5 * d[4] * g);
            add(B, rw, ow + j, 0.5 * d[4] * g);
        }
    }
}

// ---------------------------------------------------------------- A2 (rk 1)
__device__ __forceinline__ void dyn_a2(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Fwd F, Bar B,
                                       const DynAdjScratch &S, int c, int k) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c >= m.nCells || k > L) return;
    const long long oc = (long long)c * L, ow = (long long)c * Lw;
    const double ia = m.invAreaCell[c];
    const int n = m.nEdgesOnCell[c];
    const double *kd = F[kdiff];
    double *gkd = B[kdiff];
    if (k < L) {
        // theta: bf = es (theta(c2) - theta(c1)) rho_edge; sf = bf 0.5 (k1 + k2) prs
        const long long o = oc + k;
        const double gtte = bar_of(B, tend_theta_euler, o) + bar_of(B, tend_theta, o);
        const double gdt = S.g_delsq_theta[o];
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[c * ME + j];
            const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
            const double es = ia * m.edgesOnCell_sign[c * ME + j] * m.dvEdge[e] * m.invDcEdge[e];
            const double prs = P.prandtl_inv * m.meshScalingDel2[e];
            const double t2 = F[theta_m][(long long)c2 * L + k], t1 = F[theta_m][(long long)c1 * L + k];
            const double re = F[rho_edge][(long long)e * L + k];
            const double k1 = kd[(long long)c1 * L + k], k2 = kd[(long long)c2 * L + k];
            const double bf = es * (t2 - t1) * re;
            const double gbf = 0.5 * (k1 + k2) * prs * gtte + gdt;
            const double gk = bf * 0.5 * prs * gtte;
            acc(gkd + (long long)c1 * L + k, gk);
            acc(gkd + (long long)c2 * L + k, gk);
            add(B, theta_m, (long long)c2 * L + k, es * re * gbf);
            add(B, theta_m, (long long)c1 * L + k, -es * re * gbf);
            add(B, rho_edge, (long long)e * L + k, es * (t2 - t1) * gbf);
        }
    }
    if (k >= 1 && k < L) {
        // w: bf = es (re(k) + re(k-1)) (w(c2) - w(c1)); sf = bf ms2 0.25 (k1 + k2 + k1m + k2m)
        const long long i = ow + k;
        const double gtwe = bar_of(B, tend_w_euler, i) + bar_of(B, tend_w, i);
        const double gdw = S.g_delsq_w[i];
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[c * ME + j];
            const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
            const double es = 0.5 * ia * m.edgesOnCell_sign[c * ME + j] * m.dvEdge[e] * m.invDcEdge[e];
            const double r0 = F[rho_edge][(long long)e * L + k], r1 = F[rho_edge][(long long)e * L + k - 1];
            const double w2 = F[w][(long long)c2 * Lw + k], w1 = F[w][(long long)c1 * Lw + k];
            const double ksum = kd[(long long)c1 * L + k] + kd[(long long)c2 * L + k] + kd[(long long)c1 * L + k - 1] +
                                kd[(long long)c2 * L + k - 1];
            const double ms2 = m.meshScalingDel2[e];
            const double bf = es * (r0 + r1) * (w2 - w1);
            const double gbf = ms2 * 0.25 * ksum * gtwe + gdw;
            const double gk = bf * ms2 * 0.25 * gtwe;
            acc(gkd + (long long)c1 * L + k, gk);
            acc(gkd + (long long)c2 * L + k, gk);
            acc(gkd + (long long)c1 * L + k - 1, gk);
            acc(gkd + (long long)c2 * L + k - 1, gk);
            const double gr = es * (w2 - w1) * gbf;
            add(B, rho_edge, (long long)e * L + k, gr);
            add(B, rho_edge, (long long)e * L + k - 1, gr);
            add(B, w, (long long)c2 * Lw + k, es * (r0 + r1) * gbf);
            add(B, w, (long long)c1 * Lw + k, -es * (r0 + r1) * gbf);
        }
    }
}

// Gather form of A2 (no atomics).  The per-edge sums over the edge's two
// cells s (sign sg_s of the edge in s, area factor ia_s):
//   theta at level k:   A  = sum_s es_s gbf_s,   K  = sum_s bf_s 0.5 prs gtte_s
//   w at interface k:   Aw = sum_s esw_s gbfw_s, Kw = sum_s bfw_s ms2 0.25 gtwe_s
// are recomputed by the edge pass (rho_edge) and by the cell pass (kdiff,
// theta_m, w) from the forward values and the output adjoints.
struct A2Edge {
    double A, K;    // theta at level k
    double Aw, Kw;  // w at interface k (0 outside 1..L-1)
};
__device__ __forceinline__ A2Edge dyn_a2_edge(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Fwd F, Bar B,
                                              const DynAdjScratch &S, int e, double sg1, double sg2, int k,
                                              bool theta) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const double *kd = F[kdiff];
    const double ia1 = m.invAreaCell[c1], ia2 = m.invAreaCell[c2];
    A2Edge r{0.0, 0.0, 0.0, 0.0};
    if (theta) {
        const long long o1 = (long long)c1 * L + k, o2 = (long long)c2 * L + k;
        const double prs = P.prandtl_inv * m.meshScalingDel2[e];
        const double dt21 = F[theta_m][o2] - F[theta_m][o1];
        const double re = F[rho_edge][(long long)e * L + k];
        const double kk = 0.5 * (kd[o1] + kd[o2]);
        const double es1 = ia1 * sg1 * m.dvEdge[e] * m.invDcEdge[e];
        const double es2 = ia2 * sg2 * m.dvEdge[e] * m.invDcEdge[e];
        const double g1 = bar_of(B, tend_theta_euler, o1) + bar_of(B, tend_theta, o1);
        const double g2 = bar_of(B, tend_theta_euler, o2) + bar_of(B, tend_theta, o2);
        r.A = es1 * (kk * prs * g1 + S.g_delsq_theta[o1]) + es2 * (kk * prs * g2 + S.g_delsq_theta[o2]);
        r.K = (es1 * dt21 * re) * 0.5 * prs * g1 + (es2 * dt21 * re) * 0.5 * prs * g2;
    }
    if (k >= 1 && k < L) {
        const long long w1i = (long long)c1 * Lw + k, w2i = (long long)c2 * Lw + k;
        const double rr = F[rho_edge][(long long)e * L + k] + F[rho_edge][(long long)e * L + k - 1];
        const double dw = F[w][w2i] - F[w][w1i];
        const double ksum = kd[(long long)c1 * L + k] + kd[(long long)c2 * L + k] + kd[(long long)c1 * L + k - 1] +
                            kd[(long long)c2 * L + k - 1];
        const double ms2 = m.meshScalingDel2[e];
        const double es1 = 0.5 * ia1 * sg1 * m.dvEdge[e] * m.invDcEdge[e];
        const double es2 = 0.5 * ia2 * sg2 * m.dvEdge[e] * m.invDcEdge[e];
        const double g1 = bar_of(B, tend_w_euler, w1i) + bar_of(B, tend_w, w1i);
        const double g2 = bar_of(B, tend_w_euler, w2i) + bar_of(B, tend_w, w2i);
        r.Aw = es1 * (ms2 * 0.25 * ksum * g1 + S.g_delsq_w[w1i]) + es2 * (ms2 * 0.25 * ksum * g2 + S.g_delsq_w[w2i]);
        r.Kw = (es1 * rr * dw) * ms2 * 0.25 * g1 + (es2 * rr * dw) * ms2 * 0.25 * g2;
    }
    return r;
}
// A2 gather, thread (edge e, level k): rho_edge
__device__ __forceinline__ void dyn_a2_edges(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Fwd F, Bar B,
                                             const DynAdjScratch &S, int e, int k) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1;
    if (e >= m.nEdges || k >= L) return;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const double sg1 = edge_sign_in_cell(m, c1, e), sg2 = edge_sign_in_cell(m, c2, e);
    const A2Edge a = dyn_a2_edge(m, P, F, B, S, e, sg1, sg2, k, true);
    double g = (F[theta_m][(long long)c2 * L + k] - F[theta_m][(long long)c1 * L + k]) * a.A;
    if (k >= 1) g += (F[w][(long long)c2 * Lw + k] - F[w][(long long)c1 * Lw + k]) * a.Aw;
    if (k + 1 < L) {
        const A2Edge b = dyn_a2_edge(m, P, F, B, S, e, sg1, sg2, k + 1, false);
        g += (F[w][(long long)c2 * Lw + k + 1] - F[w][(long long)c1 * Lw + k + 1]) * b.Aw;
    }
    add_own(B, rho_edge, (long long)e * L + k, g);
}
// A2 gather, thread (cell c, level k): kdiff, theta_m and w (interface k)
__device__ __forceinline__ void dyn_a2_cells(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Fwd F, Bar B,
                                             const DynAdjScratch &S, int c, int k) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const int n = m.nEdgesOnCell[c];
    double gk = 0.0, gt = 0.0, gw = 0.0;
    for (int j = 0; j < n; ++j) {
        const int e = m.edgesOnCell[c * ME + j];
        const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
        const double sc = m.edgesOnCell_sign[c * ME + j];
        const double sg1 = (c1 == c) ? sc : edge_sign_in_cell(m, c1, e);
        const double sg2 = (c2 == c) ? sc : edge_sign_in_cell(m, c2, e);
        const double side = (c2 == c) ? 1.0 : -1.0;
        const A2Edge a = dyn_a2_edge(m, P, F, B, S, e, sg1, sg2, k, true);
        gk += a.K;
        gt += side * F[rho_edge][(long long)e * L + k] * a.A;
        if (k >= 1) {
            gk += a.Kw;
            gw += side * (F[rho_edge][(long long)e * L + k] + F[rho_edge][(long long)e * L + k - 1]) * a.Aw;
        }
        if (k + 1 < L) gk += dyn_a2_edge(m, P, F, B, S, e, sg1, sg2, k + 1, false).Kw;
    }
    const long long o = (long long)c * L + k;
    add_own(B, kdiff, o, gk);
    add_own(B, theta_m, o, gt);
    if (k >= 1) add_own(B, w, (long long)c * Lw + k, gw);
}

// ---------------------------------------------------------------- A3
__device__ __forceinline__ void dyn_a3(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Fwd F, Bar B,
                                       const DynAdjScratch &S, bool gather, int e, int k) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1, MA = m.maxAdv, ME2 = m.maxEdges2;
    if (e >= m.nEdges || k >= L) return;
    const bool rk1 = (P.rk_step == 1);
    const long long i = (long long)e * L + k;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const double s1 = edge_sign_in_cell(m, c1, e), s2 = edge_sign_in_cell(m, c2, e);
    const double *ru_ = F[ru] + (long long)e * L;
    const int na = m.nAdvCellsForEdge[e];
    // ---- theta edge flux: thadv(c) += -sgn(c, e) ru(e) sum_a (c0 + s c3) theta(a)
    {
        const double gf = -(s1 * S.g_thadv[(long long)c1 * L + k] + s2 * S.g_thadv[(long long)c2 * L + k]);
        const double rut = ru_[k], srt = sign_pm(rut);
        double ft = 0.0;
        for (int a = 0; a < na; ++a) {
            const int ac = m.advCellsForEdge[e * MA + a];
            const double wgt = m.adv_coefs[e * MA + a] + srt * m.adv_coefs_3rd[e * MA + a];
            const double cc = wgt * F[theta_m][(long long)ac * L + k];
            ft = (a == 0) ? cc : ft + cc;
            if (!gather) add(B, theta_m, (long long)ac * L + k, wgt * rut * gf);
        }
        if (gather) S.g_tf[i] = gf;
        add(B, ru, i, ft * gf);
        if (!rk1) {
            // thadv(c) -= sgn(c, e) dv (ru_save - ru) face
            const double face = 0.5 * (F[theta_m_save][(long long)c2 * L + k] + F[theta_m_save][(long long)c1 * L + k]);
            const double dv = m.dvEdge[e];
            const double gp = gf;  // same combination of the two cells' adjoints
            add(B, ru_save, i, dv * face * gp);
            add(B, ru, i, -dv * face * gp);
            const double gface = dv * (F[ru_save][i] - ru_[k]) * gp;
            add(B, theta_m_save, (long long)c1 * L + k, 0.5 * gface);
            add(B, theta_m_save, (long long)c2 * L + k, 0.5 * gface);
        }
    }
    // ---- w edge flux at interface k (1..L-1)
    if (k >= 1) {
        const double gf = -(s1 * S.g_hadv[(long long)c1 * Lw + k] + s2 * S.g_hadv[(long long)c2 * Lw + k]);
        const double ruw = m.fzm[k] * ru_[k] + m.fzp[k] * ru_[k - 1];
        const double srw = sign_pm(ruw);
        double fw = 0.0;
        for (int a = 0; a < na; ++a) {
            const int ac = m.advCellsForEdge[e * MA + a];
            const double wgt = m.adv_coefs[e * MA + a] + srw * m.adv_coefs_3rd[e * MA + a];
            const double cc = wgt * F[w][(long long)ac * Lw + k];
            fw = (a == 0) ? cc : fw + cc;
            if (!gather) add(B, w, (long long)ac * Lw + k, wgt * ruw * gf);
        }
        if (gather) S.g_wf[i] = gf;
        add(B, ru, i, m.fzm[k] * fw * gf);
        add(B, ru, i - 1, m.fzp[k] * fw * gf);
    }
    // ---- u tendency
    const double gtu = bar_of(B, tend_u, i);
    const double *ue = F[u] + (long long)e * L;
    const double re = F[rho_edge][i];
    {
        // base = re (q - ke_delta invDc) - u 0.5 hsum
        const int ne = m.nEdgesOnEdge[e];
        const double pvs = F[pv_edge][i];
        double q = 0.0;
        for (int j = 0; j < ne; ++j) {
            const int ee = m.edgesOnEdge[e * ME2 + j];
            const double wv = m.weightsOnEdge[e * ME2 + j];
            const double uj = F[u][(long long)ee * L + k], pj = F[pv_edge][(long long)ee * L + k];
            const double workpv = 0.5 * (pvs + pj);
            const double cc = wv * uj * workpv;
            q = (j == 0) ? cc : q + cc;
        }
        const double ked = F[ke][(long long)c2 * L + k] - F[ke][(long long)c1 * L + k];
        const double hsum = F[h_divergence][(long long)c1 * L + k] + F[h_divergence][(long long)c2 * L + k];
        add(B, rho_edge, i, (q - ked * m.invDcEdge[e]) * gtu);
        const double gq = re * gtu;
        add(B, ke, (long long)c2 * L + k, -re * m.invDcEdge[e] * gtu);
        add(B, ke, (long long)c1 * L + k, re * m.invDcEdge[e] * gtu);
        add(B, u, i, -0.5 * hsum * gtu);
        acc(B[h_divergence] + (long long)c1 * L + k, -ue[k] * 0.5 * gtu);
        acc(B[h_divergence] + (long long)c2 * L + k, -ue[k] * 0.5 * gtu);
        double gpe = 0.0;
        for (int j = 0; j < ne; ++j) {
            const int ee = m.edgesOnEdge[e * ME2 + j];
            const double wv = m.weightsOnEdge[e * ME2 + j];
            const double uj = F[u][(long long)ee * L + k], pj = F[pv_edge][(long long)ee * L + k];
            const double gp = wv * uj * 0.5 * gq;
            gpe += gp;
            if (!gather) {
                add(B, u, (long long)ee * L + k, wv * 0.5 * (pvs + pj) * gq);
                add(B, pv_edge, (long long)ee * L + k, gp);
            }
        }
        add(B, pv_edge, i, gpe);
        if (gather) S.g_q[i] = gq;
    }
    {
        // vertical flux of u at interface k (1..L-1): g_wduz(k) = g_delta(k-1) - g_delta(k)
        if (k >= 1) {
            const double g = -m.rdzw[k - 1] * bar_of(B, tend_u, i - 1) + m.rdzw[k] * gtu;
            const double erw = 0.5 * (F[rw][(long long)c1 * Lw + k] + F[rw][(long long)c2 * Lw + k]);
            double gerw;
            if (k == 1 || k == L - 1) {
                const double ui = m.fzm[k] * ue[k] + m.fzp[k] * ue[k - 1];
                gerw = ui * g;
                add(B, u, i, erw * m.fzm[k] * g);
                add(B, u, i - 1, erw * m.fzp[k] * g);
            } else {
                double d[5];
                flux3_partials(ue[k - 2], ue[k - 1], ue[k], ue[k + 1], erw, 1.0, d);
                add(B, u, i - 2, d[0] * g);
                add(B, u, i - 1, d[1] * g);
                add(B, u, i, d[2] * g);
                add(B, u, i + 1, d[3] * g);
                gerw = d[4] * g;
            }
            add(B, rw, (long long)c1 * Lw + k, 0.5 * gerw);
            add(B, rw, (long long)c2 * Lw + k, 0.5 * gerw);
        }
    }
    // ---- tend_u_euler
    const double gtue = bar_of(B, tend_u_euler, i) + gtu;
    if (!rk1) {
        add_own(B, tend_u_euler_in, i, gtue);
        return;
    }
    {
        // PGF: -cqu (ppd invDc / (0.5 zzs) - 0.5 zxu dps)
        const double zzs = m.zz[(long long)c2 * L + k] + m.zz[(long long)c1 * L + k];
        const double ppd = F[pp][(long long)c2 * L + k] - F[pp][(long long)c1 * L + k];
        const double dps = S.dpdz[(long long)c1 * L + k] + S.dpdz[(long long)c2 * L + k];
        const double cq = F[cqu][i];
        add_own(B, cqu, i, -(ppd * m.invDcEdge[e] / (0.5 * zzs) - 0.5 * m.zxu[i] * dps) * gtue);
        const double gppd = -cq * m.invDcEdge[e] / (0.5 * zzs) * gtue;
        add(B, pp, (long long)c2 * L + k, gppd);
        add(B, pp, (long long)c1 * L + k, -gppd);
        const double gdps = cq * 0.5 * m.zxu[i] * gtue;
        acc(S.g_dpdz + (long long)c1 * L + k, gdps);
        acc(S.g_dpdz + (long long)c2 * L + k, gdps);
    }
    {
        // del2: + re kdiffu ud ms2 (ud = delsq_u, adjoint finished in A5)
        const double ud = S.delsq_u[i];
        const double kd1 = F[kdiff][(long long)c1 * L + k], kd2 = F[kdiff][(long long)c2 * L + k];
        const double kdu = 0.5 * (kd1 + kd2);
        const double ms2 = m.meshScalingDel2[e];
        add(B, rho_edge, i, kdu * ud * ms2 * gtue);
        const double gk = 0.5 * re * ud * ms2 * gtue;
        acc(B[kdiff] + (long long)c1 * L + k, gk);
        acc(B[kdiff] + (long long)c2 * L + k, gk);
        acc(S.g_delsq_u + i, re * kdu * ms2 * gtue);
    }
    if (P.h_mom_eddy_visc4 > 0.0) {
        // del4: - re ((dd(c2) - dd(c1)) r_dc - (dv(v2) - dv(v1)) r_dv)
        const int v1 = m.verticesOnEdge[2 * e], v2 = m.verticesOnEdge[2 * e + 1];
        const double ums = m.meshScalingDel4[e] * P.h_mom_eddy_visc4;
        const double r_dc = ums * P.del4u_div_factor * m.invDcEdge[e];
        const double r_dv = ums * dmin(m.invDvEdge[e], 4.0 * m.invDcEdge[e]);
        const double dd = S.delsq_div[(long long)c2 * L + k] - S.delsq_div[(long long)c1 * L + k];
        const double dvv = S.delsq_vort[(long long)v2 * L + k] - S.delsq_vort[(long long)v1 * L + k];
        add(B, rho_edge, i, -(dd * r_dc - dvv * r_dv) * gtue);
        acc(S.g_delsq_div + (long long)c2 * L + k, -re * r_dc * gtue);
        acc(S.g_delsq_div + (long long)c1 * L + k, re * r_dc * gtue);
        acc(S.g_delsq_vort + (long long)v2 * L + k, re * r_dv * gtue);
        acc(S.g_delsq_vort + (long long)v1 * L + k, -re * r_dv * gtue);
    }
}

// ---------------------------------------------------------------- A3 gathers
// thread (cell c, level / interface k): adjoints of theta_m (k < L) and w
// (1 <= k < L) at the cells of the advection stencils, gathered over the
// (edge, slot) pairs naming cell c, in (edge, slot) order
__device__ __forceinline__ void dyn_a3_cells(const mpas_hip_mesh &m, Fwd F, Bar B, const DynAdjScratch &S,
                                             const mpas_hip_inverse_tables &inv, int c, int k) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1, MA = m.maxAdv;
    if (c >= m.nCells || k >= L) return;
    const int j0 = inv.advStart[c], j1 = inv.advStart[c + 1];
    double gt = 0.0, gw = 0.0;
    for (int j = j0; j < j1; ++j) {
        const int e = inv.advEdge[j], a = inv.advSlot[j];
        const long long i = (long long)e * L + k;
        const double c0 = m.adv_coefs[e * MA + a], c3 = m.adv_coefs_3rd[e * MA + a];
        const double rut = F[ru][i];
        gt += (c0 + sign_pm(rut) * c3) * rut * S.g_tf[i];
        if (k >= 1) {
            const double ruw = m.fzm[k] * rut + m.fzp[k] * F[ru][i - 1];
            gw += (c0 + sign_pm(ruw) * c3) * ruw * S.g_wf[i];
        }
    }
    add_own(B, theta_m, (long long)c * L + k, gt);
    if (k >= 1) add_own(B, w, (long long)c * Lw + k, gw);
}

// thread (edge e, level k): adjoints of u and pv_edge at the edges of the
// PV-flux stencils, gathered over the (edge, slot) pairs naming edge e
__device__ __forceinline__ void dyn_a3_edges(const mpas_hip_mesh &m, Fwd F, Bar B, const DynAdjScratch &S,
                                             const mpas_hip_inverse_tables &inv, int e, int k) {
    using namespace dyn;
    const int L = m.nLevels, ME2 = m.maxEdges2;
    if (e >= m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    const int j0 = inv.eoeStart[e], j1 = inv.eoeStart[e + 1];
    const double ue = F[u][i], pve = F[pv_edge][i];
    double gu = 0.0, gp = 0.0;
    for (int j = j0; j < j1; ++j) {
        const int eo = inv.eoeEdge[j], sl = inv.eoeSlot[j];
        const long long io = (long long)eo * L + k;
        const double wq = m.weightsOnEdge[eo * ME2 + sl] * S.g_q[io];
        gu += wq * 0.5 * (F[pv_edge][io] + pve);
        gp += wq * ue * 0.5;
    }
    add_own(B, u, i, gu);
    add_own(B, pv_edge, i, gp);
}

// ---------------------------------------------------------------- A4 (rk 1, del4)
__device__ __forceinline__ void dyn_a4(const mpas_hip_mesh &m, const DynAdjScratch &S, int i, int k) {
    const int L = m.nLevels, ME = m.maxEdges;
    if (k >= L) return;
    if (i < m.nVertices) {
        const double g = S.g_delsq_vort[(long long)i * L + k];
        for (int j = 0; j < 3; ++j) {
            const int e = m.edgesOnVertex[3 * i + j];
            const double wgt = m.invAreaTriangle[i] * m.dcEdge[e] * m.edgesOnVertex_sign[3 * i + j];
            acc(S.g_delsq_u + (long long)e * L + k, wgt * g);
        }
    }
    if (i < m.nCells) {
        const double g = S.g_delsq_div[(long long)i * L + k];
        const int n = m.nEdgesOnCell[i];
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[i * ME + j];
            const double wgt = m.invAreaCell[i] * m.dvEdge[e] * m.edgesOnCell_sign[i * ME + j];
            acc(S.g_delsq_u + (long long)e * L + k, wgt * g);
        }
    }
}

// ---------------------------------------------------------------- A5 (rk 1)
__device__ __forceinline__ void dyn_a5(const mpas_hip_mesh &m, Bar B, const DynAdjScratch &S, int e, int k) {
    using namespace dyn;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    const double g = S.g_delsq_u[(long long)e * L + k];
    if (g == 0.0) return;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const int v1 = m.verticesOnEdge[2 * e], v2 = m.verticesOnEdge[2 * e + 1];
    const double r_dc = m.invDcEdge[e], r_dv = dmin(m.invDvEdge[e], 4.0 * m.invDcEdge[e]);
    add(B, divergence, (long long)c2 * L + k, r_dc * g);
    add(B, divergence, (long long)c1 * L + k, -r_dc * g);
    add(B, vorticity, (long long)v2 * L + k, -r_dv * g);
    add(B, vorticity, (long long)v1 * L + k, r_dv * g);
}

// ---------------------------------------------------------------- A6
__device__ __forceinline__ void dyn_a6(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Fwd F, Bar B,
                                       const DynAdjScratch &S, int c, int k) {
    using namespace dyn;
    const int L = m.nLevels, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const bool rk1 = (P.rk_step == 1);
    const long long o = (long long)c * L + k;
    const int n = m.nEdgesOnCell[c];
    const double ia = m.invAreaCell[c];
    // h_divergence = ia sum sgn dv ru
    const double ghd = B[h_divergence][o];
    for (int j = 0; j < n; ++j) {
        const int e = m.edgesOnCell[c * ME + j];
        add(B, ru, (long long)e * L + k, ia * m.edgesOnCell_sign[c * ME + j] * m.dvEdge[e] * ghd);
    }
    if (!rk1) {
        add_own(B, kdiff_in, o, bar_of(B, kdiff, o));
        return;
    }
    // dpdz = -g (rb qtot + rr_save (1 + qtot))
    const double gdp = S.g_dpdz[o];
    add_own(B, qtot, o, -MPAS_GRAVITY * (m.rho_base[o] + F[rr_save][o]) * gdp);
    add_own(B, rr_save, o, -MPAS_GRAVITY * (1.0 + F[qtot][o]) * gdp);
    // kdiff = min(coef def, cap), def = sqrt(dd^2 + dof^2)
    const double gkd = B[kdiff][o];
    if (gkd == 0.0) return;
    double dd = 0.0, dof = 0.0;
    for (int j = 0; j < n; ++j) {
        const int e = m.edgesOnCell[c * ME + j];
        const double a = m.defc_a[c * ME + j], b = m.defc_b[c * ME + j];
        const double us = F[u][(long long)e * L + k], vs = F[v][(long long)e * L + k];
        dd = (dd + a * us) - b * vs;
        dof = (dof + b * us) + a * vs;
    }
    const double def = sqrt(dd * dd + dof * dof);
    double dsel, dcap;
    min_partials(P.kdiff_coef * def, P.kdiff_cap, dsel, dcap);
    if (dsel == 0.0 || def == 0.0) return;
    const double gdef = P.kdiff_coef * dsel * gkd;
    const double gdd = dd / def * gdef, gdof = dof / def * gdef;
    for (int j = 0; j < n; ++j) {
        const int e = m.edgesOnCell[c * ME + j];
        const double a = m.defc_a[c * ME + j], b = m.defc_b[c * ME + j];
        add(B, u, (long long)e * L + k, a * gdd + b * gdof);
        add(B, v, (long long)e * L + k, -b * gdd + a * gdof);
    }
}

}  // namespace adj
}  // namespace mpas
