// Adjoint of atm_compute_dyn_tend_work (dyn_tend.py:9390-10833; sections as
// in kernels_dyn.hip).  Included after adjoint_bodies.hpp.
//
// Passes (each a launch; later passes read adjoints completed by earlier ones;
// no pass writes an element another thread of the same pass writes):
//   F0  cells/edges/vertices (rk 1): forward recompute of dpdz, delsq_u,
//       delsq_div, delsq_vort
//   A1  one wavefront per cell, lane = level / interface: theta and w final
//       tendencies, their vertical transport, rthdynten, tend_rho and the
//       PGF/buoyancy part of tend_w_euler; the levels exchange what they owe
//       each other with shuffles
//   A1d cells (rk 1): the del4 w / theta Laplacian adjoints, gathered over the
//       cell's edges from both cells' tendency adjoints
//   A2  edges, then cells (rk 1): del2 w / theta mixing
//   A3  edges: edge advective fluxes of w and theta, the u tendency and the
//       non-Laplacian part of tend_u_euler; what the edge owes its cells and
//       vertices is stored per edge (x[])
//   A3c cells (+ vertices, rk 1 del4): the advection-stencil transposes over
//       the inverse table and the sums of x[] over the cell's edges
//   A3e edges: the PV-flux transposes over the inverse table; (rk 1, del4)
//       delsq_div / delsq_vort -> delsq_u
//   A6  cells (+ vertices, rk 1): delsq_u -> divergence, vorticity;
//       dpdz -> qtot, rr_save; the Smagorinsky kdiff deformation adjoints
//   A7  edges: tend_rho / h_divergence -> ru; the deformation -> u, v
#pragma once

namespace mpas {
namespace adj {

namespace dyn {
enum {
    u, v, w, ru, rw, ru_save, rw_save, theta_m, theta_m_save, rho_zz, rho_edge, divergence, vorticity, ke,
    pv_edge, pp, rr_save, cqw, cqu, rt_diabatic_tend, qtot, ur_cell, vr_cell,
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
    // the inverse-table gathers
    double *g_tf, *g_wf, *g_q;
    // what edge e owes its cells and vertices (A3 -> A3c): x[0] ke (c1 +, c2 -),
    // x[1] h_divergence, x[2] rw (interface k), x[3] pp (c1 +, c2 -; rk 1) or
    // theta_m_save (rk > 1), x[4] dpdz, x[5] kdiff, x[6] delsq_div (c1 +, c2 -),
    // x[7] delsq_vort (v1 -, v2 +), x[8] w at interfaces k and k + 1 of both
    // cells (the curvature term's w_avg / 4)
    double *x[9];
    double *g_dd, *g_dof;  // deformation adjoints of each cell (A6 -> A7)
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

// One wavefront per cell c; every lane runs (lanes past interface L hold
// nothing).  Lane k forms the shares its terms owe levels / interfaces k-2 ..
// k+2 of the column; after the exchange each lane adds its own sums once.
// kOpt builds in the curvature and vertical-mixing terms.
template <bool kOpt>
__device__ __forceinline__ void dyn_a1(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Fwd F, Bar B,
                                       const DynAdjScratch &S, int c, int k) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1;
    const bool rk1 = (P.rk_step == 1);
    const long long oc = (long long)c * L, ow = (long long)c * Lw;
    const double ia = m.invAreaCell[c];
    const double *fzm = m.fzm, *fzp = m.fzp;
    // shares by target and offset (m2 = level k-2 ... p2 = k+2): theta_m, theta_m_save,
    // pp and dpdz at levels; rw, rw_save and w at interfaces
    double th_m2 = 0.0, th_m1 = 0.0, th_0 = 0.0, th_p1 = 0.0, ts_m1 = 0.0, ts_0 = 0.0;
    double pp_m1 = 0.0, pp_0 = 0.0, dp_m1 = 0.0, dp_0 = 0.0;
    double r_0 = 0.0, r_p1 = 0.0, rs_0 = 0.0, w_m1 = 0.0, w_0 = 0.0, w_p1 = 0.0, w_p2 = 0.0;
    // rho_zz at levels k and k-1 (curvature, vertical mixing); ur_cell, vr_cell (curvature)
    double cz_0 = 0.0, cz_m1 = 0.0, cu_0 = 0.0, cu_m1 = 0.0, cv_0 = 0.0, cv_m1 = 0.0;
    const bool vmix_w = kOpt && rk1 && P.v_mom_eddy_visc2 > 0.0;
    const bool vmix_t = kOpt && rk1 && P.v_theta_eddy_visc2 > 0.0;
    // ------------------------------------------------ theta, level k
    if (k < L) {
        const long long o = oc + k;
        const double rz = F[rho_zz][o], th = F[theta_m][o];
        const double gtt = bar_of(B, tend_theta, o), grth = bar_of(B, rthdynten, o);
        const double gbase = gtt + grth / rz;
        const double rth = F[rthdynten][o];
        // rthdynten = (base - tend_rho th) / rho_zz; tt = base + rho_zz rt_diab
        const double g_trho = bar_of(B, tend_rho, o) - (th / rz) * grth;
        th_0 += -(F[tend_rho][o] / rz) * grth;
        add_own(B, rho_zz, o, -(rth / rz) * grth + F[rt_diabatic_tend][o] * gtt);
        add_own(B, rt_diabatic_tend, o, rz * gtt);
        S.g_thadv[o] = ia * gbase;
        if (vmix_t && k >= 1 && k <= L - 2) {
            // vertical mixing of theta: vt pr rho_zz ((c - b)/(zp - z0) - (b - a)/(z0 - zm)) / h
            const VmixZ z = vmix_z_cell(m.zgrid + c * Lw, k);
            const double ip = 1.0 / (z.zp - z.z0), im = 1.0 / (z.z0 - z.zm);
            const double *thc = F[theta_m] + oc, *ti = m.t_init + oc;
            const double a = P.mix_full ? thc[k - 1] : thc[k - 1] - ti[k - 1];
            const double b = P.mix_full ? thc[k] : thc[k] - ti[k];
            const double cc = P.mix_full ? thc[k + 1] : thc[k + 1] - ti[k + 1];
            const double sv = P.v_theta_eddy_visc2 * P.prandtl_inv / (0.5 * (z.zp - z.zm));
            const double gt = bar_of(B, tend_theta_euler, o) + gtt;
            cz_0 += sv * ((cc - b) * ip - (b - a) * im) * gt;
            const double gr = rz * sv * gt;
            th_p1 += gr * ip;
            th_0 += -gr * (ip + im);
            th_m1 += gr * im;
        }
        // tend_rho: computed here on the first stage, carried otherwise
        if (rk1) {
            r_p1 += -m.rdzw[k] * g_trho;
            r_0 += m.rdzw[k] * g_trho;
            add_own(B, h_divergence, o, -g_trho);
        } else {
            add_own(B, tend_rho_in, o, g_trho);
            add_own(B, tend_theta_euler_in, o, bar_of(B, tend_theta_euler, o) + gtt);
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
            rs_0 += (fzm[k] * th[k] + fzp[k] * th[k - 1]) * g;
            th_0 += rws[k] * fzm[k] * g;
            th_m1 += rws[k] * fzp[k] * g;
        } else {
            // (rws - rwc) (fzm ths(k) + fzp ths(k-1)) plus the advective part
            const double tsi = fzm[k] * ths[k] + fzp[k] * ths[k - 1];
            const double dr = rws[k] - rwc[k];
            rs_0 += tsi * g;
            r_0 += -tsi * g;
            ts_0 += dr * fzm[k] * g;
            ts_m1 += dr * fzp[k] * g;
            if (k == 1) {
                r_0 += (fzm[k] * th[k] + fzp[k] * th[k - 1]) * g;
                th_0 += rwc[k] * fzm[k] * g;
                th_m1 += rwc[k] * fzp[k] * g;
            } else {
                double d[5];
                flux3_partials(th[k - 2], th[k - 1], th[k], th[k + 1], rwc[k], P.coef_3rd_order, d);
                th_m2 += d[0] * g;
                th_m1 += d[1] * g;
                th_0 += d[2] * g;
                th_p1 += d[3] * g;
                r_0 += d[4] * g;
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
            pp_0 += -cq * m.rdzu[k] * gtwe;
            pp_m1 += cq * m.rdzu[k] * gtwe;
            dp_0 += cq * fzm[k] * gtwe;
            dp_m1 += cq * fzp[k] * gtwe;
            if (vmix_w) {
                // vertical mixing of w: visc 0.5 (rho_zz(k) + rho_zz(k-1)) (w_up - w_down) rdzu
                const double *wc = F[w] + ow, *rz = F[rho_zz] + oc;
                const double rho_half = 0.5 * (rz[k] + rz[k - 1]);
                const double w_up = (wc[k + 1] - wc[k]) * m.rdzw[k], w_down = (wc[k] - wc[k - 1]) * m.rdzw[k - 1];
                const double g = P.v_mom_eddy_visc2 * m.rdzu[k] * gtwe;
                cz_0 += 0.5 * (w_up - w_down) * g;
                cz_m1 += 0.5 * (w_up - w_down) * g;
                w_p1 += rho_half * m.rdzw[k] * g;
                w_0 += -rho_half * (m.rdzw[k] + m.rdzw[k - 1]) * g;
                w_m1 += rho_half * m.rdzw[k - 1] * g;
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
            r_p1 += 0.25 * ws * g;
            r_0 += 0.25 * ws * g;
            w_p1 += 0.25 * rs * g;
            w_0 += 0.25 * rs * g;
        } else {
            double d[5];
            flux3_partials(wc[j - 1], wc[j], wc[j + 1], wc[j + 2], 0.5 * (rwc[j + 1] + rwc[j]), 1.0, d);
            w_m1 += d[0] * g;
            w_0 += d[1] * g;
            w_p1 += d[2] * g;
            w_p2 += d[3] * g;
            r_p1 += 0.5 * d[4] * g;
            r_0 += 0.5 * d[4] * g;
        }
    }
    // curvature of w at interface k (1..L-1), added to the advective sum before
    // its 1/area scaling: rh (ur^2 + vr^2) / r_earth + curvCell ur rh, with ur, vr
    // and rh interpolated from levels k and k-1
    if (kOpt && P.curvature && k >= 1 && k < L) {
        const double gh = ia * bar_of(B, tend_w, i);
        const double *ur = F[ur_cell] + oc, *vr = F[vr_cell] + oc, *rz = F[rho_zz] + oc;
        const double uri = fzm[k] * ur[k] + fzp[k] * ur[k - 1], vri = fzm[k] * vr[k] + fzp[k] * vr[k - 1];
        const double rh = rz[k] * fzm[k] + rz[k - 1] * fzp[k];
        const double grh = gh * ((uri * uri + vri * vri) / kEarthRadius + m.curvCell[c] * uri);
        const double gur = gh * (2.0 * uri * rh / kEarthRadius + m.curvCell[c] * rh);
        const double gvr = gh * (2.0 * vri * rh / kEarthRadius);
        cz_0 += fzm[k] * grh;
        cz_m1 += fzp[k] * grh;
        cu_0 = fzm[k] * gur;
        cu_m1 = fzp[k] * gur;
        cv_0 = fzm[k] * gvr;
        cv_m1 = fzp[k] * gvr;
    }
    // ------------------------------------------------ exchange, then one write per element
    if (kOpt) {
        const double g_cz = cz_0 + from_lane(cz_m1, 1, k, L);
        const double g_cu = cu_0 + from_lane(cu_m1, 1, k, L);
        const double g_cv = cv_0 + from_lane(cv_m1, 1, k, L);
        if (P.curvature && k < L) {
            add_own(B, ur_cell, oc + k, g_cu);
            add_own(B, vr_cell, oc + k, g_cv);
        }
        if ((P.curvature || vmix_w || vmix_t) && k < L) add_own(B, rho_zz, oc + k, g_cz);
    }
    const double g_th = th_0 + from_lane(th_m1, 1, k, L) + from_lane(th_m2, 2, k, L) + from_lane(th_p1, -1, k, L);
    const double g_ts = ts_0 + from_lane(ts_m1, 1, k, L);
    const double g_pp = pp_0 + from_lane(pp_m1, 1, k, L);
    const double g_dp = dp_0 + from_lane(dp_m1, 1, k, L);
    const double g_rw = r_0 + from_lane(r_p1, -1, k, L);
    const double g_w = w_0 + from_lane(w_m1, 1, k, L) + from_lane(w_p1, -1, k, L) + from_lane(w_p2, -2, k, L);
    if (k < L) {
        add_own(B, theta_m, oc + k, g_th);
        add_own(B, theta_m_save, oc + k, g_ts);
        if (rk1) add_own(B, pp, oc + k, g_pp);
        S.g_dpdz[oc + k] = g_dp;
    }
    if (k <= L) {
        add_own(B, rw, i, g_rw);
        add_own(B, rw_save, i, rs_0);
        add_own(B, w, i, g_w);
    }
}

// A1d, thread (cell c, level / interface k), rk 1: the del4 Laplacian
// adjoints.  tend_theta_euler(s) -= sum_slots es(s, e) (delsq_theta(c2) -
// delsq_theta(c1)) over the edges e of each cell s, es = msd4 visc4 prandtl_inv
// ia(s) dv sgn(s, e) / dc, and the same for w (without prandtl_inv): cell c
// collects, edge by edge, its own term and the neighbour's.
__device__ __forceinline__ void dyn_a1_del4(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Bar B,
                                            const DynAdjScratch &S, int c, int k) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c >= m.nCells || k > L) return;
    const bool th_on = P.h_theta_eddy_visc4 > 0.0 && k < L;
    const bool w_on = P.h_mom_eddy_visc4 > 0.0 && k >= 1 && k < L;
    double gt = 0.0, gw = 0.0;
    if (th_on || w_on) {
        const int n = m.nEdgesOnCell[c];
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[c * ME + j];
            const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
            const double side = (c == c1) ? 1.0 : -1.0;  // c1 receives +es g, c2 receives -es g
            for (int q = 0; q < 2; ++q) {
                const int s_ = (q == 0) ? c : ((c == c1) ? c2 : c1);
                if (s_ >= m.nCells) continue;
                const double sg = (q == 0) ? m.edgesOnCell_sign[c * ME + j] : edge_sign_in_cell(m, s_, e);
                if (th_on) {
                    const long long o = (long long)s_ * L + k;
                    const double r_area = P.h_theta_eddy_visc4 * P.prandtl_inv * m.invAreaCell[s_];
                    const double es = m.meshScalingDel4[e] * r_area * m.dvEdge[e] * sg * m.invDcEdge[e];
                    gt += side * (es * (bar_of(B, tend_theta_euler, o) + bar_of(B, tend_theta, o)));
                }
                if (w_on) {
                    const long long o = (long long)s_ * Lw + k;
                    const double r_area = P.h_mom_eddy_visc4 * m.invAreaCell[s_];
                    const double es = m.meshScalingDel4[e] * r_area * m.dvEdge[e] * sg * m.invDcEdge[e];
                    gw += side * (es * (bar_of(B, tend_w_euler, o) + bar_of(B, tend_w, o)));
                }
            }
        }
    }
    if (k < L) S.g_delsq_theta[(long long)c * L + k] = gt;
    S.g_delsq_w[(long long)c * Lw + k] = gw;
}

// ---------------------------------------------------------------- A2 (rk 1)
// A2 as gathers.  The per-edge sums over the edge's two
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

// ---------------------------------------------------------------- A3c
// thread (cell c, level / interface k): adjoints of theta_m (k < L) and w
// (1 <= k < L) at the cells of the advection stencils, gathered over the
// (edge, slot) pairs naming cell c, in (edge, slot) order; then the sums over
// the cell's edges of what A3 stored per edge.  Launched over max(cells,
// vertices): at rk 1 with del4 the vertex threads gather delsq_vort.
__device__ __forceinline__ void dyn_a3_cells(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Fwd F, Bar B,
                                             const DynAdjScratch &S, const mpas_hip_inverse_tables &inv, int c,
                                             int k) {
    using namespace dyn;
    const int L = m.nLevels, Lw = L + 1, MA = m.maxAdv, ME = m.maxEdges;
    if (k >= L) return;
    const bool rk1 = (P.rk_step == 1);
    const bool del4 = rk1 && P.h_mom_eddy_visc4 > 0.0;
    if (del4 && c < m.nVertices) {
        double gv = 0.0;
        for (int j = 0; j < 3; ++j) {
            const int e = m.edgesOnVertex[3 * c + j];
            if (e >= m.nEdges) continue;
            const double x = S.x[7][(long long)e * L + k];
            gv += (m.verticesOnEdge[2 * e + 1] == c) ? x : -x;
        }
        S.g_delsq_vort[(long long)c * L + k] = gv;
    }
    if (c >= m.nCells) return;
    // The cell's (edge, slot) entries, L at a time across the lanes: lane t
    // loads entry t of the chunk and its two stencil coefficients, and the
    // entry loop takes them back by readlane, so its iterations depend on no
    // load but their own column gathers.  Lanes 0..L-1 are all present here.
    const int j0 = inv.advStart[c], cnt = inv.advStart[c + 1] - j0;
    const double fzm_k = m.fzm[k], fzp_k = m.fzp[k];
    double gt = 0.0, gw = 0.0;
    for (int b0 = 0; b0 < cnt; b0 += L) {
        const int t_l = b0 + k;
        int e_l = 0;
        double c0_l = 0.0, c3_l = 0.0;
        if (t_l < cnt) {
            e_l = inv.advEdge[j0 + t_l];
            const int a_l = inv.advSlot[j0 + t_l];
            c0_l = m.adv_coefs[e_l * MA + a_l];
            c3_l = m.adv_coefs_3rd[e_l * MA + a_l];
        }
        const int nt = (cnt - b0 < L) ? cnt - b0 : L;
        // four entries at a time: their gathers are all issued before the
        // first is used
        for (int t0 = 0; t0 < nt; t0 += 4) {
            double rv[4], gtf[4], gwf[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                rv[u] = gtf[u] = gwf[u] = 0.0;
                if (t0 + u < nt) {  // wave-uniform
                    const long long i = (long long)bcast(e_l, t0 + u) * L + k;
                    rv[u] = F[ru][i];
                    gtf[u] = S.g_tf[i];
                    gwf[u] = S.g_wf[i];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (t0 + u >= nt) break;
                const double c0 = bcast(c0_l, t0 + u), c3 = bcast(c3_l, t0 + u);
                const double rut = rv[u];
                // ru at level k - 1 of the same edge column: lane k - 1's value
                const double rut_m1 = __shfl_up(rut, 1);
                gt += (c0 + sign_pm(rut) * c3) * rut * gtf[u];
                if (k >= 1) {
                    const double ruw = fzm_k * rut + fzp_k * rut_m1;
                    gw += (c0 + sign_pm(ruw) * c3) * ruw * gwf[u];
                }
            }
        }
    }
    const long long o = (long long)c * L + k;
    add_own(B, theta_m, o, gt);
    if (k >= 1) add_own(B, w, (long long)c * Lw + k, gw);
    const int n = m.nEdgesOnCell[c];
    double gke = 0.0, ghd = 0.0, grw = 0.0, g3 = 0.0, gdp = 0.0, gkd = 0.0, gdd = 0.0;
    for (int j = 0; j < n; ++j) {
        const int e = m.edgesOnCell[c * ME + j];
        const long long i = (long long)e * L + k;
        const double side = (m.cellsOnEdge[2 * e] == c) ? 1.0 : -1.0;
        gke += side * S.x[0][i];
        ghd += S.x[1][i];
        if (k >= 1) grw += S.x[2][i];
        if (rk1) {
            g3 += side * S.x[3][i];
            gdp += S.x[4][i];
            gkd += S.x[5][i];
            if (del4) gdd += side * S.x[6][i];
        } else {
            g3 += S.x[3][i];
        }
    }
    add_own(B, ke, o, gke);
    add_own(B, h_divergence, o, ghd);
    if (k >= 1) add_own(B, rw, (long long)c * Lw + k, grw);
    if (P.curvature) {
        // w_avg of each edge reads interfaces k and k + 1 of both its cells:
        // interface k collects the shares of levels k and k - 1, lane L - 1 also
        // interface L's
        double gcw = 0.0, gtop = 0.0;
        for (int j = 0; j < n; ++j) {
            const long long i = (long long)m.edgesOnCell[c * ME + j] * L + k;
            gcw += S.x[8][i];
            if (k >= 1) gcw += S.x[8][i - 1];
            if (k == L - 1) gtop += S.x[8][i];
        }
        add_own(B, w, (long long)c * Lw + k, gcw);
        if (k == L - 1) add_own(B, w, (long long)c * Lw + L, gtop);
    }
    if (rk1) {
        add_own(B, pp, o, g3);
        S.g_dpdz[o] += gdp;
        add_own(B, kdiff, o, gkd);
        if (del4) S.g_delsq_div[o] = gdd;
    } else {
        add_own(B, theta_m_save, o, g3);
    }
}

// ---------------------------------------------------------------- A3e
// thread (edge e, level k): adjoints of u and pv_edge at the edges of the
// PV-flux stencils, gathered over the (edge, slot) pairs naming edge e; at
// rk 1 with del4, delsq_u(e) through delsq_div of its cells and delsq_vort of
// its vertices
__device__ __forceinline__ void dyn_a3_edges(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Fwd F, Bar B,
                                             const DynAdjScratch &S, const mpas_hip_inverse_tables &inv, int e,
                                             int k) {
    using namespace dyn;
    const int L = m.nLevels, ME2 = m.maxEdges2, ME = m.maxEdges;
    if (e >= m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    const int j0 = inv.eoeStart[e], cnt = inv.eoeStart[e + 1] - j0;
    const double ue = F[u][i], pve = F[pv_edge][i];
    double gu = 0.0, gp = 0.0;
    // the (edge, slot) entries naming e, L at a time across the lanes (lane t
    // holds entry t of the chunk and its weight), taken back by readlane; four
    // entries' gathers are issued before the first is used
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
            double gq[4], pv[4];
#pragma unroll
            for (int u4 = 0; u4 < 4; ++u4) {
                gq[u4] = pv[u4] = 0.0;
                if (t0 + u4 < nt) {  // wave-uniform
                    const long long io = (long long)bcast(eo_l, t0 + u4) * L + k;
                    gq[u4] = S.g_q[io];
                    pv[u4] = F[pv_edge][io];
                }
            }
#pragma unroll
            for (int u4 = 0; u4 < 4; ++u4) {
                if (t0 + u4 >= nt) break;
                const double wq = bcast(w_l, t0 + u4) * gq[u4];
                gu += wq * 0.5 * (pv[u4] + pve);
                gp += wq * ue * 0.5;
            }
        }
    }
    add_own(B, u, i, gu);
    add_own(B, pv_edge, i, gp);
    if (P.rk_step == 1 && P.h_mom_eddy_visc4 > 0.0) {
        // delsq_vort(v) = iat sum dc sgn delsq_u; delsq_div(c) = ia sum dv sgn delsq_u
        double g = S.g_delsq_u[i];
        for (int q = 0; q < 2; ++q) {
            const int vx = m.verticesOnEdge[2 * e + q];
            if (vx >= m.nVertices) continue;
            for (int j = 0; j < 3; ++j) {
                if (m.edgesOnVertex[3 * vx + j] != e) continue;
                const double wgt = m.invAreaTriangle[vx] * m.dcEdge[e] * m.edgesOnVertex_sign[3 * vx + j];
                g += wgt * S.g_delsq_vort[(long long)vx * L + k];
            }
        }
        for (int q = 0; q < 2; ++q) {
            const int c = m.cellsOnEdge[2 * e + q];
            if (c >= m.nCells) continue;
            const int n = m.nEdgesOnCell[c];
            for (int j = 0; j < n; ++j) {
                if (m.edgesOnCell[c * ME + j] != e) continue;
                const double wgt = m.invAreaCell[c] * m.dvEdge[e] * m.edgesOnCell_sign[c * ME + j];
                g += wgt * S.g_delsq_div[(long long)c * L + k];
            }
        }
        S.g_delsq_u[i] = g;
    }
}

// ---------------------------------------------------------------- A6
// thread (cell c, level k), launched over max(cells, vertices) at rk 1: the
// vertex threads take vorticity(v) through delsq_u of the vertex's edges
// (delsq_u(e) = (div(c2) - div(c1)) r_dc - (vort(v2) - vort(v1)) r_dv)
__device__ __forceinline__ void dyn_a6(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Fwd F, Bar B,
                                       const DynAdjScratch &S, int c, int k) {
    using namespace dyn;
    const int L = m.nLevels, ME = m.maxEdges;
    if (k >= L) return;
    const bool rk1 = (P.rk_step == 1);
    if (rk1 && c < m.nVertices) {
        double gz = 0.0;
        for (int j = 0; j < 3; ++j) {
            const int e = m.edgesOnVertex[3 * c + j];
            if (e >= m.nEdges) continue;
            const double r_dv = dmin(m.invDvEdge[e], 4.0 * m.invDcEdge[e]);
            const double g = S.g_delsq_u[(long long)e * L + k];
            gz += (m.verticesOnEdge[2 * e + 1] == c) ? -r_dv * g : r_dv * g;
        }
        add_own(B, vorticity, (long long)c * L + k, gz);
    }
    if (c >= m.nCells) return;
    const long long o = (long long)c * L + k;
    if (!rk1) {
        add_own(B, kdiff_in, o, bar_of(B, kdiff, o));
        return;
    }
    const int n = m.nEdgesOnCell[c];
    {
        double gd = 0.0;
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[c * ME + j];
            const double g = S.g_delsq_u[(long long)e * L + k];
            gd += (m.cellsOnEdge[2 * e + 1] == c) ? m.invDcEdge[e] * g : -m.invDcEdge[e] * g;
        }
        add_own(B, divergence, o, gd);
    }
    // dpdz = -g (rb qtot + rr_save (1 + qtot))
    const double gdp = S.g_dpdz[o];
    add_own(B, qtot, o, -MPAS_GRAVITY * (m.rho_base[o] + F[rr_save][o]) * gdp);
    add_own(B, rr_save, o, -MPAS_GRAVITY * (1.0 + F[qtot][o]) * gdp);
    // kdiff = min(coef def, cap), def = sqrt(dd^2 + dof^2), floored on the CAM
    // levels by max(kdiff, cam): the adjoints of dd and dof, spread over the
    // cell's edges by A7.  A fixed kdiff depends on nothing.
    double gdd = 0.0, gdof = 0.0;
    double gkd = B[kdiff][o];
    if (gkd != 0.0 && !P.mix_fixed) {
        double dd = 0.0, dof = 0.0;
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[c * ME + j];
            const double a = m.defc_a[c * ME + j], b = m.defc_b[c * ME + j];
            const double us = F[u][(long long)e * L + k], vs = F[v][(long long)e * L + k];
            dd = (dd + a * us) - b * vs;
            dof = (dof + b * us) + a * vs;
        }
        const double def = sqrt(dd * dd + dof * dof);
        const double smag = dmin(P.kdiff_coef * def, P.kdiff_cap);
        if (k >= L - P.cam_levels) {
            double dk, dcam;
            max_partials(smag, cam_kdiff(P, L, k), dk, dcam);
            gkd *= dk;
        }
        double dsel, dcap;
        min_partials(P.kdiff_coef * def, P.kdiff_cap, dsel, dcap);
        if (dsel != 0.0 && def != 0.0) {
            const double gdef = P.kdiff_coef * dsel * gkd;
            gdd = dd / def * gdef;
            gdof = dof / def * gdef;
        }
    }
    S.g_dd[o] = gdd;
    S.g_dof[o] = gdof;
}

// ---------------------------------------------------------------- A7
// thread (edge e, level k): ru(e) through h_divergence = ia sum sgn dv ru of
// its two cells; at rk 1, u and v through the cells' deformation
__device__ __forceinline__ void dyn_a7(const mpas_hip_mesh &m, const mpas_hip_dyn_params &P, Bar B,
                                       const DynAdjScratch &S, int e, int k) {
    using namespace dyn;
    const int L = m.nLevels, ME = m.maxEdges;
    if (e >= m.nEdges || k >= L) return;
    const bool rk1 = (P.rk_step == 1);
    double gru = 0.0, gu = 0.0, gv = 0.0;
    for (int q = 0; q < 2; ++q) {
        const int c = m.cellsOnEdge[2 * e + q];
        if (c >= m.nCells) continue;
        const int n = m.nEdgesOnCell[c];
        const long long o = (long long)c * L + k;
        for (int j = 0; j < n; ++j) {
            if (m.edgesOnCell[c * ME + j] != e) continue;
            gru += m.invAreaCell[c] * m.edgesOnCell_sign[c * ME + j] * m.dvEdge[e] * B[h_divergence][o];
            if (rk1) {
                const double a = m.defc_a[c * ME + j], b = m.defc_b[c * ME + j];
                gu += a * S.g_dd[o] + b * S.g_dof[o];
                gv += -b * S.g_dd[o] + a * S.g_dof[o];
            }
        }
    }
    const long long i = (long long)e * L + k;
    add_own(B, ru, i, gru);
    if (rk1) {
        add_own(B, u, i, gu);
        add_own(B, v, i, gv);
    }
}

}  // namespace adj
}  // namespace mpas
