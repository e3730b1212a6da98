// Tangent-linear (Jacobian-vector product) bodies of the split-explicit
// step's kernels.  Included after adjoint_bodies.hpp, whose slot enums they
// share: fwd[] holds a forward call's inputs then outputs, dot[] their
// tangents in the same slots (input tangents read, null meaning zero; output
// tangents written).  Each thread computes the tangent of the outputs it owns
// from the tangents of the inputs its forward stencil reads, so no two
// threads write one element.  Derivatives follow the same conventions as the
// adjoints (the hard primal, branches frozen, torch's values at switching
// points).
#pragma once

namespace mpas {
namespace adj {

using Dot = double *const *;

__device__ __forceinline__ double dot_of(Dot T, int slot, long long i) { return T[slot] ? T[slot][i] : 0.0; }
__device__ __forceinline__ void dot_put(Dot T, int slot, long long i, double v) {
    if (T[slot]) T[slot][i] = v;
}

// flux3 directional derivative
__device__ __forceinline__ double flux3_dot(double qm2, double qm1, double q0, double qp1, double ua, double c3,
                                            double dqm2, double dqm1, double dq0, double dqp1, double dua) {
    double d[5];
    flux3_partials(qm2, qm1, q0, qp1, ua, c3, d);
    return d[0] * dqm2 + d[1] * dqm1 + d[2] * dq0 + d[3] * dqp1 + d[4] * dua;
}
// min / max with the tie split
__device__ __forceinline__ double min_dot(double a, double b, double da, double db) {
    double pa, pb;
    min_partials(a, b, pa, pb);
    return pa * da + pb * db;
}
__device__ __forceinline__ double max_dot(double a, double b, double da, double db) {
    double pa, pb;
    max_partials(a, b, pa, pb);
    return pa * da + pb * db;
}

// ===========================================================================
// smlstep, thread (cell c, interface k)
// ===========================================================================
__device__ __forceinline__ void smlstep_jvp(const mpas_hip_mesh &m, Fwd F, Dot T, int c, int k) {
    using namespace smlstep;
    const int L = m.nLevels, ME = m.maxEdges, Lw = L + 1;
    const long long i = (long long)c * Lw + k;
    double dw = dot_of(T, w_tend_in, i);
    if (c >= m.nCells || k < 1 || k >= L) {
        dot_put(T, w_tend_out, i, dw);
        return;
    }
    // terrain terms only: a flat mesh has zb_cell = zb3_cell = 0 and adds nothing
    const int n = m.flatTerrain ? 0 : m.nEdgesOnCell[c];
    for (int j = 0; j < n; ++j) {
        const int e = m.edgesOnCell[c * ME + j];
        const double uk = F[u_tend][(long long)e * L + k];
        const long long zo = ((long long)c * ME + j) * Lw + k;
        const double coef = m.zb_cell[zo] + sign_pm(uk) * m.zb3_cell[zo];
        const double dui = m.fzm[k] * dot_of(T, u_tend, (long long)e * L + k) +
                           m.fzp[k] * dot_of(T, u_tend, (long long)e * L + k - 1);
        dw -= coef * m.edgesOnCell_sign[c * ME + j] * dui;
    }
    const double zi = m.fzm[k] * m.zz[c * L + k] + m.fzp[k] * m.zz[c * L + k - 1];
    dot_put(T, w_tend_out, i, zi * dw);
}

// ===========================================================================
// divergence damping, thread (edge e, level k)
// ===========================================================================
__device__ __forceinline__ void divdamp_jvp(const mpas_hip_mesh &m, double coef, Fwd F, Dot T, int e, int k) {
    using namespace divdamp;
    const int L = m.nLevels;
    const long long i = (long long)e * L + k;
    double d = dot_of(T, ru_p_in, i);
    if (e < m.nEdges) {
        const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
        const long long o1 = (long long)c1 * L + k, o2 = (long long)c2 * L + k;
        const double d1 = -(F[rtheta_pp][o1] - F[rtheta_pp_old][o1]);
        const double d2 = -(F[rtheta_pp][o2] - F[rtheta_pp_old][o2]);
        const double ts = F[theta_m][o1] + F[theta_m][o2];
        if (fabs(ts) >= 1.0e-30) {
            const double dd1 = -(dot_of(T, rtheta_pp, o1) - dot_of(T, rtheta_pp_old, o1));
            const double dd2 = -(dot_of(T, rtheta_pp, o2) - dot_of(T, rtheta_pp_old, o2));
            const double dts = dot_of(T, theta_m, o1) + dot_of(T, theta_m, o2);
            d += coef * (dd2 - dd1) / ts - (coef * (d2 - d1) / ts) / ts * dts;
        }
    }
    dot_put(T, ru_p_out, i, d);
}

// ===========================================================================
// moist coefficients
// ===========================================================================
__device__ __forceinline__ double moist_dqtot(int S, int q0, int q1, Dot T, long long o) {
    using namespace moist;
    double d = 0.0;
    for (int q = q0; q < q1; ++q) d += dot_of(T, scalars, o * S + q);
    return d;
}
// thread (cell c, interface/level k): qtot(k) and cqw(k)
__device__ __forceinline__ void moist_cell_jvp(const mpas_hip_mesh &m, int S, int q0, int q1, Fwd F, Dot T, int c,
                                               int k) {
    using namespace moist;
    const int L = m.nLevels, Lw = L + 1;
    if (c > m.nCells || k > L) return;
    if (c == m.nCells) {
        if (k < L) dot_put(T, qtot, (long long)c * L + k, 0.0);
        dot_put(T, cqw, (long long)c * Lw + k, 0.0);
        return;
    }
    if (k < L) dot_put(T, qtot, (long long)c * L + k, moist_dqtot(S, q0, q1, T, (long long)c * L + k));
    double dcw = 0.0;
    if (k >= 1 && k < L) {
        const double cw = F[cqw][(long long)c * Lw + k];
        const double dq = 0.5 * (moist_dqtot(S, q0, q1, T, (long long)c * L + k) +
                                 moist_dqtot(S, q0, q1, T, (long long)c * L + k - 1));
        dcw = -(cw * cw) * dq;
    }
    dot_put(T, cqw, (long long)c * Lw + k, dcw);
}
// thread (edge e, level k): cqu
__device__ __forceinline__ void moist_edge_jvp(const mpas_hip_mesh &m, int S, int q0, int q1, Fwd F, Dot T, int e,
                                               int k) {
    using namespace moist;
    const int L = m.nLevels;
    if (e > m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    if (e == m.nEdges) {
        dot_put(T, cqu, i, 0.0);
        return;
    }
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    double dq = 0.0;
    for (int q = q0; q < q1; ++q)
        dq += 0.5 * (dot_of(T, scalars, ((long long)c1 * L + k) * S + q) +
                     dot_of(T, scalars, ((long long)c2 * L + k) * S + q));
    const double cu = F[cqu][i];
    dot_put(T, cqu, i, -(cu * cu) * dq);
}

// ===========================================================================
// reconstruction, thread (cell c, level k)
// ===========================================================================
__device__ __forceinline__ void reconstruct_jvp(const mpas_hip_mesh &m, Fwd F, Dot T, int c, int k) {
    using namespace recon;
    (void)F;
    const int L = m.nLevels, ME = m.maxEdges;
    if (c > m.nCells || k >= L) return;
    const long long o = (long long)c * L + k;
    double x = 0.0, y = 0.0, z = 0.0;
    if (c < m.nCells) {
        const int n = m.nEdgesOnCell[c];
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[c * ME + j];
            const double *cf = m.coeffs_reconstruct + ((long long)c * ME + j) * 3;
            const double du = dot_of(T, u, (long long)e * L + k);
            x += cf[0] * du;
            y += cf[1] * du;
            z += cf[2] * du;
        }
    }
    const double clat = (c < m.nCells) ? cos(m.latCell[c]) : 0.0, slat = (c < m.nCells) ? sin(m.latCell[c]) : 0.0;
    const double clon = (c < m.nCells) ? cos(m.lonCell[c]) : 0.0, slon = (c < m.nCells) ? sin(m.lonCell[c]) : 0.0;
    dot_put(T, ux, o, x);
    dot_put(T, uy, o, y);
    dot_put(T, uz, o, z);
    dot_put(T, uzonal, o, -x * slon + y * clon);
    dot_put(T, umerid, o, -(x * clon + y * slon) * slat + z * clat);
}

// ===========================================================================
// recovery
// ===========================================================================
// pass 1, thread (cell c, level k): densities, theta_m, exner, pressure
__device__ __forceinline__ void recover_cell_jvp(const mpas_hip_mesh &m, int rk_step, double dt, Fwd F, Dot T, int c,
                                                 int k) {
    using namespace recover;
    const int L = m.nLevels;
    if (c > m.nCells || k >= L) return;
    const long long o = (long long)c * L + k;
    if (c == m.nCells) {
        for (int s : {rho_p, rho_zz, rtheta_p, theta_m, exner, pressure_p}) dot_put(T, s, o, 0.0);
        return;
    }
    const double rcv = MPAS_RGAS / (MPAS_CP - MPAS_RGAS);
    const double drp = dot_of(T, rho_p_save, o) + dot_of(T, rho_pp, o);
    const double drz = drp;
    const double rz = F[rho_zz][o];
    double drtp = dot_of(T, rtheta_p_save, o) + dot_of(T, rtheta_pp, o);
    if (rk_step == 3)
        drtp -= dt * (drz * F[rt_diabatic_tend][o] + rz * dot_of(T, rt_diabatic_tend, o));
    dot_put(T, rho_p, o, drp);
    dot_put(T, rho_zz, o, drz);
    dot_put(T, rtheta_p, o, drtp);
    dot_put(T, theta_m, o, drtp / rz - F[theta_m][o] / rz * drz);
    if (rk_step == 3) {
        const double rtp = F[rtheta_p][o], rtb = m.rtheta_base[o], ex = F[exner][o];
        const double dex = rcv * ex / (rtp + rtb) * drtp;
        dot_put(T, exner, o, dex);
        dot_put(T, pressure_p, o, m.zz[o] * MPAS_RGAS * (dex * rtp + ex * drtp + rtb * dex));
    } else {
        dot_put(T, exner, o, dot_of(T, exner_in, o));
        dot_put(T, pressure_p, o, dot_of(T, pressure_p_in, o));
    }
}
// pass 2, thread (edge e, level k): ruAvg, ru, u (needs the rho_zz tangent of pass 1)
__device__ __forceinline__ void recover_edge_jvp(const mpas_hip_mesh &m, double invNs, Fwd F, Dot T, int e, int k) {
    using namespace recover;
    const int L = m.nLevels;
    if (e > m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    if (e == m.nEdges) {
        dot_put(T, ruAvg, i, dot_of(T, ruAvg_in, i));
        dot_put(T, ru, i, 0.0);
        dot_put(T, u, i, 0.0);
        return;
    }
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    dot_put(T, ruAvg, i, dot_of(T, ru_save, i) + dot_of(T, ruAvg_in, i) * invNs);
    const double drn = dot_of(T, ru_save, i) + dot_of(T, ru_p, i);
    dot_put(T, ru, i, drn);
    const double s = F[rho_zz][(long long)c1 * L + k] + F[rho_zz][(long long)c2 * L + k];
    const double ds = dot_of(T, rho_zz, (long long)c1 * L + k) + dot_of(T, rho_zz, (long long)c2 * L + k);
    dot_put(T, u, i, 2.0 * drn / s - F[u][i] / s * ds);
}
// pass 3, thread (cell c, interface k): wwAvg, rw, w (needs the ru and rho_zz tangents)
__device__ __forceinline__ void recover_w_jvp(const mpas_hip_mesh &m, double invNs, Fwd F, Dot T, int c, int k) {
    using namespace recover;
    const int L = m.nLevels, ME = m.maxEdges, Lw = L + 1;
    if (c > m.nCells || k > L) return;
    const long long ow = (long long)c * Lw, oc = (long long)c * L;
    if (c == m.nCells || k == L) {
        dot_put(T, wwAvg, ow + k, dot_of(T, wwAvg_in, ow + k));
        dot_put(T, rw, ow + k, 0.0);
        dot_put(T, w, ow + k, 0.0);
        return;
    }
    // the edge loops carry the terrain terms only (nothing on a flat mesh)
    const int n = m.flatTerrain ? 0 : m.nEdgesOnCell[c];
    const double *rz = F[rho_zz] + oc;
    auto drz = [&](int kk) { return dot_of(T, rho_zz, oc + kk); };
    if (k == 0) {
        dot_put(T, wwAvg, ow, dot_of(T, wwAvg_in, ow));
        dot_put(T, rw, ow, 0.0);
        double dsacc = 0.0;
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[c * ME + j];
            const double *re = F[ru] + (long long)e * L;
            const double fs = m.cf1 * re[0] + m.cf2 * re[1] + m.cf3 * re[2];
            const long long zo = ((long long)c * ME + j) * Lw;
            const double coef = m.zb_cell[zo] + sign_pm(fs) * m.zb3_cell[zo];
            const double dfs = m.cf1 * dot_of(T, ru, (long long)e * L) + m.cf2 * dot_of(T, ru, (long long)e * L + 1) +
                               m.cf3 * dot_of(T, ru, (long long)e * L + 2);
            dsacc += m.edgesOnCell_sign[c * ME + j] * coef * dfs;
        }
        const double den = m.cf1 * rz[0] + m.cf2 * rz[1] + m.cf3 * rz[2];
        const double dden = m.cf1 * drz(0) + m.cf2 * drz(1) + m.cf3 * drz(2);
        dot_put(T, w, ow, dsacc / den - F[w][ow] / den * dden);
        return;
    }
    dot_put(T, wwAvg, ow + k, dot_of(T, rw_save, ow + k) + dot_of(T, wwAvg_in, ow + k) * invNs);
    const double drw = dot_of(T, rw_save, ow + k) + dot_of(T, rw_p, ow + k);
    dot_put(T, rw, ow + k, drw);
    const double zi = m.fzm[k] * m.zz[oc + k] + m.fzp[k] * m.zz[oc + k - 1];
    double dacc = drw / zi;
    for (int j = 0; j < n; ++j) {
        const int e = m.edgesOnCell[c * ME + j];
        const double *re = F[ru] + (long long)e * L;
        const double fi = m.fzm[k] * re[k] + m.fzp[k] * re[k - 1];
        const long long zo = ((long long)c * ME + j) * Lw + k;
        const double coef = m.zb_cell[zo] + sign_pm(fi) * m.zb3_cell[zo];
        const double dfi = m.fzm[k] * dot_of(T, ru, (long long)e * L + k) + m.fzp[k] * dot_of(T, ru, (long long)e * L + k - 1);
        dacc += m.edgesOnCell_sign[c * ME + j] * coef * dfi;
    }
    const double den = m.fzm[k] * rz[k] + m.fzp[k] * rz[k - 1];
    const double dden = m.fzm[k] * drz(k) + m.fzp[k] * drz(k - 1);
    dot_put(T, w, ow + k, dacc / den - F[w][ow + k] / den * dden);
}

// ===========================================================================
// diagnostics: pass 1 vertices, pass 2 cells, pass 3 edges
// scratch tangents: pv_vertex and ke_vertex (V), pv_cell (C)
// ===========================================================================
struct DiagDot {
    double *d_pvv, *d_kev, *d_pvc;
    double *pvc;  // pv_cell values, formed once by the cell pass when apvm is on (else null)
};
// thread (edge e, level k): d ke_edge = dc dv 2 u du (inline helper)
__device__ __forceinline__ double diag_dke_edge(const mpas_hip_mesh &m, Fwd F, Dot T, int e, int k) {
    using namespace diag;
    const long long i = (long long)e * m.nLevels + k;
    return (m.dcEdge[e] * m.dvEdge[e]) * 2.0 * F[u][i] * dot_of(T, u, i);
}
__device__ __forceinline__ void diag_vertex_jvp(const mpas_hip_mesh &m, Fwd F, Dot T, const DiagDot &S, int vx,
                                                int k) {
    using namespace diag;
    const int L = m.nLevels;
    if (vx > m.nVertices || k >= L) return;
    const long long o = (long long)vx * L + k;
    if (vx == m.nVertices) {
        dot_put(T, vorticity, o, 0.0);
        return;
    }
    double s = 0.0, sk = 0.0;
    for (int j = 0; j < 3; ++j) {
        const int e = m.edgesOnVertex[3 * vx + j];
        s += m.edgesOnVertex_sign[3 * vx + j] * m.dcEdge[e] * dot_of(T, u, (long long)e * L + k);
        sk += diag_dke_edge(m, F, T, e, k);
    }
    const double dz = s * m.invAreaTriangle[vx];
    dot_put(T, vorticity, o, dz);
    S.d_pvv[o] = dz;
    S.d_kev[o] = 0.25 * m.invAreaTriangle[vx] * sk;
}
__device__ __forceinline__ void diag_cell_jvp(const mpas_hip_mesh &m, Fwd F, Dot T, const DiagDot &S, int c, int k) {
    using namespace diag;
    const int L = m.nLevels, ME = m.maxEdges;
    if (c > m.nCells || k >= L) return;
    const long long o = (long long)c * L + k;
    if (c == m.nCells) {
        dot_put(T, divergence, o, 0.0);
        dot_put(T, ke, o, 0.0);
        return;
    }
    const int n = m.nEdgesOnCell[c];
    const double ia = m.invAreaCell[c];
    const double ke_fact = 1.0 - 0.375;
    double d = 0.0, kc = 0.0, pc = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i) {
        const int e = m.edgesOnCell[c * ME + i];
        d += m.edgesOnCell_sign[c * ME + i] * m.dvEdge[e] * dot_of(T, u, (long long)e * L + k);
        kc += 0.25 * diag_dke_edge(m, F, T, e, k);
    }
    tot = ke_fact * (kc * ia);
    for (int i = 0; i < n; ++i) {
        const int vx = m.verticesOnCell[c * ME + i];
        const double ka = m.kiteAreasOnVertex[3 * vx + m.kiteForCell[c * ME + i]];
        tot += (1.0 - ke_fact) * ka * S.d_kev[(long long)vx * L + k] * ia;
        pc += ka * S.d_pvv[(long long)vx * L + k] * ia;
    }
    dot_put(T, divergence, o, d * ia);
    dot_put(T, ke, o, tot);
    S.d_pvc[o] = pc;
    if (S.pvc) S.pvc[o] = diag_pv_cell(m, F, c, k);
}
__device__ __forceinline__ void diag_edge_jvp(const mpas_hip_mesh &m, double dt, double apvm, int rk_step, Fwd F,
                                              Dot T, const DiagDot &S, int e, int k) {
    using namespace diag;
    const int L = m.nLevels;
    if (e > m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    if (e == m.nEdges) {
        for (int s : {h_edge, v, pv_edge}) dot_put(T, s, i, 0.0);
        return;
    }
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const int v1 = m.verticesOnEdge[2 * e], v2 = m.verticesOnEdge[2 * e + 1];
    dot_put(T, h_edge, i, 0.5 * (dot_of(T, h, (long long)c1 * L + k) + dot_of(T, h, (long long)c2 * L + k)));
    double dvv;
    if (rk_step <= 0 || rk_step == 3) {
        dvv = 0.0;
        const int ne = m.nEdgesOnEdge[e];
        for (int j = 0; j < ne; ++j) {
            const int ee = m.edgesOnEdge[e * m.maxEdges2 + j];
            dvv += m.weightsOnEdge[e * m.maxEdges2 + j] * dot_of(T, u, (long long)ee * L + k);
        }
    } else {
        dvv = dot_of(T, v_in, i);
    }
    dot_put(T, v, i, dvv);
    const double dp1 = S.d_pvv[(long long)v1 * L + k], dp2 = S.d_pvv[(long long)v2 * L + k];
    double dpe = 0.5 * (dp1 + dp2);
    if (apvm > 0.0) {
        const double p1 = m.fVertex[v1] + F[vorticity][(long long)v1 * L + k];
        const double p2 = m.fVertex[v2] + F[vorticity][(long long)v2 * L + k];
        const double gpt = (p2 - p1) * m.invDvEdge[e];
        const double dgpt = (dp2 - dp1) * m.invDvEdge[e];
        const double pc1 = S.pvc[(long long)c1 * L + k], pc2 = S.pvc[(long long)c2 * L + k];
        const double gpn = (pc2 - pc1) * m.invDcEdge[e];
        const double dgpn = (S.d_pvc[(long long)c2 * L + k] - S.d_pvc[(long long)c1 * L + k]) * m.invDcEdge[e];
        dpe -= (apvm * dt) * (dvv * gpt + F[v][i] * dgpt + dot_of(T, u, i) * gpn + F[u][i] * dgpn);
    }
    dot_put(T, pv_edge, i, dpe);
}

// ===========================================================================
// vertical-implicit coefficients and Thomas factor, thread per column
// ===========================================================================
__device__ __forceinline__ void vimp_col_jvp(const mpas_hip_mesh &m, double dts, double epssm, Fwd F, Dot T, int c) {
    using namespace vimp;
    const int L = m.nLevels, Lw = L + 1;
    if (c > m.nCells) return;
    const long long o = (long long)c * L, ow = (long long)c * Lw;
    if (c == m.nCells) {
        for (int k = 0; k < L; ++k)
            for (int s : {cofwr, cofwz, cofwt, a_tri, alpha_tri, gamma_tri}) dot_put(T, s, o + k, 0.0);
        for (int k = 0; k <= L; ++k) dot_put(T, coftz, ow + k, 0.0);
        return;
    }
    if (c == 0 && T[cofrz])
        for (int k = 0; k < L; ++k) T[cofrz][k] = 0.0;
    const double dtseps = 0.5 * dts * (1.0 + epssm);
    const double rcv = MPAS_RGAS / (MPAS_CP - MPAS_RGAS);
    const double c2 = MPAS_CP * rcv;
    const double *zz = m.zz + o, *rdzw = m.rdzw, *cofrz_ = F[cofrz];
    const double *WZ = F[cofwz] + o, *TZ = F[coftz] + ow, *WT = F[cofwt] + o, *WR = F[cofwr] + o;
    const double *P = F[p] + o, *Q = F[qtot] + o, *RT = F[rt] + o, *CQ = F[cqw] + ow;
    auto dP = [&](int k) { return dot_of(T, p, o + k); };
    auto dTt = [&](int k) { return dot_of(T, t, o + k); };
    // coefficient tangents (stored in the output tangent slots, read back below)
    double *dWZ = T[cofwz] + o, *dTZ = T[coftz] + ow, *dWT = T[cofwt] + o;
    dWZ[0] = 0.0;
    dTZ[0] = 0.0;
    dTZ[L] = 0.0;
    if (T[cofwr])
        for (int k = 0; k < L; ++k) T[cofwr][o + k] = 0.0;
    for (int k = 1; k < L; ++k) {
        const double zi = m.fzm[k] * zz[k] + m.fzp[k] * zz[k - 1];
        const double pi = m.fzm[k] * P[k] + m.fzp[k] * P[k - 1];
        const double dpi = m.fzm[k] * dP(k) + m.fzp[k] * dP(k - 1);
        const double base = dtseps * c2 * zi * m.rdzu[k];
        dWZ[k] = base * (dot_of(T, cqw, ow + k) * pi + CQ[k] * dpi);
        dTZ[k] = dtseps * (m.fzm[k] * dTt(k) + m.fzp[k] * dTt(k - 1));
    }
    for (int k = 0; k < L; ++k) {
        const double A0 = 0.5 * dtseps * rcv * zz[k] * MPAS_GRAVITY * m.rho_base[o + k];
        const double q1 = 1.0 + Q[k], rr = m.rtheta_base[o + k] + RT[k], pb = m.exner_base[o + k];
        dWT[k] = A0 / q1 / (rr * pb) * dP(k) - WT[k] / q1 * dot_of(T, qtot, o + k) - WT[k] / rr * dot_of(T, rt, o + k);
    }
    const double *a_ = F[a_tri] + o, *al_ = F[alpha_tri] + o, *ga_ = F[gamma_tri] + o;
    dot_put(T, a_tri, o, 0.0);
    dot_put(T, alpha_tri, o, 0.0);
    dot_put(T, gamma_tri, o, 0.0);
    double dgprev = 0.0;
    for (int k = 1; k < L; ++k) {
        const double da = -dWZ[k] * TZ[k - 1] * rdzw[k - 1] * zz[k - 1] - WZ[k] * dTZ[k - 1] * rdzw[k - 1] * zz[k - 1] -
                          dWT[k - 1] * TZ[k - 1] * rdzw[k - 1] - WT[k - 1] * dTZ[k - 1] * rdzw[k - 1];
        const double db = dWZ[k] * (TZ[k] * rdzw[k] * zz[k] + TZ[k] * rdzw[k - 1] * zz[k - 1]) +
                          WZ[k] * dTZ[k] * (rdzw[k] * zz[k] + rdzw[k - 1] * zz[k - 1]) -
                          dTZ[k] * (WT[k] * rdzw[k] - WT[k - 1] * rdzw[k - 1]) -
                          TZ[k] * (dWT[k] * rdzw[k] - dWT[k - 1] * rdzw[k - 1]);
        const double dc = -dWZ[k] * TZ[k + 1] * rdzw[k] * zz[k] - WZ[k] * dTZ[k + 1] * rdzw[k] * zz[k] +
                          dWT[k] * TZ[k + 1] * rdzw[k] + WT[k] * dTZ[k + 1] * rdzw[k];
        const double cc = -WZ[k] * TZ[k + 1] * rdzw[k] * zz[k] - WR[k] * cofrz_[k] + WT[k] * TZ[k + 1] * rdzw[k];
        const double gprev = (k > 1) ? ga_[k - 1] : 0.0;
        const double al = al_[k];
        const double dden = db - da * gprev - a_[k] * dgprev;
        const double dal = -(al * al) * dden;
        const double dg = dc * al + cc * dal;
        dot_put(T, a_tri, o + k, da);
        dot_put(T, alpha_tri, o + k, dal);
        dot_put(T, gamma_tri, o + k, dg);
        dgprev = dg;
    }
}

// ===========================================================================
// acoustic substep: edge pass, then one thread per column.  Column scratch
// as for the adjoint: element (a, k) of column c at ((a (L+1) + k) ncol + c).
// ===========================================================================
__device__ __forceinline__ void acoustic_edge_jvp(const mpas_hip_mesh &m, const mpas_hip_acoustic_params &P, Fwd F,
                                                  Dot T, int e, int k) {
    using namespace acoustic;
    const int L = m.nLevels;
    if (e > m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    const double dts = P.dts;
    if (e == m.nEdges) {
        dot_put(T, ru_p, i, dot_of(T, ru_p_in, i));
        dot_put(T, ruAvg, i, dot_of(T, ruAvg_in, i));
        return;
    }
    if (P.small_step == 1) {
        const double d = dts * dot_of(T, tend_ru, i);
        dot_put(T, ru_p, i, d);
        dot_put(T, ruAvg, i, d);
        return;
    }
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const long long o1 = (long long)c1 * L + k, o2 = (long long)c2 * L + k;
    const double zs = 0.5 * (m.zz[o2] + m.zz[o1]);
    const double pg1 = ((F[rtheta_pp_in][o2] - F[rtheta_pp_in][o1]) * m.invDcEdge[e]) / zs;
    const double dpg1 = (dot_of(T, rtheta_pp_in, o2) - dot_of(T, rtheta_pp_in, o1)) * m.invDcEdge[e] / zs;
    const double exs = F[exner][o1] + F[exner][o2];
    const double dexs = dot_of(T, exner, o1) + dot_of(T, exner, o2);
    const double cq = F[cqu][i];
    const double dpg = dot_of(T, cqu, i) * 0.5 * P.c2 * exs * pg1 + cq * 0.5 * P.c2 * dexs * pg1 +
                       cq * 0.5 * P.c2 * exs * dpg1 +
                       0.5 * m.zxu[i] * MPAS_GRAVITY * (dot_of(T, rho_pp_in, o1) + dot_of(T, rho_pp_in, o2));
    const double drup = dot_of(T, ru_p_in, i) + dts * (dot_of(T, tend_ru, i) - dpg);
    dot_put(T, ru_p, i, drup);
    dot_put(T, ruAvg, i, dot_of(T, ruAvg_in, i) + drup);
}

__device__ __forceinline__ void acoustic_col_jvp(const mpas_hip_mesh &m, const mpas_hip_acoustic_params &P, Fwd F,
                                                 Dot T, double *scr, int c) {
    using namespace acoustic;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges, nC = m.nCells;
    if (c > nC) return;
    const long long oc = (long long)c * L, ow = (long long)c * Lw;
    if (c == nC) {
        for (int k = 0; k < L; ++k) {
            dot_put(T, rho_pp, oc + k, dot_of(T, rho_pp_in, oc + k));
            dot_put(T, rtheta_pp, oc + k, dot_of(T, rtheta_pp_in, oc + k));
            dot_put(T, rtheta_pp_old, oc + k, dot_of(T, rtheta_pp_old_in, oc + k));
        }
        for (int k = 0; k <= L; ++k) {
            dot_put(T, rw_p, ow + k, dot_of(T, rw_p_in, ow + k));
            dot_put(T, wwAvg, ow + k, dot_of(T, wwAvg_in, ow + k));
        }
        return;
    }
    const long long ncol = nC;
    auto S = [&](int a, int k) -> double & { return scr[((long long)a * Lw + k) * ncol + c]; };
    enum { RS, TS, Y, X, DRS, DTS, DY, DX };
    const double dts = P.dts, epssm = P.epssm, resm = P.resm;
    const bool first = (P.small_step == 1);
    const double *zz = m.zz + oc, *rdzw = m.rdzw, *fzm = m.fzm, *fzp = m.fzp, *cofrz_ = F[cofrz];
    const double *WZ = F[cofwz] + oc, *WR = F[cofwr] + oc, *WT = F[cofwt] + oc, *TZ = F[coftz] + ow;
    const double *a_ = F[a_tri] + oc, *al_ = F[alpha_tri] + oc, *ga_ = F[gamma_tri] + oc;
    auto RPP = [&](int k) { return first ? 0.0 : F[rho_pp_in][oc + k]; };
    auto RTPP = [&](int k) { return first ? 0.0 : F[rtheta_pp_in][oc + k]; };
    auto RWP = [&](int k) { return first ? 0.0 : F[rw_p_in][ow + k]; };
    auto WWA = [&](int k) { return first ? 0.0 : F[wwAvg_in][ow + k]; };
    auto dRPP = [&](int k) { return first ? 0.0 : dot_of(T, rho_pp_in, oc + k); };
    auto dRTPP = [&](int k) { return first ? 0.0 : dot_of(T, rtheta_pp_in, oc + k); };
    auto dRWP = [&](int k) { return first ? 0.0 : dot_of(T, rw_p_in, ow + k); };
    auto dWWA = [&](int k) { return first ? 0.0 : dot_of(T, wwAvg_in, ow + k); };
    auto dWZ = [&](int k) { return dot_of(T, cofwz, oc + k); };
    auto dWR = [&](int k) { return dot_of(T, cofwr, oc + k); };
    auto dWT = [&](int k) { return dot_of(T, cofwt, oc + k); };
    auto dTZ = [&](int k) { return dot_of(T, coftz, ow + k); };
    const int n = m.nEdgesOnCell[c];
    const double ia = m.invAreaCell[c];
    (void)WWA;
    for (int k = 0; k < L; ++k) {
        double rsum = 0.0, tsum = 0.0, drsum = 0.0, dtsum = 0.0;
        for (int i = 0; i < n; ++i) {
            const int e = m.edgesOnCell[c * ME + i];
            const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
            const double metric = m.edgesOnCell_sign[c * ME + i] * m.dvEdge[e];
            const double flux = dts * metric * F[ru_p][(long long)e * L + k] * ia;
            const double dflux = dts * metric * dot_of(T, ru_p, (long long)e * L + k) * ia;
            const double thedge = 0.5 * (F[theta_m][(long long)c2 * L + k] + F[theta_m][(long long)c1 * L + k]);
            const double dthedge =
                0.5 * (dot_of(T, theta_m, (long long)c2 * L + k) + dot_of(T, theta_m, (long long)c1 * L + k));
            const double rcon = -flux, tcon = -(flux * thedge);
            rsum = (i == 0) ? rcon : rsum + rcon;
            tsum = (i == 0) ? tcon : tsum + tcon;
            drsum += -dflux;
            dtsum += -(dflux * thedge + flux * dthedge);
        }
        S(RS, k) = RPP(k) + dts * F[tend_rho][oc + k] + rsum - cofrz_[k] * resm * (RWP(k + 1) - RWP(k));
        S(TS, k) = RTPP(k) + dts * F[tend_rt][oc + k] + tsum - resm * rdzw[k] * (TZ[k + 1] * RWP(k + 1) - TZ[k] * RWP(k));
        S(DRS, k) = dRPP(k) + dts * dot_of(T, tend_rho, oc + k) + drsum - cofrz_[k] * resm * (dRWP(k + 1) - dRWP(k));
        S(DTS, k) = dRTPP(k) + dts * dot_of(T, tend_rt, oc + k) + dtsum -
                    resm * rdzw[k] * (dTZ(k + 1) * RWP(k + 1) + TZ[k + 1] * dRWP(k + 1) - dTZ(k) * RWP(k) - TZ[k] * dRWP(k));
    }
    // Thomas solve with its tangent
    S(Y, 0) = RWP(0);
    S(DY, 0) = dRWP(0);
    for (int k = 1; k < L; ++k) {
        const double A = zz[k] * S(TS, k) - zz[k - 1] * S(TS, k - 1);
        const double dA = zz[k] * S(DTS, k) - zz[k - 1] * S(DTS, k - 1);
        const double Bz = zz[k] * RTPP(k) - zz[k - 1] * RTPP(k - 1);
        const double dBz = zz[k] * dRTPP(k) - zz[k - 1] * dRTPP(k - 1);
        const double R2 = (S(RS, k) + S(RS, k - 1)) + resm * (RPP(k) + RPP(k - 1));
        const double dR2 = (S(DRS, k) + S(DRS, k - 1)) + resm * (dRPP(k) + dRPP(k - 1));
        const double r = RWP(k) + dts * F[tend_rw][ow + k] - WZ[k] * (A + resm * Bz) - WR[k] * R2 +
                         WT[k] * (S(TS, k) + resm * RTPP(k)) + WT[k - 1] * (S(TS, k - 1) + resm * RTPP(k - 1));
        const double dr = dRWP(k) + dts * dot_of(T, tend_rw, ow + k) - dWZ(k) * (A + resm * Bz) -
                          WZ[k] * (dA + resm * dBz) - dWR(k) * R2 - WR[k] * dR2 +
                          dWT(k) * (S(TS, k) + resm * RTPP(k)) + WT[k] * (S(DTS, k) + resm * dRTPP(k)) +
                          dWT(k - 1) * (S(TS, k - 1) + resm * RTPP(k - 1)) +
                          WT[k - 1] * (S(DTS, k - 1) + resm * dRTPP(k - 1));
        const double rm = r - a_[k] * S(Y, k - 1);
        S(Y, k) = rm * al_[k];
        S(DY, k) = (dr - dot_of(T, a_tri, oc + k) * S(Y, k - 1) - a_[k] * S(DY, k - 1)) * al_[k] +
                   rm * dot_of(T, alpha_tri, oc + k);
    }
    S(X, L) = RWP(L);
    S(DX, L) = dRWP(L);
    for (int k = L - 1; k >= 0; --k) {
        S(X, k) = S(Y, k) - ga_[k] * S(X, k + 1);
        S(DX, k) = S(DY, k) - dot_of(T, gamma_tri, oc + k) * S(X, k + 1) - ga_[k] * S(DX, k + 1);
    }
    // damping and the averaged flux; final rw_p tangent kept in DX
    for (int k = 0; k <= L; ++k) {
        double dww = dWWA(k);
        if (k >= 1 && k < L) {
            dww += 0.5 * (1.0 - epssm) * dRWP(k);
            const double D = 1.0 + dts * m.dss[oc + k];
            const double zi = fzm[k] * zz[k] + fzp[k] * zz[k - 1];
            const double ri = fzm[k] * F[rho_zz][oc + k] + fzp[k] * F[rho_zz][oc + k - 1];
            const double dri = fzm[k] * dot_of(T, rho_zz, oc + k) + fzp[k] * dot_of(T, rho_zz, oc + k - 1);
            const double drwd = dot_of(T, rw_save, ow + k) - dot_of(T, rw, ow + k);
            const double drr = (S(DX, k) + drwd - dts * m.dss[oc + k] * zi *
                                                      (dri * F[w][ow + k] + ri * dot_of(T, w, ow + k))) / D - drwd;
            S(DX, k) = drr;
            dww += 0.5 * (1.0 + epssm) * drr;
        }
        dot_put(T, wwAvg, ow + k, dww);
        dot_put(T, rw_p, ow + k, S(DX, k));
    }
    for (int k = 0; k < L; ++k) {
        const double rwf1 = F[rw_p][ow + k + 1], rwf0 = F[rw_p][ow + k];
        dot_put(T, rho_pp, oc + k, S(DRS, k) - cofrz_[k] * (S(DX, k + 1) - S(DX, k)));
        dot_put(T, rtheta_pp, oc + k,
                S(DTS, k) - rdzw[k] * (dTZ(k + 1) * rwf1 + TZ[k + 1] * S(DX, k + 1) - dTZ(k) * rwf0 - TZ[k] * S(DX, k)));
        dot_put(T, rtheta_pp_old, oc + k, dRTPP(k));
    }
}

}  // namespace adj
}  // namespace mpas
