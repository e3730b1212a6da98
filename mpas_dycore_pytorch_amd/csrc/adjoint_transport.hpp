// Adjoints of the scalar transport: atm_advance_scalars_work (explicit RK
// stages, advance_scalars.py:128-700) and atm_advance_scalars_mono_work
// (monotonic FCT, advance_scalars_mono.py:3332-4037 with the tile numeric
// :1476-1818).  Included after adjoint_bodies.hpp.
//
// The FCT adjoint freezes every branch the forward took (upwind selections,
// flux signs, the limiter's min selections, the clamps) and differentiates
// the hard primal with torch's conventions as the reference's primitives
// define them (soft_limiter.py:72-186): clamp and clamp_min pass the adjoint
// where the argument lies inside the closed interval, minimum / maximum split
// it on a tie, upwind selections are where-selects.
#pragma once

namespace mpas {
namespace adj {

// pattern-C high-order edge value (pattern_c.py:582-733) of scalar s
__device__ __forceinline__ double patternc(const mpas_hip_mesh &m, int e, double sgn, const double *q, int S, int k,
                                           int s) {
    const int L = m.nLevels, MA = m.maxAdv;
    const int na = m.nAdvCellsForEdge[e];
    double f = 0.0;
    for (int j = 0; j < na; ++j) {
        const int ac = m.advCellsForEdge[e * MA + j];
        const double wgt = m.adv_coefs[e * MA + j] + sgn * m.adv_coefs_3rd[e * MA + j];
        const double cc = wgt * q[((long long)ac * L + k) * S + s];
        f = (j == 0) ? cc : f + cc;
    }
    return f;
}
// pattern-C transpose at cell c, level k: the sum over the (edge, slot)
// pairs of the advection stencils naming c, in (edge, slot) order, of the
// slot weight (upwind choice from uh(e, k)) times the edge value's adjoint
// gv[(e L + k) S + s]
__device__ __forceinline__ double patternc_gather(const mpas_hip_mesh &m, const mpas_hip_inverse_tables &inv,
                                                  const double *uh, const double *gv, int S, int c, int k, int s) {
    const int L = m.nLevels, MA = m.maxAdv;
    double g = 0.0;
    const int j1 = inv.advStart[c + 1];
    for (int j = inv.advStart[c]; j < j1; ++j) {
        const int e = inv.advEdge[j], a = inv.advSlot[j];
        const long long i = (long long)e * L + k;
        const double wgt = m.adv_coefs[e * MA + a] + sign_pm(uh[i]) * m.adv_coefs_3rd[e * MA + a];
        g += wgt * gv[i * S + s];
    }
    return g;
}

// vertical high-order scalar flux wdtn at interface k (0 at k = 0 and L)
__device__ __forceinline__ double scalar_wdtn(const mpas_hip_mesh &m, double coef3, const double *q, const double *ww,
                                              int S, int c, int k, int s) {
    const int L = m.nLevels;
    auto Q = [&](int kk) { return q[((long long)c * L + kk) * S + s]; };
    const double *wc = ww + (long long)c * (L + 1);
    if (k == L - 1 && L > 2) return wc[k] * (m.fnm[k] * Q(k) + m.fnp[k] * Q(k - 1));
    if (k == 1) return wc[1] * (m.fnm[1] * Q(1) + m.fnp[1] * Q(0));
    if (k >= 2 && k <= L - 2) return flux3(Q(k - 2), Q(k - 1), Q(k), Q(k + 1), wc[k], coef3);
    return 0.0;
}
// Adjoint of wdtn at interface k (1..L-1) for adjoint g, as shares: sq[0..3]
// for the scalar at levels k-2..k+1 and gww for wwAvg(k)
__device__ __forceinline__ void scalar_wdtn_shares(const mpas_hip_mesh &m, double coef3, const double *q,
                                                   const double *ww, int S, int c, int k, int s, double g,
                                                   double sq[4], double &gww) {
    const int L = m.nLevels, Lw = L + 1;
    sq[0] = sq[1] = sq[2] = sq[3] = 0.0;
    gww = 0.0;
    if (g == 0.0 || k < 1 || k > L - 1) return;
    auto qi = [&](int kk) { return ((long long)c * L + kk) * S + s; };
    const double wk = ww[(long long)c * Lw + k];
    if ((k == L - 1 && L > 2) || k == 1) {
        gww = (m.fnm[k] * q[qi(k)] + m.fnp[k] * q[qi(k - 1)]) * g;
        sq[2] = wk * m.fnm[k] * g;
        sq[1] = wk * m.fnp[k] * g;
        return;
    }
    double d[5];
    flux3_partials(q[qi(k - 2)], q[qi(k - 1)], q[qi(k)], q[qi(k + 1)], wk, coef3, d);
    sq[0] = d[0] * g;
    sq[1] = d[1] * g;
    sq[2] = d[2] * g;
    sq[3] = d[3] * g;
    gww = d[4] * g;
}
// Level k's sum of the wdtn shares of interfaces k + 2, k + 1, k and k - 1,
// exchanged across the column's lanes (every lane must run it).
__device__ __forceinline__ double wdtn_shares_at(const double sq[4], int k, int L) {
    return sq[2] + from_lane(sq[1], 1, k, L) + from_lane(sq[0], 2, k, L) + from_lane(sq[3], -1, k, L);
}

// ===========================================================================
// explicit stage
// ===========================================================================
namespace scal {
enum { scalar_old, scalar_new_in, rho_zz_old, rho_zz_new, uhAvg, wwAvg, scalar_new_out, horiz_flux, N };
}
struct ScalAdjScratch {
    double *g_tc;  // adjoint of the horizontal flux divergence, (C, S)
    double *g_hf;  // adjoint of the pattern-C edge values, (E, S)
};

// pass 1, thread (cell c, level k): final update
__device__ __forceinline__ void scal_cell_vjp(const mpas_hip_mesh &m, int S, double dt, double coef3, double w_old,
                                              double w_new, Fwd F, Bar B, const ScalAdjScratch &X, int c, int k) {
    using namespace scal;
    const int L = m.nLevels, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const long long o = (long long)c * L + k;
    const double rzo = F[rho_zz_old][o], rzn = F[rho_zz_new][o];
    const double rho_i = w_old * rzo + w_new * rzn;
    const double inv = 1.0 / rho_i;
    const int n = m.nEdgesOnCell[c];
    const double ia = m.invAreaCell[c];
    double g_rho_i = 0.0, g_rzo = 0.0;
    for (int s = 0; s < S; ++s) {
        const long long os = o * S + s;
        const double g = bar_of(B, scalar_new_out, os);
        // recompute the numerator
        double hd = 0.0;
        for (int i = 0; i < n; ++i) {
            const int e = m.edgesOnCell[c * ME + i];
            const double cc = -m.edgesOnCell_sign[c * ME + i] * F[uhAvg][(long long)e * L + k] *
                              F[horiz_flux][((long long)e * L + k) * S + s];
            hd = (i == 0) ? cc : hd + cc;
        }
        const double tc = hd * ia + 0.0;
        const double wlo = scalar_wdtn(m, coef3, F[scalar_new_in], F[wwAvg], S, c, k, s);
        const double whi = scalar_wdtn(m, coef3, F[scalar_new_in], F[wwAvg], S, c, k + 1, s);
        const double vd = m.rdnw[k] * (whi - wlo);
        const double so = F[scalar_old][os];
        const double num = so * rzo + dt * (tc - vd);
        const double gnum = inv * g;
        g_rho_i += -(inv * inv) * num * g;
        add_own(B, scalar_old, os, rzo * gnum);
        g_rzo += so * gnum;
        X.g_tc[os] = dt * gnum;
    }
    add_own(B, rho_zz_old, o, g_rzo + w_old * g_rho_i);
    add_own(B, rho_zz_new, o, w_new * g_rho_i);
}

// pass 2, thread (edge e, level k): horizontal flux
__device__ __forceinline__ void scal_edge_vjp(const mpas_hip_mesh &m, int S, Fwd F, Bar B, const ScalAdjScratch &X,
                                              int e, int k) {
    using namespace scal;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const double s1 = edge_sign_in_cell(m, c1, e) * m.invAreaCell[c1];
    const double s2 = edge_sign_in_cell(m, c2, e) * m.invAreaCell[c2];
    const double uh = F[uhAvg][i];
    double guh = 0.0;
    for (int s = 0; s < S; ++s) {
        const double gp = -(s1 * X.g_tc[((long long)c1 * L + k) * S + s] + s2 * X.g_tc[((long long)c2 * L + k) * S + s]);
        const double hf = F[horiz_flux][i * S + s];
        guh += hf * gp;
        X.g_hf[i * S + s] = uh * gp + bar_of(B, horiz_flux, i * S + s);
    }
    add_own(B, uhAvg, i, guh);
}

// pass 3, one wavefront per cell, every lane present, lane k = interface k
// and level k: the vertical-flux adjoints (shares exchanged across lanes),
// wwAvg(k), and the scalar at level k through both flux directions
__device__ __forceinline__ void scal_col_vjp(const mpas_hip_mesh &m, const mpas_hip_inverse_tables &inv, int S,
                                             double dt, double coef3, double w_old, double w_new, Fwd F, Bar B,
                                             const ScalAdjScratch &X, int c, int k) {
    using namespace scal;
    const int L = m.nLevels, Lw = L + 1;
    const bool iface = (k >= 1 && k <= L - 1);
    double rho_lo = 1.0, rho_hi = 1.0;
    if (iface) {
        rho_lo = w_old * F[rho_zz_old][(long long)c * L + k - 1] + w_new * F[rho_zz_new][(long long)c * L + k - 1];
        rho_hi = w_old * F[rho_zz_old][(long long)c * L + k] + w_new * F[rho_zz_new][(long long)c * L + k];
    }
    double gww_all = 0.0;
    for (int s = 0; s < S; ++s) {
        double g = 0.0;
        if (iface) {
            // wdtn(k) enters vd(k-1) with +rdnw(k-1) and vd(k) with -rdnw(k); g_vd = -dt g_num
            const double gn_lo = bar_of(B, scalar_new_out, ((long long)c * L + k - 1) * S + s) / rho_lo;
            const double gn_hi = bar_of(B, scalar_new_out, ((long long)c * L + k) * S + s) / rho_hi;
            g = m.rdnw[k - 1] * (-dt * gn_lo) - m.rdnw[k] * (-dt * gn_hi);
        }
        double sq[4], gww;
        scalar_wdtn_shares(m, coef3, F[scalar_new_in], F[wwAvg], S, c, k, s, g, sq, gww);
        gww_all += gww;
        const double gv = wdtn_shares_at(sq, k, L);
        if (k < L) {
            const double gh = patternc_gather(m, inv, F[uhAvg], X.g_hf, S, c, k, s);
            add_own(B, scalar_new_in, ((long long)c * L + k) * S + s, gv + gh);
        }
    }
    if (iface) add_own(B, wwAvg, (long long)c * Lw + k, gww_all);
}

// ===========================================================================
// monotonic FCT stage
// ===========================================================================
namespace mono {
enum { scalars_old, scalars_new_in, rho_zz_old, uhAvg, wwAvg, scalars_new_out, rho_zz_int, N };
}
// Per-scalar forward values and adjoints, each (N+1)-row: cell arrays (C),
// interface arrays (W), edge arrays (E).
struct MonoScratch {
    double *smax, *smin, *wdtn, *wup, *wmid, *srp, *sinr, *soutr, *sinl, *soutl;  // C or W
    double *fho, *fup, *ftmp;                                                      // E
    double *g_srf, *g_sinl, *g_soutl, *g_sinr, *g_soutr, *g_wmid, *g_wup, *g_fup, *g_ftmp, *g_smax, *g_smin;
    double *g_wdtn;
    // adjoint only: the final adjoint of the upwind sum (C); the bound
    // shares of each (cell, level) for levels k-1, k, k+1 (r5[0..2]) and for
    // the neighbour of each slot (r5[3 + slot]) (C); the edge adjoints of the
    // pattern-C value and of the upwind value (E)
    double *g_sr, *r5[3 + kMaxEdgesAdj], *g_hv, *g_up;
};

__device__ __forceinline__ double mono_so(const double *so, int S, int s, long long o) { return so[o * S + s]; }

// F1, thread (cell c, level k): local + neighbour bounds, wdtn at interface k
__device__ __forceinline__ void mono_f1(const mpas_hip_mesh &m, int S, int s, double coef3, Fwd F,
                                        const MonoScratch &X, int c, int k) {
    using namespace mono;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c >= m.nCells || k > L) return;
    const long long oc = (long long)c * L;
    X.wdtn[(long long)c * Lw + k] = scalar_wdtn(m, coef3, F[scalars_new_in], F[wwAvg], S, c, k, s);
    if (k >= L) return;
    const double *so = F[scalars_old];
    auto SO = [&](int kk) { return mono_so(so, S, s, oc + kk); };
    double mx, mn;
    if (k == 0) { mx = dmax(SO(0), SO(1)); mn = dmin(SO(0), SO(1)); }
    else if (k == L - 1) { mx = dmax(SO(k), SO(k - 1)); mn = dmin(SO(k), SO(k - 1)); }
    else { mx = dmax(dmax(SO(k - 1), SO(k)), SO(k + 1)); mn = dmin(dmin(SO(k - 1), SO(k)), SO(k + 1)); }
    const int n = m.nEdgesOnCell[c];
    for (int i = 0; i < n; ++i) {
        const int nb = m.cellsOnCell[c * ME + i];
        mx = dmax(mx, mono_so(so, S, s, (long long)nb * L + k));
        mn = dmin(mn, mono_so(so, S, s, (long long)nb * L + k));
    }
    X.smax[oc + k] = mx;
    X.smin[oc + k] = mn;
}

// F2, thread (edge e, level k): high-order, upwind and antidiffusive fluxes
__device__ __forceinline__ void mono_f2(const mpas_hip_mesh &m, int S, int s, double dt, Fwd F, const MonoScratch &X,
                                        int e, int k) {
    using namespace mono;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    const long long o = (long long)e * L + k;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const double uh = F[uhAvg][o];
    const double hv = patternc(m, e, sign_pm(uh), F[scalars_new_in], S, k, s);
    const double fh = uh * hv;
    const double up = (uh >= 0.0) ? mono_so(F[scalars_old], S, s, (long long)c1 * L + k)
                                  : mono_so(F[scalars_old], S, s, (long long)c2 * L + k);
    const double fu = m.dvEdge[e] * dt * (uh * up);
    X.fho[o] = fh;
    X.fup[o] = fu;
    X.ftmp[o] = dt * fh - fu;
}

// F3, thread (cell c, interface k): upwind and antidiffusive vertical fluxes
__device__ __forceinline__ void mono_f3(const mpas_hip_mesh &m, int S, int s, double dt, Fwd F, const MonoScratch &X,
                                        int c, int k) {
    using namespace mono;
    const int L = m.nLevels, Lw = L + 1;
    if (c >= m.nCells || k < 1 || k > L - 1) return;
    const long long oc = (long long)c * L, iw = (long long)c * Lw + k;
    const double vel = F[wwAvg][iw];
    const double val = (vel >= 0.0) ? mono_so(F[scalars_old], S, s, oc + k - 1) : mono_so(F[scalars_old], S, s, oc + k);
    const double wu = dt * (vel * val);
    X.wup[iw] = wu;
    X.wmid[iw] = dt * X.wdtn[iw] - wu;
}

// F4, thread (cell c, level k): upwind update, in/out sums, scale factors
__device__ __forceinline__ void mono_f4(const mpas_hip_mesh &m, int S, int s, Fwd F, const MonoScratch &X, int c,
                                        int k) {
    using namespace mono;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const long long o = (long long)c * L + k, ow = (long long)c * Lw;
    double sr = mono_so(F[scalars_old], S, s, o) * F[rho_zz_old][o];
    if (k < L - 1) sr = sr - X.wup[ow + k + 1] * m.rdnw[k];
    if (k >= 1) sr = sr + X.wup[ow + k] * m.rdnw[k];
    const double prev = (k == 0) ? 0.0 : X.wmid[ow + k];
    const double next = (k == L - 1) ? 0.0 : X.wmid[ow + k + 1];
    double sin_ = -m.rdnw[k] * ((-max0(-next)) - max0(prev));
    double sout = -m.rdnw[k] * (max0(next) - (-max0(-prev)));
    const int n = m.nEdgesOnCell[c];
    const double ia = m.invAreaCell[c];
    for (int i = 0; i < n; ++i) {
        const int e = m.edgesOnCell[c * ME + i];
        const double sgn = m.edgesOnCell_sign[c * ME + i];
        const long long oe = (long long)e * L + k;
        const double sc_ = X.fup[oe] * (sgn * ia);
        const double sf = sgn * X.ftmp[oe];
        sr = sr - sc_;
        sout = sout - max0(sf) * ia;
        sin_ = sin_ - (-max0(-sf)) * ia;
    }
    X.srp[o] = sr;
    X.sinr[o] = sin_;
    X.soutr[o] = sout;
    const double rzi = F[rho_zz_int][o];
    const double fi = (X.smax[o] * rzi - sr) / (sin_ + 1.0e-20);
    const double fo = (X.smin[o] * rzi - sr) / (sout - 1.0e-20);
    X.sinl[o] = dmin(dmax(fi, 0.0), 1.0);
    X.soutl[o] = dmin(dmax(fo, 0.0), 1.0);
}

__device__ __forceinline__ double mono_edge_lim(const MonoScratch &X, double f, long long o1, long long o2) {
    return (f >= 0.0) ? dmin(X.soutl[o1], X.sinl[o2]) : dmin(X.sinl[o1], X.soutl[o2]);
}
// vertical interface k: lower cell level k-1, upper level k
__device__ __forceinline__ double mono_vert_lim(const MonoScratch &X, double f, long long olo, long long ohi) {
    return (f >= 0.0) ? dmin(X.soutl[olo], X.sinl[ohi]) : dmin(X.soutl[ohi], X.sinl[olo]);
}

// ---------------------------------------------------------------------------
// Reverse passes, per scalar.  Every pass writes only its own entity's
// elements; what an entity owes its neighbours is recomputed by the
// neighbour (the limiter selections, from g_srf) or stored per entity and
// gathered (the bounds shares r5, the edge adjoints g_hv / g_up).
//   R1 cells:      positivity and the final update -> g_srf, rho_zz_int
//   R2 interfaces: adjoint of the unlimited vertical antidiffusive flux
//   R4 cells:      scale factors (their adjoints gathered from the edge and
//                  interface limiters) -> bounds, rho_zz_int, the in/out
//                  sums, the upwind update
//   R5 cells:      bounds -> shares of the old scalar
//   R6 columns:    vertical fluxes -> wwAvg, old and new scalars
//   R7 edges:      horizontal fluxes -> uhAvg; g_hv, g_up for R7g
//   R7g cells:     the pattern-C and upwind transposes, the bounds shares
// then, after every scalar, R8 (columns) and R8e (edges): rho_zz_int.
// ---------------------------------------------------------------------------

// adjoint of the limited vertical flux at interface kk (1..L-1), from R1
__device__ __forceinline__ double mono_gl_vert(const mpas_hip_mesh &m, const MonoScratch &X, long long oc, int kk) {
    // interface kk enters delta(kk-1) with + and delta(kk) with -, delta(k) through -rdnw(k) g_srf(k)
    return (-m.rdnw[kk - 1] * X.g_srf[oc + kk - 1]) + (-(-m.rdnw[kk] * X.g_srf[oc + kk]));
}
// adjoint of the limited horizontal flux of edge e at level k, from R1
__device__ __forceinline__ double mono_gl_edge(const mpas_hip_mesh &m, const MonoScratch &X, int e, int k) {
    const int L = m.nLevels;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const double s1 = edge_sign_in_cell(m, c1, e) * m.invAreaCell[c1];
    const double s2 = edge_sign_in_cell(m, c2, e) * m.invAreaCell[c2];
    return -(s1 * X.g_srf[(long long)c1 * L + k] + s2 * X.g_srf[(long long)c2 * L + k]);
}

// R1, thread (cell c, level k): positivity and final update -> g_srf, rho_zz_int
__device__ __forceinline__ void mono_r1(const mpas_hip_mesh &m, int S, int s, Fwd F, Bar B, const MonoScratch &X,
                                        int c, int k) {
    using namespace mono;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const long long o = (long long)c * L + k, ow = (long long)c * Lw;
    // recompute st2
    double sr = X.srp[o];
    const int n = m.nEdgesOnCell[c];
    for (int i = 0; i < n; ++i) {
        const int e = m.edgesOnCell[c * ME + i];
        const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
        const long long oe = (long long)e * L + k;
        const double f = X.ftmp[oe];
        const double fl = f * mono_edge_lim(X, f, (long long)c1 * L + k, (long long)c2 * L + k);
        sr = sr - fl * (m.edgesOnCell_sign[c * ME + i] * m.invAreaCell[c]);
    }
    auto wl = [&](int kk) -> double {  // limited vertical flux at interface kk
        if (kk <= 0 || kk >= L) return 0.0;
        const double f = X.wmid[ow + kk];
        return f * mono_vert_lim(X, f, (long long)c * L + kk - 1, (long long)c * L + kk);
    };
    const double delta = wl(k + 1) - wl(k);
    const double rzi = F[rho_zz_int][o];
    const double st2 = (sr - delta * m.rdnw[k]) / rzi;
    const double g = (st2 >= 0.0) ? bar_of(B, scalars_new_out, o * S + s) : 0.0;
    add_own(B, rho_zz_int, o, -(st2 / rzi) * g);
    X.g_srf[o] = g / rzi;
}

// R2, thread (cell c, interface k in 1..L-1): limited -> unlimited vertical flux
__device__ __forceinline__ void mono_r2(const mpas_hip_mesh &m, const MonoScratch &X, int c, int k) {
    const int L = m.nLevels, Lw = L + 1;
    if (c >= m.nCells || k < 1 || k > L - 1) return;
    const long long iw = (long long)c * Lw + k, oc = (long long)c * L;
    const double f = X.wmid[iw];
    X.g_wmid[iw] = mono_vert_lim(X, f, oc + k - 1, oc + k) * mono_gl_vert(m, X, oc, k);
}

// R4, thread (cell c, level k): scale factors -> bounds, rho_zz_int, sums;
// upwind update -> old scalar, rho_zz_old.  The scale factors' adjoints are
// gathered from the limiter selections min(scale(c1), scale(c2)) of the
// interfaces above and below and of the cell's edges (torch.minimum tie split).
__device__ __forceinline__ void mono_r4(const mpas_hip_mesh &m, int S, int s, Fwd F, Bar B, const MonoScratch &X,
                                        int c, int k) {
    using namespace mono;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const long long oc = (long long)c * L, o = oc + k, ow = (long long)c * Lw;
    double gsinl = 0.0, gsoutl = 0.0;
    double da, db;
    if (k + 1 <= L - 1) {  // interface k + 1: this level is the lower cell
        const double f = X.wmid[ow + k + 1];
        const double glim = f * mono_gl_vert(m, X, oc, k + 1);
        if (f >= 0.0) {
            min_partials(X.soutl[o], X.sinl[o + 1], da, db);
            gsoutl += da * glim;
        } else {
            min_partials(X.soutl[o + 1], X.sinl[o], da, db);
            gsinl += db * glim;
        }
    }
    if (k >= 1) {  // interface k: this level is the upper cell
        const double f = X.wmid[ow + k];
        const double glim = f * mono_gl_vert(m, X, oc, k);
        if (f >= 0.0) {
            min_partials(X.soutl[o - 1], X.sinl[o], da, db);
            gsinl += db * glim;
        } else {
            min_partials(X.soutl[o], X.sinl[o - 1], da, db);
            gsoutl += da * glim;
        }
    }
    const int n = m.nEdgesOnCell[c];
    for (int i = 0; i < n; ++i) {
        const int e = m.edgesOnCell[c * ME + i];
        const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
        const long long o1 = (long long)c1 * L + k, o2 = (long long)c2 * L + k;
        const double f = X.ftmp[(long long)e * L + k];
        const double glim = f * mono_gl_edge(m, X, e, k);
        if (f >= 0.0) {
            min_partials(X.soutl[o1], X.sinl[o2], da, db);
            if (c == c1) gsoutl += da * glim;
            if (c == c2) gsinl += db * glim;
        } else {
            min_partials(X.sinl[o1], X.soutl[o2], da, db);
            if (c == c1) gsinl += da * glim;
            if (c == c2) gsoutl += db * glim;
        }
    }
    const double rzi = F[rho_zz_int][o];
    const double sr = X.srp[o];
    double gsr = X.g_srf[o];
    double grzi = 0.0;
    {
        const double den = X.sinr[o] + 1.0e-20;
        const double fi = (X.smax[o] * rzi - sr) / den;
        const double g = (fi >= 0.0 && fi <= 1.0) ? gsinl : 0.0;
        X.g_smax[o] = rzi * g / den;
        grzi += X.smax[o] * g / den;
        gsr += -g / den;
        X.g_sinr[o] = -(fi / den) * g;
    }
    {
        const double den = X.soutr[o] - 1.0e-20;
        const double fo = (X.smin[o] * rzi - sr) / den;
        const double g = (fo >= 0.0 && fo <= 1.0) ? gsoutl : 0.0;
        X.g_smin[o] = rzi * g / den;
        grzi += X.smin[o] * g / den;
        gsr += -g / den;
        X.g_soutr[o] = -(fo / den) * g;
    }
    add_own(B, rho_zz_int, o, grzi);
    X.g_sr[o] = gsr;
    // sr = so rzo - wup(k+1) rdnw (k < L-1) + wup(k) rdnw (k >= 1) - sum_edges fup sgn ia
    const double so = mono_so(F[scalars_old], S, s, o);
    add_own(B, scalars_old, o * S + s, F[rho_zz_old][o] * gsr);
    add_own(B, rho_zz_old, o, so * gsr);
}

// R5, thread (cell c, level k): bounds -> shares of the old scalar
// (torch.maximum/minimum chains: local vertical chain, then the ordered
// neighbour chain, then the pair)
__device__ __forceinline__ void mono_r5(const mpas_hip_mesh &m, int S, int s, Fwd F, Bar B, const MonoScratch &X,
                                        int c, int k) {
    using namespace mono;
    (void)B;
    const int L = m.nLevels, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const long long oc = (long long)c * L, o = oc + k;
    const double *so = F[scalars_old];
    const int n = m.nEdgesOnCell[c];
    double sh_loc[3] = {0.0, 0.0, 0.0};  // levels k-1, k, k+1
    double sh_nb[kMaxEdgesAdj];
    for (int i = 0; i < kMaxEdgesAdj; ++i) sh_nb[i] = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
        const bool is_max = (pass == 0);
        const double g = is_max ? X.g_smax[o] : X.g_smin[o];
        if (g == 0.0) continue;
        auto better = [&](double a, double b) { return is_max ? (a > b) : (a < b); };
        // local chain over up to three levels
        int lk[3];
        int nl;
        if (k == 0) { lk[0] = 0; lk[1] = 1; nl = 2; }
        else if (k == L - 1) { lk[0] = k; lk[1] = k - 1; nl = 2; }
        else { lk[0] = k - 1; lk[1] = k; lk[2] = k + 1; nl = 3; }
        double wloc[3] = {1.0, 0.0, 0.0};
        double vloc = mono_so(so, S, s, oc + lk[0]);
        for (int j = 1; j < nl; ++j) {
            const double y = mono_so(so, S, s, oc + lk[j]);
            if (better(vloc, y)) continue;
            if (better(y, vloc)) {
                for (int q = 0; q < j; ++q) wloc[q] = 0.0;
                wloc[j] = 1.0;
                vloc = y;
            } else {
                for (int q = 0; q < j; ++q) wloc[q] *= 0.5;
                wloc[j] = 0.5;
            }
        }
        // neighbour chain
        double wn[kMaxEdgesAdj];
        double vn = 0.0;
        for (int i = 0; i < n; ++i) {
            const int nb = m.cellsOnCell[c * ME + i];
            const double y = mono_so(so, S, s, (long long)nb * L + k);
            if (i == 0) { wn[0] = 1.0; vn = y; continue; }
            if (better(vn, y)) { wn[i] = 0.0; continue; }
            if (better(y, vn)) {
                for (int q = 0; q < i; ++q) wn[q] = 0.0;
                wn[i] = 1.0;
                vn = y;
            } else {
                for (int q = 0; q < i; ++q) wn[q] *= 0.5;
                wn[i] = 0.5;
            }
        }
        // pair
        double wl_, wr_;
        if (n == 0) { wl_ = 1.0; wr_ = 0.0; }
        else if (is_max) max_partials(vloc, vn, wl_, wr_);
        else min_partials(vloc, vn, wl_, wr_);
        for (int j = 0; j < nl; ++j)
            if (wloc[j] != 0.0) sh_loc[lk[j] - k + 1] += wl_ * wloc[j] * g;
        for (int i = 0; i < n; ++i)
            if (wn[i] != 0.0) sh_nb[i] += wr_ * wn[i] * g;
    }
    for (int j = 0; j < 3; ++j) X.r5[j][o] = sh_loc[j];
    for (int i = 0; i < n; ++i) X.r5[3 + i][o] = sh_nb[i];
}

// R6, one wavefront per cell, every lane present, lane k = interface k and
// level k: vertical fluxes -> wdtn, wwAvg, old and new scalars.  The
// adjoints of the unlimited antidiffusive flux (from R2 and the in/out sums
// of R4) and of the upwind flux (the upwind update of R4) are formed here.
__device__ __forceinline__ void mono_r6(const mpas_hip_mesh &m, int S, int s, double dt, double coef3, Fwd F, Bar B,
                                        const MonoScratch &X, int c, int k) {
    using namespace mono;
    const int L = m.nLevels, Lw = L + 1;
    const long long oc = (long long)c * L, iw = (long long)c * Lw + k;
    auto pd = [](double x) { return (x >= 0.0) ? 1.0 : 0.0; };  // d max0(x)
    auto nd = [](double x) { return (x <= 0.0) ? 1.0 : 0.0; };  // d (-max0(-x))
    double so_m1 = 0.0, so_0 = 0.0, gww = 0.0, sq[4] = {0.0, 0.0, 0.0, 0.0};
    if (k >= 1 && k <= L - 1) {
        const double wm = X.wmid[iw];
        // sums of level k-1 (this interface is its "next") and of level k (its "prev")
        const double gm = X.g_wmid[iw] +
                          (-m.rdnw[k - 1] * (nd(wm) * X.g_sinr[oc + k - 1] + pd(wm) * X.g_soutr[oc + k - 1])) +
                          m.rdnw[k] * (pd(wm) * X.g_sinr[oc + k] + nd(wm) * X.g_soutr[oc + k]);
        const double gwup = (-m.rdnw[k - 1] * X.g_sr[oc + k - 1]) + m.rdnw[k] * X.g_sr[oc + k];
        // wmid = dt wdtn - wup; wup = dt vel val
        const double gwu = gwup - gm;
        const double vel = F[wwAvg][iw];
        const bool lo = (vel >= 0.0);
        const double val = mono_so(F[scalars_old], S, s, lo ? oc + k - 1 : oc + k);
        (lo ? so_m1 : so_0) = dt * vel * gwu;
        double gwd;
        scalar_wdtn_shares(m, coef3, F[scalars_new_in], F[wwAvg], S, c, k, s, dt * gm, sq, gwd);
        gww = dt * val * gwu + gwd;
    }
    const double g_old = so_0 + from_lane(so_m1, 1, k, L);
    const double g_new = wdtn_shares_at(sq, k, L);
    if (k < L) {
        add_own(B, scalars_old, (oc + k) * S + s, g_old);
        add_own(B, scalars_new_in, (oc + k) * S + s, g_new);
    }
    if (k >= 1 && k <= L - 1) add_own(B, wwAvg, iw, gww);
}

// R7, thread (edge e, level k): antidiffusive / upwind / high-order
// horizontal fluxes.  The antidiffusive flux's adjoint is its limited use
// (from R1) plus the in/out sums of both cells (R4); the upwind flux's is the
// upwind update of both cells (R4).
__device__ __forceinline__ void mono_r7(const mpas_hip_mesh &m, int S, int s, double dt, Fwd F, Bar B,
                                        const MonoScratch &X, int e, int k) {
    using namespace mono;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    auto pd = [](double x) { return (x >= 0.0) ? 1.0 : 0.0; };
    auto nd = [](double x) { return (x <= 0.0) ? 1.0 : 0.0; };
    const long long o = (long long)e * L + k;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const double f = X.ftmp[o];
    double gt = mono_edge_lim(X, f, (long long)c1 * L + k, (long long)c2 * L + k) * mono_gl_edge(m, X, e, k);
    double gfup = 0.0;
    for (int q = 0; q < 2; ++q) {
        const int c = m.cellsOnEdge[2 * e + q];
        if (c >= m.nCells) continue;
        const long long oc = (long long)c * L + k;
        const double sgn = edge_sign_in_cell(m, c, e), ia = m.invAreaCell[c];
        const double sf = sgn * f;
        gt += sgn * (-ia) * (pd(sf) * X.g_soutr[oc] + nd(sf) * X.g_sinr[oc]);
        gfup += -(sgn * ia) * X.g_sr[oc];
    }
    const double gfh = dt * gt;
    const double gfu = gfup - gt;
    const double uh = F[uhAvg][o];
    const double sgn = sign_pm(uh);
    const double hv = patternc(m, e, sgn, F[scalars_new_in], S, k, s);
    const long long osel = (uh >= 0.0) ? (long long)c1 * L + k : (long long)c2 * L + k;
    const double up = mono_so(F[scalars_old], S, s, osel);
    const double cdt = m.dvEdge[e] * dt;
    add_own(B, uhAvg, o, hv * gfh + cdt * up * gfu);
    X.g_hv[o] = uh * gfh;
    X.g_up[o] = cdt * uh * gfu;
}

// R7g, thread (cell c, level k): the new scalar through the pattern-C values
// of the edges naming c; the old scalar through the upwind values of the
// cell's edges and the bounds shares of this column and of the cells naming c
__device__ __forceinline__ void mono_r7g(const mpas_hip_mesh &m, const mpas_hip_inverse_tables &inv, int S, int s,
                                         Fwd F, Bar B, const MonoScratch &X, int c, int k) {
    using namespace mono;
    const int L = m.nLevels, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const long long oc = (long long)c * L, o = oc + k;
    add_own(B, scalars_new_in, o * S + s, patternc_gather(m, inv, F[uhAvg], X.g_hv, 1, c, k, 0));
    double g = X.r5[1][o];
    if (k + 1 <= L - 1) g += X.r5[0][o + 1];
    if (k >= 1) g += X.r5[2][o - 1];
    const int j1 = inv.cocStart[c + 1];
    for (int j = inv.cocStart[c]; j < j1; ++j) g += X.r5[3 + inv.cocSlot[j]][(long long)inv.cocCell[j] * L + k];
    const int n = m.nEdgesOnCell[c];
    for (int i = 0; i < n; ++i) {
        const int e = m.edgesOnCell[c * ME + i];
        const long long oe = (long long)e * L + k;
        const int up = (F[uhAvg][oe] >= 0.0) ? m.cellsOnEdge[2 * e] : m.cellsOnEdge[2 * e + 1];
        if (up == c) g += X.g_up[oe];
    }
    add_own(B, scalars_old, o * S + s, g);
}

// R8, thread (cell c, interface / level k), after all scalars:
// rho_zz_int -> rho_zz_old, wwAvg
__device__ __forceinline__ void mono_r8(const mpas_hip_mesh &m, double dt, Fwd F, Bar B, int c, int k) {
    using namespace mono;
    (void)F;
    const int L = m.nLevels, Lw = L + 1;
    if (c >= m.nCells || k > L) return;
    const long long oc = (long long)c * L;
    if (k < L) add_own(B, rho_zz_old, oc + k, B[rho_zz_int][oc + k]);
    // rho_int(k) = ... - dt rdnw(k) (ww(k+1) - ww(k))
    double gw = 0.0;
    if (k >= 1) gw += -m.rdnw[k - 1] * (dt * B[rho_zz_int][oc + k - 1]);
    if (k <= L - 1) gw += m.rdnw[k] * (dt * B[rho_zz_int][oc + k]);
    add_own(B, wwAvg, (long long)c * Lw + k, gw);
}

// R8e, thread (edge e, level k), after all scalars: rho_zz_int -> uhAvg
// (rho_int(c) = ... - dt ia sum_edges sgn dv uh)
__device__ __forceinline__ void mono_r8e(const mpas_hip_mesh &m, double dt, Bar B, int e, int k) {
    using namespace mono;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    double g = 0.0;
    for (int q = 0; q < 2; ++q) {
        const int c = m.cellsOnEdge[2 * e + q];
        if (c >= m.nCells) continue;
        const double metric = -edge_sign_in_cell(m, c, e) * m.dvEdge[e];
        g += metric * m.invAreaCell[c] * (dt * B[rho_zz_int][(long long)c * L + k]);
    }
    add_own(B, uhAvg, (long long)e * L + k, g);
}

}  // namespace adj
}  // namespace mpas
