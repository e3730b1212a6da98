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
__device__ __forceinline__ void patternc_adj(const mpas_hip_mesh &m, int e, double sgn, Bar B, int slot, int S, int k,
                                             int s, double g) {
    const int L = m.nLevels, MA = m.maxAdv;
    const int na = m.nAdvCellsForEdge[e];
    for (int j = 0; j < na; ++j) {
        const int ac = m.advCellsForEdge[e * MA + j];
        const double wgt = m.adv_coefs[e * MA + j] + sgn * m.adv_coefs_3rd[e * MA + j];
        add(B, slot, ((long long)ac * L + k) * S + s, wgt * g);
    }
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
__device__ __forceinline__ void scalar_wdtn_adj(const mpas_hip_mesh &m, double coef3, const double *q,
                                                const double *ww, Bar B, int q_slot, int ww_slot, int S, int c, int k,
                                                int s, double g) {
    const int L = m.nLevels, Lw = L + 1;
    if (g == 0.0 || k < 1 || k > L - 1) return;
    auto qi = [&](int kk) { return ((long long)c * L + kk) * S + s; };
    const long long iw = (long long)c * Lw + k;
    const double wk = ww[iw];
    if ((k == L - 1 && L > 2) || k == 1) {
        add(B, ww_slot, iw, (m.fnm[k] * q[qi(k)] + m.fnp[k] * q[qi(k - 1)]) * g);
        add(B, q_slot, qi(k), wk * m.fnm[k] * g);
        add(B, q_slot, qi(k - 1), wk * m.fnp[k] * g);
        return;
    }
    double d[5];
    flux3_partials(q[qi(k - 2)], q[qi(k - 1)], q[qi(k)], q[qi(k + 1)], wk, coef3, d);
    add(B, q_slot, qi(k - 2), d[0] * g);
    add(B, q_slot, qi(k - 1), d[1] * g);
    add(B, q_slot, qi(k), d[2] * g);
    add(B, q_slot, qi(k + 1), d[3] * g);
    add(B, ww_slot, iw, d[4] * g);
}

// ===========================================================================
// explicit stage
// ===========================================================================
namespace scal {
enum { scalar_old, scalar_new_in, rho_zz_old, rho_zz_new, uhAvg, wwAvg, scalar_new_out, horiz_flux, N };
}
struct ScalAdjScratch {
    double *g_tc;  // adjoint of the horizontal flux divergence, (C, S)
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
    const double sgn = sign_pm(uh);
    double guh = 0.0;
    for (int s = 0; s < S; ++s) {
        const double gp = -(s1 * X.g_tc[((long long)c1 * L + k) * S + s] + s2 * X.g_tc[((long long)c2 * L + k) * S + s]);
        const double hf = F[horiz_flux][i * S + s];
        guh += hf * gp;
        const double ghf = uh * gp + bar_of(B, horiz_flux, i * S + s);
        patternc_adj(m, e, sgn, B, scalar_new_in, S, k, s, ghf);
    }
    add_own(B, uhAvg, i, guh);
}

// pass 3, thread (cell c, interface k): vertical flux
__device__ __forceinline__ void scal_vert_vjp(const mpas_hip_mesh &m, int S, double dt, double coef3, double w_old,
                                              double w_new, Fwd F, Bar B, int c, int k) {
    using namespace scal;
    const int L = m.nLevels;
    if (c >= m.nCells || k < 1 || k > L - 1) return;
    const double rho_lo = w_old * F[rho_zz_old][(long long)c * L + k - 1] + w_new * F[rho_zz_new][(long long)c * L + k - 1];
    const double rho_hi = w_old * F[rho_zz_old][(long long)c * L + k] + w_new * F[rho_zz_new][(long long)c * L + k];
    for (int s = 0; s < S; ++s) {
        // wdtn(k) enters vd(k-1) with +rdnw(k-1) and vd(k) with -rdnw(k); g_vd = -dt g_num
        const double gn_lo = bar_of(B, scalar_new_out, ((long long)c * L + k - 1) * S + s) / rho_lo;
        const double gn_hi = bar_of(B, scalar_new_out, ((long long)c * L + k) * S + s) / rho_hi;
        const double g = m.rdnw[k - 1] * (-dt * gn_lo) - m.rdnw[k] * (-dt * gn_hi);
        scalar_wdtn_adj(m, coef3, F[scalar_new_in], F[wwAvg], B, scalar_new_in, wwAvg, S, c, k, s, g);
    }
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

// limiter selections (with the torch.minimum tie split)
__device__ __forceinline__ void mono_edge_limit(const MonoScratch &X, double f, long long o1, long long o2,
                                                double g_lim) {
    double da, db;
    if (f >= 0.0) {
        min_partials(X.soutl[o1], X.sinl[o2], da, db);
        acc(X.g_soutl + o1, da * g_lim);
        acc(X.g_sinl + o2, db * g_lim);
    } else {
        min_partials(X.sinl[o1], X.soutl[o2], da, db);
        acc(X.g_sinl + o1, da * g_lim);
        acc(X.g_soutl + o2, db * g_lim);
    }
}
__device__ __forceinline__ double mono_edge_lim(const MonoScratch &X, double f, long long o1, long long o2) {
    return (f >= 0.0) ? dmin(X.soutl[o1], X.sinl[o2]) : dmin(X.sinl[o1], X.soutl[o2]);
}
// vertical interface k: lower cell level k-1, upper level k
__device__ __forceinline__ double mono_vert_lim(const MonoScratch &X, double f, long long olo, long long ohi) {
    return (f >= 0.0) ? dmin(X.soutl[olo], X.sinl[ohi]) : dmin(X.soutl[ohi], X.sinl[olo]);
}

// R1, thread (cell c, level k): positivity and final update -> g_srf, g_wmid
// (limited), rho_zz_int adjoint
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
    const double gsr = g / rzi;
    acc(B[rho_zz_int] + o, -(st2 / rzi) * g);
    X.g_srf[o] = gsr;
    const double gdelta = -m.rdnw[k] * gsr;
    if (k + 1 <= L - 1) acc(X.g_wmid + ow + k + 1, gdelta);  // adjoint of the limited flux (k+1)
    if (k >= 1) acc(X.g_wmid + ow + k, -gdelta);
}

// R2, thread (cell c, interface k): limited vertical flux -> antidiffusive flux and scales
// (g_wmid holds the adjoint of the limited flux on entry, of the unlimited one on exit)
__device__ __forceinline__ void mono_r2(const mpas_hip_mesh &m, const MonoScratch &X, int c, int k) {
    const int L = m.nLevels, Lw = L + 1;
    if (c >= m.nCells || k < 1 || k > L - 1) return;
    const long long iw = (long long)c * Lw + k;
    const double gl = X.g_wmid[iw];
    const double f = X.wmid[iw];
    const long long olo = (long long)c * L + k - 1, ohi = (long long)c * L + k;
    const double lim = mono_vert_lim(X, f, olo, ohi);
    X.g_wmid[iw] = lim * gl;
    const double glim = f * gl;
    double da, db;
    if (f >= 0.0) {
        min_partials(X.soutl[olo], X.sinl[ohi], da, db);
        acc(X.g_soutl + olo, da * glim);
        acc(X.g_sinl + ohi, db * glim);
    } else {
        min_partials(X.soutl[ohi], X.sinl[olo], da, db);
        acc(X.g_soutl + ohi, da * glim);
        acc(X.g_sinl + olo, db * glim);
    }
}

// R3, thread (edge e, level k): limited horizontal flux -> antidiffusive flux and scales
__device__ __forceinline__ void mono_r3(const mpas_hip_mesh &m, const MonoScratch &X, int e, int k) {
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    const long long o = (long long)e * L + k;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const long long o1 = (long long)c1 * L + k, o2 = (long long)c2 * L + k;
    const double s1 = edge_sign_in_cell(m, c1, e) * m.invAreaCell[c1];
    const double s2 = edge_sign_in_cell(m, c2, e) * m.invAreaCell[c2];
    const double gl = -(s1 * X.g_srf[o1] + s2 * X.g_srf[o2]);
    const double f = X.ftmp[o];
    const double lim = mono_edge_lim(X, f, o1, o2);
    acc(X.g_ftmp + o, lim * gl);
    mono_edge_limit(X, f, o1, o2, f * gl);
}

// R4, thread (cell c, level k): scale factors -> bounds, rho_zz_int, sums;
// sums -> antidiffusive fluxes; upwind update -> old scalar, rho_zz_old
__device__ __forceinline__ void mono_r4(const mpas_hip_mesh &m, int S, int s, Fwd F, Bar B, const MonoScratch &X,
                                        int c, int k) {
    using namespace mono;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const long long o = (long long)c * L + k, ow = (long long)c * Lw;
    const double rzi = F[rho_zz_int][o];
    const double sr = X.srp[o];
    double gsr = X.g_srf[o];
    double grzi = 0.0;
    {
        const double den = X.sinr[o] + 1.0e-20;
        const double fi = (X.smax[o] * rzi - sr) / den;
        const double g = (fi >= 0.0 && fi <= 1.0) ? X.g_sinl[o] : 0.0;
        X.g_smax[o] = rzi * g / den;
        grzi += X.smax[o] * g / den;
        gsr += -g / den;
        X.g_sinr[o] = -(fi / den) * g;
    }
    {
        const double den = X.soutr[o] - 1.0e-20;
        const double fo = (X.smin[o] * rzi - sr) / den;
        const double g = (fo >= 0.0 && fo <= 1.0) ? X.g_soutl[o] : 0.0;
        X.g_smin[o] = rzi * g / den;
        grzi += X.smin[o] * g / den;
        gsr += -g / den;
        X.g_soutr[o] = -(fo / den) * g;
    }
    acc(B[rho_zz_int] + o, grzi);
    // vertical part of the sums: next = wmid(k+1), prev = wmid(k)
    const double gin = X.g_sinr[o], gout = X.g_soutr[o];
    auto pd = [](double x) { return (x >= 0.0) ? 1.0 : 0.0; };  // d max0(x)
    auto nd = [](double x) { return (x <= 0.0) ? 1.0 : 0.0; };  // d (-max0(-x))
    if (k < L - 1) {
        const double next = X.wmid[ow + k + 1];
        acc(X.g_wmid + ow + k + 1, -m.rdnw[k] * (nd(next) * gin + pd(next) * gout));
    }
    if (k >= 1) {
        const double prev = X.wmid[ow + k];
        acc(X.g_wmid + ow + k, m.rdnw[k] * (pd(prev) * gin + nd(prev) * gout));
    }
    // horizontal part of the sums and the upwind update
    const int n = m.nEdgesOnCell[c];
    const double ia = m.invAreaCell[c];
    for (int i = 0; i < n; ++i) {
        const int e = m.edgesOnCell[c * ME + i];
        const double sgn = m.edgesOnCell_sign[c * ME + i];
        const long long oe = (long long)e * L + k;
        const double sf = sgn * X.ftmp[oe];
        acc(X.g_ftmp + oe, sgn * (-ia) * (pd(sf) * gout + nd(sf) * gin));
        acc(X.g_fup + oe, -(sgn * ia) * gsr);
    }
    // sr = so rzo - wup(k+1) rdnw (k < L-1) + wup(k) rdnw (k >= 1)
    const double so = mono_so(F[scalars_old], S, s, o);
    add(B, scalars_old, o * S + s, F[rho_zz_old][o] * gsr);
    add(B, rho_zz_old, o, so * gsr);
    if (k < L - 1) acc(X.g_wup + ow + k + 1, -m.rdnw[k] * gsr);
    if (k >= 1) acc(X.g_wup + ow + k, m.rdnw[k] * gsr);
}

// R5, thread (cell c, level k): bounds -> old scalar (torch.maximum/minimum
// chains: local vertical chain, then the ordered neighbour chain, then the pair)
__device__ __forceinline__ void mono_r5(const mpas_hip_mesh &m, int S, int s, Fwd F, Bar B, const MonoScratch &X,
                                        int c, int k) {
    using namespace mono;
    const int L = m.nLevels, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const long long oc = (long long)c * L, o = oc + k;
    const double *so = F[scalars_old];
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
        const int n = m.nEdgesOnCell[c];
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
            if (wloc[j] != 0.0) add(B, scalars_old, (oc + lk[j]) * S + s, wl_ * wloc[j] * g);
        for (int i = 0; i < n; ++i) {
            if (wn[i] == 0.0) continue;
            const int nb = m.cellsOnCell[c * ME + i];
            add(B, scalars_old, ((long long)nb * L + k) * S + s, wr_ * wn[i] * g);
        }
    }
}

// R6, thread (cell c, interface k): vertical fluxes -> wdtn, wwAvg, old scalar
__device__ __forceinline__ void mono_r6(const mpas_hip_mesh &m, int S, int s, double dt, double coef3, Fwd F, Bar B,
                                        const MonoScratch &X, int c, int k) {
    using namespace mono;
    const int L = m.nLevels, Lw = L + 1;
    if (c >= m.nCells || k < 1 || k > L - 1) return;
    const long long oc = (long long)c * L, iw = (long long)c * Lw + k;
    // wmid = dt wdtn - wup; wup = dt vel val
    const double gm = X.g_wmid[iw];
    const double gwu = X.g_wup[iw] - gm;
    const double gw = dt * gm;
    const double vel = F[wwAvg][iw];
    const long long osel = (vel >= 0.0) ? oc + k - 1 : oc + k;
    const double val = mono_so(F[scalars_old], S, s, osel);
    add(B, wwAvg, iw, dt * val * gwu);
    add(B, scalars_old, osel * S + s, dt * vel * gwu);
    scalar_wdtn_adj(m, coef3, F[scalars_new_in], F[wwAvg], B, scalars_new_in, wwAvg, S, c, k, s, gw);
}

// R7, thread (edge e, level k): antidiffusive / upwind / high-order horizontal fluxes
__device__ __forceinline__ void mono_r7(const mpas_hip_mesh &m, int S, int s, double dt, Fwd F, Bar B,
                                        const MonoScratch &X, int e, int k) {
    using namespace mono;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    const long long o = (long long)e * L + k;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const double gt = X.g_ftmp[o];
    const double gfh = dt * gt;
    const double gfu = X.g_fup[o] - gt;
    const double uh = F[uhAvg][o];
    const double sgn = sign_pm(uh);
    const double hv = patternc(m, e, sgn, F[scalars_new_in], S, k, s);
    const long long osel = (uh >= 0.0) ? (long long)c1 * L + k : (long long)c2 * L + k;
    const double up = mono_so(F[scalars_old], S, s, osel);
    const double c = m.dvEdge[e] * dt;
    add_own(B, uhAvg, o, hv * gfh + c * up * gfu);
    patternc_adj(m, e, sgn, B, scalars_new_in, S, k, s, uh * gfh);
    add(B, scalars_old, osel * S + s, c * uh * gfu);
}

// R8, thread (cell c, level k), after all scalars: rho_zz_int -> rho_zz_old, uhAvg, wwAvg
__device__ __forceinline__ void mono_r8(const mpas_hip_mesh &m, double dt, Fwd F, Bar B, int c, int k) {
    using namespace mono;
    (void)F;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const long long o = (long long)c * L + k, ow = (long long)c * Lw;
    const double g = B[rho_zz_int][o];
    add(B, rho_zz_old, o, g);
    const double gd = dt * g;
    const int n = m.nEdgesOnCell[c];
    for (int i = 0; i < n; ++i) {
        const int e = m.edgesOnCell[c * ME + i];
        const double metric = -m.edgesOnCell_sign[c * ME + i] * m.dvEdge[e];
        add(B, uhAvg, (long long)e * L + k, metric * m.invAreaCell[c] * gd);
    }
    add(B, wwAvg, ow + k + 1, -m.rdnw[k] * gd);
    add(B, wwAvg, ow + k, m.rdnw[k] * gd);
}

}  // namespace adj
}  // namespace mpas
