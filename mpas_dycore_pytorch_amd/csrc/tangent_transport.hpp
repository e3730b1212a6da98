// Tangent-linear models of the scalar transport: atm_advance_scalars_work
// (explicit RK stages, advance_scalars.py:128-700) and
// atm_advance_scalars_mono_work (monotonic FCT, advance_scalars_mono.py:
// 3332-4037, tile numeric :1476-1818).  Included after adjoint_transport.hpp,
// whose slot enums, forward recompute passes (mono_f1..mono_f4) and scratch
// layout they share; the g_* arrays of MonoScratch hold tangents here.
// Branches are frozen and switching points take the adjoint's conventions,
// so the tangent is the exact transpose of the adjoint.
#pragma once

namespace mpas {
namespace adj {

// tangent of the vertical high-order scalar flux wdtn at interface k
__device__ __forceinline__ double scalar_wdtn_dot(const mpas_hip_mesh &m, double coef3, const double *q,
                                                  const double *ww, Dot T, int q_slot, int ww_slot, int S, int c,
                                                  int k, int s) {
    const int L = m.nLevels, Lw = L + 1;
    if (k < 1 || k > L - 1) return 0.0;
    auto qi = [&](int kk) { return ((long long)c * L + kk) * S + s; };
    const long long iw = (long long)c * Lw + k;
    const double wk = ww[iw], dwk = dot_of(T, ww_slot, iw);
    if ((k == L - 1 && L > 2) || k == 1)
        return dwk * (m.fnm[k] * q[qi(k)] + m.fnp[k] * q[qi(k - 1)]) +
               wk * (m.fnm[k] * dot_of(T, q_slot, qi(k)) + m.fnp[k] * dot_of(T, q_slot, qi(k - 1)));
    return flux3_dot(q[qi(k - 2)], q[qi(k - 1)], q[qi(k)], q[qi(k + 1)], wk, coef3, dot_of(T, q_slot, qi(k - 2)),
                     dot_of(T, q_slot, qi(k - 1)), dot_of(T, q_slot, qi(k)), dot_of(T, q_slot, qi(k + 1)), dwk);
}
// tangent of the pattern-C edge value (sign frozen)
__device__ __forceinline__ double patternc_dot(const mpas_hip_mesh &m, int e, double sgn, Dot T, int slot, int S,
                                               int k, int s) {
    const int L = m.nLevels, MA = m.maxAdv;
    const int na = m.nAdvCellsForEdge[e];
    double f = 0.0;
    for (int j = 0; j < na; ++j) {
        const int ac = m.advCellsForEdge[e * MA + j];
        const double wgt = m.adv_coefs[e * MA + j] + sgn * m.adv_coefs_3rd[e * MA + j];
        f += wgt * dot_of(T, slot, ((long long)ac * L + k) * S + s);
    }
    return f;
}

// ===========================================================================
// explicit stage
// ===========================================================================
// pass 1, thread (edge e, level k): horizontal flux tangents of all scalars
__device__ __forceinline__ void scal_edge_jvp(const mpas_hip_mesh &m, int S, Fwd F, Dot T, int e, int k) {
    using namespace scal;
    const int L = m.nLevels;
    if (e > m.nEdges || k >= L) return;
    const long long i = (long long)e * L + k;
    if (e == m.nEdges) {
        for (int s = 0; s < S; ++s) dot_put(T, horiz_flux, i * S + s, 0.0);
        return;
    }
    const double sgn = sign_pm(F[uhAvg][i]);
    for (int s = 0; s < S; ++s) dot_put(T, horiz_flux, i * S + s, patternc_dot(m, e, sgn, T, scalar_new_in, S, k, s));
}

// pass 2, thread (cell c, level k): the update
__device__ __forceinline__ void scal_cell_jvp(const mpas_hip_mesh &m, int S, double dt, double coef3, double w_old,
                                              double w_new, Fwd F, Dot T, int c, int k) {
    using namespace scal;
    const int L = m.nLevels, ME = m.maxEdges;
    if (c > m.nCells || k >= L) return;
    const long long o = (long long)c * L + k;
    if (c == m.nCells) {
        for (int s = 0; s < S; ++s) dot_put(T, scalar_new_out, o * S + s, 0.0);
        return;
    }
    const double rzo = F[rho_zz_old][o], rzn = F[rho_zz_new][o];
    const double drzo = dot_of(T, rho_zz_old, o), drzn = dot_of(T, rho_zz_new, o);
    const double inv = 1.0 / (w_old * rzo + w_new * rzn);
    const double drho_i = w_old * drzo + w_new * drzn;
    const int n = m.nEdgesOnCell[c];
    const double ia = m.invAreaCell[c];
    for (int s = 0; s < S; ++s) {
        const long long os = o * S + s;
        double dhd = 0.0;
        for (int i = 0; i < n; ++i) {
            const int e = m.edgesOnCell[c * ME + i];
            const long long ie = (long long)e * L + k;
            dhd -= m.edgesOnCell_sign[c * ME + i] *
                   (dot_of(T, uhAvg, ie) * F[horiz_flux][ie * S + s] + F[uhAvg][ie] * dot_of(T, horiz_flux, ie * S + s));
        }
        const double dlo = scalar_wdtn_dot(m, coef3, F[scalar_new_in], F[wwAvg], T, scalar_new_in, wwAvg, S, c, k, s);
        const double dhi =
            scalar_wdtn_dot(m, coef3, F[scalar_new_in], F[wwAvg], T, scalar_new_in, wwAvg, S, c, k + 1, s);
        const double dvd = m.rdnw[k] * (dhi - dlo);
        const double dnum = dot_of(T, scalar_old, os) * rzo + F[scalar_old][os] * drzo + dt * (dhd * ia - dvd);
        dot_put(T, scalar_new_out, os, dnum * inv - F[scalar_new_out][os] * inv * drho_i);
    }
}

// ===========================================================================
// monotonic FCT stage
// ===========================================================================
// T0, thread (cell c, level k), once: rho_zz_int tangent
__device__ __forceinline__ void mono_rho_jvp(const mpas_hip_mesh &m, double dt, Dot T, int c, int k) {
    using namespace mono;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c > m.nCells || k >= L) return;
    const long long o = (long long)c * L + k, ow = (long long)c * Lw;
    if (c == m.nCells) {
        dot_put(T, rho_zz_int, o, 0.0);
        return;
    }
    double hd = 0.0;
    const int n = m.nEdgesOnCell[c];
    for (int i = 0; i < n; ++i) {
        const int e = m.edgesOnCell[c * ME + i];
        hd += -m.edgesOnCell_sign[c * ME + i] * m.dvEdge[e] * m.invAreaCell[c] * dot_of(T, uhAvg, (long long)e * L + k);
    }
    const double vd = m.rdnw[k] * (dot_of(T, wwAvg, ow + k + 1) - dot_of(T, wwAvg, ow + k));
    dot_put(T, rho_zz_int, o, dot_of(T, rho_zz_old, o) + dt * (hd - vd));
}

// tangent of the bound chain (local levels, then neighbours, then the pair)
// with the tie split of torch.maximum / torch.minimum
__device__ __forceinline__ double mono_bound_dot(const mpas_hip_mesh &m, int S, int s, const double *so, Dot T,
                                                 bool is_max, int c, int k) {
    using namespace mono;
    const int L = m.nLevels, ME = m.maxEdges;
    const long long oc = (long long)c * L;
    auto better = [&](double a, double b) { return is_max ? (a > b) : (a < b); };
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
    double wl_, wr_;
    if (n == 0) { wl_ = 1.0; wr_ = 0.0; }
    else if (is_max) max_partials(vloc, vn, wl_, wr_);
    else min_partials(vloc, vn, wl_, wr_);
    double d = 0.0;
    for (int j = 0; j < nl; ++j) d += wl_ * wloc[j] * dot_of(T, scalars_old, (oc + lk[j]) * S + s);
    for (int i = 0; i < n; ++i) {
        if (wn[i] == 0.0) continue;
        const int nb = m.cellsOnCell[c * ME + i];
        d += wr_ * wn[i] * dot_of(T, scalars_old, ((long long)nb * L + k) * S + s);
    }
    return d;
}

// M1, thread (cell c, interface k): wdtn, smax, smin tangents
__device__ __forceinline__ void mono_m1(const mpas_hip_mesh &m, int S, int s, double coef3, Fwd F, Dot T,
                                        const MonoScratch &X, int c, int k) {
    using namespace mono;
    const int L = m.nLevels, Lw = L + 1;
    if (c >= m.nCells || k > L) return;
    X.g_wdtn[(long long)c * Lw + k] =
        scalar_wdtn_dot(m, coef3, F[scalars_new_in], F[wwAvg], T, scalars_new_in, wwAvg, S, c, k, s);
    if (k >= L) return;
    const long long o = (long long)c * L + k;
    X.g_smax[o] = mono_bound_dot(m, S, s, F[scalars_old], T, true, c, k);
    X.g_smin[o] = mono_bound_dot(m, S, s, F[scalars_old], T, false, c, k);
}

// M2, thread (edge e, level k): upwind and antidiffusive flux tangents
__device__ __forceinline__ void mono_m2(const mpas_hip_mesh &m, int S, int s, double dt, Fwd F, Dot T,
                                        const MonoScratch &X, int e, int k) {
    using namespace mono;
    const int L = m.nLevels;
    if (e >= m.nEdges || k >= L) return;
    const long long o = (long long)e * L + k;
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const double uh = F[uhAvg][o], duh = dot_of(T, uhAvg, o);
    const double sgn = sign_pm(uh);
    const double hv = patternc(m, e, sgn, F[scalars_new_in], S, k, s);
    const double dhv = patternc_dot(m, e, sgn, T, scalars_new_in, S, k, s);
    const long long osel = (uh >= 0.0) ? (long long)c1 * L + k : (long long)c2 * L + k;
    const double up = mono_so(F[scalars_old], S, s, osel), dup = dot_of(T, scalars_old, osel * S + s);
    const double dfu = m.dvEdge[e] * dt * (duh * up + uh * dup);
    X.g_fup[o] = dfu;
    X.g_ftmp[o] = dt * (duh * hv + uh * dhv) - dfu;
}

// M3, thread (cell c, interface k): upwind and antidiffusive vertical flux tangents
__device__ __forceinline__ void mono_m3(const mpas_hip_mesh &m, int S, int s, double dt, Fwd F, Dot T,
                                        const MonoScratch &X, int c, int k) {
    using namespace mono;
    const int L = m.nLevels, Lw = L + 1;
    if (c >= m.nCells || k < 1 || k > L - 1) return;
    const long long oc = (long long)c * L, iw = (long long)c * Lw + k;
    const double vel = F[wwAvg][iw], dvel = dot_of(T, wwAvg, iw);
    const long long osel = (vel >= 0.0) ? oc + k - 1 : oc + k;
    const double val = mono_so(F[scalars_old], S, s, osel), dval = dot_of(T, scalars_old, osel * S + s);
    const double dwu = dt * (dvel * val + vel * dval);
    X.g_wup[iw] = dwu;
    X.g_wmid[iw] = dt * X.g_wdtn[iw] - dwu;
}

// M4, thread (cell c, level k): upwind update, in/out sums and scale factor tangents
__device__ __forceinline__ void mono_m4(const mpas_hip_mesh &m, int S, int s, Fwd F, Dot T, const MonoScratch &X,
                                        int c, int k) {
    using namespace mono;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c >= m.nCells || k >= L) return;
    const long long o = (long long)c * L + k, ow = (long long)c * Lw;
    auto pd = [](double x) { return (x >= 0.0) ? 1.0 : 0.0; };  // d max0(x)
    auto nd = [](double x) { return (x <= 0.0) ? 1.0 : 0.0; };  // d (-max0(-x))
    double dsr = dot_of(T, scalars_old, o * S + s) * F[rho_zz_old][o] +
                 mono_so(F[scalars_old], S, s, o) * dot_of(T, rho_zz_old, o);
    if (k < L - 1) dsr -= X.g_wup[ow + k + 1] * m.rdnw[k];
    if (k >= 1) dsr += X.g_wup[ow + k] * m.rdnw[k];
    double dsin = 0.0, dsout = 0.0;
    if (k < L - 1) {
        const double next = X.wmid[ow + k + 1], dnext = X.g_wmid[ow + k + 1];
        dsin += -m.rdnw[k] * nd(next) * dnext;
        dsout += -m.rdnw[k] * pd(next) * dnext;
    }
    if (k >= 1) {
        const double prev = X.wmid[ow + k], dprev = X.g_wmid[ow + k];
        dsin += m.rdnw[k] * pd(prev) * dprev;
        dsout += m.rdnw[k] * nd(prev) * dprev;
    }
    const int n = m.nEdgesOnCell[c];
    const double ia = m.invAreaCell[c];
    for (int i = 0; i < n; ++i) {
        const int e = m.edgesOnCell[c * ME + i];
        const double sgn = m.edgesOnCell_sign[c * ME + i];
        const long long oe = (long long)e * L + k;
        const double sf = sgn * X.ftmp[oe], dsf = sgn * X.g_ftmp[oe];
        dsr -= X.g_fup[oe] * (sgn * ia);
        dsout -= pd(sf) * dsf * ia;
        dsin -= nd(sf) * dsf * ia;
    }
    X.g_srf[o] = dsr;
    const double rzi = F[rho_zz_int][o], drzi = dot_of(T, rho_zz_int, o);
    const double sr = X.srp[o];
    {
        const double den = X.sinr[o] + 1.0e-20;
        const double fi = (X.smax[o] * rzi - sr) / den;
        const double dfi = (X.g_smax[o] * rzi + X.smax[o] * drzi - dsr) / den - fi / den * dsin;
        X.g_sinl[o] = (fi >= 0.0 && fi <= 1.0) ? dfi : 0.0;
    }
    {
        const double den = X.soutr[o] - 1.0e-20;
        const double fo = (X.smin[o] * rzi - sr) / den;
        const double dfo = (X.g_smin[o] * rzi + X.smin[o] * drzi - dsr) / den - fo / den * dsout;
        X.g_soutl[o] = (fo >= 0.0 && fo <= 1.0) ? dfo : 0.0;
    }
}

// M5, thread (cell c, level k): limited fluxes, final update and positivity
__device__ __forceinline__ void mono_m5(const mpas_hip_mesh &m, int S, int s, Fwd F, Dot T, const MonoScratch &X,
                                        int c, int k) {
    using namespace mono;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    if (c > m.nCells || k >= L) return;
    const long long o = (long long)c * L + k, ow = (long long)c * Lw;
    if (c == m.nCells) {
        dot_put(T, scalars_new_out, o * S + s, 0.0);
        return;
    }
    // tangent of a limiter min over (a, b) with the tie split
    auto lim_dot = [&](long long oa, const double *va, const double *dva, long long ob, const double *vb,
                       const double *dvb) { return min_dot(va[oa], vb[ob], dva[oa], dvb[ob]); };
    double sr = X.srp[o], dsr = X.g_srf[o];
    const int n = m.nEdgesOnCell[c];
    for (int i = 0; i < n; ++i) {
        const int e = m.edgesOnCell[c * ME + i];
        const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
        const long long oe = (long long)e * L + k, o1 = (long long)c1 * L + k, o2 = (long long)c2 * L + k;
        const double f = X.ftmp[oe], df = X.g_ftmp[oe];
        const double lim = mono_edge_lim(X, f, o1, o2);
        const double dlim = (f >= 0.0) ? lim_dot(o1, X.soutl, X.g_soutl, o2, X.sinl, X.g_sinl)
                                       : lim_dot(o1, X.sinl, X.g_sinl, o2, X.soutl, X.g_soutl);
        const double si = m.edgesOnCell_sign[c * ME + i] * m.invAreaCell[c];
        sr = sr - (f * lim) * si;
        dsr -= (df * lim + f * dlim) * si;
    }
    auto wl = [&](int kk, double &d) -> double {  // limited vertical flux at interface kk and its tangent
        d = 0.0;
        if (kk <= 0 || kk >= L) return 0.0;
        const double f = X.wmid[ow + kk], df = X.g_wmid[ow + kk];
        const long long olo = (long long)c * L + kk - 1, ohi = (long long)c * L + kk;
        const double lim = mono_vert_lim(X, f, olo, ohi);
        const double dlim = (f >= 0.0) ? lim_dot(olo, X.soutl, X.g_soutl, ohi, X.sinl, X.g_sinl)
                                       : lim_dot(ohi, X.soutl, X.g_soutl, olo, X.sinl, X.g_sinl);
        d = df * lim + f * dlim;
        return f * lim;
    };
    double dhi, dlo;
    const double delta = wl(k + 1, dhi) - wl(k, dlo);
    const double rzi = F[rho_zz_int][o];
    const double st2 = (sr - delta * m.rdnw[k]) / rzi;
    const double dst2 = (dsr - (dhi - dlo) * m.rdnw[k]) / rzi - st2 / rzi * dot_of(T, rho_zz_int, o);
    dot_put(T, scalars_new_out, o * S + s, (st2 >= 0.0) ? dst2 : 0.0);
}

}  // namespace adj
}  // namespace mpas
