// Split scalar transport: explicit high-order stages
// (atm_advance_scalars_work, advance_scalars.py:128-700) and the monotonic
// flux-corrected stage (atm_advance_scalars_mono_work,
// advance_scalars_mono.py:3332-4037, tile numeric :1476-1818).
// Scalars are (cells+1, levels, nScalars) with the species index innermost.
#include "common.hpp"

namespace mpas {

// Pattern-C high-order edge value (pattern_c.py:582-733), slot order
__device__ __forceinline__ double patternc_value(const mpas_hip_mesh &m, int e, double sgn,
                                                 const double *__restrict__ q, int L, int S, int k, int s) {
    const int na = m.nAdvCellsForEdge[e], MA = m.maxAdv;
    double f = 0.0;
    for (int j = 0; j < na; ++j) {
        int ac = m.advCellsForEdge[e * MA + j];
        double wgt = m.adv_coefs[e * MA + j] + sgn * m.adv_coefs_3rd[e * MA + j];
        double cc = wgt * q[((size_t)ac * L + k) * S + s];
        f = (j == 0) ? cc : f + cc;
    }
    return f;
}

// ---------------------------------------------------------------------------
// explicit stages
// ---------------------------------------------------------------------------
__global__ void k_scalar_hflux(mpas_hip_mesh m, int S, const hreal *__restrict__ uhAvg,
                               const double *__restrict__ scalar_new, hreal *__restrict__ hflux) {
    const int L = m.nLevels, nE = m.nEdges;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nE * L) return;
    int e = i / L, k = i - e * L;
    double sgn = sign_pm(uhAvg[i]);
    for (int s = 0; s < S; ++s) hflux[(size_t)i * S + s] = patternc_value(m, e, sgn, scalar_new, L, S, k, s);
}

__global__ void k_scalar_update(mpas_hip_mesh m, int S, double dt, double coef3, double w_old, double w_new,
                                const double *__restrict__ scalar_old, const double *__restrict__ sn,
                                const double *__restrict__ rho_old, const double *__restrict__ rho_new,
                                const hreal *__restrict__ uhAvg, const double *__restrict__ wwAvg,
                                const hreal *__restrict__ hflux, double *__restrict__ out) {
    const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nC * L) return;
    int c = i / L, k = i - c * L;
    const int n = m.nEdgesOnCell[c];
    const double *ww = wwAvg + c * Lw;
    for (int s = 0; s < S; ++s) {
        auto SN = [&](int kk) { return sn[((size_t)c * L + kk) * S + s]; };
        auto wdtn = [&](int kk) -> double {
            if (kk <= 0 || kk >= L) return 0.0;
            if (kk == 1) return ww[1] * (m.fnm[1] * SN(1) + m.fnp[1] * SN(0));
            if (kk == L - 1) return ww[L - 1] * (m.fnm[L - 1] * SN(L - 1) + m.fnp[L - 1] * SN(L - 2));
            return flux3(SN(kk - 2), SN(kk - 1), SN(kk), SN(kk + 1), ww[kk], coef3);
        };
        double hd = 0.0;
        for (int j = 0; j < n; ++j) {
            int e = m.edgesOnCell[c * ME + j];
            double cc = -m.edgesOnCell_sign[c * ME + j] * uhAvg[e * L + k] * hflux[((size_t)e * L + k) * S + s];
            hd = (j == 0) ? cc : hd + cc;
        }
        double tend_col = hd * m.invAreaCell[c] + 0.0;
        double vdiv;
        if (k == 0) vdiv = m.rdnw[0] * wdtn(1);
        else if (k == L - 1) vdiv = m.rdnw[L - 1] * (-wdtn(L - 1));
        else vdiv = m.rdnw[k] * (wdtn(k + 1) - wdtn(k));
        double rho_i = w_old * rho_old[i] + w_new * rho_new[i];
        double inv = 1.0 / rho_i;
        out[(size_t)i * S + s] = (scalar_old[(size_t)i * S + s] * rho_old[i] + dt * (tend_col - vdiv)) * inv;
    }
}

void launch_advance_scalars(const mpas_hip_mesh &m, int S, double dt, double coef3, double w_old, double w_new,
                            const double *scalar_old, const double *scalar_new_in, const double *rho_zz_old,
                            const double *rho_zz_new, const hreal *uhAvg, const double *wwAvg,
                            double *scalar_new_out, hreal *horiz_flux, hipStream_t s) {
    const int L = m.nLevels;
    MPAS_LAUNCH(k_scalar_hflux, dim3(blocks_for((long long)m.nEdges * L)), dim3(kBlock), 0, s, m, S, uhAvg,
                       scalar_new_in, horiz_flux);
    MPAS_LAUNCH(k_scalar_update, dim3(blocks_for((long long)m.nCells * L)), dim3(kBlock), 0, s, m, S, dt,
                       coef3, w_old, w_new, scalar_old, scalar_new_in, rho_zz_old, rho_zz_new, uhAvg, wwAvg,
                       horiz_flux, scalar_new_out);
}

// ---------------------------------------------------------------------------
// monotonic FCT stage
// ---------------------------------------------------------------------------
size_t mono_scratch_doubles(const mpas_hip_mesh &m) {
    const size_t L = m.nLevels, C = m.nCells + 1, E = m.nEdges + 1;
    // so, smax, smin, scell, sin, sout (C*L); wdtn, wmid (C*(L+1)); fup, ftmp, flim (E*L)
    return 6 * slice_doubles<double>(C * L) + 2 * slice_doubles<double>(C * (L + 1)) + 3 * slice_doubles<hreal>(E * L);
}

struct MonoScratch {
    double *so, *smax, *smin, *scell, *sin, *sout, *wdtn, *wmid;
    hreal *fup, *ftmp, *flim;
};

static MonoScratch carve_mono(const mpas_hip_mesh &m, double *base) {
    const size_t L = m.nLevels, C = m.nCells + 1, E = m.nEdges + 1;
    Carver k{reinterpret_cast<char *>(base)};
    MonoScratch s;
    s.so = k.take<double>(C * L);
    s.smax = k.take<double>(C * L);
    s.smin = k.take<double>(C * L);
    s.scell = k.take<double>(C * L);
    s.sin = k.take<double>(C * L);
    s.sout = k.take<double>(C * L);
    s.wdtn = k.take<double>(C * (L + 1));
    s.wmid = k.take<double>(C * (L + 1));
    s.fup = k.take<hreal>(E * L);
    s.ftmp = k.take<hreal>(E * L);
    s.flim = k.take<hreal>(E * L);
    return s;
}

// density with the averaged fluxes: rho_zz_int = rho_old + dt*(hdiv - vdiv)
__global__ void k_mono_density(mpas_hip_mesh m, double dt, const hreal *__restrict__ uhAvg,
                               const double *__restrict__ wwAvg, const double *__restrict__ rho_old,
                               double *__restrict__ rho_int) {
    const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nC * L) return;
    int c = i / L, k = i - c * L;
    const int n = m.nEdgesOnCell[c];
    double hd = 0.0;
    for (int j = 0; j < n; ++j) {
        int e = m.edgesOnCell[c * ME + j];
        double metric = -m.edgesOnCell_sign[c * ME + j] * m.dvEdge[e];
        double base = uhAvg[e * L + k] * metric;
        double cc = base * m.invAreaCell[c];
        hd = (j == 0) ? cc : hd + cc;
    }
    double vd = m.rdnw[k] * (wwAvg[c * Lw + k + 1] - wwAvg[c * Lw + k]);
    rho_int[i] = rho_old[i] + dt * (hd - vd);
}

// old scalar plus (zero) source tendency
__global__ void k_mono_old(int nC, int L, int S, int s, double dt, const double *__restrict__ scalars_old,
                           const double *__restrict__ rho_old, double *__restrict__ so) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nC * L) return;
    so[i] = scalars_old[(size_t)i * S + s] + dt * 0.0 / rho_old[i];
}

// local vertical + neighbour bounds, high-order vertical flux
__global__ void k_mono_bounds(mpas_hip_mesh m, int S, int s, double coef3, const double *__restrict__ so,
                              const double *__restrict__ sn, const double *__restrict__ wwAvg,
                              double *__restrict__ smax, double *__restrict__ smin, double *__restrict__ wdtn) {
    const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nC * Lw) return;
    int c = i / Lw, k = i - c * Lw;
    const double *ww = wwAvg + c * Lw;
    auto SN = [&](int kk) { return sn[((size_t)c * L + kk) * S + s]; };
    double wd = 0.0;
    if (k == 1) wd = ww[1] * (m.fnm[1] * SN(1) + m.fnp[1] * SN(0));
    else if (k == L - 1) wd = ww[L - 1] * (m.fnm[L - 1] * SN(L - 1) + m.fnp[L - 1] * SN(L - 2));
    else if (k >= 2 && k <= L - 2) wd = flux3(SN(k - 2), SN(k - 1), SN(k), SN(k + 1), ww[k], coef3);
    wdtn[i] = wd;
    if (k >= L) return;
    const double *sc = so + c * L;
    double mx, mn;
    if (k == 0) {
        mx = dmax(sc[0], sc[1]);
        mn = dmin(sc[0], sc[1]);
    } else if (k == L - 1) {
        mx = dmax(sc[k], sc[k - 1]);
        mn = dmin(sc[k], sc[k - 1]);
    } else {
        mx = dmax(dmax(sc[k - 1], sc[k]), sc[k + 1]);
        mn = dmin(dmin(sc[k - 1], sc[k]), sc[k + 1]);
    }
    const int n = m.nEdgesOnCell[c];
    for (int j = 0; j < n; ++j) {
        int nb = m.cellsOnCell[c * ME + j];
        mx = dmax(mx, so[nb * L + k]);
        mn = dmin(mn, so[nb * L + k]);
    }
    smax[c * L + k] = mx;
    smin[c * L + k] = mn;
}

// horizontal upwind and antidiffusive fluxes
__global__ void k_mono_hflux(mpas_hip_mesh m, int S, int s, double dt, const hreal *__restrict__ uhAvg,
                             const double *__restrict__ sn, const double *__restrict__ so,
                             hreal *__restrict__ fup, hreal *__restrict__ ftmp) {
    const int L = m.nLevels, nE = m.nEdges;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nE * L) return;
    int e = i / L, k = i - e * L;
    int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    double uh = uhAvg[i];
    double hv = patternc_value(m, e, sign_pm(uh), sn, L, S, k, s);
    double fedge = uh * hv;
    double up = (uh >= 0.0) ? so[c1 * L + k] : so[c2 * L + k];
    double fu = m.dvEdge[e] * dt * (uh * up);
    fup[i] = fu;
    ftmp[i] = dt * fedge - fu;
}

// vertical upwind update, limiter accumulators and scale factors
__global__ void k_mono_scale(mpas_hip_mesh m, double dt, const double *__restrict__ so,
                             const double *__restrict__ rho_old, const double *__restrict__ rho_int,
                             const double *__restrict__ wwAvg, const double *__restrict__ wdtn,
                             const double *__restrict__ smax, const double *__restrict__ smin,
                             const hreal *__restrict__ fup, const hreal *__restrict__ ftmp,
                             double *__restrict__ scell, double *__restrict__ wmid, double *__restrict__ sin_,
                             double *__restrict__ sout) {
    const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nC * L) return;
    int c = i / L, k = i - c * L;
    const double eps = 1.0e-20;
    const double *sc = so + c * L, *ww = wwAvg + c * Lw, *wd = wdtn + c * Lw;
    auto fu = [&](int kk) -> double {  // vertical upwind flux at interface kk (1..L-1)
        double vel = ww[kk];
        double val = (vel >= 0.0) ? sc[kk - 1] : sc[kk];
        return dt * (vel * val);
    };
    double sr = sc[k] * rho_old[i];
    if (k < L - 1) sr = sr - fu(k + 1) * m.rdnw[k];
    if (k >= 1) sr = sr + fu(k) * m.rdnw[k];
    double prev = (k == 0) ? wd[0] : dt * wd[k] - fu(k);
    double next = (k == L - 1) ? wd[L] : dt * wd[k + 1] - fu(k + 1);
    if (k >= 1) wmid[c * Lw + k] = prev;
    double si = -m.rdnw[k] * ((-max0(-next)) - max0(prev));
    double so_ = -m.rdnw[k] * (max0(next) - (-max0(-prev)));
    const int n = m.nEdgesOnCell[c];
    const double w_area = m.invAreaCell[c];
    for (int j = 0; j < n; ++j) {
        int e = m.edgesOnCell[c * ME + j];
        double sgn = m.edgesOnCell_sign[c * ME + j];
        double signed_inv = sgn * w_area;
        double sf = sgn * ftmp[e * L + k];
        sr = sr - fup[e * L + k] * signed_inv;
        so_ = so_ - max0(sf) * w_area;
        si = si - (-max0(-sf)) * w_area;
    }
    scell[i] = sr;
    double fi = (smax[i] * rho_int[i] - sr) / (si + eps);
    double fo = (smin[i] * rho_int[i] - sr) / (so_ - eps);
    sin_[i] = dmin(dmax(fi, 0.0), 1.0);
    sout[i] = dmin(dmax(fo, 0.0), 1.0);
}

// limited horizontal flux
__global__ void k_mono_hlimit(mpas_hip_mesh m, const hreal *__restrict__ ftmp, const double *__restrict__ sin_,
                              const double *__restrict__ sout, hreal *__restrict__ flim) {
    const int L = m.nLevels, nE = m.nEdges;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nE * L) return;
    int e = i / L, k = i - e * L;
    int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    double f = ftmp[i];
    double lim = (f >= 0.0) ? dmin(sout[c1 * L + k], sin_[c2 * L + k]) : dmin(sin_[c1 * L + k], sout[c2 * L + k]);
    flim[i] = f * lim;
}

// limited vertical flux, final update, positivity
__global__ void k_mono_final(mpas_hip_mesh m, int S, int s, const double *__restrict__ scell,
                             const double *__restrict__ wdtn, const double *__restrict__ wmid,
                             const double *__restrict__ sin_, const double *__restrict__ sout,
                             const hreal *__restrict__ flim, const double *__restrict__ rho_int,
                             double *__restrict__ out) {
    const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nC * L) return;
    int c = i / L, k = i - c * L;
    const int oc = c * L, ow = c * Lw;
    auto wl = [&](int kk) -> double {  // limited antidiffusive vertical flux at interface kk
        double f = wmid[ow + kk];
        double lim = (f >= 0.0) ? dmin(sout[oc + kk - 1], sin_[oc + kk]) : dmin(sout[oc + kk], sin_[oc + kk - 1]);
        return f * lim;
    };
    double sr = scell[i];
    const int n = m.nEdgesOnCell[c];
    for (int j = 0; j < n; ++j) {
        int e = m.edgesOnCell[c * ME + j];
        double signed_inv = m.edgesOnCell_sign[c * ME + j] * m.invAreaCell[c];
        sr = sr - flim[e * L + k] * signed_inv;
    }
    double lo = (k == 0) ? wdtn[ow + 0] : wl(k);
    double hi = (k == L - 1) ? wdtn[ow + L] : wl(k + 1);
    double delta = hi - lo;
    double st2 = (sr - delta * m.rdnw[k]) / rho_int[i];
    out[(size_t)i * S + s] = max0(st2);
}

void launch_advance_scalars_mono(const mpas_hip_mesh &m, int S, double dt, double coef3, const double *scalars_old,
                                 const double *scalars_new_in, const double *rho_zz_old, const hreal *uhAvg,
                                 const double *wwAvg, double *scalars_new_out, double *rho_zz_int, double *scratch,
                                 hipStream_t s) {
    const int L = m.nLevels, nC = m.nCells, nE = m.nEdges;
    MonoScratch w = carve_mono(m, scratch);
    const unsigned bc = blocks_for((long long)nC * L), bcw = blocks_for((long long)nC * (L + 1)),
                   be = blocks_for((long long)nE * L);
    MPAS_LAUNCH(k_mono_density, dim3(bc), dim3(kBlock), 0, s, m, dt, uhAvg, wwAvg, rho_zz_old, rho_zz_int);
    for (int q = 0; q < S; ++q) {
        MPAS_LAUNCH(k_mono_old, dim3(bc), dim3(kBlock), 0, s, nC, L, S, q, dt, scalars_old, rho_zz_old, w.so);
        MPAS_LAUNCH(k_mono_bounds, dim3(bcw), dim3(kBlock), 0, s, m, S, q, coef3, w.so, scalars_new_in, wwAvg,
                           w.smax, w.smin, w.wdtn);
        MPAS_LAUNCH(k_mono_hflux, dim3(be), dim3(kBlock), 0, s, m, S, q, dt, uhAvg, scalars_new_in, w.so,
                           w.fup, w.ftmp);
        MPAS_LAUNCH(k_mono_scale, dim3(bc), dim3(kBlock), 0, s, m, dt, w.so, rho_zz_old, rho_zz_int, wwAvg,
                           w.wdtn, w.smax, w.smin, w.fup, w.ftmp, w.scell, w.wmid, w.sin, w.sout);
        MPAS_LAUNCH(k_mono_hlimit, dim3(be), dim3(kBlock), 0, s, m, w.ftmp, w.sin, w.sout, w.flim);
        MPAS_LAUNCH(k_mono_final, dim3(bc), dim3(kBlock), 0, s, m, S, q, w.scell, w.wdtn, w.wmid, w.sin,
                           w.sout, w.flim, rho_zz_int, scalars_new_out);
    }
}

}  // namespace mpas
