// Split scalar transport: explicit high-order stages
// (atm_advance_scalars_work, advance_scalars.py:128-700) and the monotonic
// flux-corrected stage (atm_advance_scalars_mono_work,
// advance_scalars_mono.py:3332-4037, tile numeric :1476-1818).
// Scalars are (cells+1, levels, nScalars) with the species index innermost.
#include "common.hpp"

namespace mpas {


// ---------------------------------------------------------------------------
// explicit stages
// ---------------------------------------------------------------------------
// Pattern-C high-order value (pattern_c.py:582-733) of column level k of
// scalar s at an edge, one wavefront per edge (lane = level): advv / c0v / c3v
// hold the edge's stencil row (lane j = slot j); all slot gathers are issued
// before the first is used and the sum runs in slot order.
// The sum is binary32 with a compensation term in the mixed build.
__device__ __forceinline__ freal patternc_wave(int na, int advv, double c0v, double c3v, double sgn,
                                               const double *__restrict__ q, int L, int S, int k, int s) {
    SlotSum f;
    f.first(0.0);
#pragma unroll
    for (int a0 = 0; a0 < kMaxAdv; a0 += 8) {
        if (a0 >= na) break;
        freal qv[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            qv[t] = 0.0;
            if (a0 + t < na) qv[t] = (freal)q[((size_t)bcast(advv, a0 + t) * L + k) * S + s];
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int a = a0 + t;
            const freal wgt = (freal)bcast(c0v, a) + (freal)sgn * (freal)bcast(c3v, a);
            const freal cc = wgt * qv[t];
            f.slot(a, na, cc);
        }
    }
    return f.value();
}

// One wavefront per edge (hflux) or cell (update), lane = level.  The
// connectivity is read with scalar loads, each slot loop issues all its
// gathers before the first is used, and the vertical stencil takes the
// neighbouring levels of the wavefront's own column from lane shuffles.  The
// sums run in slot order.
__global__ __launch_bounds__(kMaxWaveBlock) void k_scalar_hflux(long long nb, mpas_hip_mesh m, int S,
                                                                const double *__restrict__ uhAvg,
                                                                const double *__restrict__ scalar_new,
                                                                hreal *__restrict__ hflux) {
    const int L = m.nLevels, nE = m.nEdges, MA = m.maxAdv;
    const int e = wave_entity_base(nb);
    const int lane = wave_lane();
    if (e >= nE) return;  // wave-uniform
    const int k = lane < L ? lane : L - 1;
    const int i = e * L + k;
    const int na = m.nAdvCellsForEdge[e];
    const int sl = lane < na ? lane : na - 1;
    const int advv = m.advCellsForEdge[e * MA + sl];
    const double c0v = m.adv_coefs[e * MA + sl], c3v = m.adv_coefs_3rd[e * MA + sl];
    const double sgn = sign_pm(uhAvg[i]);
    for (int s = 0; s < S; ++s) {
        const freal f = patternc_wave(na, advv, c0v, c3v, sgn, scalar_new, L, S, k, s);
        if (lane < L) hflux[(size_t)i * S + s] = f;
    }
}

__global__ __launch_bounds__(kMaxWaveBlock) void k_scalar_update(
    long long nb, mpas_hip_mesh m, int S, double dt, double coef3, double w_old, double w_new,
    const double *__restrict__ scalar_old, const double *__restrict__ sn, const double *__restrict__ rho_old,
    const double *__restrict__ rho_new, const double *__restrict__ uhAvg, const double *__restrict__ wwAvg,
    const hreal *__restrict__ hflux, double *__restrict__ out) {
    const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
    const int c = wave_entity_base(nb);
    const int lane = wave_lane();
    if (c >= nC) return;  // wave-uniform
    const int k = lane < L ? lane : L - 1;
    const int i = c * L + k;
    const int n = m.nEdgesOnCell[c];
    const double *ww = wwAvg + c * Lw;
    int eo[kMaxEdges];
    double sg[kMaxEdges], uh[kMaxEdges];
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        const int jj = clamp_slot(j, n);
        eo[j] = m.edgesOnCell[c * ME + jj];
        sg[j] = m.edgesOnCell_sign[c * ME + jj];
    }
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        uh[j] = 0.0;
        if (j < n) uh[j] = uhAvg[eo[j] * L + k];
    }
    const double rho_i = w_old * rho_old[i] + w_new * rho_new[i];
    const double inv = 1.0 / rho_i;
    for (int s = 0; s < S; ++s) {
        // vertical flux at this lane's interface k (1..L-1); its neighbours'
        // column values come from the neighbouring lanes
        const double sn0 = sn[((size_t)c * L + k) * S + s];
        const double sn_m1 = __shfl_up(sn0, 1), sn_m2 = __shfl_up(sn0, 2), sn_p1 = __shfl_down(sn0, 1);
        double wd = 0.0;
        if (k == 1) wd = ww[1] * (m.fnm[1] * sn0 + m.fnp[1] * sn_m1);
        else if (k == L - 1) wd = ww[L - 1] * (m.fnm[L - 1] * sn0 + m.fnp[L - 1] * sn_m1);
        else if (k >= 2 && k <= L - 2) wd = flux3(sn_m2, sn_m1, sn0, sn_p1, ww[k], coef3);
        const double wd_p1 = __shfl_down(wd, 1);  // the flux at interface k + 1
        freal hf[kMaxEdges];
#pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            hf[j] = 0.0;
            if (j < n) hf[j] = (freal)hflux[((size_t)eo[j] * L + k) * S + s];
        }
        // the flux divergence: a compensated binary32 sum in the mixed build
        SlotSum hd;
        hd.first(0.0);
#pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            const freal cc = -(freal)sg[j] * (freal)uh[j] * hf[j];
            hd.slot(j, n, cc);
        }
        const double tend_col = (double)(hd.value() * (freal)m.invAreaCell[c]) + 0.0;
        double vdiv;
        if (k == 0) vdiv = m.rdnw[0] * wd_p1;
        else if (k == L - 1) vdiv = m.rdnw[L - 1] * (-wd);
        else vdiv = m.rdnw[k] * (wd_p1 - wd);
        if (lane < L) out[(size_t)i * S + s] = (scalar_old[(size_t)i * S + s] * rho_old[i] + dt * (tend_col - vdiv)) * inv;
    }
}

void launch_advance_scalars(const mpas_hip_mesh &m, int S, double dt, double coef3, double w_old, double w_new,
                            const double *scalar_old, const double *scalar_new_in, const double *rho_zz_old,
                            const double *rho_zz_new, const double *uhAvg, const double *wwAvg,
                            double *scalar_new_out, hreal *horiz_flux, hipStream_t s) {
    const int L = m.nLevels;
    (void)L;
    MPAS_WAVE_LAUNCH(k_scalar_hflux, m.nEdges, s, m, S, uhAvg, scalar_new_in, horiz_flux);
    MPAS_WAVE_LAUNCH(k_scalar_update, m.nCells, s, m, S, dt, coef3, w_old, w_new, scalar_old, scalar_new_in,
                     rho_zz_old, rho_zz_new, uhAvg, wwAvg, horiz_flux, scalar_new_out);
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
// The FCT passes below run one wavefront per cell or edge (lane = level), like
// the explicit passes above: scalar connectivity loads, slot gathers issued
// before use, and neighbouring levels of the own column by lane shuffles.
__global__ __launch_bounds__(kMaxWaveBlock) void k_mono_density(long long nb, mpas_hip_mesh m, double dt,
                                                                const double *__restrict__ uhAvg,
                                                                const double *__restrict__ wwAvg,
                                                                const double *__restrict__ rho_old,
                                                                double *__restrict__ rho_int) {
    const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
    const int c = wave_entity_base(nb);
    const int k = wave_lane();
    if (c >= nC || k >= L) return;
    const int i = c * L + k;
    const int n = m.nEdgesOnCell[c];
    int eo[kMaxEdges];
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) eo[j] = m.edgesOnCell[c * ME + clamp_slot(j, n)];
    double uh[kMaxEdges];
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        uh[j] = 0.0;
        if (j < n) uh[j] = uhAvg[eo[j] * L + k];
    }
    double hd = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        const double metric = -m.edgesOnCell_sign[c * ME + clamp_slot(j, n)] * m.dvEdge[eo[j]];
        const double base = uh[j] * metric;
        const double cc = base * m.invAreaCell[c];
        hd = (j < n) ? ((j == 0) ? cc : hd + cc) : hd;
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
__global__ __launch_bounds__(kMaxWaveBlock) void k_mono_bounds(long long nb, mpas_hip_mesh m, int S, int s,
                                                               double coef3, const double *__restrict__ so,
                                                               const double *__restrict__ sn,
                                                               const double *__restrict__ wwAvg,
                                                               double *__restrict__ smax, double *__restrict__ smin,
                                                               double *__restrict__ wdtn) {
    const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
    const int c = wave_entity_base(nb);
    const int lane = wave_lane();
    if (c >= nC) return;  // wave-uniform
    const int k = lane;   // interface (0..L) or level (0..L-1)
    const int kl = k < L ? k : L - 1;
    const double *ww = wwAvg + c * Lw;
    // own column values of this lane's level; neighbouring levels by shuffles
    const double sn0 = sn[((size_t)c * L + kl) * S + s];
    const double sn_m1 = __shfl_up(sn0, 1), sn_m2 = __shfl_up(sn0, 2), sn_p1 = __shfl_down(sn0, 1);
    const double so0 = so[c * L + kl];
    const double so_m1 = __shfl_up(so0, 1), so_p1 = __shfl_down(so0, 1);
    if (k >= Lw) return;
    double wd = 0.0;
    if (k == 1) wd = ww[1] * (m.fnm[1] * sn0 + m.fnp[1] * sn_m1);
    else if (k == L - 1) wd = ww[L - 1] * (m.fnm[L - 1] * sn0 + m.fnp[L - 1] * sn_m1);
    else if (k >= 2 && k <= L - 2) wd = flux3(sn_m2, sn_m1, sn0, sn_p1, ww[k], coef3);
    wdtn[c * Lw + k] = wd;
    if (k >= L) return;
    double mx, mn;
    if (k == 0) {
        mx = dmax(so0, so_p1);
        mn = dmin(so0, so_p1);
    } else if (k == L - 1) {
        mx = dmax(so0, so_m1);
        mn = dmin(so0, so_m1);
    } else {
        mx = dmax(dmax(so_m1, so0), so_p1);
        mn = dmin(dmin(so_m1, so0), so_p1);
    }
    const int n = m.nEdgesOnCell[c];
    int nbc[kMaxEdges];
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) nbc[j] = m.cellsOnCell[c * ME + clamp_slot(j, n)];
    double sv[kMaxEdges];
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        sv[j] = 0.0;
        if (j < n) sv[j] = so[nbc[j] * L + k];
    }
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        if (j >= n) break;
        mx = dmax(mx, sv[j]);
        mn = dmin(mn, sv[j]);
    }
    smax[c * L + k] = mx;
    smin[c * L + k] = mn;
}

// horizontal upwind and antidiffusive fluxes
__global__ __launch_bounds__(kMaxWaveBlock) void k_mono_hflux(long long nb, mpas_hip_mesh m, int S, int s, double dt,
                                                              const double *__restrict__ uhAvg,
                                                              const double *__restrict__ sn,
                                                              const double *__restrict__ so, hreal *__restrict__ fup,
                                                              hreal *__restrict__ ftmp) {
    const int L = m.nLevels, nE = m.nEdges, MA = m.maxAdv;
    const int e = wave_entity_base(nb);
    const int lane = wave_lane();
    if (e >= nE) return;  // wave-uniform
    const int k = lane < L ? lane : L - 1;
    const int i = e * L + k;
    const int na = m.nAdvCellsForEdge[e];
    const int sl = lane < na ? lane : na - 1;
    const int advv = m.advCellsForEdge[e * MA + sl];
    const double c0v = m.adv_coefs[e * MA + sl], c3v = m.adv_coefs_3rd[e * MA + sl];
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const double uh = uhAvg[i];
    const double hv = patternc_wave(na, advv, c0v, c3v, sign_pm(uh), sn, L, S, k, s);
    if (lane >= L) return;
    const double fedge = uh * hv;
    const double up = (uh >= 0.0) ? so[c1 * L + k] : so[c2 * L + k];
    const double fu = m.dvEdge[e] * dt * (uh * up);
    fup[i] = fu;
    ftmp[i] = dt * fedge - fu;
}

// vertical upwind update, limiter accumulators and scale factors
__global__ __launch_bounds__(kMaxWaveBlock) void k_mono_scale(
    long long nb, mpas_hip_mesh m, double dt, const double *__restrict__ so, const double *__restrict__ rho_old,
    const double *__restrict__ rho_int, const double *__restrict__ wwAvg, const double *__restrict__ wdtn,
    const double *__restrict__ smax, const double *__restrict__ smin, const hreal *__restrict__ fup,
    const hreal *__restrict__ ftmp, double *__restrict__ scell, double *__restrict__ wmid, double *__restrict__ sin_,
    double *__restrict__ sout) {
    const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
    const int c = wave_entity_base(nb);
    const int lane = wave_lane();
    if (c >= nC) return;  // wave-uniform
    const int k = lane < L ? lane : L - 1;
    const int kw = lane < Lw ? lane : L;
    const int i = c * L + k;
    const double eps = 1.0e-20;
    const double *ww = wwAvg + c * Lw;
    // own column values; the neighbouring levels by shuffles
    const double sc0 = so[c * L + k];
    const double sc_m1 = __shfl_up(sc0, 1);
    const double wdk = wdtn[c * Lw + kw];
    const double wd_p1 = __shfl_down(wdk, 1);
    // vertical upwind flux at this lane's interface k (1..L-1)
    double fuk = 0.0;
    if (k >= 1) {
        const double vel = ww[k];
        const double val = (vel >= 0.0) ? sc_m1 : sc0;
        fuk = dt * (vel * val);
    }
    const double fu_p1 = __shfl_down(fuk, 1);  // at interface k + 1
    if (lane >= L) return;
    double sr = sc0 * rho_old[i];
    if (k < L - 1) sr = sr - fu_p1 * m.rdnw[k];
    if (k >= 1) sr = sr + fuk * m.rdnw[k];
    const double prev = (k == 0) ? wdtn[c * Lw + 0] : dt * wdk - fuk;
    const double next = (k == L - 1) ? wdtn[c * Lw + L] : dt * wd_p1 - fu_p1;
    if (k >= 1) wmid[c * Lw + k] = prev;
    double si = -m.rdnw[k] * ((-max0(-next)) - max0(prev));
    double so_ = -m.rdnw[k] * (max0(next) - (-max0(-prev)));
    const int n = m.nEdgesOnCell[c];
    const double w_area = m.invAreaCell[c];
    int eo[kMaxEdges];
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) eo[j] = m.edgesOnCell[c * ME + clamp_slot(j, n)];
    double fu_e[kMaxEdges], ft_e[kMaxEdges];
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        fu_e[j] = ft_e[j] = 0.0;
        if (j < n) {
            fu_e[j] = fup[eo[j] * L + k];
            ft_e[j] = ftmp[eo[j] * L + k];
        }
    }
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        if (j >= n) break;
        const double sgn = m.edgesOnCell_sign[c * ME + j];
        const double signed_inv = sgn * w_area;
        const double sf = sgn * ft_e[j];
        sr = sr - fu_e[j] * signed_inv;
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
__global__ __launch_bounds__(kMaxWaveBlock) void k_mono_final(long long nb, mpas_hip_mesh m, int S, int s,
                                                              const double *__restrict__ scell,
                                                              const double *__restrict__ wdtn,
                                                              const double *__restrict__ wmid,
                                                              const double *__restrict__ sin_,
                                                              const double *__restrict__ sout,
                                                              const hreal *__restrict__ flim,
                                                              const double *__restrict__ rho_int,
                                                              double *__restrict__ out) {
    const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
    const int c = wave_entity_base(nb);
    const int lane = wave_lane();
    if (c >= nC) return;  // wave-uniform
    const int k = lane < L ? lane : L - 1;
    const int i = c * L + k;
    const int oc = c * L, ow = c * Lw;
    // limited antidiffusive vertical flux at this lane's interface k (1..L-1)
    const double sin_k = sin_[oc + k], sout_k = sout[oc + k];
    const double sin_km1 = __shfl_up(sin_k, 1), sout_km1 = __shfl_up(sout_k, 1);
    double wlk = 0.0;
    if (k >= 1) {
        const double f = wmid[ow + k];
        const double lim = (f >= 0.0) ? dmin(sout_km1, sin_k) : dmin(sout_k, sin_km1);
        wlk = f * lim;
    }
    const double wl_p1 = __shfl_down(wlk, 1);  // at interface k + 1
    if (lane >= L) return;
    double sr = scell[i];
    const int n = m.nEdgesOnCell[c];
    int eo[kMaxEdges];
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) eo[j] = m.edgesOnCell[c * ME + clamp_slot(j, n)];
    double fl[kMaxEdges];
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        fl[j] = 0.0;
        if (j < n) fl[j] = flim[eo[j] * L + k];
    }
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        if (j >= n) break;
        const double signed_inv = m.edgesOnCell_sign[c * ME + j] * m.invAreaCell[c];
        sr = sr - fl[j] * signed_inv;
    }
    const double lo = (k == 0) ? wdtn[ow + 0] : wlk;
    const double hi = (k == L - 1) ? wdtn[ow + L] : wl_p1;
    const double delta = hi - lo;
    const double st2 = (sr - delta * m.rdnw[k]) / rho_int[i];
    out[(size_t)i * S + s] = max0(st2);
}

void launch_advance_scalars_mono(const mpas_hip_mesh &m, int S, double dt, double coef3, const double *scalars_old,
                                 const double *scalars_new_in, const double *rho_zz_old, const double *uhAvg,
                                 const double *wwAvg, double *scalars_new_out, double *rho_zz_int, double *scratch,
                                 hipStream_t s) {
    const int L = m.nLevels, nC = m.nCells, nE = m.nEdges;
    MonoScratch w = carve_mono(m, scratch);
    const unsigned bc = blocks_for((long long)nC * L), be = blocks_for((long long)nE * L);
    MPAS_WAVE_LAUNCH(k_mono_density, nC, s, m, dt, uhAvg, wwAvg, rho_zz_old, rho_zz_int);
    for (int q = 0; q < S; ++q) {
        MPAS_LAUNCH(k_mono_old, dim3(bc), dim3(kBlock), 0, s, nC, L, S, q, dt, scalars_old, rho_zz_old, w.so);
        MPAS_WAVE_LAUNCH(k_mono_bounds, nC, s, m, S, q, coef3, w.so, scalars_new_in, wwAvg, w.smax, w.smin,
                         w.wdtn);
        MPAS_WAVE_LAUNCH(k_mono_hflux, nE, s, m, S, q, dt, uhAvg, scalars_new_in, w.so, w.fup, w.ftmp);
        MPAS_WAVE_LAUNCH(k_mono_scale, nC, s, m, dt, w.so, rho_zz_old, rho_zz_int, wwAvg, w.wdtn, w.smax, w.smin,
                         w.fup, w.ftmp, w.scell, w.wmid, w.sin, w.sout);
        MPAS_LAUNCH(k_mono_hlimit, dim3(be), dim3(kBlock), 0, s, m, w.ftmp, w.sin, w.sout, w.flim);
        MPAS_WAVE_LAUNCH(k_mono_final, nC, s, m, S, q, w.scell, w.wdtn, w.wmid, w.sin, w.sout, w.flim,
                         rho_zz_int, scalars_new_out);
    }
}

}  // namespace mpas
