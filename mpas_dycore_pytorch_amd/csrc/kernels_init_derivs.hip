// Derivatives of the cold-start coupling atm_init_coupled_diagnostics
// (mpas_pytorch/dynamics/init_coupled_diagnostics.py:147-380), so that a
// differentiable run can start from raw fields as the reference's
// run_simulation does (atmosphere_model.py:1589-1597 -> _prepare_timestep_state
// :782): the reverse mode (mpas_hip_init_coupled_diagnostics_vjp) and the
// forward mode (mpas_hip_init_coupled_diagnostics_jvp).
//
// Forward, per active cell and level (kernels_init.hip):
//   tm  = theta (1 + rvord qv)   or the given theta_m (derived mode)
//   rz  = rho / zz               or the given rho_zz  (derived mode)
//   rp  = rz - rb,  rtp = tm rp + rb (tm - tb),  ex = (zz R/p0 (rtp + rtb))^rcv
//   pp  = zz R (ex rtp + rtb (ex - exb))
// per edge: ru = 0.5 u (rz1 + rz2); per interior interface:
//   rw = w ri zi - sum_j sgn_j (zb_j + sign(f_j) zb3_j) f_j zi,
//   ri = fzp rz[k-1] + fzm rz[k], f_j = fzm ru_j[k] + fzp ru_j[k-1].
// The base-state outputs (rtheta_base, exner_base, pressure_base) depend on
// the mesh-bound base state only and carry no derivative.  sign(f) is frozen
// at its forward value (soft_limiter.py:153-164, derivative 0).
//
// Slots of fwd[] / bar[] / dot[]:
//   0 theta, 1 rho, 2 qv, 3 theta_m_in, 4 rho_zz_in, 5 u, 6 w |
//   7 theta_m, 8 rho_zz, 9 ru, 10 rw, 11 rho_p, 12 rtheta_p, 13 exner, 14 pressure_p
// Raw mode reads slots 0-2, derived mode slots 3-4.  Every transpose is a
// gather (each thread owns the adjoint it writes), so results are
// deterministic.
#include "common.hpp"

namespace mpas {
namespace icd {

enum {
    THETA, RHO, QV, TM_IN, RZ_IN, U, W, TM, RZ, RU, RW, RP, RTP, EX, PP, NSLOT
};
constexpr double kRvord = 461.6 / 287.0;

struct Ptrs {
    const double *f[NSLOT];
    double *b[NSLOT];
};

__device__ __forceinline__ double tm_of(const double *const *f, int derived, long i) {
    return derived ? f[TM_IN][i] : f[THETA][i] * (1.0 + kRvord * f[QV][i]);
}
__device__ __forceinline__ double rz_of(const double *const *f, int derived, const double *zz, long i) {
    return derived ? f[RZ_IN][i] : f[RHO][i] / zz[i];
}

// adjoint of ru carried by one edge at one level, with the terrain terms of
// rw gathered from both cells of the edge
__device__ double ru_bar_total(const mpas_hip_mesh &m, const Ptrs &P, int e, int k) {
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    double g = P.b[RU] ? P.b[RU][(long)e * L + k] : 0.0;
    if (m.flatTerrain || !P.b[RW]) return g;
    for (int side = 0; side < 2; ++side) {
        const int c = m.cellsOnEdge[2 * e + side];
        const int n = m.nEdgesOnCell[c];
        int j = -1;
        for (int s = 0; s < n; ++s)
            if (m.edgesOnCell[c * ME + s] == e) j = s;
        if (j < 0) continue;
        const double sgn = m.edgesOnCell_sign[c * ME + j];
        // rw at interface kk reads ru[k] with fzm[kk] (kk = k) and ru[k-1] with fzp[kk] (kk = k + 1)
        for (int t = 0; t < 2; ++t) {
            const int kk = k + t;
            if (kk < 1 || kk > L - 1) continue;
            const long ic = (long)c * Lw + kk;
            const double rwb = P.b[RW][ic];
            if (rwb == 0.0) continue;
            const double *zz = m.zz;
            const long o = (long)c * L;
            const double zi = m.fzp[kk] * zz[o + kk - 1] + m.fzm[kk] * zz[o + kk];
            const double f = m.fzm[kk] * P.f[RU][(long)e * L + kk] + m.fzp[kk] * P.f[RU][(long)e * L + kk - 1];
            const size_t zo = ((size_t)c * ME + j) * Lw + kk;
            const double coef = sgn * (m.zb_cell[zo] + sign_pm(f) * m.zb3_cell[zo]) * zi;
            const double fb = -rwb * coef;
            g = g + (t == 0 ? m.fzm[kk] : m.fzp[kk]) * fb;
        }
    }
    return g;
}

// pass 1: per (edge, level) the total adjoint of ru, u's adjoint, and the
// total kept in scratch for the cell pass
__global__ void k_icd_vjp_edges(mpas_hip_mesh m, Ptrs P, int derived, double *rub) {
    const int L = m.nLevels;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)m.nEdges * L) return;
    const int e = (int)(i / L), k = (int)(i - (long)e * L);
    const double g = ru_bar_total(m, P, e, k);
    rub[i] = g;
    if (P.b[U]) {
        const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
        const double rz1 = rz_of(P.f, derived, m.zz, (long)c1 * L + k), rz2 = rz_of(P.f, derived, m.zz, (long)c2 * L + k);
        P.b[U][i] += g * (0.5 * (rz1 + rz2));
    }
}

// pass 2: per (cell, interface) the adjoint of w
__global__ void k_icd_vjp_w(mpas_hip_mesh m, Ptrs P, int derived) {
    const int L = m.nLevels, Lw = L + 1;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)m.nCells * Lw) return;
    const int c = (int)(i / Lw), k = (int)(i - (long)c * Lw);
    if (k == 0 || k == L || !P.b[W] || !P.b[RW]) return;
    const long o = (long)c * L;
    const double *zz = m.zz;
    const double ri = m.fzp[k] * rz_of(P.f, derived, zz, o + k - 1) + m.fzm[k] * rz_of(P.f, derived, zz, o + k);
    const double zi = m.fzp[k] * zz[o + k - 1] + m.fzm[k] * zz[o + k];
    P.b[W][i] += P.b[RW][i] * (ri * zi);
}

// pass 3: per (cell, level) the thermodynamic chain, the density adjoint
// gathered from rw (interfaces k and k+1) and from ru (the cell's edges)
__global__ void k_icd_vjp_cells(mpas_hip_mesh m, Ptrs P, int derived, const double *theta_base, const double *rub) {
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)m.nCells * L) return;
    const int c = (int)(i / L), k = (int)(i - (long)c * L);
    const double R = MPAS_RGAS, rgas_over_p0 = MPAS_RGAS / MPAS_P0, rcv = MPAS_RGAS / (MPAS_CP - MPAS_RGAS);
    const double *zz = m.zz;
    const double z = zz[i], rb = m.rho_base[i], tb = theta_base[i];
    const double tm = tm_of(P.f, derived, i), rz = rz_of(P.f, derived, zz, i);
    const double rp = rz - rb, rtb = tb * rb;
    const double rtp = tm * rp + rb * (tm - tb);
    const double base = z * rgas_over_p0 * (rtp + rtb);
    const double ex = pow(base, rcv);
    auto bar = [&](int s) { return P.b[s] ? P.b[s][i] : 0.0; };
    const double ppb = bar(PP);
    const double exb_tot = bar(EX) + ppb * (z * R * (rtp + rtb));
    const double dex = rcv * pow(base, rcv - 1.0) * (z * rgas_over_p0);
    const double rtpb = bar(RTP) + ppb * (z * R * ex) + exb_tot * dex;
    const double tmb = bar(TM) + rtpb * (rp + rb);
    double rzb = bar(RZ) + bar(RP) + rtpb * tm;
    // rw: ri(kk) = fzp[kk] rz[kk-1] + fzm[kk] rz[kk]
    if (P.b[RW]) {
        for (int t = 0; t < 2; ++t) {
            const int kk = k + t;
            if (kk < 1 || kk > L - 1) continue;
            const long o = (long)c * L;
            const double zi = m.fzp[kk] * zz[o + kk - 1] + m.fzm[kk] * zz[o + kk];
            const double rib = P.b[RW][(long)c * Lw + kk] * (P.f[W][(long)c * Lw + kk] * zi);
            rzb = rzb + (t == 0 ? m.fzm[kk] : m.fzp[kk]) * rib;
        }
    }
    // ru = 0.5 u (rz1 + rz2) on each edge of the cell
    const int n = m.nEdgesOnCell[c];
    for (int j = 0; j < n; ++j) {
        const int e = m.edgesOnCell[c * ME + j];
        const long ie = (long)e * L + k;
        rzb = rzb + rub[ie] * (0.5 * P.f[U][ie]);
    }
    if (derived) {
        if (P.b[TM_IN]) P.b[TM_IN][i] += tmb;
        if (P.b[RZ_IN]) P.b[RZ_IN][i] += rzb;
    } else {
        const double th = P.f[THETA][i], qv = P.f[QV][i];
        if (P.b[THETA]) P.b[THETA][i] += tmb * (1.0 + kRvord * qv);
        if (P.b[QV]) P.b[QV][i] += tmb * (th * kRvord);
        if (P.b[RHO]) P.b[RHO][i] += rzb / z;
    }
}

// ---- forward mode ---------------------------------------------------------
struct TPtrs {
    const double *f[NSLOT];
    const double *d[7];  // tangents of the inputs (slots 0-6, null = none)
    double *o[NSLOT];    // tangents of the outputs (slots 7-14)
};

__device__ __forceinline__ double tdot(const TPtrs &T, int s, long i) { return T.d[s] ? T.d[s][i] : 0.0; }

__global__ void k_icd_jvp_cells(mpas_hip_mesh m, TPtrs T, int derived, const double *theta_base) {
    const int L = m.nLevels;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)(m.nCells + 1) * L) return;
    if (i >= (long)m.nCells * L) {
        for (int s = TM; s < NSLOT; ++s)
            if (s != RU && s != RW) T.o[s][i] = 0.0;
        return;
    }
    const double R = MPAS_RGAS, rgas_over_p0 = MPAS_RGAS / MPAS_P0, rcv = MPAS_RGAS / (MPAS_CP - MPAS_RGAS);
    const double z = m.zz[i], rb = m.rho_base[i], tb = theta_base[i];
    double tm, rz, tmd, rzd;
    if (derived) {
        tm = T.f[TM_IN][i];
        rz = T.f[RZ_IN][i];
        tmd = tdot(T, TM_IN, i);
        rzd = tdot(T, RZ_IN, i);
    } else {
        const double th = T.f[THETA][i], qv = T.f[QV][i];
        tm = th * (1.0 + kRvord * qv);
        rz = T.f[RHO][i] / z;
        tmd = tdot(T, THETA, i) * (1.0 + kRvord * qv) + th * (kRvord * tdot(T, QV, i));
        rzd = tdot(T, RHO, i) / z;
    }
    const double rp = rz - rb, rtb = tb * rb;
    const double rtp = tm * rp + rb * (tm - tb);
    const double base = z * rgas_over_p0 * (rtp + rtb);
    const double ex = pow(base, rcv);
    const double rtpd = tmd * rp + tm * rzd + rb * tmd;
    const double exd = rcv * pow(base, rcv - 1.0) * (z * rgas_over_p0) * rtpd;
    T.o[TM][i] = tmd;
    T.o[RZ][i] = rzd;
    T.o[RP][i] = rzd;
    T.o[RTP][i] = rtpd;
    T.o[EX][i] = exd;
    T.o[PP][i] = z * R * (exd * rtp + ex * rtpd + rtb * exd);
}

__global__ void k_icd_jvp_edges(mpas_hip_mesh m, TPtrs T, int derived) {
    const int L = m.nLevels;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)(m.nEdges + 1) * L) return;
    const int e = (int)(i / L), k = (int)(i - (long)e * L);
    if (e == m.nEdges) {
        T.o[RU][i] = 0.0;
        return;
    }
    const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    const long i1 = (long)c1 * L + k, i2 = (long)c2 * L + k;
    const double rz1 = rz_of(T.f, derived, m.zz, i1);
    const double rz2 = rz_of(T.f, derived, m.zz, i2);
    const double rzd = T.o[RZ][i1] + T.o[RZ][i2];
    T.o[RU][i] = 0.5 * (tdot(T, U, i) * (rz1 + rz2) + T.f[U][i] * rzd);
}

__global__ void k_icd_jvp_rw(mpas_hip_mesh m, TPtrs T, int derived) {
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)(m.nCells + 1) * Lw) return;
    const int c = (int)(i / Lw), k = (int)(i - (long)c * Lw);
    if (c == m.nCells || k == 0 || k == L) {
        T.o[RW][i] = 0.0;
        return;
    }
    const long o = (long)c * L;
    const double *zz = m.zz;
    const double ri = m.fzp[k] * rz_of(T.f, derived, zz, o + k - 1) + m.fzm[k] * rz_of(T.f, derived, zz, o + k);
    const double rid = m.fzp[k] * T.o[RZ][o + k - 1] + m.fzm[k] * T.o[RZ][o + k];
    const double zi = m.fzp[k] * zz[o + k - 1] + m.fzm[k] * zz[o + k];
    double r = tdot(T, W, i) * (ri * zi) + T.f[W][i] * (rid * zi);
    if (!m.flatTerrain) {
        const int n = m.nEdgesOnCell[c];
        for (int j = 0; j < n; ++j) {
            const int e = m.edgesOnCell[c * ME + j];
            const double f = m.fzm[k] * T.f[RU][(long)e * L + k] + m.fzp[k] * T.f[RU][(long)e * L + k - 1];
            const double fd = m.fzm[k] * T.o[RU][(long)e * L + k] + m.fzp[k] * T.o[RU][(long)e * L + k - 1];
            const size_t zo = ((size_t)c * ME + j) * Lw + k;
            r = r - m.edgesOnCell_sign[c * ME + j] * (m.zb_cell[zo] + sign_pm(f) * m.zb3_cell[zo]) * fd * zi;
        }
    }
    T.o[RW][i] = r;
}

}  // namespace icd
}  // namespace mpas

using namespace mpas;

static int icd_check(const mpas_hip_mesh *m, int derived, const double *const *fwd, const char *what) {
    if (!m || !fwd) {
        set_error(std::string(what) + ": null mesh or fwd");
        return 1;
    }
    const bool raw_ok = fwd[icd::THETA] && fwd[icd::RHO] && fwd[icd::QV];
    const bool der_ok = fwd[icd::TM_IN] && fwd[icd::RZ_IN];
    if ((derived && !der_ok) || (!derived && !raw_ok) || !fwd[icd::U] || !fwd[icd::W] || !fwd[icd::RU]) {
        set_error(std::string(what) + ": the forward inputs of the selected mode are missing");
        return 1;
    }
    return 0;
}

extern "C" {

int64_t mpas_hip_init_coupled_vjp_scratch_doubles(int32_t nEdges, int32_t nLevels) {
    return (int64_t)(nEdges + 1) * nLevels;
}

int mpas_hip_init_coupled_diagnostics_vjp(const mpas_hip_mesh *m, int32_t derived, const double *theta_base,
                                          const double *const *fwd, double *const *bar, double *scratch,
                                          void *stream) {
    if (icd_check(m, derived, fwd, "init_coupled_diagnostics_vjp")) return 1;
    if (!bar || !scratch || !theta_base) {
        set_error("init_coupled_diagnostics_vjp: null bar, scratch or theta_base");
        return 1;
    }
    icd::Ptrs P{};
    for (int s = 0; s < icd::NSLOT; ++s) {
        P.f[s] = fwd[s];
        P.b[s] = bar[s];
    }
    const int L = m->nLevels;
    hipStream_t s = (hipStream_t)stream;
    MPAS_LAUNCH(icd::k_icd_vjp_edges, dim3(blocks_for((long long)m->nEdges * L)), dim3(kBlock), 0, s, *m, P,
                (int)derived, scratch);
    MPAS_LAUNCH(icd::k_icd_vjp_w, dim3(blocks_for((long long)m->nCells * (L + 1))), dim3(kBlock), 0, s, *m, P,
                (int)derived);
    MPAS_LAUNCH(icd::k_icd_vjp_cells, dim3(blocks_for((long long)m->nCells * L)), dim3(kBlock), 0, s, *m, P,
                (int)derived, theta_base, (const double *)scratch);
    return check_launch("init_coupled_diagnostics_vjp");
}

int mpas_hip_init_coupled_diagnostics_jvp(const mpas_hip_mesh *m, int32_t derived, const double *theta_base,
                                          const double *const *fwd, double *const *dot, void *stream) {
    if (icd_check(m, derived, fwd, "init_coupled_diagnostics_jvp")) return 1;
    if (!dot || !theta_base) {
        set_error("init_coupled_diagnostics_jvp: null dot or theta_base");
        return 1;
    }
    for (int s = icd::TM; s < icd::NSLOT; ++s)
        if (!dot[s]) {
            set_error("init_coupled_diagnostics_jvp: every output tangent slot must be given");
            return 1;
        }
    icd::TPtrs T{};
    for (int s = 0; s < icd::NSLOT; ++s) {
        T.f[s] = fwd[s];
        T.o[s] = s >= icd::TM ? dot[s] : nullptr;
    }
    for (int s = 0; s < 7; ++s) T.d[s] = dot[s];
    const int L = m->nLevels;
    hipStream_t s = (hipStream_t)stream;
    MPAS_LAUNCH(icd::k_icd_jvp_cells, dim3(blocks_for((long long)(m->nCells + 1) * L)), dim3(kBlock), 0, s, *m, T,
                (int)derived, theta_base);
    MPAS_LAUNCH(icd::k_icd_jvp_edges, dim3(blocks_for((long long)(m->nEdges + 1) * L)), dim3(kBlock), 0, s, *m, T,
                (int)derived);
    MPAS_LAUNCH(icd::k_icd_jvp_rw, dim3(blocks_for((long long)(m->nCells + 1) * (L + 1))), dim3(kBlock), 0, s, *m,
                T, (int)derived);
    return check_launch("init_coupled_diagnostics_jvp");
}

}  // extern "C"
