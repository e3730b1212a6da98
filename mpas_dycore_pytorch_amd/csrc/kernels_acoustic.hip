// Small-step kernels: small-step perturbation, acoustic substep with the
// vertically implicit tridiagonal solve, divergence damping, large-step
// recovery.
#include "common.hpp"

namespace mpas {

// ---------------------------------------------------------------------------
// atm_set_smlstep_pert_variables_work (smlstep_pert.py:197-249)
// ---------------------------------------------------------------------------
__global__ void k_smlstep(mpas_hip_mesh m, const hreal *__restrict__ u_tend, double *__restrict__ w_tend) {
    const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nC * Lw) return;
    int c = i / Lw, k = i - c * Lw;
    if (k < 1 || k >= L) return;
    double wv = w_tend[i];
    if (!m.flatTerrain) {
        const int n = m.nEdgesOnCell[c];
        for (int j = 0; j < n; ++j) {
            int e = m.edgesOnCell[c * ME + j];
            double uk = u_tend[e * L + k];
            double ui = m.fzm[k] * uk + m.fzp[k] * u_tend[e * L + k - 1];
            double flux = m.edgesOnCell_sign[c * ME + j] * ui;
            size_t zo = ((size_t)c * ME + j) * Lw + k;
            double coef = m.zb_cell[zo] + sign_pm(uk) * m.zb3_cell[zo];
            wv = wv - coef * flux;
        }
    }
    // flat terrain: coef == 0 and wv - 0*flux == wv exactly
    double zi = m.fzm[k] * m.zz[c * L + k] + m.fzp[k] * m.zz[c * L + k - 1];
    w_tend[i] = zi * wv;
}

void launch_smlstep(const mpas_hip_mesh &m, const hreal *u_tend, double *w_tend, hipStream_t s) {
    MPAS_LAUNCH(k_smlstep, dim3(blocks_for((long long)m.nCells * (m.nLevels + 1))), dim3(kBlock), 0, s, m,
                       u_tend, w_tend);
}

// ---------------------------------------------------------------------------
// atm_advance_acoustic_step_work (acoustic_step.py:456-889, 1297-1951)
// Edge pass: pressure-gradient update of ru_p and ruAvg (:576-649).
// ---------------------------------------------------------------------------
// Divergence-damping increment of ru_p at (edge, level) (divergence_damping.py:187-209):
// coef (d2 - d1) / (theta1 + theta2), d = -(rtheta_pp - rtheta_pp_old), zero
// when |theta1 + theta2| < 1e-30.  Shared by the stand-alone damping pass and
// the passes that fold it in (the next substep's edge pass, the recovery).
__device__ __forceinline__ double divdamp_increment(double coef, const double *__restrict__ theta_m,
                                                   const double *__restrict__ rtheta_pp,
                                                   const double *__restrict__ rtheta_pp_old, int c1, int c2,
                                                   int L, int k) {
    double d1 = -(rtheta_pp[c1 * L + k] - rtheta_pp_old[c1 * L + k]);
    double d2 = -(rtheta_pp[c2 * L + k] - rtheta_pp_old[c2 * L + k]);
    double ts = theta_m[c1 * L + k] + theta_m[c2 * L + k];
    return (fabs(ts) >= 1.0e-30) ? coef * (d2 - d1) / ts : 0.0;
}

// With fold.on, the damping of the previous acoustic substep is applied to
// ru_p first (same rounding sequence as the stand-alone pass followed by
// this one).
__global__ __launch_bounds__(kMaxWaveBlock) void k_acoustic_edges(long long nb, mpas_hip_mesh m, AcousticParams P, const double *__restrict__ rtheta_pp,
                                 const double *__restrict__ rho_pp, const double *__restrict__ exner,
                                 const hreal *__restrict__ cqu, const hreal *__restrict__ tend_ru,
                                 double *__restrict__ ru_p, double *__restrict__ ruAvg, DampFold fold) {
    auto body_ = [&](const int e) {
        const int L = m.nLevels, nE = m.nEdges;
        const int k = wave_lane();
        if (e >= nE || k >= L) return;
        const int i = e * L + k;
        const double dts = P.dts;
        if (P.small_step == 1) {
            double r = dts * tend_ru[i];
            ru_p[i] = r;
            ruAvg[i] = r;
            return;
        }
        int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
        double rup = ru_p[i];
        if (fold.on) rup = rup + divdamp_increment(fold.coef, fold.theta_m, rtheta_pp, fold.rtheta_pp_old, c1, c2, L, k);
        double pg = ((rtheta_pp[c2 * L + k] - rtheta_pp[c1 * L + k]) * m.invDcEdge[e]) /
                    (0.5 * (m.zz[c2 * L + k] + m.zz[c1 * L + k]));
        pg = cqu[i] * 0.5 * P.c2 * (exner[c1 * L + k] + exner[c2 * L + k]) * pg;
        pg = pg + 0.5 * m.zxu[i] * MPAS_GRAVITY * (rho_pp[c1 * L + k] + rho_pp[c2 * L + k]);
        double r = rup + dts * (tend_ru[i] - pg);
        ru_p[i] = r;
        ruAvg[i] = ruAvg[i] + r;
    };
    const int base_ = wave_entity_base(nb);
#pragma unroll 1
    for (int t_ = 0; t_ < kEntPerWave; ++t_) body_(base_ + t_);
}

// Cell pass: one wavefront per column, lane = level.  Flux sums and the
// right-hand side are level parallel; the Thomas sweeps
// (tridiagonal_solve.py:50-63) are the serial recurrence of lane 0 over the
// column in LDS; Rayleigh damping and the rho''/rtheta'' back-substitution
// are level parallel again.
//
// The column index is wave-uniform (readfirstlane), so the cell's
// connectivity and metrics are scalar loads.  Each lane loads its own level
// and interface of the column's arrays once; the value at the level or
// interface above or below comes from the neighbouring lane by a shuffle.
// The flux loop loads the cell's own theta once and, per live edge, the
// edge's ru_p and the other cell's theta; slots past the cell's edge count
// load nothing.
constexpr int kAcWaves = 4;

__global__ __launch_bounds__(64 * kAcWaves) void k_acoustic_cells(mpas_hip_mesh m, AcousticParams P,
                                                                   AcousticInputs in, AcousticState io) {
    __shared__ double s_rs[kAcWaves][64], s_ts[kAcWaves][64], s_rw[kAcWaves][65];
    __shared__ double s_a[kAcWaves][64], s_al[kAcWaves][64], s_g[kAcWaves][64];
    const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.y), k = threadIdx.x;
    const int c = blockIdx.x * kAcWaves + wv;
    const bool live = c < nC;  // wave-uniform
    const int oc = c * L, ow = c * Lw;
    const int kl = k < L ? k : L - 1, kw = k < Lw ? k : L;  // clamped into the column
    const bool first = (P.small_step == 1);
    const double dts = P.dts, resm = P.resm, epssm = P.epssm;
    // the cell's edges, the other cell of each and the flux metric, read
    // before the kernel's first store so that they stay scalar loads
    int n = 0;
    double ia = 0.0;
    int eo[kMaxEdges], nbr[kMaxEdges];
    bool own2[kMaxEdges];
    double met[kMaxEdges];
    if (live) {
        n = m.nEdgesOnCell[c];
        ia = m.invAreaCell[c];
#pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            const int jj = clamp_slot(j, n);
            const int e = m.edgesOnCell[c * ME + jj];
            const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
            eo[j] = e;
            own2[j] = (c2 == c);
            nbr[j] = (c2 == c) ? c1 : c2;
            met[j] = m.edgesOnCell_sign[c * ME + jj] * m.dvEdge[e];
        }
    }
    // old values (zero on the first small step, acoustic_step.py:1702-1761)
    double rho_old = 0.0, rt_old = 0.0, rw_k = 0.0, cz0 = 0.0;
    if (live) {
        if (!first) {
            rho_old = io.rho_pp[oc + kl];
            rt_old = io.rtheta_pp[oc + kl];
            rw_k = io.rw_p[ow + kw];
        }
        cz0 = in.coftz[ow + kw];
        if (k < L) {
            io.rtheta_pp_old[oc + k] = first ? 0.0 : rt_old;
            // tridiagonal factors of this column, staged for the serial sweeps
            s_a[wv][k] = in.a_tri[oc + k];
            s_al[wv][k] = in.alpha_tri[oc + k];
            s_g[wv][k] = in.gamma_tri[oc + k];
        }
    }
    // the interface above (k + 1), every lane
    const double rw_k1 = __shfl_down(rw_k, 1), cz1 = __shfl_down(cz0, 1);
    // phase A: horizontal flux sums, rs and ts (:456-511, :700-716)
    if (live && k < L) {
        const double tc = in.theta_m[oc + k];
        double rp[kMaxEdges], tn[kMaxEdges];
#pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            rp[j] = 0.0;
            tn[j] = 0.0;
            if (j < n) {  // wave-uniform
                rp[j] = io.ru_p[eo[j] * L + k];
                tn[j] = in.theta_m[nbr[j] * L + k];
            }
        }
        double rsum = 0.0, tsum = 0.0;
#pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            const double flux = dts * met[j] * rp[j] * ia;
            const double t2 = own2[j] ? tc : tn[j], t1 = own2[j] ? tn[j] : tc;
            const double thedge = 0.5 * (t2 + t1);
            const double rcon = -flux, tcon = -(flux * thedge);
            rsum = (j < n) ? ((j == 0) ? rcon : rsum + rcon) : rsum;
            tsum = (j < n) ? ((j == 0) ? tcon : tsum + tcon) : tsum;
        }
        double rs = rho_old + dts * in.tend_rho[oc + k] + rsum - in.cofrz[k] * resm * (rw_k1 - rw_k);
        double ts = rt_old + dts * in.tend_rt[oc + k] + tsum - resm * m.rdzw[k] * (cz1 * rw_k1 - cz0 * rw_k);
        s_rs[wv][k] = rs;
        s_ts[wv][k] = ts;
    }
    __syncthreads();
    // the columns phases B and D read, and the level below (k - 1), every lane
    double z0 = 0.0, cwt0 = 0.0, rz0 = 0.0, ww_k = 0.0;
    if (live) {
        if (!first) ww_k = io.wwAvg[ow + kw];
        z0 = m.zz[oc + kl];
        cwt0 = in.cofwt[oc + kl];
        rz0 = in.rho_zz[oc + kl];
    }
    const double rt_km = __shfl_up(rt_old, 1), rho_km = __shfl_up(rho_old, 1);
    const double z_m1 = __shfl_up(z0, 1), cwt_m1 = __shfl_up(cwt0, 1), rz_m1 = __shfl_up(rz0, 1);
    // phase B: wwAvg half-update and the implicit rw_p right-hand side (:719-755)
    if (live && k <= L) {
        if (k >= 1 && k < L) {
            ww_k = ww_k + 0.5 * (1.0 - epssm) * rw_k;
            const double tsk = s_ts[wv][k], tskm = s_ts[wv][k - 1];
            const double rsk = s_rs[wv][k], rskm = s_rs[wv][k - 1];
            double r = rw_k + dts * in.tend_rw[ow + k] -
                       in.cofwz[oc + k] * ((z0 * tsk - z_m1 * tskm) + resm * (z0 * rt_old - z_m1 * rt_km)) -
                       in.cofwr[oc + k] * ((rsk + rskm) + resm * (rho_old + rho_km)) +
                       cwt0 * (tsk + resm * rt_old) + cwt_m1 * (tskm + resm * rt_km);
            s_rw[wv][k] = r;
        } else {
            s_rw[wv][k] = rw_k;  // boundary rows pass through the elimination
        }
    }
    __syncthreads();
    // phase C: Thomas sweeps, lane 0 of each column
    if (live && k == 0) {
        double *x = s_rw[wv];
        const double *a = s_a[wv], *al = s_al[wv], *g = s_g[wv];
        double xp = x[0];
#pragma unroll 8
        for (int kk = 1; kk < L; ++kk) {
            xp = (x[kk] - a[kk] * xp) * al[kk];
            x[kk] = xp;
        }
        double xn = x[L];
#pragma unroll 8
        for (int kk = L - 1; kk >= 0; --kk) {
            xn = x[kk] - g[kk] * xn;
            x[kk] = xn;
        }
    }
    __syncthreads();
    // phase D: implicit Rayleigh damping of w, wwAvg (:838-870)
    double rwn = 0.0;
    if (live && k <= L) {
        rwn = s_rw[wv][k];
        if (k >= 1 && k < L) {
            double zi = m.fzm[k] * z0 + m.fzp[k] * z_m1;
            double ri = m.fzm[k] * rz0 + m.fzp[k] * rz_m1;
            double rwd = in.rw_save[ow + k] - in.rw[ow + k];
            double dss = m.dss[oc + k];
            rwn = (rwn + rwd - dts * dss * zi * ri * in.w[ow + k]) / (1.0 + dts * dss) - rwd;
            ww_k = ww_k + 0.5 * (1.0 + epssm) * rwn;
        }
        io.wwAvg[ow + k] = ww_k;
        io.rw_p[ow + k] = rwn;
    }
    __syncthreads();
    if (live && k <= L) s_rw[wv][k] = rwn;
    __syncthreads();
    // phase E: rho'' and rtheta'' from the new rw_p (:872-876)
    if (live && k < L) {
        double rw0 = s_rw[wv][k], rw1 = s_rw[wv][k + 1];
        io.rho_pp[oc + k] = s_rs[wv][k] - in.cofrz[k] * (rw1 - rw0);
        io.rtheta_pp[oc + k] = s_ts[wv][k] - m.rdzw[k] * (cz1 * rw1 - cz0 * rw0);
    }
}

void launch_acoustic(const mpas_hip_mesh &m, const AcousticParams &P, const AcousticInputs &in,
                     const AcousticState &io, hipStream_t s, const DampFold &fold) {
    MPAS_WAVE_LAUNCH(k_acoustic_edges, m.nEdges, s, m,
                       P, io.rtheta_pp, io.rho_pp, in.exner, in.cqu, in.tend_ru, io.ru_p, io.ruAvg, fold);
    MPAS_LAUNCH(k_acoustic_cells, dim3((m.nCells + kAcWaves - 1) / kAcWaves), dim3(64, kAcWaves), 0, s, m,
                       P, in, io);
}

// ---------------------------------------------------------------------------
// atm_divergence_damping_3d (divergence_damping.py:187-209, coefficient :267)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kMaxWaveBlock) void k_divdamp(long long nb, mpas_hip_mesh m, double coef, const double *__restrict__ theta_m,
                          const double *__restrict__ rtheta_pp, const double *__restrict__ rtheta_pp_old,
                          double *__restrict__ ru_p) {
    auto body_ = [&](const int e) {
        const int L = m.nLevels, nE = m.nEdges;
        const int k = wave_lane();
        if (e >= nE || k >= L) return;
        const int i = e * L + k;
        int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
        ru_p[i] = ru_p[i] + divdamp_increment(coef, theta_m, rtheta_pp, rtheta_pp_old, c1, c2, L, k);
    };
    const int base_ = wave_entity_base(nb);
#pragma unroll 1
    for (int t_ = 0; t_ < kEntPerWave; ++t_) body_(base_ + t_);
}

void launch_divdamp(const mpas_hip_mesh &m, double dts, double smdiv, double len_disp, const double *theta_m,
                    const double *rtheta_pp, const double *rtheta_pp_old, double *ru_p, hipStream_t s) {
    const double coef = 2.0 * smdiv * len_disp / dts;
    MPAS_WAVE_LAUNCH(k_divdamp, m.nEdges, s, m, coef,
                       theta_m, rtheta_pp, rtheta_pp_old, ru_p);
}

// ---------------------------------------------------------------------------
// atm_recover_large_step_variables_work (recover_large_step.py:311-711)
// ---------------------------------------------------------------------------
__global__ void k_recover_cells(mpas_hip_mesh m, int rk_step, double dt, double invNs, RecoverInputs in,
                                RecoverOutputs out) {
    const int L = m.nLevels, nC = m.nCells, Lw = L + 1;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (nC + 1) * Lw) return;
    int c = i / Lw, k = i - c * Lw;
    if (c == nC) {
        if (k < L) out.rho_zz[c * L + k] = 1.0;  // padding sentinel (:985-998)
        return;
    }
    if (k < L) {
        const int o = c * L + k;
        double rp = in.rho_p_save[o] + in.rho_pp[o];
        double rz = rp + m.rho_base[o];
        double rtp;
        if (rk_step == 3)
            rtp = in.rtheta_p_save[o] + in.rtheta_pp[o] - dt * rz * in.rt_diabatic_tend[o];
        else
            rtp = in.rtheta_p_save[o] + in.rtheta_pp[o];
        out.rho_p[o] = rp;
        out.rho_zz[o] = rz;
        out.rtheta_p[o] = rtp;
        out.theta_m[o] = (rtp + m.rtheta_base[o]) / rz;
        if (rk_step == 3) {
            const double rcv = MPAS_RGAS / (MPAS_CP - MPAS_RGAS);
            double ex = pow(m.zz[o] * (MPAS_RGAS / MPAS_P0) * (rtp + m.rtheta_base[o]), rcv);
            out.exner[o] = ex;
            out.pressure_p[o] = m.zz[o] * MPAS_RGAS * (ex * rtp + m.rtheta_base[o] * (ex - m.exner_base[o]));
        }
    }
    if (k >= 1 && k < L) {
        out.wwAvg[i] = in.rw_save[i] + (out.wwAvg[i] * invNs);
        out.rw[i] = in.rw_save[i] + in.rw_p[i];
    }
}

__global__ __launch_bounds__(kMaxWaveBlock) void k_recover_edges(long long nb, mpas_hip_mesh m, double invNs, const double *__restrict__ ru_save,
                                const double *__restrict__ ru_p, const double *__restrict__ rho_zz,
                                double *__restrict__ ruAvg, double *__restrict__ ru, double *__restrict__ u,
                                DampFold fold) {
    auto body_ = [&](const int e) {
        const int L = m.nLevels, nE = m.nEdges;
        const int k = wave_lane();
        if (e >= nE || k >= L) return;
        const int i = e * L + k;
        int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
        ruAvg[i] = ru_save[i] + (ruAvg[i] * invNs);
        double rup = ru_p[i];
        if (fold.on)
            rup = rup + divdamp_increment(fold.coef, fold.theta_m, fold.rtheta_pp, fold.rtheta_pp_old, c1, c2, L, k);
        double rn = ru_save[i] + rup;
        ru[i] = rn;
        u[i] = 2.0 * rn / (rho_zz[c1 * L + k] + rho_zz[c2 * L + k]);
    };
    const int base_ = wave_entity_base(nb);
#pragma unroll 1
    for (int t_ = 0; t_ < kEntPerWave; ++t_) body_(base_ + t_);
}

__global__ __launch_bounds__(kMaxWaveBlock) void k_recover_w(long long nb, mpas_hip_mesh m, const double *__restrict__ ru, const double *__restrict__ rw,
                            const double *__restrict__ rho_zz, double *__restrict__ w) {
    auto body_ = [&](const int c) {
        const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
        const int k = wave_lane();
        if (c >= nC || k >= Lw) return;
        const int i = c * Lw + k;
        const int oc = c * L;
        const int n = m.nEdgesOnCell[c];
        // Flat terrain: zb_cell = zb3_cell = 0, every terrain term is a signed
        // zero and adding it leaves the value unchanged (up to the sign of a zero
        // result), so the edge loops are skipped.
        const int nt = m.flatTerrain ? 0 : n;
        if (k == 0) {
            double sacc = 0.0;
            for (int j = 0; j < nt; ++j) {
                int e = m.edgesOnCell[c * ME + j];
                const double *re = ru + e * L;
                double fs = m.cf1 * re[0] + m.cf2 * re[1] + m.cf3 * re[2];
                size_t zo = ((size_t)c * ME + j) * Lw;
                double coef = m.zb_cell[zo] + sign_pm(fs) * m.zb3_cell[zo];
                double con = m.edgesOnCell_sign[c * ME + j] * coef * fs;
                sacc = (j == 0) ? con : sacc + con;
            }
            w[i] = sacc / (m.cf1 * rho_zz[oc + 0] + m.cf2 * rho_zz[oc + 1] + m.cf3 * rho_zz[oc + 2]);
        } else if (k < L) {
            double zi = m.fzm[k] * m.zz[oc + k] + m.fzp[k] * m.zz[oc + k - 1];
            double acc = rw[i] / zi;
            for (int j = 0; j < nt; ++j) {
                int e = m.edgesOnCell[c * ME + j];
                const double *re = ru + e * L;
                double fi = m.fzm[k] * re[k] + m.fzp[k] * re[k - 1];
                size_t zo = ((size_t)c * ME + j) * Lw + k;
                double coef = m.zb_cell[zo] + sign_pm(fi) * m.zb3_cell[zo];
                acc = acc + m.edgesOnCell_sign[c * ME + j] * coef * fi;
            }
            w[i] = acc / (m.fzm[k] * rho_zz[oc + k] + m.fzp[k] * rho_zz[oc + k - 1]);
        } else {
            w[i] = 0.0;
        }
    };
    const int base_ = wave_entity_base(nb);
#pragma unroll 1
    for (int t_ = 0; t_ < kEntPerWave; ++t_) body_(base_ + t_);
}

void launch_recover(const mpas_hip_mesh &m, int rk_step, double dt, int ns, const RecoverInputs &in,
                    const RecoverOutputs &out, hipStream_t s, const DampFold &fold) {
    const int L = m.nLevels, nC = m.nCells, nE = m.nEdges;
    const double invNs = 1.0 / (double)ns;
    MPAS_LAUNCH(k_recover_cells, dim3(blocks_for((long long)(nC + 1) * (L + 1))), dim3(kBlock), 0, s, m,
                       rk_step, dt, invNs, in, out);
    MPAS_WAVE_LAUNCH(k_recover_edges, nE, s, m, invNs,
                       in.ru_save, in.ru_p, out.rho_zz, out.ruAvg, out.ru, out.u, fold);
    MPAS_WAVE_LAUNCH(k_recover_w, nC, s, m, out.ru,
                       out.rw, out.rho_zz, out.w);
}

}  // namespace mpas
