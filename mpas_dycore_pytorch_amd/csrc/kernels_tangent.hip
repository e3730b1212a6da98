// Tangent-linear kernels of the large step: launch wrappers over the bodies
// in tangent_bodies.hpp / tangent_dyn.hpp / tangent_transport.hpp and the
// extern "C" entry points mpas_hip_*_jvp.
//
// Every entry point takes fwd[] (the forward inputs, then the forward
// outputs, of one call) and dot[] (their tangents, same slots; input
// tangents are read, a null input meaning a zero tangent; output tangents
// are written and must all be given).  Both are host arrays of device
// pointers.  Every thread gathers the tangents of the elements it owns, so
// there are no atomics and results are deterministic.
#include "adjoint_launch.hpp"

namespace mpas {
namespace adj {

ADJ_WAVE_KERNEL(k_tan_smlstep, smlstep_jvp(m, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_divdamp, divdamp_jvp(m, A.coef, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_moist_cell, moist_cell_jvp(m, A.S, A.q0, A.q1, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_moist_edge, moist_edge_jvp(m, A.S, A.q0, A.q1, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_reconstruct, reconstruct_jvp(m, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_recover_cell, recover_cell_jvp(m, A.rk, A.dt, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_recover_edge, recover_edge_jvp(m, A.coef, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_recover_w, recover_w_jvp(m, A.coef, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_diag_vertex, diag_vertex_jvp(m, ADJ_F, ADJ_B, A.dd, ent, k))
ADJ_WAVE_KERNEL(k_tan_diag_cell, diag_cell_jvp(m, ADJ_F, ADJ_B, A.dd, ent, k))
ADJ_WAVE_KERNEL(k_tan_diag_edge, diag_edge_jvp(m, A.dt, A.apvm, A.rk, ADJ_F, ADJ_B, A.dd, ent, k))
ADJ_WAVE_KERNEL(k_tan_acoustic_edge, acoustic_edge_jvp(m, A.ap, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_j1, dyn_j1(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_j2, dyn_j2(m, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_j3, dyn_j3(m, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_j4, dyn_j4<false>(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_j5, dyn_j5<false>(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_j4_opt, dyn_j4<true>(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_j5_opt, dyn_j5<true>(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_fwd_cells, dyn_fwd_cells(m, ADJ_F, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_fwd_edges, dyn_fwd_edges(m, ADJ_F, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_fwd_del4, dyn_fwd_del4(m, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_scal_edge, scal_edge_jvp(m, A.S, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_scal_cell, scal_cell_jvp(m, A.S, A.dt, A.coef3, A.w_old, A.w_new, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_rho, mono_rho_jvp(m, A.dt, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_f1, mono_f1(m, A.S, A.s, A.coef3, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_f2, mono_f2(m, A.S, A.s, A.dt, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_f3, mono_f3(m, A.S, A.s, A.dt, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_f4, mono_f4(m, A.S, A.s, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_m1, mono_m1(m, A.S, A.s, A.coef3, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_m2, mono_m2(m, A.S, A.s, A.dt, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_m3, mono_m3(m, A.S, A.s, A.dt, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_m4, mono_m4(m, A.S, A.s, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_m5, mono_m5(m, A.S, A.s, ADJ_F, ADJ_B, A.mo, ent, k))

// ---------------------------------------------------------------------------
// Column tangents, one wavefront per column (lane = level or interface): the
// algebra of vimp_col_jvp / acoustic_col_jvp, with the level-parallel parts
// on all lanes and the Thomas recurrences on lane 0 over the column staged in
// LDS.  The grid covers the padding column (c == nCells), whose outputs are
// the pass-through of the inputs (or zero) as in the per-column bodies.
// ---------------------------------------------------------------------------
constexpr int kTanColWaves = 4;

__global__ __launch_bounds__(64 * kTanColWaves) void k_tan_vimp_colw(mpas_hip_mesh m, Slots S_, KArgs A) {
    using namespace vimp;
    constexpr int W = kTanColWaves;
    __shared__ double s_dwz[W][65], s_dtz[W][66], s_dwt[W][65], s_da[W][65], s_db[W][65], s_dc[W][65], s_cc[W][65];
    __shared__ double s_fa[W][65], s_fal[W][65], s_fga[W][65], s_dal[W][65], s_dg[W][65];
    const double *const *F = S_.f;
    double *const *T = S_.b;
    const int L = m.nLevels, nC = m.nCells, Lw = L + 1;
    // the column index is wave-uniform, so the cell's connectivity and
    // metrics are scalar loads
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.y), k = threadIdx.x;
    const int c = blockIdx.x * W + wv;  // columns past the padding one do nothing but reach the barriers
    const long long o = (long long)c * L, ow = (long long)c * Lw;
    if (c == nC) {
        if (k < L)
            for (int sl : {cofwr, cofwz, cofwt, a_tri, alpha_tri, gamma_tri}) dot_put(T, sl, o + k, 0.0);
        if (k <= L) dot_put(T, coftz, ow + k, 0.0);
    }
    const bool live = c < nC;
    const double dtseps = 0.5 * A.dt * (1.0 + A.epssm);
    const double rcv = MPAS_RGAS / (MPAS_CP - MPAS_RGAS);
    const double c2 = MPAS_CP * rcv;
    const double *zz = m.zz + o, *rdzw = m.rdzw;
    auto WZ = [&](int kk) { return F[cofwz][o + kk]; };
    auto TZ = [&](int kk) { return F[coftz][ow + kk]; };
    auto WT = [&](int kk) { return F[cofwt][o + kk]; };
    auto WR = [&](int kk) { return F[cofwr][o + kk]; };
    // ---- phase 1: coefficient tangents per level / interface; stage the factor
    if (live && k <= L) {
        if (c == 0 && k < L) dot_put(T, cofrz, k, 0.0);
        if (k < L) dot_put(T, cofwr, o + k, 0.0);
        double dwz = 0.0, dtz = 0.0;
        if (k >= 1 && k < L) {
            const double zi = m.fzm[k] * zz[k] + m.fzp[k] * zz[k - 1];
            const double pi = m.fzm[k] * F[p][o + k] + m.fzp[k] * F[p][o + k - 1];
            const double dpi = m.fzm[k] * dot_of(T, p, o + k) + m.fzp[k] * dot_of(T, p, o + k - 1);
            const double base = dtseps * c2 * zi * m.rdzu[k];
            dwz = base * (dot_of(T, cqw, ow + k) * pi + F[cqw][ow + k] * dpi);
            dtz = dtseps * (m.fzm[k] * dot_of(T, t, o + k) + m.fzp[k] * dot_of(T, t, o + k - 1));
        }
        if (k < L) dot_put(T, cofwz, o + k, dwz);
        dot_put(T, coftz, ow + k, dtz);
        s_dwz[wv][k] = dwz;
        s_dtz[wv][k] = dtz;
        if (k < L) {
            const double A0 = 0.5 * dtseps * rcv * zz[k] * MPAS_GRAVITY * m.rho_base[o + k];
            const double q1 = 1.0 + F[qtot][o + k], rr = m.rtheta_base[o + k] + F[rt][o + k],
                         pb = m.exner_base[o + k];
            const double dwt = A0 / q1 / (rr * pb) * dot_of(T, p, o + k) - WT(k) / q1 * dot_of(T, qtot, o + k) -
                               WT(k) / rr * dot_of(T, rt, o + k);
            dot_put(T, cofwt, o + k, dwt);
            s_dwt[wv][k] = dwt;
            s_fa[wv][k] = F[a_tri][o + k];
            s_fal[wv][k] = F[alpha_tri][o + k];
            s_fga[wv][k] = F[gamma_tri][o + k];
        }
    }
    __syncthreads();
    // ---- phase 2: tangents of the tridiagonal rows a, b, c (and c itself)
    if (live && k >= 1 && k < L) {
        const double *cofrz_ = F[cofrz];
        const double dWZk = s_dwz[wv][k], dTZk = s_dtz[wv][k], dTZm = s_dtz[wv][k - 1], dTZp = s_dtz[wv][k + 1];
        const double dWTk = s_dwt[wv][k], dWTm = s_dwt[wv][k - 1];
        s_da[wv][k] = -dWZk * TZ(k - 1) * rdzw[k - 1] * zz[k - 1] - WZ(k) * dTZm * rdzw[k - 1] * zz[k - 1] -
                      dWTm * TZ(k - 1) * rdzw[k - 1] - WT(k - 1) * dTZm * rdzw[k - 1];
        s_db[wv][k] = dWZk * (TZ(k) * rdzw[k] * zz[k] + TZ(k) * rdzw[k - 1] * zz[k - 1]) +
                      WZ(k) * dTZk * (rdzw[k] * zz[k] + rdzw[k - 1] * zz[k - 1]) -
                      dTZk * (WT(k) * rdzw[k] - WT(k - 1) * rdzw[k - 1]) -
                      TZ(k) * (dWTk * rdzw[k] - dWTm * rdzw[k - 1]);
        s_dc[wv][k] = -dWZk * TZ(k + 1) * rdzw[k] * zz[k] - WZ(k) * dTZp * rdzw[k] * zz[k] +
                      dWTk * TZ(k + 1) * rdzw[k] + WT(k) * dTZp * rdzw[k];
        s_cc[wv][k] = -WZ(k) * TZ(k + 1) * rdzw[k] * zz[k] - WR(k) * cofrz_[k] + WT(k) * TZ(k + 1) * rdzw[k];
    }
    __syncthreads();
    // ---- phase 3: the factor's tangent recurrence (lane 0, over LDS)
    if (live && k == 0) {
        double dgprev = 0.0;
        s_dal[wv][0] = 0.0;
        s_dg[wv][0] = 0.0;
        for (int kk = 1; kk < L; ++kk) {
            const double gprev = (kk > 1) ? s_fga[wv][kk - 1] : 0.0;
            const double al = s_fal[wv][kk];
            const double dden = s_db[wv][kk] - s_da[wv][kk] * gprev - s_fa[wv][kk] * dgprev;
            const double dal = -(al * al) * dden;
            const double dg = s_dc[wv][kk] * al + s_cc[wv][kk] * dal;
            s_dal[wv][kk] = dal;
            s_dg[wv][kk] = dg;
            dgprev = dg;
        }
    }
    __syncthreads();
    if (live && k < L) {
        dot_put(T, a_tri, o + k, k >= 1 ? s_da[wv][k] : 0.0);
        dot_put(T, alpha_tri, o + k, s_dal[wv][k]);
        dot_put(T, gamma_tri, o + k, s_dg[wv][k]);
    }
}

__global__ __launch_bounds__(64 * kTanColWaves) void k_tan_acoustic_colw(mpas_hip_mesh m, Slots S_, KArgs A) {
    using namespace acoustic;
    constexpr int W = kTanColWaves;
    __shared__ double s_rs[W][65], s_ts[W][65], s_drs[W][65], s_dts[W][65], s_r[W][65], s_dr[W][65];
    __shared__ double s_y[W][65], s_dy[W][65], s_x[W][66], s_dx[W][66];
    __shared__ double s_a[W][65], s_al[W][65], s_ga[W][65], s_da[W][65], s_dal[W][65], s_dga[W][65];
    const double *const *F = S_.f;
    double *const *T = S_.b;
    const int L = m.nLevels, Lw = L + 1, ME = m.maxEdges, nC = m.nCells;
    // the column index is wave-uniform, so the cell's connectivity and
    // metrics are scalar loads
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.y), k = threadIdx.x;
    const int c = blockIdx.x * W + wv;  // columns past the padding one do nothing but reach the barriers
    const long long oc = (long long)c * L, ow = (long long)c * Lw;
    if (c == nC) {
        if (k < L) {
            dot_put(T, rho_pp, oc + k, dot_of(T, rho_pp_in, oc + k));
            dot_put(T, rtheta_pp, oc + k, dot_of(T, rtheta_pp_in, oc + k));
            dot_put(T, rtheta_pp_old, oc + k, dot_of(T, rtheta_pp_old_in, oc + k));
        }
        if (k <= L) {
            dot_put(T, rw_p, ow + k, dot_of(T, rw_p_in, ow + k));
            dot_put(T, wwAvg, ow + k, dot_of(T, wwAvg_in, ow + k));
        }
    }
    const bool live = c < nC;
    const mpas_hip_acoustic_params &P = A.ap;
    const double dts = P.dts, epssm = P.epssm, resm = P.resm;
    const bool first = (P.small_step == 1);
    const double *zz = m.zz + oc, *rdzw = m.rdzw, *fzm = m.fzm, *fzp = m.fzp, *cofrz_ = F[cofrz];
    auto WZ = [&](int kk) { return F[cofwz][oc + kk]; };
    auto WR = [&](int kk) { return F[cofwr][oc + kk]; };
    auto WT = [&](int kk) { return F[cofwt][oc + kk]; };
    auto TZ = [&](int kk) { return F[coftz][ow + kk]; };
    auto RPP = [&](int kk) { return first ? 0.0 : F[rho_pp_in][oc + kk]; };
    auto RTPP = [&](int kk) { return first ? 0.0 : F[rtheta_pp_in][oc + kk]; };
    auto RWP = [&](int kk) { return first ? 0.0 : F[rw_p_in][ow + kk]; };
    auto dRPP = [&](int kk) { return first ? 0.0 : dot_of(T, rho_pp_in, oc + kk); };
    auto dRTPP = [&](int kk) { return first ? 0.0 : dot_of(T, rtheta_pp_in, oc + kk); };
    auto dRWP = [&](int kk) { return first ? 0.0 : dot_of(T, rw_p_in, ow + kk); };
    auto dWWA = [&](int kk) { return first ? 0.0 : dot_of(T, wwAvg_in, ow + kk); };
    auto dWZ = [&](int kk) { return dot_of(T, cofwz, oc + kk); };
    auto dWR = [&](int kk) { return dot_of(T, cofwr, oc + kk); };
    auto dWT = [&](int kk) { return dot_of(T, cofwt, oc + kk); };
    auto dTZ = [&](int kk) { return dot_of(T, coftz, ow + kk); };
    // ---- phase 1: rs, ts and their tangents at level k; stage the factor
    if (live && k < L) {
        const int n = m.nEdgesOnCell[c];
        const double ia = m.invAreaCell[c];
        // indices first; then the values of every slot (the cell's own theta
        // once, the other cell's per edge) and their sums in slot order; then
        // the tangents the same way
        int eo[kMaxEdges], nbr[kMaxEdges];
        bool own2[kMaxEdges];
#pragma unroll
        for (int i = 0; i < kMaxEdges; ++i) {
            eo[i] = m.edgesOnCell[c * ME + clamp_slot(i, n)];
            const int c1 = m.cellsOnEdge[2 * eo[i]], c2 = m.cellsOnEdge[2 * eo[i] + 1];
            own2[i] = (c2 == c);
            nbr[i] = (c2 == c) ? c1 : c2;
        }
        double rsum = 0.0, tsum = 0.0, drsum = 0.0, dtsum = 0.0;
        double fl[kMaxEdges], th[kMaxEdges];
        {
            const double tc = F[theta_m][oc + k];
            double rp[kMaxEdges], tn[kMaxEdges];
#pragma unroll
            for (int i = 0; i < kMaxEdges; ++i) {
                rp[i] = tn[i] = 0.0;
                if (i < n) {  // wave-uniform
                    rp[i] = F[ru_p][(long long)eo[i] * L + k];
                    tn[i] = F[theta_m][(long long)nbr[i] * L + k];
                }
            }
#pragma unroll
            for (int i = 0; i < kMaxEdges; ++i) {
                fl[i] = th[i] = 0.0;
                if (i >= n) break;
                const double metric = m.edgesOnCell_sign[c * ME + i] * m.dvEdge[eo[i]];
                const double flux = dts * metric * rp[i] * ia;
                const double t2 = own2[i] ? tc : tn[i], t1 = own2[i] ? tn[i] : tc;
                const double thedge = 0.5 * (t2 + t1);
                const double rcon = -flux, tcon = -(flux * thedge);
                rsum = (i == 0) ? rcon : rsum + rcon;
                tsum = (i == 0) ? tcon : tsum + tcon;
                fl[i] = flux;
                th[i] = thedge;
            }
        }
        {
            const double dtc = dot_of(T, theta_m, oc + k);
            double drp[kMaxEdges], dtn[kMaxEdges];
#pragma unroll
            for (int i = 0; i < kMaxEdges; ++i) {
                drp[i] = dtn[i] = 0.0;
                if (i < n) {  // wave-uniform
                    drp[i] = dot_of(T, ru_p, (long long)eo[i] * L + k);
                    dtn[i] = dot_of(T, theta_m, (long long)nbr[i] * L + k);
                }
            }
#pragma unroll
            for (int i = 0; i < kMaxEdges; ++i) {
                if (i >= n) break;
                const double metric = m.edgesOnCell_sign[c * ME + i] * m.dvEdge[eo[i]];
                const double dflux = dts * metric * drp[i] * ia;
                const double dt2 = own2[i] ? dtc : dtn[i], dt1 = own2[i] ? dtn[i] : dtc;
                const double dthedge = 0.5 * (dt2 + dt1);
                drsum += -dflux;
                dtsum += -(dflux * th[i] + fl[i] * dthedge);
            }
        }
        s_rs[wv][k] = RPP(k) + dts * F[tend_rho][oc + k] + rsum - cofrz_[k] * resm * (RWP(k + 1) - RWP(k));
        s_ts[wv][k] =
            RTPP(k) + dts * F[tend_rt][oc + k] + tsum - resm * rdzw[k] * (TZ(k + 1) * RWP(k + 1) - TZ(k) * RWP(k));
        s_drs[wv][k] =
            dRPP(k) + dts * dot_of(T, tend_rho, oc + k) + drsum - cofrz_[k] * resm * (dRWP(k + 1) - dRWP(k));
        s_dts[wv][k] = dRTPP(k) + dts * dot_of(T, tend_rt, oc + k) + dtsum -
                       resm * rdzw[k] *
                           (dTZ(k + 1) * RWP(k + 1) + TZ(k + 1) * dRWP(k + 1) - dTZ(k) * RWP(k) - TZ(k) * dRWP(k));
        s_a[wv][k] = F[a_tri][oc + k];
        s_al[wv][k] = F[alpha_tri][oc + k];
        s_ga[wv][k] = F[gamma_tri][oc + k];
        s_da[wv][k] = dot_of(T, a_tri, oc + k);
        s_dal[wv][k] = dot_of(T, alpha_tri, oc + k);
        s_dga[wv][k] = dot_of(T, gamma_tri, oc + k);
    }
    __syncthreads();
    // ---- phase 2: right-hand side r(k) and its tangent, 1 <= k < L
    if (live && k >= 1 && k < L) {
        const double tsk = s_ts[wv][k], tskm = s_ts[wv][k - 1], rsk = s_rs[wv][k], rskm = s_rs[wv][k - 1];
        const double dtsk = s_dts[wv][k], dtskm = s_dts[wv][k - 1], drsk = s_drs[wv][k], drskm = s_drs[wv][k - 1];
        const double Az = zz[k] * tsk - zz[k - 1] * tskm;
        const double dA = zz[k] * dtsk - zz[k - 1] * dtskm;
        const double Bz = zz[k] * RTPP(k) - zz[k - 1] * RTPP(k - 1);
        const double dBz = zz[k] * dRTPP(k) - zz[k - 1] * dRTPP(k - 1);
        const double R2 = (rsk + rskm) + resm * (RPP(k) + RPP(k - 1));
        const double dR2 = (drsk + drskm) + resm * (dRPP(k) + dRPP(k - 1));
        s_r[wv][k] = RWP(k) + dts * F[tend_rw][ow + k] - WZ(k) * (Az + resm * Bz) - WR(k) * R2 +
                     WT(k) * (tsk + resm * RTPP(k)) + WT(k - 1) * (tskm + resm * RTPP(k - 1));
        s_dr[wv][k] = dRWP(k) + dts * dot_of(T, tend_rw, ow + k) - dWZ(k) * (Az + resm * Bz) -
                      WZ(k) * (dA + resm * dBz) - dWR(k) * R2 - WR(k) * dR2 + dWT(k) * (tsk + resm * RTPP(k)) +
                      WT(k) * (dtsk + resm * dRTPP(k)) + dWT(k - 1) * (tskm + resm * RTPP(k - 1)) +
                      WT(k - 1) * (dtskm + resm * dRTPP(k - 1));
    }
    __syncthreads();
    // ---- phase 3: Thomas solve and its tangent (lane 0, over LDS)
    if (live && k == 0) {
        double y = RWP(0), dy = dRWP(0);
        s_y[wv][0] = y;
        s_dy[wv][0] = dy;
        for (int kk = 1; kk < L; ++kk) {
            const double a = s_a[wv][kk], al = s_al[wv][kk];
            const double rm = s_r[wv][kk] - a * y;
            const double ny = rm * al;
            dy = (s_dr[wv][kk] - s_da[wv][kk] * y - a * dy) * al + rm * s_dal[wv][kk];
            y = ny;
            s_y[wv][kk] = y;
            s_dy[wv][kk] = dy;
        }
        double x = RWP(L), dx = dRWP(L);
        s_x[wv][L] = x;
        s_dx[wv][L] = dx;
        for (int kk = L - 1; kk >= 0; --kk) {
            const double ga = s_ga[wv][kk];
            dx = s_dy[wv][kk] - s_dga[wv][kk] * x - ga * dx;
            x = s_y[wv][kk] - ga * x;
            s_x[wv][kk] = x;
            s_dx[wv][kk] = dx;
        }
    }
    __syncthreads();
    // ---- phase 4: damping and the averaged flux at interface k
    if (live && k <= L) {
        double dww = dWWA(k);
        double dxk = s_dx[wv][k];
        if (k >= 1 && k < L) {
            dww += 0.5 * (1.0 - epssm) * dRWP(k);
            const double D = 1.0 + dts * m.dss[oc + k];
            const double zi = fzm[k] * zz[k] + fzp[k] * zz[k - 1];
            const double ri = fzm[k] * F[rho_zz][oc + k] + fzp[k] * F[rho_zz][oc + k - 1];
            const double dri = fzm[k] * dot_of(T, rho_zz, oc + k) + fzp[k] * dot_of(T, rho_zz, oc + k - 1);
            const double drwd = dot_of(T, rw_save, ow + k) - dot_of(T, rw, ow + k);
            dxk = (dxk + drwd - dts * m.dss[oc + k] * zi * (dri * F[w][ow + k] + ri * dot_of(T, w, ow + k))) / D -
                  drwd;
            dww += 0.5 * (1.0 + epssm) * dxk;
        }
        dot_put(T, wwAvg, ow + k, dww);
        dot_put(T, rw_p, ow + k, dxk);
        s_dx[wv][k] = dxk;  // each lane rewrites only its own interface
    }
    __syncthreads();
    // ---- phase 5: final density and rho-theta tangents at level k
    if (live && k < L) {
        const double rwf1 = F[rw_p][ow + k + 1], rwf0 = F[rw_p][ow + k];
        const double dx1 = s_dx[wv][k + 1], dx0 = s_dx[wv][k];
        dot_put(T, rho_pp, oc + k, s_drs[wv][k] - cofrz_[k] * (dx1 - dx0));
        dot_put(T, rtheta_pp, oc + k,
                s_dts[wv][k] - rdzw[k] * (dTZ(k + 1) * rwf1 + TZ(k + 1) * dx1 - dTZ(k) * rwf0 - TZ(k) * dx0));
        dot_put(T, rtheta_pp_old, oc + k, dRTPP(k));
    }
}

}  // namespace adj
}  // namespace mpas

using namespace mpas;
using namespace mpas::adj;

static int jvp_check(const mpas_hip_mesh *m, const char *what) {
    if (!m) {
        set_error(std::string(what) + ": null mesh");
        return 1;
    }
    if (m->nLevels < 3 || m->nLevels > 63 || m->maxEdges > kMaxEdges || m->maxEdges2 > kMaxEdges2 ||
        m->maxAdv > kMaxAdv) {
        set_error(std::string(what) + ": mesh outside the kernels' limits");
        return 1;
    }
    return 0;
}
// every output tangent in slots [first, n) must be given
static int need_outputs(double *const *dot, int first, int n, const char *what) {
    for (int i = first; i < n; ++i)
        if (!dot[i]) {
            set_error(std::string(what) + ": every output tangent must be provided (slot " + std::to_string(i) + ")");
            return 1;
        }
    return 0;
}

extern "C" {

int mpas_hip_smlstep_pert_jvp(const mpas_hip_mesh *mesh, const double *const *fwd, double *const *dot, void *stream) {
    if (jvp_check(mesh, "smlstep_pert_jvp") || need_outputs(dot, smlstep::w_tend_out, smlstep::N, "smlstep_pert_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(smlstep::N, fwd, dot);
    KArgs A{};
    ADJ_LAUNCH(k_tan_smlstep, m.nCells + 1, m.nLevels + 1, STREAM(stream));
    return check_launch("smlstep_pert_jvp");
}

int mpas_hip_divergence_damping_jvp(const mpas_hip_mesh *mesh, double dts, double smdiv, double len_disp,
                                    const double *const *fwd, double *const *dot, void *stream) {
    if (jvp_check(mesh, "divergence_damping_jvp") ||
        need_outputs(dot, divdamp::ru_p_out, divdamp::N, "divergence_damping_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(divdamp::N, fwd, dot);
    KArgs A{};
    A.coef = 2.0 * smdiv * len_disp / dts;
    ADJ_LAUNCH(k_tan_divdamp, m.nEdges + 1, m.nLevels, STREAM(stream));
    return check_launch("divergence_damping_jvp");
}

int mpas_hip_moist_coefficients_jvp(const mpas_hip_mesh *mesh, int32_t nScalars, int32_t moist_start,
                                    int32_t moist_end, const double *const *fwd, double *const *dot, void *stream) {
    if (jvp_check(mesh, "moist_coefficients_jvp") ||
        need_outputs(dot, moist::cqw, moist::N, "moist_coefficients_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(moist::N, fwd, dot);
    KArgs A{};
    A.S = nScalars;
    A.q0 = moist_start - 1;
    A.q1 = moist_end;
    ADJ_LAUNCH(k_tan_moist_cell, m.nCells + 1, m.nLevels + 1, STREAM(stream));
    ADJ_LAUNCH(k_tan_moist_edge, m.nEdges + 1, m.nLevels, STREAM(stream));
    return check_launch("moist_coefficients_jvp");
}

int mpas_hip_reconstruct_jvp(const mpas_hip_mesh *mesh, const double *const *fwd, double *const *dot, void *stream) {
    if (jvp_check(mesh, "reconstruct_jvp") || need_outputs(dot, recon::ux, recon::N, "reconstruct_jvp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(recon::N, fwd, dot);
    KArgs A{};
    ADJ_LAUNCH(k_tan_reconstruct, m.nCells + 1, m.nLevels, STREAM(stream));
    return check_launch("reconstruct_jvp");
}

int mpas_hip_recover_large_step_jvp(const mpas_hip_mesh *mesh, int32_t rk_step, double dt, int32_t ns,
                                    const double *const *fwd, double *const *dot, void *stream) {
    if (jvp_check(mesh, "recover_large_step_jvp") ||
        need_outputs(dot, recover::wwAvg, recover::N, "recover_large_step_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(recover::N, fwd, dot);
    KArgs A{};
    A.rk = rk_step;
    A.dt = dt;
    A.coef = 1.0 / (double)ns;
    ADJ_LAUNCH(k_tan_recover_cell, m.nCells + 1, m.nLevels, STREAM(stream));
    ADJ_LAUNCH(k_tan_recover_edge, m.nEdges + 1, m.nLevels, STREAM(stream));
    ADJ_LAUNCH(k_tan_recover_w, m.nCells + 1, m.nLevels + 1, STREAM(stream));
    return check_launch("recover_large_step_jvp");
}

int mpas_hip_solve_diagnostics_jvp(const mpas_hip_mesh *mesh, double dt, double apvm, int32_t rk_step,
                                   const double *const *fwd, double *const *dot, double *scratch, void *stream) {
    if (jvp_check(mesh, "solve_diagnostics_jvp") ||
        need_outputs(dot, diag::h_edge, diag::N, "solve_diagnostics_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    const size_t L = m.nLevels, V = (m.nVertices + 1) * L, C = (m.nCells + 1) * L;
    Slots S_ = make_slots(diag::N, fwd, dot);
    KArgs A{};
    A.dt = dt;
    A.apvm = apvm;
    A.rk = rk_step;
    A.dd = DiagDot{scratch, scratch + V, scratch + 2 * V, apvm > 0.0 ? scratch + 2 * V + C : nullptr};
    ADJ_LAUNCH(k_tan_diag_vertex, m.nVertices + 1, m.nLevels, s);
    ADJ_LAUNCH(k_tan_diag_cell, m.nCells + 1, m.nLevels, s);
    ADJ_LAUNCH(k_tan_diag_edge, m.nEdges + 1, m.nLevels, s);
    return check_launch("solve_diagnostics_jvp");
}

int mpas_hip_vert_imp_coefs_jvp(const mpas_hip_mesh *mesh, double dts, double epssm, const double *const *fwd,
                                double *const *dot, void *stream) {
    if (jvp_check(mesh, "vert_imp_coefs_jvp") || need_outputs(dot, vimp::cofrz, vimp::N, "vert_imp_coefs_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(vimp::N, fwd, dot);
    KArgs A{};
    A.dt = dts;
    A.epssm = epssm;
    MPAS_LAUNCH(k_tan_vimp_colw, dim3((m.nCells + 1 + kTanColWaves - 1) / kTanColWaves), dim3(64, kTanColWaves), 0,
                STREAM(stream), m, S_, A);
    return check_launch("vert_imp_coefs_jvp");
}

int mpas_hip_acoustic_step_jvp(const mpas_hip_mesh *mesh, const mpas_hip_acoustic_params *p, const double *const *fwd,
                               double *const *dot, double *scratch, void *stream) {
    if (jvp_check(mesh, "acoustic_step_jvp") || need_outputs(dot, acoustic::ru_p, acoustic::N, "acoustic_step_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(acoustic::N, fwd, dot);
    KArgs A{};
    A.ap = *p;
    A.col = scratch;
    ADJ_LAUNCH(k_tan_acoustic_edge, m.nEdges + 1, m.nLevels, s);
    MPAS_LAUNCH(k_tan_acoustic_colw, dim3((m.nCells + 1 + kTanColWaves - 1) / kTanColWaves), dim3(64, kTanColWaves),
                0, s, m, S_, A);
    return check_launch("acoustic_step_jvp");
}

int mpas_hip_dyn_tend_jvp(const mpas_hip_mesh *mesh, const mpas_hip_dyn_params *p, const double *const *fwd,
                          double *const *dot, double *scratch, void *stream) {
    if (jvp_check(mesh, "dyn_tend_jvp") || need_outputs(dot, dyn::tend_rho, dyn::N, "dyn_tend_jvp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(dyn::N, fwd, dot);
    KArgs A{};
    A.dp = *p;
    A.dy = carve_dyn(m, scratch);
    const int L = m.nLevels, nC = m.nCells, nE = m.nEdges, nV = m.nVertices;
    const bool rk1 = (p->rk_step == 1);
    const int ncv = nC > nV ? nC : nV;
    if (rk1) {
        ADJ_LAUNCH(k_tan_dyn_fwd_cells, nC, L, s);
        ADJ_LAUNCH(k_tan_dyn_fwd_edges, nE, L, s);
        ADJ_LAUNCH(k_tan_dyn_fwd_del4, ncv, L, s);
    }
    ADJ_LAUNCH(k_tan_dyn_j1, nC + 1, L + 1, s);
    if (rk1) ADJ_LAUNCH(k_tan_dyn_j2, nE + 1, L, s);
    if (rk1 && p->h_mom_eddy_visc4 > 0.0) ADJ_LAUNCH(k_tan_dyn_j3, ncv, L, s);
    if (dyn_options(*p)) {
        ADJ_LAUNCH(k_tan_dyn_j4_opt, nE + 1, L, s);
        ADJ_LAUNCH(k_tan_dyn_j5_opt, nC + 1, L + 1, s);
    } else {
        ADJ_LAUNCH(k_tan_dyn_j4, nE + 1, L, s);
        ADJ_LAUNCH(k_tan_dyn_j5, nC + 1, L + 1, s);
    }
    return check_launch("dyn_tend_jvp");
}

int mpas_hip_advance_scalars_jvp(const mpas_hip_mesh *mesh, int32_t nScalars, double dt, double coef3,
                                 double weight_old, double weight_new, const double *const *fwd, double *const *dot,
                                 void *stream) {
    if (jvp_check(mesh, "advance_scalars_jvp") ||
        need_outputs(dot, scal::scalar_new_out, scal::N, "advance_scalars_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(scal::N, fwd, dot);
    KArgs A{};
    A.S = nScalars;
    A.dt = dt;
    A.coef3 = coef3;
    A.w_old = weight_old;
    A.w_new = weight_new;
    ADJ_LAUNCH(k_tan_scal_edge, m.nEdges + 1, m.nLevels, s);
    ADJ_LAUNCH(k_tan_scal_cell, m.nCells + 1, m.nLevels, s);
    return check_launch("advance_scalars_jvp");
}

int mpas_hip_advance_scalars_mono_jvp(const mpas_hip_mesh *mesh, int32_t nScalars, double dt, double coef3,
                                      const double *const *fwd, double *const *dot, double *scratch, void *stream) {
    if (jvp_check(mesh, "advance_scalars_mono_jvp") ||
        need_outputs(dot, mono::scalars_new_out, mono::N, "advance_scalars_mono_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(mono::N, fwd, dot);
    KArgs A{};
    A.S = nScalars;
    A.dt = dt;
    A.coef3 = coef3;
    A.mo = carve_mono(m, scratch);
    const int L = m.nLevels, nC = m.nCells, nE = m.nEdges;
    ADJ_LAUNCH(k_tan_mono_rho, nC + 1, L, s);
    for (int sc = 0; sc < nScalars; ++sc) {
        A.s = sc;
        ADJ_LAUNCH(k_tan_mono_f1, nC, L + 1, s);
        ADJ_LAUNCH(k_tan_mono_f2, nE, L, s);
        ADJ_LAUNCH(k_tan_mono_f3, nC, L + 1, s);
        ADJ_LAUNCH(k_tan_mono_f4, nC, L, s);
        ADJ_LAUNCH(k_tan_mono_m1, nC, L + 1, s);
        ADJ_LAUNCH(k_tan_mono_m2, nE, L, s);
        ADJ_LAUNCH(k_tan_mono_m3, nC, L + 1, s);
        ADJ_LAUNCH(k_tan_mono_m4, nC, L, s);
        ADJ_LAUNCH(k_tan_mono_m5, nC + 1, L, s);
    }
    return check_launch("advance_scalars_mono_jvp");
}

}  // extern "C"
