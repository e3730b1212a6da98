// atm_compute_dyn_tend_work (dyn_tend.py:9390-10833) as seven gather passes,
// one thread per (entity, level).  Section references:
//   Smagorinsky kdiff                :3299-3465
//   density / h_divergence / dpdz    :1045-1094
//   u: vertical flux, PV flux, KE grad, divergence coupling, PGF
//                                    :1097-1131, :1455-1791, :2178-2282
//   u: del2 / del4 mixing, final     :2508-2730, :2956-2999
//   w/theta horizontal advection     :3627-3747, :4174-4380
//   theta RK>1 perturbation flux     :5901-6014
//   w/theta horizontal mixing        :5315-5900, ordered del4 :6485-6553
//   w vertical flux, PGF, buoyancy   :4415-4657
//   theta vertical flux, diabatic    :4803-5057
// Active configuration: 2d_smagorinsky, no vertical eddy viscosity, no u
// Rayleigh damping, no curvature term, no CAM damping, no physics tendencies.
#include "common.hpp"

namespace mpas {

size_t dyn_tend_scratch_doubles(const mpas_hip_mesh &m) {
    const size_t L = m.nLevels, C = m.nCells + 1, E = m.nEdges + 1, V = m.nVertices + 1;
    // dpdz, delsq_u, delsq_div, delsq_vort, delsq_w, delsq_theta, edge fluxes (w, theta, theta pert.)
    return slice_doubles<double>(C * L) + slice_doubles<hreal>(E * L) + slice_doubles<double>(C * L) +
           slice_doubles<hreal>(V * L) + slice_doubles<double>(C * (L + 1)) + slice_doubles<double>(C * L) +
           3 * slice_doubles<double>(E * L);
}

struct DynScratch {
    double *dpdz;
    hreal *delsq_u;
    double *delsq_div;
    hreal *delsq_vort;
    double *delsq_w, *delsq_theta;
    double *wflux, *tflux, *tpert;
};

static DynScratch carve(const mpas_hip_mesh &m, double *base) {
    const size_t L = m.nLevels, C = m.nCells + 1, E = m.nEdges + 1, V = m.nVertices + 1;
    Carver k{reinterpret_cast<char *>(base)};
    DynScratch s;
    s.dpdz = k.take<double>(C * L);
    s.delsq_u = k.take<hreal>(E * L);
    s.delsq_div = k.take<double>(C * L);
    s.delsq_vort = k.take<hreal>(V * L);
    s.delsq_w = k.take<double>(C * (L + 1));
    s.delsq_theta = k.take<double>(C * L);
    s.wflux = k.take<double>(E * L);
    s.tflux = k.take<double>(E * L);
    s.tpert = k.take<double>(E * L);
    return s;
}

// Smagorinsky eddy viscosity of cell c at level k: min(coef def, cap), def the
// horizontal deformation summed over the cell's edges in slot order
__device__ __forceinline__ double smagorinsky_kdiff(mpas_hip_mesh m, DynParams P,
                                                    const double *__restrict__ u, const hreal *__restrict__ v,
                                                    int c, int k) {
    const int L = m.nLevels, ME = m.maxEdges;
    const int n = m.nEdgesOnCell[c];
    int eo[kMaxEdges];
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) eo[j] = m.edgesOnCell[c * ME + clamp_slot(j, n)];
    double uv[kMaxEdges], vv[kMaxEdges];
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        uv[j] = vv[j] = 0.0;
        if (j < n) {  // wave-uniform: slots past the live count skip the loads
            uv[j] = u[eo[j] * L + k];
            vv[j] = v[eo[j] * L + k];
        }
    }
    double dd = 0.0, dof = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxEdges; ++j) {
        const int jj = clamp_slot(j, n);
        const double a = m.defc_a[c * ME + jj], b = m.defc_b[c * ME + jj];
        const double us = uv[j], vs = vv[j];
        const double ndd = (dd + a * us) - b * vs;
        const double ndof = (dof + b * us) + a * vs;
        dd = (j < n) ? ndd : dd;
        dof = (j < n) ? ndof : dof;
    }
    double def = sqrt(dd * dd + dof * dof);
    return dmin(P.kdiff_coef * def, P.kdiff_cap);
}

// P1: eddy viscosity of the first RK stage: Smagorinsky or fixed, then the
// CAM floor on the top levels
__global__ __launch_bounds__(kMaxWaveBlock) void k_dyn_kdiff(long long nb, mpas_hip_mesh m, DynParams P, const double *__restrict__ u,
                            const hreal *__restrict__ v, double *__restrict__ kdiff) {
    auto body_ = [&](const int c) {
        const int L = m.nLevels, nC = m.nCells;
        const int k = wave_lane();
        if (c > nC || k >= L) return;
        const int i = c * L + k;
        if (c == nC) {
            kdiff[i] = 0.0;
            return;
        }
        double kd;
        if (P.mix_fixed) {
            kd = P.kdiff_fixed;
        } else {
            kd = smagorinsky_kdiff(m, P, u, v, c, k);
        }
        if (k >= L - P.cam_levels) kd = dmax(kd, cam_kdiff(P, L, k));
        kdiff[i] = kd;
    };
    const int base_ = wave_entity_base(nb);
#pragma unroll 1
    for (int t_ = 0; t_ < kEntPerWave; ++t_) body_(base_ + t_);
}

// P2: mass-flux divergence, density tendency, buoyancy dpdz; first stage also
// the del2 w / theta mixing sums and their del2 fields for the del4 pass.
__global__ __launch_bounds__(kMaxWaveBlock) void k_dyn_cells_a(long long nb, mpas_hip_mesh m, DynParams P, const double *__restrict__ ru,
                              const double *__restrict__ rw, const double *__restrict__ rr_save,
                              const double *__restrict__ qtot, const double *__restrict__ w,
                              const double *__restrict__ theta_m, const hreal *__restrict__ rho_edge,
                              const double *__restrict__ kdiff, double *__restrict__ h_div,
                              double *__restrict__ tend_rho, double *__restrict__ dpdz,
                              double *__restrict__ delsq_w, double *__restrict__ tend_w_euler,
                              double *__restrict__ delsq_theta, double *__restrict__ tend_theta_euler) {
    auto body_ = [&](const int c) {
        const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
        const int k = wave_lane();
        if (c > nC) return;  // wave-uniform
        const int i = c * Lw + k;
        const bool rk1 = (P.rk_step == 1);
        if (c == nC) {
            if (k < L) h_div[c * L + k] = 0.0;
            return;
        }
        const int n = m.nEdgesOnCell[c];
        const double ia = m.invAreaCell[c];
        const int oc = c * L, ow = c * Lw;
        const int kl = k < L ? k : L - 1, kw = k < Lw ? k : L;  // clamped into the column
        int eo[kMaxEdges], nb_[kMaxEdges];
        double sg[kMaxEdges];
        bool own2[kMaxEdges];  // this cell is the edge's second cell
    #pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            const int jj = clamp_slot(j, n);
            eo[j] = m.edgesOnCell[c * ME + jj];
            sg[j] = m.edgesOnCell_sign[c * ME + jj];
            const int c1 = m.cellsOnEdge[2 * eo[j]], c2 = m.cellsOnEdge[2 * eo[j] + 1];
            own2[j] = (c2 == c);
            nb_[j] = own2[j] ? c1 : c2;
        }
        // slots past the live count skip their loads (wave-uniform guards)
        double rv[kMaxEdges];
    #pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            rv[j] = 0.0;
            if (j < n) rv[j] = ru[eo[j] * L + kl];
        }
        if (k < L) {
            // the mass-flux divergence stays binary64 in both builds: its terms nearly cancel
            double s = 0.0;
    #pragma unroll
            for (int j = 0; j < kMaxEdges; ++j) {
                const double wgt = sg[j] * m.dvEdge[eo[j]];
                const double cc = wgt * rv[j];
                s = (j < n) ? ((j == 0) ? cc : s + cc) : s;
            }
            double hd = s * ia;
            h_div[oc + k] = hd;
            if (rk1) {
                const int o = oc + k;
                tend_rho[o] = -hd - m.rdzw[k] * (rw[ow + k + 1] - rw[ow + k]);
                dpdz[o] = -MPAS_GRAVITY * (m.rho_base[o] * qtot[o] + rr_save[o] * (1.0 + qtot[o]));
            }
        }
        if (!rk1) return;  // wave-uniform
        // First stage: del2 of theta (levels, dyn_tend.py:5563-5640) and of w
        // (interfaces, :5315-5420).  Each slot loads the other cell's theta,
        // kdiff and w and the edge's rho_edge at this lane's level; the own
        // cell's values are loaded once, and the level k-1 values the
        // interface stencil needs come from lane k-1.  All lanes run the loads
        // and shuffles.
        const double th0 = theta_m[oc + kl], kd0 = kdiff[oc + kl], w0 = w[ow + kw];
        double re[kMaxEdges], thn[kMaxEdges], kn[kMaxEdges], wn[kMaxEdges];
    #pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            re[j] = thn[j] = kn[j] = wn[j] = 0.0;
            if (j < n) {
                re[j] = rho_edge[eo[j] * L + kl];
                thn[j] = theta_m[nb_[j] * L + kl];
                kn[j] = kdiff[nb_[j] * L + kl];
                wn[j] = w[nb_[j] * Lw + kw];
            }
        }
        const double kd0_m1 = __shfl_up(kd0, 1);
        double re_m1[kMaxEdges], kn_m1[kMaxEdges];
    #pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            re_m1[j] = kn_m1[j] = 0.0;
            if (j < n) {
                re_m1[j] = __shfl_up(re[j], 1);
                kn_m1[j] = __shfl_up(kn[j], 1);
            }
        }
        if (k < L) {
            const int o = oc + k;
            double dsum = 0.0, tsum = 0.0;
    #pragma unroll
            for (int j = 0; j < kMaxEdges; ++j) {
                const int e = eo[j];
                const double t2 = own2[j] ? th0 : thn[j], t1 = own2[j] ? thn[j] : th0;
                const double k1 = own2[j] ? kn[j] : kd0, k2 = own2[j] ? kd0 : kn[j];
                const double es = ia * sg[j] * m.dvEdge[e] * m.invDcEdge[e];
                const double prs = P.prandtl_inv * m.meshScalingDel2[e];
                const double bf = es * (t2 - t1) * re[j];
                const double sf = bf * 0.5 * (k1 + k2) * prs;
                dsum = (j < n) ? ((j == 0) ? bf : dsum + bf) : dsum;
                tsum = (j < n) ? ((j == 0) ? sf : tsum + sf) : tsum;
            }
            delsq_theta[o] = dsum;
            tend_theta_euler[o] = tsum;
        }
        if (k >= 1 && k < L) {
            double dsum = 0.0, tsum = 0.0;
    #pragma unroll
            for (int j = 0; j < kMaxEdges; ++j) {
                const int e = eo[j];
                const double r0 = re[j], r1 = re_m1[j];
                const double w2 = own2[j] ? w0 : wn[j], w1 = own2[j] ? wn[j] : w0;
                const double k1 = own2[j] ? kn[j] : kd0, k2 = own2[j] ? kd0 : kn[j];
                const double k1m = own2[j] ? kn_m1[j] : kd0_m1, k2m = own2[j] ? kd0_m1 : kn_m1[j];
                const double es = 0.5 * ia * sg[j] * m.dvEdge[e] * m.invDcEdge[e];
                const double bf = es * (r0 + r1) * (w2 - w1);
                const double sf = bf * m.meshScalingDel2[e] * 0.25 * (k1 + k2 + k1m + k2m);
                dsum = (j < n) ? ((j == 0) ? bf : dsum + bf) : dsum;
                tsum = (j < n) ? ((j == 0) ? sf : tsum + sf) : tsum;
            }
            delsq_w[i] = dsum;
            tend_w_euler[i] = tsum;
        } else if (k < Lw) {
            tend_w_euler[i] = 0.0;
        }
    };
    const int base_ = wave_entity_base(nb);
#pragma unroll 1
    for (int t_ = 0; t_ < kEntPerWave; ++t_) body_(base_ + t_);
}

// Vertical mixing of u at level k (1..L-2) of edge e (dyn_tend.py:2731-2781),
// from u at levels k-1, k, k+1; added to tend_u_euler after the del4 mixing
__device__ __forceinline__ double u_vertical_mix(mpas_hip_mesh m, DynParams P, int e, int k, double re,
                                                 double u_m1, double u_0, double u_p1) {
    const int Lw = m.nLevels + 1;
    const VmixZ z = vmix_z_edge(m.zgrid + m.cellsOnEdge[2 * e] * Lw, m.zgrid + m.cellsOnEdge[2 * e + 1] * Lw, k);
    const double a = vmix_u(m, P, u_m1, e, k - 1), b = vmix_u(m, P, u_0, e, k), c = vmix_u(m, P, u_p1, e, k + 1);
    return re * P.v_mom_eddy_visc2 * ((c - b) / (z.zp - z.z0) - (b - a) / (z.z0 - z.zm)) / (0.5 * (z.zp - z.zm));
}

// P4: normal-velocity tendency (vertical flux, PV flux, KE gradient,
// divergence coupling); first stage also PGF and del2 mixing into tend_u_euler.
// kOpt builds in the curvature and vertical-mixing terms
template <bool kOpt>
__device__ __forceinline__ void dyn_edges_a_body(long long nb, mpas_hip_mesh m, DynParams P, const double *__restrict__ u,
                              const double *__restrict__ rw, const double *__restrict__ w,
                              const hreal *__restrict__ pv_edge,
                              const double *__restrict__ ke, const double *__restrict__ h_div,
                              const hreal *__restrict__ rho_edge, const double *__restrict__ pp,
                              const double *__restrict__ dpdz, const hreal *__restrict__ cqu,
                              const double *__restrict__ divergence, const hreal *__restrict__ vorticity,
                              const double *__restrict__ kdiff, hreal *__restrict__ tend_u,
                              hreal *__restrict__ tend_u_euler, hreal *__restrict__ delsq_u, int add_euler) {
    auto body_ = [&](const int e) {
        const int L = m.nLevels, nE = m.nEdges, ME2 = m.maxEdges2, Lw = L + 1;
        const int lane = wave_lane();
        if (e > nE) return;
        if (e == nE) {
            if (lane < L) tend_u[e * L + lane] = 0.0;
            return;
        }
        // lanes past the column repeat the top level for their loads, store nothing
        const int k = lane < L ? lane : L - 1;
        const int i = e * L + k;
        const int ne = m.nEdgesOnEdge[e];
        const int sl = lane < ne ? lane : ne - 1;
        const int eev = m.edgesOnEdge[e * ME2 + sl];
        const double wv = m.weightsOnEdge[e * ME2 + sl];
        const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
        const double *ue = u + e * L;
        const double *fzm = m.fzm, *fzp = m.fzp;
        // vertical flux of u at this lane's interface k (1..L-1); the stencil
        // values u(k-2), u(k-1), u(k+1) are the neighbouring lanes' own u,
        // exchanged by shuffles instead of re-read through the vector cache
        const double u0 = ue[k];
        const double um1 = __shfl_up(u0, 1), um2 = __shfl_up(u0, 2), up1 = __shfl_down(u0, 1);
        double wz = 0.0;
        if (k >= 1) {
            const double erw = 0.5 * (rw[c1 * Lw + k] + rw[c2 * Lw + k]);
            wz = (k == 1 || k == L - 1) ? erw * (fzm[k] * u0 + fzp[k] * um1) : flux3(um2, um1, u0, up1, erw, 1.0);
        }
        const double wz_up = __shfl_down(wz, 1);  // the flux at interface k + 1
        double delta;
        if (k == 0) delta = wz_up;
        else if (k == L - 1) delta = -wz;
        else delta = wz_up - wz;
        double vert = -m.rdzw[k] * delta;
        // horizontal terms (binary32 in the mixed build): the PV flux over
        // edgesOnEdge in slot order, the KE gradient and the divergence coupling
        SlotSum q;
        q.first(0.0);
        const freal pvs = (freal)pv_edge[i];
    #pragma unroll
        for (int j0 = 0; j0 < kMaxEdges2; j0 += 8) {
            if (j0 >= ne) break;
            freal pj[8], uj[8];
    #pragma unroll
            for (int t = 0; t < 8; ++t) {
                // slots past the live count skip the load (wave-uniform branch)
                pj[t] = 0.0;
                uj[t] = 0.0;
                if (j0 + t < ne) {
                    const int ee = bcast(eev, j0 + t);
                    pj[t] = (freal)pv_edge[ee * L + k];
                    uj[t] = (freal)u[ee * L + k];
                }
            }
    #pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int j = j0 + t;
                const freal workpv = (freal)0.5 * (pvs + pj[t]);
                const freal cc = (freal)bcast(wv, j) * uj[t] * workpv;
                q.slot(j, ne, cc);
            }
        }
        const freal ke_delta = (freal)ke[c2 * L + k] - (freal)ke[c1 * L + k];
        const freal hsum = (freal)h_div[c1 * L + k] + (freal)h_div[c2 * L + k];
        const double base = (double)((freal)rho_edge[i] * (q.value() - ke_delta * (freal)m.invDcEdge[e]) -
                                     (freal)ue[k] * (freal)0.5 * hsum);
        if (lane >= L) return;
        double tu = vert + base;
        if (kOpt && P.curvature) {
            // curvature (dyn_tend.py:2471-2505): tu -= curvEdge rho_edge w_avg + u w_avg rho_edge / r_earth
            const double *w1 = w + c1 * Lw, *w2 = w + c2 * Lw;
            const double w_avg = 0.25 * (((w1[k] + w1[k + 1]) + w2[k]) + w2[k + 1]);
            const double re = (double)rho_edge[i];
            tu = tu - (m.curvEdge[e] * re * w_avg + ue[k] * w_avg * re * (1.0 / kEarthRadius));
        }
        // Rayleigh damping of u on the top levels, before the Euler part (dyn_tend.py:2969-2994)
        const int k0 = L - P.rayleigh_levels;
        if (P.rayleigh_levels > 0 && k >= k0) {
            const double damp = (((double)k + 1.0) - (double)k0) * P.rayleigh_coef_inverse;
            tu = tu - (double)rho_edge[i] * ue[k] * damp;
        }
        double eu = 0.0;
        if (P.rk_step == 1) {
            double zzs = m.zz[c2 * L + k] + m.zz[c1 * L + k];
            double ppd = pp[c2 * L + k] - pp[c1 * L + k];
            double dps = dpdz[c1 * L + k] + dpdz[c2 * L + k];
            eu = -cqu[i] * (ppd * m.invDcEdge[e] / (0.5 * zzs) - 0.5 * m.zxu[i] * dps);
            const int v1 = m.verticesOnEdge[2 * e], v2 = m.verticesOnEdge[2 * e + 1];
            double r_dc = m.invDcEdge[e];
            double r_dv = dmin(m.invDvEdge[e], 4.0 * m.invDcEdge[e]);
            double ud = (divergence[c2 * L + k] - divergence[c1 * L + k]) * r_dc -
                        (vorticity[v2 * L + k] - vorticity[v1 * L + k]) * r_dv;
            delsq_u[i] = ud;
            double kdiffu = 0.5 * (kdiff[c1 * L + k] + kdiff[c2 * L + k]);
            eu = eu + rho_edge[i] * kdiffu * ud * m.meshScalingDel2[e];
            // without a del4 pass the vertical mixing closes tend_u_euler here
            if (kOpt && add_euler && P.v_mom_eddy_visc2 > 0.0 && k >= 1 && k <= L - 2)
                eu = eu + u_vertical_mix(m, P, e, k, (double)rho_edge[i], um1, u0, up1);
            tend_u_euler[i] = eu;
        } else if (add_euler) {
            eu = tend_u_euler[i];
        }
        // without a del4 pass the final tendency is completed here (same
        // rounding as the separate pass: tend_u = (vert + base) + euler)
        tend_u[i] = add_euler ? tu + eu : tu;
    };
    const int base_ = wave_entity_base(nb);
#pragma unroll 1
    for (int t_ = 0; t_ < kEntPerWave; ++t_) body_(base_ + t_);
}
__global__ __launch_bounds__(kMaxWaveBlock) void k_dyn_edges_a(long long nb, mpas_hip_mesh m, DynParams P, const double *__restrict__ u,
                              const double *__restrict__ rw, const double *__restrict__ w,
                              const hreal *__restrict__ pv_edge,
                              const double *__restrict__ ke, const double *__restrict__ h_div,
                              const hreal *__restrict__ rho_edge, const double *__restrict__ pp,
                              const double *__restrict__ dpdz, const hreal *__restrict__ cqu,
                              const double *__restrict__ divergence, const hreal *__restrict__ vorticity,
                              const double *__restrict__ kdiff, hreal *__restrict__ tend_u,
                              hreal *__restrict__ tend_u_euler, hreal *__restrict__ delsq_u, int add_euler) {
    dyn_edges_a_body<false>(nb, m, P, u, rw, w, pv_edge, ke, h_div, rho_edge, pp, dpdz, cqu, divergence,
                            vorticity, kdiff, tend_u, tend_u_euler, delsq_u, add_euler);
}
__global__ __launch_bounds__(kMaxWaveBlock) void k_dyn_edges_a_opt(long long nb, mpas_hip_mesh m, DynParams P, const double *__restrict__ u,
                              const double *__restrict__ rw, const double *__restrict__ w,
                              const hreal *__restrict__ pv_edge,
                              const double *__restrict__ ke, const double *__restrict__ h_div,
                              const hreal *__restrict__ rho_edge, const double *__restrict__ pp,
                              const double *__restrict__ dpdz, const hreal *__restrict__ cqu,
                              const double *__restrict__ divergence, const hreal *__restrict__ vorticity,
                              const double *__restrict__ kdiff, hreal *__restrict__ tend_u,
                              hreal *__restrict__ tend_u_euler, hreal *__restrict__ delsq_u, int add_euler) {
    dyn_edges_a_body<true>(nb, m, P, u, rw, w, pv_edge, ke, h_div, rho_edge, pp, dpdz, cqu, divergence,
                           vorticity, kdiff, tend_u, tend_u_euler, delsq_u, add_euler);
}

// P5: del2 of the u Laplacian at vertices and cells (del4 first half); one
// wavefront per index (vertex and cell of the same number), lane = level
__global__ __launch_bounds__(kMaxWaveBlock) void k_dyn_del4_vc(long long nb, mpas_hip_mesh m,
                                                               const hreal *__restrict__ delsq_u,
                                                               hreal *__restrict__ delsq_vort,
                                                               double *__restrict__ delsq_div) {
    const int L = m.nLevels, nC = m.nCells, nV = m.nVertices, ME = m.maxEdges;
    const int id = wave_entity_base(nb);
    const int k = wave_lane();
    if (k >= L) return;
    if (id < nV) {
        const int vx = id;
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            int e = m.edgesOnVertex[3 * vx + j];
            double wgt = m.invAreaTriangle[vx] * m.dcEdge[e] * m.edgesOnVertex_sign[3 * vx + j];
            double cc = wgt * delsq_u[e * L + k];
            s = (j == 0) ? cc : s + cc;
        }
        delsq_vort[vx * L + k] = s;
    }
    if (id < nC) {
        const int c = id;
        const int n = m.nEdgesOnCell[c];
        int eo[kMaxEdges];
#pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) eo[j] = m.edgesOnCell[c * ME + clamp_slot(j, n)];
        double du[kMaxEdges];
#pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            du[j] = 0.0;
            if (j < n) du[j] = delsq_u[eo[j] * L + k];  // wave-uniform guard
        }
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            double wgt = m.invAreaCell[c] * m.dvEdge[eo[j]] * m.edgesOnCell_sign[c * ME + clamp_slot(j, n)];
            double cc = wgt * du[j];
            s = (j < n) ? ((j == 0) ? cc : s + cc) : s;
        }
        delsq_div[c * L + k] = s;
    }
}

// P6: del4 u mixing (first stage) and the final u tendency; one wavefront per
// edge, lane = level
__global__ __launch_bounds__(kMaxWaveBlock) void k_dyn_edges_b(long long nb, mpas_hip_mesh m, DynParams P,
                                                               int do_del4, const double *__restrict__ u,
                                                               const hreal *__restrict__ rho_edge,
                                                               const double *__restrict__ delsq_div,
                                                               const hreal *__restrict__ delsq_vort,
                                                               hreal *__restrict__ tend_u_euler,
                                                               hreal *__restrict__ tend_u) {
    const int L = m.nLevels, nE = m.nEdges;
    const int e = wave_entity_base(nb);
    const int k = wave_lane();
    if (e >= nE || k >= L) return;
    const int i = e * L + k;
    double eu = tend_u_euler[i];
    if (do_del4) {
        const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
        const int v1 = m.verticesOnEdge[2 * e], v2 = m.verticesOnEdge[2 * e + 1];
        double ums = m.meshScalingDel4[e] * P.h_mom_eddy_visc4;
        double r_dc = ums * P.del4u_div_factor * m.invDcEdge[e];
        double r_dv = ums * dmin(m.invDvEdge[e], 4.0 * m.invDcEdge[e]);
        double ud = rho_edge[i] * ((delsq_div[c2 * L + k] - delsq_div[c1 * L + k]) * r_dc -
                                   (delsq_vort[v2 * L + k] - delsq_vort[v1 * L + k]) * r_dv);
        eu = eu - ud;
        if (P.v_mom_eddy_visc2 > 0.0 && k >= 1 && k <= L - 2)
            eu = eu + u_vertical_mix(m, P, e, k, (double)rho_edge[i], u[i - 1], u[i], u[i + 1]);
        tend_u_euler[i] = eu;
    }
    tend_u[i] = tend_u[i] + eu;
}

// P6b: horizontal advective fluxes of w (at interfaces) and theta_m (at
// levels) through each edge, once per edge: flux = ru_e * sum_a (c_a + s c3_a)
// q_a over advCellsForEdge in slot order (dyn_tend.py:3627-3747, :4174-4380);
// for rk_step > 1 also the theta perturbation flux
// dv (ru_save - ru) 0.5 (theta_m_save1 + theta_m_save2) (:5901-6014).  The
// cell pass applies the edge sign, an exact sign flip of these products.
__global__ __launch_bounds__(kMaxWaveBlock) void k_dyn_edge_adv(long long nb, mpas_hip_mesh m, DynParams P, const double *__restrict__ ru,
                               const double *__restrict__ ru_save, const double *__restrict__ w,
                               const double *__restrict__ theta_m, const double *__restrict__ theta_m_save,
                               double *__restrict__ wflux, double *__restrict__ tflux, double *__restrict__ tpert) {
    auto body_ = [&](const int e) {
        const int L = m.nLevels, nE = m.nEdges, MA = m.maxAdv, Lw = L + 1;
        const int lane = wave_lane();
        if (e >= nE) return;
        const int k = lane < L ? lane : L - 1;  // lanes past the column: loads only
        const int i = e * L + k;
        const int na = m.nAdvCellsForEdge[e];
        const int sl = lane < na ? lane : na - 1;
        const int advv = m.advCellsForEdge[e * MA + sl];
        const double c0v = m.adv_coefs[e * MA + sl], c3v = m.adv_coefs_3rd[e * MA + sl];
        const double *rue_col = ru + e * L;
        const double rut = rue_col[k];
        const double srt = sign_pm(rut);
        const bool wl = (k >= 1);
        const double ruw = wl ? m.fzm[k] * rue_col[k] + m.fzp[k] * rue_col[k - 1] : 0.0;
        const double srw = sign_pm(ruw);
        // binary64 in both builds: the theta and w fluxes feed flux divergences whose
        // terms nearly cancel (theta ~ 300 K against its gradients, w against its mean)
        double ft = 0.0, fw = 0.0;
    #pragma unroll
        for (int a0 = 0; a0 < kMaxAdv; a0 += 8) {
            if (a0 >= na) break;
            double qt[8], qw[8];
    #pragma unroll
            for (int t = 0; t < 8; ++t) {
                qt[t] = 0.0;
                qw[t] = 0.0;
                if (a0 + t < na) {  // wave-uniform: slots past the live count skip the load
                    const int ac = bcast(advv, a0 + t);
                    qt[t] = theta_m[ac * L + k];
                    qw[t] = w[ac * Lw + k];
                }
            }
    #pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int a = a0 + t;
                const double c0 = bcast(c0v, a), c3 = bcast(c3v, a);
                const double ct = (c0 + srt * c3) * qt[t];
                const double cw = (c0 + srw * c3) * qw[t];
                ft = (a < na) ? ((a == 0) ? ct : ft + ct) : ft;
                fw = (a < na) ? ((a == 0) ? cw : fw + cw) : fw;
            }
        }
        if (!wl) fw = 0.0;
        if (lane >= L) return;
        tflux[i] = rut * ft;
        wflux[i] = ruw * fw;
        if (P.rk_step != 1) {
            const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
            const double face = 0.5 * (theta_m_save[c2 * L + k] + theta_m_save[c1 * L + k]);
            tpert[i] = m.dvEdge[e] * (ru_save[i] - rut) * face;
        }
    };
    const int base_ = wave_entity_base(nb);
#pragma unroll 1
    for (int t_ = 0; t_ < kEntPerWave; ++t_) body_(base_ + t_);
}

// P7: w and theta tendencies (advection, mixing finish, PGF/buoyancy,
// vertical transport, diabatic forcing), one thread per (cell, w-level).
// kOpt builds in the curvature and vertical-mixing terms
template <bool kOpt>
__device__ __forceinline__ void dyn_cells_b_body(long long nb, mpas_hip_mesh m, DynParams P, const double *__restrict__ wflux,
                              const double *__restrict__ rw, const double *__restrict__ tflux,
                              const double *__restrict__ rw_save, const double *__restrict__ tpert,
                              const double *__restrict__ w,
                              const double *__restrict__ theta_m, const double *__restrict__ theta_m_save,
                              const double *__restrict__ rho_zz, const double *__restrict__ ur_cell,
                              const double *__restrict__ vr_cell, const double *__restrict__ pp,
                              const double *__restrict__ cqw, const double *__restrict__ rt_diab,
                              const double *__restrict__ dpdz, const double *__restrict__ delsq_w,
                              const double *__restrict__ delsq_theta, const double *__restrict__ tend_rho,
                              double *__restrict__ tend_w_euler, double *__restrict__ tend_theta_euler,
                              double *__restrict__ tend_w, double *__restrict__ tend_theta,
                              double *__restrict__ rthdynten) {
    auto body_ = [&](const int c) {
        const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
        const int k = wave_lane();
        if (c > nC) return;  // wave-uniform
        const int i = c * Lw + k;
        if (c == nC) {
            if (k < Lw) tend_w[i] = 0.0;
            if (k < L) {
                tend_theta[c * L + k] = 0.0;
                rthdynten[c * L + k] = 0.0;
            }
            return;
        }
        const bool rk1 = (P.rk_step == 1);
        const int n = m.nEdgesOnCell[c];
        const double ia = m.invAreaCell[c];
        const double *fzm = m.fzm, *fzp = m.fzp;
        const int oc = c * L, ow = c * Lw;
        // Each lane loads its own level / interface of the column (clamped into
        // the column for the spare lanes); the neighbouring levels the vertical
        // stencils need come from the neighbouring lanes by shuffles, and each
        // lane forms the vertical flux at its own interface only.  All lanes
        // run this part.
        const int kl = k < L ? k : L - 1, kw = k < Lw ? k : L;
        const double w0 = w[ow + kw], rw0 = rw[ow + kw], rws0 = rw_save[ow + kw];
        const double th0 = theta_m[oc + kl], ths0 = theta_m_save[oc + kl];
        const double w_m1 = __shfl_up(w0, 1), w_p1 = __shfl_down(w0, 1), w_p2 = __shfl_down(w0, 2);
        const double rw_p1 = __shfl_down(rw0, 1);
        const double th_m2 = __shfl_up(th0, 2), th_m1 = __shfl_up(th0, 1), th_p1 = __shfl_down(th0, 1);
        const double ths_m1 = __shfl_up(ths0, 1);
        // vertical flux of w at mass level k (0..L-2)
        double wz = 0.0;
        if (k <= L - 2) {
            wz = (k == 0 || k == L - 2) ? 0.25 * (rw_p1 + rw0) * (w_p1 + w0)
                                        : flux3(w_m1, w0, w_p1, w_p2, 0.5 * (rw_p1 + rw0), 1.0);
        }
        const double wz_m1 = __shfl_up(wz, 1);
        // vertical flux of theta at interface k (1..L-1)
        double tz = 0.0;
        if (k >= 1 && k <= L - 1) {
            if (k == 1)
                tz = rw0 * (fzm[k] * th0 + fzp[k] * th_m1) + (rws0 - rw0) * (fzm[k] * ths0 + fzp[k] * ths_m1);
            else if (k == L - 1)
                tz = rws0 * (fzm[k] * th0 + fzp[k] * th_m1);
            else
                tz = flux3(th_m2, th_m1, th0, th_p1, rw0, P.coef_3rd_order) +
                     (rws0 - rw0) * (fzm[k] * ths0 + fzp[k] * ths_m1);
        }
        const double tz_p1 = __shfl_down(tz, 1);
        double pp0 = 0.0, dp0 = 0.0;
        if (rk1) {
            pp0 = pp[oc + kl];
            dp0 = dpdz[oc + kl];
        }
        const double pp_m1 = __shfl_up(pp0, 1), dp_m1 = __shfl_up(dp0, 1);
        if (k >= Lw) return;
        int eo[kMaxEdges];
        double sg[kMaxEdges];
        int nb_[kMaxEdges];  // the other cell of each edge
    #pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            const int jj = clamp_slot(j, n);
            eo[j] = m.edgesOnCell[c * ME + jj];
            sg[j] = m.edgesOnCell_sign[c * ME + jj];
            const int c1 = m.cellsOnEdge[2 * eo[j]], c2 = m.cellsOnEdge[2 * eo[j] + 1];
            nb_[j] = (c1 == c) ? c2 : c1;
        }
        // ---------------- w at interface k
        if (k >= 1 && k < L) {
            double fv[kMaxEdges];
    #pragma unroll
            for (int j = 0; j < kMaxEdges; ++j) {
                fv[j] = 0.0;
                if (j < n) fv[j] = wflux[eo[j] * L + k];  // wave-uniform guard
            }
            double hadv = 0.0;
    #pragma unroll
            for (int j = 0; j < kMaxEdges; ++j) {
                const double cc = -sg[j] * fv[j];
                hadv = (j < n) ? ((j == 0) ? cc : hadv + cc) : hadv;
            }
            if (kOpt && P.curvature) {
                // curvature (dyn_tend.py:4241-4267), added to the advective sum
                // before its 1/area scaling
                const double ur = fzm[k] * ur_cell[oc + k] + fzp[k] * ur_cell[oc + k - 1];
                const double vr = fzm[k] * vr_cell[oc + k] + fzp[k] * vr_cell[oc + k - 1];
                const double rh = rho_zz[oc + k] * fzm[k] + rho_zz[oc + k - 1] * fzp[k];
                hadv = hadv + rh * (ur * ur + vr * vr) / kEarthRadius;
                hadv = hadv + m.curvCell[c] * ur * rh;
            }
            double eu = tend_w_euler[i];
            if (rk1) {
                if (P.h_mom_eddy_visc4 > 0.0) {
                    double r_area = P.h_mom_eddy_visc4 * ia;
                    const double dsw = delsq_w[ow + k];
                    double dn[kMaxEdges];
    #pragma unroll
                    for (int j = 0; j < kMaxEdges; ++j) {
                        dn[j] = 0.0;
                        if (j < n) dn[j] = delsq_w[nb_[j] * Lw + k];
                    }
    #pragma unroll
                    for (int j = 0; j < kMaxEdges; ++j) {
                        const int e = eo[j];
                        const bool own2 = (m.cellsOnEdge[2 * e + 1] == c);
                        const double d2 = own2 ? dsw : dn[j], d1 = own2 ? dn[j] : dsw;
                        const double es = m.meshScalingDel4[e] * r_area * m.dvEdge[e] * sg[j] * m.invDcEdge[e];
                        const double nx = eu - es * (d2 - d1);
                        eu = (j < n) ? nx : eu;
                    }
                }
                eu = eu - cqw[i] * (m.rdzu[k] * (pp0 - pp_m1) - (fzm[k] * dp0 + fzp[k] * dp_m1));
                if (kOpt && P.v_mom_eddy_visc2 > 0.0) {
                    // vertical mixing of w (dyn_tend.py:4633-4647)
                    const double rho_half = 0.5 * (rho_zz[oc + k] + rho_zz[oc + k - 1]);
                    const double w_up = (w_p1 - w0) * m.rdzw[k], w_down = (w0 - w_m1) * m.rdzw[k - 1];
                    eu = eu + P.v_mom_eddy_visc2 * rho_half * (w_up - w_down) * m.rdzu[k];
                }
                tend_w_euler[i] = eu;
            }
            const double delta = (k == L - 1) ? -wz_m1 : wz - wz_m1;
            double inner = hadv * ia - m.rdzu[k] * delta;
            tend_w[i] = inner + eu;
        } else {
            tend_w[i] = 0.0;
        }
        // ---------------- theta at level k
        if (k < L) {
            const int o = c * L + k;
            double fv[kMaxEdges];
    #pragma unroll
            for (int j = 0; j < kMaxEdges; ++j) {
                fv[j] = 0.0;
                if (j < n) fv[j] = tflux[eo[j] * L + k];
            }
            double thadv = 0.0;
    #pragma unroll
            for (int j = 0; j < kMaxEdges; ++j) {
                const double cc = -sg[j] * fv[j];
                thadv = (j < n) ? ((j == 0) ? cc : thadv + cc) : thadv;
            }
            if (!rk1) {
    #pragma unroll
                for (int j = 0; j < kMaxEdges; ++j) {
                    fv[j] = 0.0;
                    if (j < n) fv[j] = tpert[eo[j] * L + k];
                }
    #pragma unroll
                for (int j = 0; j < kMaxEdges; ++j) {
                    const double nx = thadv - sg[j] * fv[j];
                    thadv = (j < n) ? nx : thadv;
                }
            }
            double tte = tend_theta_euler[o];
            if (rk1 && P.h_theta_eddy_visc4 > 0.0) {
                double r_area = P.h_theta_eddy_visc4 * P.prandtl_inv * ia;
                const double dst = delsq_theta[o];
                double dn[kMaxEdges];
    #pragma unroll
                for (int j = 0; j < kMaxEdges; ++j) {
                    dn[j] = 0.0;
                    if (j < n) dn[j] = delsq_theta[nb_[j] * L + k];
                }
    #pragma unroll
                for (int j = 0; j < kMaxEdges; ++j) {
                    const int e = eo[j];
                    const bool own2 = (m.cellsOnEdge[2 * e + 1] == c);
                    const double d2 = own2 ? dst : dn[j], d1 = own2 ? dn[j] : dst;
                    const double es = m.meshScalingDel4[e] * r_area * m.dvEdge[e] * sg[j] * m.invDcEdge[e];
                    const double nx = tte - es * (d2 - d1);
                    tte = (j < n) ? nx : tte;
                }
                tend_theta_euler[o] = tte;
            }
            if (kOpt && rk1 && P.v_theta_eddy_visc2 > 0.0 && k >= 1 && k <= L - 2) {
                // vertical mixing of theta_m, or of theta_m - t_init (dyn_tend.py:5060-5097)
                const VmixZ z = vmix_z_cell(m.zgrid + c * Lw, k);
                const double *ti = m.t_init + oc;
                const double a = P.mix_full ? th_m1 : th_m1 - ti[k - 1], b = P.mix_full ? th0 : th0 - ti[k];
                const double cth = P.mix_full ? th_p1 : th_p1 - ti[k + 1];
                tte = tte + P.v_theta_eddy_visc2 * P.prandtl_inv * rho_zz[o] *
                                ((cth - b) / (z.zp - z.z0) - (b - a) / (z.z0 - z.zm)) / (0.5 * (z.zp - z.zm));
                tend_theta_euler[o] = tte;
            }
            double delta;
            if (k == 0) delta = tz_p1;
            else if (k == L - 1) delta = -tz;
            else delta = tz_p1 - tz;
            double base = thadv * ia - m.rdzw[k] * delta;
            rthdynten[o] = (base - tend_rho[o] * th0) / rho_zz[o];
            double tt = base + rho_zz[o] * rt_diab[o];
            tend_theta[o] = tt + tte;
        }
    };
    const int base_ = wave_entity_base(nb);
#pragma unroll 1
    for (int t_ = 0; t_ < kEntPerWave; ++t_) body_(base_ + t_);
}
__global__ __launch_bounds__(kMaxWaveBlock) void k_dyn_cells_b(long long nb, mpas_hip_mesh m, DynParams P, const double *__restrict__ wflux,
                              const double *__restrict__ rw, const double *__restrict__ tflux,
                              const double *__restrict__ rw_save, const double *__restrict__ tpert,
                              const double *__restrict__ w,
                              const double *__restrict__ theta_m, const double *__restrict__ theta_m_save,
                              const double *__restrict__ rho_zz, const double *__restrict__ ur_cell,
                              const double *__restrict__ vr_cell, const double *__restrict__ pp,
                              const double *__restrict__ cqw, const double *__restrict__ rt_diab,
                              const double *__restrict__ dpdz, const double *__restrict__ delsq_w,
                              const double *__restrict__ delsq_theta, const double *__restrict__ tend_rho,
                              double *__restrict__ tend_w_euler, double *__restrict__ tend_theta_euler,
                              double *__restrict__ tend_w, double *__restrict__ tend_theta,
                              double *__restrict__ rthdynten) {
    dyn_cells_b_body<false>(nb, m, P, wflux, rw, tflux, rw_save, tpert, w, theta_m, theta_m_save, rho_zz,
                            ur_cell, vr_cell, pp, cqw, rt_diab, dpdz, delsq_w, delsq_theta, tend_rho,
                            tend_w_euler, tend_theta_euler, tend_w, tend_theta, rthdynten);
}
__global__ __launch_bounds__(kMaxWaveBlock) void k_dyn_cells_b_opt(long long nb, mpas_hip_mesh m, DynParams P, const double *__restrict__ wflux,
                              const double *__restrict__ rw, const double *__restrict__ tflux,
                              const double *__restrict__ rw_save, const double *__restrict__ tpert,
                              const double *__restrict__ w,
                              const double *__restrict__ theta_m, const double *__restrict__ theta_m_save,
                              const double *__restrict__ rho_zz, const double *__restrict__ ur_cell,
                              const double *__restrict__ vr_cell, const double *__restrict__ pp,
                              const double *__restrict__ cqw, const double *__restrict__ rt_diab,
                              const double *__restrict__ dpdz, const double *__restrict__ delsq_w,
                              const double *__restrict__ delsq_theta, const double *__restrict__ tend_rho,
                              double *__restrict__ tend_w_euler, double *__restrict__ tend_theta_euler,
                              double *__restrict__ tend_w, double *__restrict__ tend_theta,
                              double *__restrict__ rthdynten) {
    dyn_cells_b_body<true>(nb, m, P, wflux, rw, tflux, rw_save, tpert, w, theta_m, theta_m_save, rho_zz,
                           ur_cell, vr_cell, pp, cqw, rt_diab, dpdz, delsq_w, delsq_theta, tend_rho,
                           tend_w_euler, tend_theta_euler, tend_w, tend_theta, rthdynten);
}

void launch_dyn_tend(const mpas_hip_mesh &m, const DynParams &P, const DynInputs &in, const DynCarried &io,
                     const DynOutputs &out, double *scratch, hipStream_t s) {
    const int L = m.nLevels, nC = m.nCells, nE = m.nEdges, nV = m.nVertices;
    DynScratch sc = carve(m, scratch);
    const bool rk1 = (P.rk_step == 1);
    if (rk1)
        MPAS_WAVE_LAUNCH(k_dyn_kdiff, nC + 1, s, m, P, in.u,
                           in.v, io.kdiff);
    MPAS_WAVE_LAUNCH(k_dyn_cells_a, nC + 1, s, m, P,
                       in.ru, in.rw, in.rr_save, in.qtot, in.w, in.theta_m, in.rho_edge, io.kdiff, out.h_divergence,
                       io.tend_rho, sc.dpdz, sc.delsq_w, io.tend_w_euler, sc.delsq_theta, io.tend_theta_euler);
    // Without a del4 pass (rk_step > 1, or no fourth-order mixing) the edge
    // pass completes tend_u itself and the del4/finish pass is skipped.
    const int do_del4 = (rk1 && P.h_mom_eddy_visc4 > 0.0) ? 1 : 0;
    const bool opt = dyn_options(P);
    if (opt)
        MPAS_WAVE_LAUNCH(k_dyn_edges_a_opt, nE + 1, s, m, P, in.u, in.rw, in.w, in.pv_edge, in.ke, out.h_divergence,
                         in.rho_edge, in.pp, sc.dpdz, in.cqu, in.divergence, in.vorticity, io.kdiff, out.tend_u,
                         io.tend_u_euler, sc.delsq_u, do_del4 ? 0 : 1);
    else
        MPAS_WAVE_LAUNCH(k_dyn_edges_a, nE + 1, s, m, P, in.u, in.rw, in.w, in.pv_edge, in.ke, out.h_divergence,
                         in.rho_edge, in.pp, sc.dpdz, in.cqu, in.divergence, in.vorticity, io.kdiff, out.tend_u,
                         io.tend_u_euler, sc.delsq_u, do_del4 ? 0 : 1);
    if (do_del4) {
        MPAS_WAVE_LAUNCH(k_dyn_del4_vc, nC > nV ? nC : nV, s, m, sc.delsq_u, sc.delsq_vort, sc.delsq_div);
        MPAS_WAVE_LAUNCH(k_dyn_edges_b, nE, s, m, P, do_del4, in.u, in.rho_edge, sc.delsq_div, sc.delsq_vort,
                         io.tend_u_euler, out.tend_u);
    }
    MPAS_WAVE_LAUNCH(k_dyn_edge_adv, nE, s, m, P, in.ru, in.ru_save,
                in.w, in.theta_m, in.theta_m_save, sc.wflux, sc.tflux, sc.tpert);
    if (opt)
        MPAS_WAVE_LAUNCH(k_dyn_cells_b_opt, nC + 1, s, m, P, sc.wflux, in.rw, sc.tflux, in.rw_save, sc.tpert, in.w,
                         in.theta_m, in.theta_m_save, in.rho_zz, in.ur_cell, in.vr_cell, in.pp, in.cqw,
                         in.rt_diabatic_tend, sc.dpdz, sc.delsq_w, sc.delsq_theta, io.tend_rho, io.tend_w_euler,
                         io.tend_theta_euler, out.tend_w, out.tend_theta, out.rthdynten);
    else
        MPAS_WAVE_LAUNCH(k_dyn_cells_b, nC + 1, s, m, P, sc.wflux, in.rw, sc.tflux, in.rw_save, sc.tpert, in.w,
                         in.theta_m, in.theta_m_save, in.rho_zz, in.ur_cell, in.vr_cell, in.pp, in.cqw,
                         in.rt_diabatic_tend, sc.dpdz, sc.delsq_w, sc.delsq_theta, io.tend_rho, io.tend_w_euler,
                         io.tend_theta_euler, out.tend_w, out.tend_theta, out.rthdynten);
}

}  // namespace mpas
