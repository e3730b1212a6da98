// atm_compute_solve_diagnostics_work (solve_diagnostics.py:531-756):
// h_edge, kinetic energy with the 0.375 vertex blend, vorticity, divergence,
// APVM-upwinded pv_edge and (rk_step 3 / first call) the TRiSK tangential
// velocity.  Three gather passes (vertices, cells, edges), one wavefront per
// entity with lane = level; every slot sum runs left to right in the
// reference's slot order.  The edge kinetic energy dc dv u^2 is recomputed
// where it is used instead of being stored.
#include "common.hpp"

namespace mpas {

size_t diagnostics_scratch_doubles(const mpas_hip_mesh &m) {
    const size_t L = m.nLevels;
    return 2 * slice_doubles<hreal>((size_t)m.nVertices * L) + slice_doubles<double>((size_t)m.nCells * L);
}

__global__ __launch_bounds__(kMaxWaveBlock) void k_diag_vertices(long long nb, mpas_hip_mesh m, const double *__restrict__ u,
                                hreal *__restrict__ vort,
                                hreal *__restrict__ ke_vertex, hreal *__restrict__ pv_vertex) {
    auto body_ = [&](const int vx) {
        const int L = m.nLevels, nV = m.nVertices;
        const int k = wave_lane();
        if (vx > nV || k >= L) return;
        const int i = vx * L + k;
        if (vx == nV) {
            vort[i] = 0.0;
            return;
        }
        // the circulation (terms that nearly cancel) in binary64; the kinetic
        // energy (positive terms) in binary32 in the mixed build
        double s = 0.0;
        SlotSum sk;
    #pragma unroll
        for (int j = 0; j < 3; ++j) {
            int e = m.edgesOnVertex[3 * vx + j];
            double w = m.edgesOnVertex_sign[3 * vx + j] * m.dcEdge[e];
            double c = w * u[e * L + k];
            const freal ue = (freal)u[e * L + k];
            const freal kk = (freal)(m.dcEdge[e] * m.dvEdge[e]) * (ue * ue);  // ke_edge of the edge
            s = (j == 0) ? c : s + c;
            sk.slot(j, 3, kk);
        }
        double vv = s * m.invAreaTriangle[vx];
        vort[i] = vv;
        ke_vertex[i] = (freal)(0.25 * m.invAreaTriangle[vx]) * sk.value();
        pv_vertex[i] = m.fVertex[vx] + vv;
    };
    const int base_ = wave_entity_base(nb);
#pragma unroll 1
    for (int t_ = 0; t_ < kEntPerWave; ++t_) body_(base_ + t_);
}

__global__ __launch_bounds__(kMaxWaveBlock) void k_diag_cells(long long nb, mpas_hip_mesh m, const double *__restrict__ u,
                             const hreal *__restrict__ ke_vertex, const hreal *__restrict__ pv_vertex,
                             double *__restrict__ div, double *__restrict__ ke, double *__restrict__ pv_cell) {
    auto body_ = [&](const int c) {
        const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges;
        const int k = wave_lane();
        if (c > nC || k >= L) return;
        const int i = c * L + k;
        if (c == nC) {
            div[i] = 0.0;
            ke[i] = 0.0;
            return;
        }
        const int n = m.nEdgesOnCell[c];
        const double ia = m.invAreaCell[c];
        int eo[kMaxEdges], vo[kMaxEdges];
    #pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {  // indices first: one wait for all scalar loads
            eo[j] = m.edgesOnCell[c * ME + clamp_slot(j, n)];
            vo[j] = m.verticesOnCell[c * ME + clamp_slot(j, n)];
        }
        double uo[kMaxEdges];
    #pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            uo[j] = 0.0;
            if (j < n) uo[j] = u[eo[j] * L + k];  // wave-uniform guard
        }
        // the divergence (terms that nearly cancel) in binary64; the kinetic
        // energy in binary32 in the mixed build
        double d = 0.0;
        SlotSum kc;
    #pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            const int jj = clamp_slot(j, n);
            const int e = m.edgesOnCell[c * ME + jj];
            const double w = m.edgesOnCell_sign[c * ME + jj] * m.dvEdge[e];
            const freal ue = (freal)uo[j], kee = (freal)(m.dcEdge[e] * m.dvEdge[e]) * (ue * ue);
            const double cd = w * uo[j];
            const freal ck = (freal)0.25 * kee;
            d = (j < n) ? ((j == 0) ? cd : d + cd) : d;
            kc.slot(j, n, ck);
        }
        const freal fia = (freal)ia;
        div[i] = d * ia;
        const double ke_fact = 1.0 - 0.375;
        SlotSum tot, pc;
        tot.first((freal)ke_fact * (kc.value() * fia));
        freal kvv[kMaxEdges], pvv[kMaxEdges];
    #pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            kvv[j] = pvv[j] = 0.0;
            if (j < n) {
                kvv[j] = (freal)ke_vertex[vo[j] * L + k];
                pvv[j] = (freal)pv_vertex[vo[j] * L + k];
            }
        }
    #pragma unroll
        for (int j = 0; j < kMaxEdges; ++j) {
            const int jj = clamp_slot(j, n);
            const int vx = m.verticesOnCell[c * ME + jj];
            const double ka = m.kiteAreasOnVertex[3 * vx + m.kiteForCell[c * ME + jj]];
            const freal kv = kvv[j], pv = pvv[j];
            freal contrib = (freal)((1.0 - ke_fact) * ka);
            contrib = contrib * kv;
            contrib = contrib * fia;
            if (j < n) tot.add(contrib);
            const freal cp = (freal)ka * pv * fia;
            pc.slot(j, n, cp);
        }
        ke[i] = tot.value();
        pv_cell[i] = pc.value();
    };
    const int base_ = wave_entity_base(nb);
#pragma unroll 1
    for (int t_ = 0; t_ < kEntPerWave; ++t_) body_(base_ + t_);
}

__global__ __launch_bounds__(kMaxWaveBlock) void k_diag_edges2(long long nb, mpas_hip_mesh m, double apvm_dt, int apvm_on, int reconstruct_v,
                              const double *__restrict__ u, const hreal *__restrict__ v_in,
                              const hreal *__restrict__ pv_vertex, const double *__restrict__ pv_cell,
                              const double *__restrict__ h, hreal *__restrict__ h_edge,
                              hreal *__restrict__ v, hreal *__restrict__ pv_edge) {
    auto body_ = [&](const int e) {
        const int L = m.nLevels, nE = m.nEdges, ME2 = m.maxEdges2;
        const int lane = wave_lane();
        if (e > nE) return;
        if (e == nE) {
            if (lane < L) {
                h_edge[e * L + lane] = 0.0;
                v[e * L + lane] = 0.0;
                pv_edge[e * L + lane] = 0.0;
            }
            return;
        }
        const int k = lane < L ? lane : L - 1;  // lanes past the column: loads only
        const int i = e * L + k;
        freal vv;
        if (reconstruct_v) {
            SlotSum vs;
            vs.first(0.0);
            const int ne = m.nEdgesOnEdge[e];
            const int sl = lane < ne ? lane : ne - 1;
            const int eev = m.edgesOnEdge[e * ME2 + sl];
            const double wv = m.weightsOnEdge[e * ME2 + sl];
    #pragma unroll
            for (int j0 = 0; j0 < kMaxEdges2; j0 += 8) {
                if (j0 >= ne) break;
                freal uj[8];
    #pragma unroll
                for (int t = 0; t < 8; ++t) {
                    uj[t] = 0.0;
                    if (j0 + t < ne) uj[t] = (freal)u[bcast(eev, j0 + t) * L + k];  // wave-uniform guard
                }
    #pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const int j = j0 + t;
                    const freal c = (freal)bcast(wv, j) * uj[t];
                    vs.slot(j, ne, c);
                }
            }
            vv = vs.value();
        } else {
            vv = (freal)v_in[i];
        }
        if (lane >= L) return;
        v[i] = vv;
        int v1 = m.verticesOnEdge[2 * e], v2 = m.verticesOnEdge[2 * e + 1];
        const freal p1 = (freal)pv_vertex[v1 * L + k], p2 = (freal)pv_vertex[v2 * L + k];
        freal pe = (freal)0.5 * (p1 + p2);
        {
            const int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
            h_edge[i] = 0.5 * (h[c1 * L + k] + h[c2 * L + k]);
        }
        if (apvm_on) {
            int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
            const freal grad_pvt = (p2 - p1) * (freal)m.invDvEdge[e];
            const freal grad_pvn = ((freal)pv_cell[c2 * L + k] - (freal)pv_cell[c1 * L + k]) * (freal)m.invDcEdge[e];
            const freal delta = (freal)apvm_dt * (vv * grad_pvt + (freal)u[i] * grad_pvn);
            pe = pe - delta;
        }
        pv_edge[i] = pe;
    };
    const int base_ = wave_entity_base(nb);
#pragma unroll 1
    for (int t_ = 0; t_ < kEntPerWave; ++t_) body_(base_ + t_);
}

void launch_diagnostics(const mpas_hip_mesh &m, double dt, double apvm, int rk_step, const double *u,
                        const double *h, const hreal *v_in, hreal *h_edge, hreal *v, double *div,
                        hreal *vort, double *ke, hreal *pv_edge, double *scratch, hipStream_t s) {
    const int L = m.nLevels, nC = m.nCells, nE = m.nEdges, nV = m.nVertices;
    Carver k{reinterpret_cast<char *>(scratch)};
    hreal *ke_vertex = k.take<hreal>((size_t)nV * L);
    hreal *pv_vertex = k.take<hreal>((size_t)nV * L);
    double *pv_cell = k.take<double>((size_t)nC * L);
    // ke_vertex / pv_vertex / pv_cell have no padding row: the padding-row
    // wavefronts of the pass that fills them only zero that pass's outputs.
    MPAS_WAVE_LAUNCH(k_diag_vertices, nV + 1, s, m, u, vort, ke_vertex, pv_vertex);
    MPAS_WAVE_LAUNCH(k_diag_cells, nC + 1, s, m, u, ke_vertex, pv_vertex, div, ke,
                       pv_cell);
    const int reconstruct_v = (rk_step <= 0 || rk_step == 3) ? 1 : 0;
    MPAS_WAVE_LAUNCH(k_diag_edges2, nE + 1, s, m, apvm * dt, apvm > 0.0 ? 1 : 0,
                       reconstruct_v, u, v_in, pv_vertex, pv_cell, h, h_edge, v, pv_edge);
}

}  // namespace mpas
