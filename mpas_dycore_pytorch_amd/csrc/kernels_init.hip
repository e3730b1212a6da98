// Cold-start coupling of the raw initial fields: atm_init_coupled_diagnostics
// (mpas_pytorch/dynamics/init_coupled_diagnostics.py:147-380).
//   theta_m = theta (1 + rvord qv), rho_zz = rho / zz           (:240-262)
//   (or, with theta null, the theta_m and rho_zz given in the outputs)
//   ru = 0.5 u (rho_zz1 + rho_zz2)                             (:283)
//   rw = w rho_zz_i zz_i - sum_slots sgn (zb + sign(f) zb3) f zz_i (:285-343)
//   rho_p, rtheta_base, rtheta_p, exner, exner_base, pressure_p,
//   pressure_base                                              (:345-372)
// Three passes: cells (thermodynamics), edges (ru), cell interfaces (rw).
#include "common.hpp"

namespace mpas {

constexpr double kRvord = 461.6 / 287.0;

__global__ void k_icd_cells(mpas_hip_mesh m, const double *theta, const double *rho, const double *qv, const double *__restrict__ theta_base,
                            InitCoupledOutputs o) {
    const int L = m.nLevels, nC = m.nCells;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (nC + 1) * L) return;
    if (i >= nC * L) {  // padding row: zero, with the rho_zz sentinel 1
        o.theta_m[i] = 0.0;
        o.rho_zz[i] = 1.0;
        o.rho_p[i] = o.rtheta_base[i] = o.rtheta_p[i] = 0.0;
        o.exner[i] = o.exner_base[i] = o.pressure_p[i] = o.pressure_base[i] = 0.0;
        return;
    }
    const double rgas_over_p0 = MPAS_RGAS / MPAS_P0;
    const double rcv = MPAS_RGAS / (MPAS_CP - MPAS_RGAS);
    const double zz = m.zz[i], rb = m.rho_base[i], tb = theta_base[i];
    // raw fields, or (theta null) the theta_m / rho_zz already held in the
    // outputs: the reference's derived path (init_coupled_diagnostics.py:256-262)
    const double tm = theta ? theta[i] * (1.0 + kRvord * qv[i]) : o.theta_m[i];
    const double rz = theta ? rho[i] / zz : o.rho_zz[i];
    const double rp = rz - rb;
    const double rtb = tb * rb;
    const double rtp = tm * rp + rb * (tm - tb);
    const double ex = pow(zz * rgas_over_p0 * (rtp + rtb), rcv);
    const double exb = pow(zz * rgas_over_p0 * rtb, rcv);
    o.theta_m[i] = tm;
    o.rho_zz[i] = rz;
    o.rho_p[i] = rp;
    o.rtheta_base[i] = rtb;
    o.rtheta_p[i] = rtp;
    o.exner[i] = ex;
    o.exner_base[i] = exb;
    o.pressure_p[i] = zz * MPAS_RGAS * (ex * rtp + rtb * (ex - exb));
    o.pressure_base[i] = zz * MPAS_RGAS * exb * rtb;
}

__global__ void k_icd_edges(mpas_hip_mesh m, const double *__restrict__ u, const double *__restrict__ rho_zz,
                            double *__restrict__ ru) {
    const int L = m.nLevels, nE = m.nEdges;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (nE + 1) * L) return;
    int e = i / L, k = i - e * L;
    if (e == nE) {
        ru[i] = 0.0;
        return;
    }
    int c1 = m.cellsOnEdge[2 * e], c2 = m.cellsOnEdge[2 * e + 1];
    ru[i] = 0.5 * u[i] * (rho_zz[c1 * L + k] + rho_zz[c2 * L + k]);
}

__global__ void k_icd_rw(mpas_hip_mesh m, const double *__restrict__ w, const double *__restrict__ rho_zz,
                         const double *__restrict__ ru, double *__restrict__ rw) {
    const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges, Lw = L + 1;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (nC + 1) * Lw) return;
    int c = i / Lw, k = i - c * Lw;
    if (c == nC || k == 0 || k == L) {
        rw[i] = 0.0;
        return;
    }
    const int o = c * L;
    const double *zz = m.zz;
    const double ri = m.fzp[k] * rho_zz[o + k - 1] + m.fzm[k] * rho_zz[o + k];
    const double zi = m.fzp[k] * zz[o + k - 1] + m.fzm[k] * zz[o + k];
    double r = w[i] * ri * zi;
    if (!m.flatTerrain) {
        const int n = m.nEdgesOnCell[c];
        for (int j = 0; j < n; ++j) {
            int e = m.edgesOnCell[c * ME + j];
            double f = m.fzm[k] * ru[e * L + k] + m.fzp[k] * ru[e * L + k - 1];
            size_t zo = ((size_t)c * ME + j) * Lw + k;
            double contrib = m.edgesOnCell_sign[c * ME + j] * (m.zb_cell[zo] + sign_pm(f) * m.zb3_cell[zo]) * f * zi;
            r = r - contrib;
        }
    }
    rw[i] = r;
}

void launch_init_coupled(const mpas_hip_mesh &m, const double *theta, const double *rho, const double *qv,
                         const double *u, const double *w, const double *theta_base, const InitCoupledOutputs &o,
                         hipStream_t s) {
    const int L = m.nLevels, nC = m.nCells, nE = m.nEdges;
    MPAS_LAUNCH(k_icd_cells, dim3(blocks_for((long long)(nC + 1) * L)), dim3(kBlock), 0, s, m, theta, rho, qv,
                theta_base, o);
    MPAS_LAUNCH(k_icd_edges, dim3(blocks_for((long long)(nE + 1) * L)), dim3(kBlock), 0, s, m, u, o.rho_zz, o.ru);
    MPAS_LAUNCH(k_icd_rw, dim3(blocks_for((long long)(nC + 1) * (L + 1))), dim3(kBlock), 0, s, m, w, o.rho_zz, o.ru,
                o.rw);
}

// History diagnostics (output_writer.py:264-282) on the active rows:
// theta = theta_m / (1 + rvord qv), rho = rho_zz zz, pressure = p' + p_base.
__global__ __launch_bounds__(256) void k_history_fields(int nC, int L, int S, const double *__restrict__ theta_m,
                                                        const double *__restrict__ scalars,
                                                        const double *__restrict__ rho_zz,
                                                        const double *__restrict__ zz,
                                                        const double *__restrict__ pressure_p,
                                                        const double *__restrict__ pressure_base,
                                                        double *__restrict__ theta, double *__restrict__ rho,
                                                        double *__restrict__ pressure) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nC * L) return;
    const double qv = (scalars && S > 0) ? scalars[i * S] : 0.0;
    if (theta) theta[i] = theta_m[i] / (1.0 + kRvord * qv);
    if (rho) rho[i] = rho_zz[i] * zz[i];
    if (pressure) pressure[i] = pressure_p[i] + pressure_base[i];
}

void launch_history_fields(int nC, int L, int S, const double *theta_m, const double *scalars, const double *rho_zz,
                           const double *zz, const double *pressure_p, const double *pressure_base, double *theta,
                           double *rho, double *pressure, hipStream_t s) {
    const long long n = (long long)nC * L;
    if (n == 0) return;
    MPAS_LAUNCH(k_history_fields, dim3(blocks_for(n)), dim3(kBlock), 0, s, nC, L, S, theta_m, scalars, rho_zz, zz,
                pressure_p, pressure_base, theta, rho, pressure);
}

}  // namespace mpas
