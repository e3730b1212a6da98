// Column-local kernels: moist coefficients, vertical implicit coefficients
// with the Thomas factorisation, cell-vector reconstruction.
#include "common.hpp"

namespace mpas {

// ---------------------------------------------------------------------------
// atm_compute_moist_coefficients (moist_coefficients.py:124-291)
// ---------------------------------------------------------------------------
__global__ void k_moist_cells(int nC, int L, int S, int q0, int q1, const double *__restrict__ sc,
                              double *__restrict__ qtot) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nC * L) return;
    double t = 0.0;
    if (q1 > q0) {
        t = sc[(size_t)i * S + q0];
        for (int q = q0 + 1; q < q1; ++q) t = t + sc[(size_t)i * S + q];
    }
    qtot[i] = t;
}

__global__ void k_moist_cqw(int nC, int L, const double *__restrict__ qtot, double *__restrict__ cqw) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int Lw = L + 1;
    if (i >= (nC + 1) * Lw) return;
    int c = i / Lw, k = i - c * Lw;
    double r = 0.0;
    if (c < nC && k >= 1 && k < L) {
        double qw = 0.5 * (qtot[c * L + k] + qtot[c * L + k - 1]);
        r = 1.0 / (1.0 + qw);
    }
    cqw[i] = r;
}

__global__ void k_moist_cqu(int nE, int L, int S, int q0, int q1, const int *__restrict__ coe,
                            const double *__restrict__ sc, hreal *__restrict__ cqu) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (nE + 1) * L) return;
    int e = i / L, k = i - e * L;
    double r = 0.0;
    if (e < nE) {
        int c1 = coe[2 * e], c2 = coe[2 * e + 1];
        double qu = 0.0;
        if (q1 > q0) {
            qu = 0.5 * (sc[((size_t)c1 * L + k) * S + q0] + sc[((size_t)c2 * L + k) * S + q0]);
            for (int q = q0 + 1; q < q1; ++q)
                qu = qu + 0.5 * (sc[((size_t)c1 * L + k) * S + q] + sc[((size_t)c2 * L + k) * S + q]);
        }
        r = 1.0 / (1.0 + qu);
    }
    cqu[i] = r;
}

void launch_moist(const mpas_hip_mesh &m, const double *scalars, int S, int ms, int me, double *cqw,
                  hreal *cqu, double *qtot, hipStream_t s) {
    const int nC = m.nCells, nE = m.nEdges, L = m.nLevels;
    fill_zero(qtot + (size_t)nC * L, L, s);
    MPAS_LAUNCH(k_moist_cells, dim3(blocks_for((long long)nC * L)), dim3(kBlock), 0, s, nC, L, S,
                       ms - 1, me, scalars, qtot);
    MPAS_LAUNCH(k_moist_cqw, dim3(blocks_for((long long)(nC + 1) * (L + 1))), dim3(kBlock), 0, s,
                       nC, L, qtot, cqw);
    MPAS_LAUNCH(k_moist_cqu, dim3(blocks_for((long long)(nE + 1) * L)), dim3(kBlock), 0, s, nE, L,
                       S, ms - 1, me, m.cellsOnEdge, scalars, cqu);
}

// ---------------------------------------------------------------------------
// atm_compute_vert_imp_coefs_work (vert_imp_coefs.py:188-474): one wavefront
// per column, lane = level; the Thomas factorisation (:100-121) is the
// serial recurrence of lane 0 over the column held in LDS.
// ---------------------------------------------------------------------------
constexpr int kColWaves = 4;  // columns per 256-thread block

__global__ __launch_bounds__(64 * kColWaves) void k_vimp(
    mpas_hip_mesh m, double dtseps, double rcv, double c2, const double *__restrict__ cqw,
    const double *__restrict__ p, const double *__restrict__ t, const double *__restrict__ rt,
    const double *__restrict__ qtot, double *__restrict__ cofwr, double *__restrict__ cofwz,
    double *__restrict__ coftz, double *__restrict__ cofwt, double *__restrict__ a_tri,
    double *__restrict__ alpha_tri, double *__restrict__ gamma_tri) {
    __shared__ double s_coftz[kColWaves][65], s_cofwt[kColWaves][64], s_a[kColWaves][64],
        s_b[kColWaves][64], s_c[kColWaves][64];
    const int L = m.nLevels, nC = m.nCells;
    const int w = threadIdx.y, k = threadIdx.x;
    const int c = blockIdx.x * kColWaves + w;
    const bool live = c < nC;
    const int o = c * L, ow = c * (L + 1);
    const double *zz = m.zz, *rdzw = m.rdzw, *fzm = m.fzm, *fzp = m.fzp, *rdzu = m.rdzu;
    // stage 1: per-level coefficients
    if (live) {
        if (k < L) {
            double cwr = 0.0, cwz = 0.0, ctz = 0.0;
            if (k >= 1) {
                double zi = fzm[k] * zz[o + k] + fzp[k] * zz[o + k - 1];
                double pi = fzm[k] * p[o + k] + fzp[k] * p[o + k - 1];
                double ti = fzm[k] * t[o + k] + fzp[k] * t[o + k - 1];
                cwr = 0.5 * dtseps * MPAS_GRAVITY * zi;
                cwz = dtseps * c2 * zi * rdzu[k] * cqw[ow + k] * pi;
                ctz = dtseps * ti;
            }
            double cwt = 0.5 * dtseps * rcv * zz[o + k] * MPAS_GRAVITY * m.rho_base[o + k] /
                         (1.0 + qtot[o + k]) * p[o + k] /
                         ((m.rtheta_base[o + k] + rt[o + k]) * m.exner_base[o + k]);
            cofwr[o + k] = cwr;
            cofwz[o + k] = cwz;
            cofwt[o + k] = cwt;
            coftz[ow + k] = ctz;
            s_coftz[w][k] = ctz;
            s_cofwt[w][k] = cwt;
            s_a[w][k] = cwz;  // stash cofwz
            s_b[w][k] = cwr;  // stash cofwr
        } else if (k == L) {
            coftz[ow + L] = 0.0;
            s_coftz[w][L] = 0.0;
        }
    }
    __syncthreads();
    // stage 2: tridiagonal rows a, b, c (vert_imp_coefs.py:388-445)
    double a = 0.0, b = 1.0, cc = 0.0;
    if (live && k >= 1 && k < L) {
        const double cwz = s_a[w][k], cwr = s_b[w][k];
        const double cofrz_k = dtseps * rdzw[k], cofrz_km = dtseps * rdzw[k - 1];
        a = -cwz * s_coftz[w][k - 1] * rdzw[k - 1] * zz[o + k - 1] + cwr * cofrz_km -
            s_cofwt[w][k - 1] * s_coftz[w][k - 1] * rdzw[k - 1];
        b = 1.0 + cwz * (s_coftz[w][k] * rdzw[k] * zz[o + k] + s_coftz[w][k] * rdzw[k - 1] * zz[o + k - 1]) -
            s_coftz[w][k] * (s_cofwt[w][k] * rdzw[k] - s_cofwt[w][k - 1] * rdzw[k - 1]) +
            cwr * (cofrz_k - cofrz_km);
        cc = -cwz * s_coftz[w][k + 1] * rdzw[k] * zz[o + k] - cwr * cofrz_k +
             s_cofwt[w][k] * s_coftz[w][k + 1] * rdzw[k];
    }
    __syncthreads();
    if (live && k < L) {
        s_a[w][k] = a;
        s_b[w][k] = b;
        s_c[w][k] = cc;
        a_tri[o + k] = a;
    }
    __syncthreads();
    // stage 3: Thomas factorisation, lane 0 of each column.  Each factor is
    // written over the LDS row just read (b -> alpha, c -> gamma) and the
    // wavefront stores the column once, coalesced, instead of lane 0
    // storing one element at a time.
    if (live && k == 0) {
        double g = 0.0;
        for (int kk = 1; kk < L; ++kk) {
            double al = 1.0 / (s_b[w][kk] - s_a[w][kk] * g);
            g = s_c[w][kk] * al;
            s_b[w][kk] = al;
            s_c[w][kk] = g;
        }
    }
    __syncthreads();
    if (live && k < L) {
        alpha_tri[o + k] = (k == 0) ? 0.0 : s_b[w][k];
        gamma_tri[o + k] = (k == 0) ? 0.0 : s_c[w][k];
    }
}

__global__ void k_cofrz(int L, double dtseps, const double *__restrict__ rdzw, double *__restrict__ cofrz) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < L) cofrz[k] = dtseps * rdzw[k];
}

void launch_vimp(const mpas_hip_mesh &m, double dts, double epssm, const double *cqw, const double *p,
                 const double *t, const double *rt, const double *qtot, double *cofrz, double *cofwr,
                 double *cofwz, double *coftz, double *cofwt, double *a_tri, double *alpha_tri,
                 double *gamma_tri, hipStream_t s) {
    const int nC = m.nCells, L = m.nLevels;
    const double dtseps = 0.5 * dts * (1.0 + epssm);
    const double rcv = MPAS_RGAS / (MPAS_CP - MPAS_RGAS);
    const double c2 = MPAS_CP * rcv;
    MPAS_LAUNCH(k_cofrz, dim3(1), dim3(64), 0, s, L, dtseps, m.rdzw, cofrz);
    // padding rows stay zero
    fill_zero(cofwr + (size_t)nC * L, L, s);
    fill_zero(cofwz + (size_t)nC * L, L, s);
    fill_zero(cofwt + (size_t)nC * L, L, s);
    fill_zero(coftz + (size_t)nC * (L + 1), L + 1, s);
    fill_zero(a_tri + (size_t)nC * L, L, s);
    fill_zero(alpha_tri + (size_t)nC * L, L, s);
    fill_zero(gamma_tri + (size_t)nC * L, L, s);
    MPAS_LAUNCH(k_vimp, dim3((nC + kColWaves - 1) / kColWaves), dim3(64, kColWaves), 0, s, m, dtseps,
                       rcv, c2, cqw, p, t, rt, qtot, cofwr, cofwz, coftz, cofwt, a_tri, alpha_tri, gamma_tri);
}

// ---------------------------------------------------------------------------
// mpas_reconstruct_2d (vector_reconstruction.py:311-441), on a sphere
// ---------------------------------------------------------------------------
__global__ void k_reconstruct(mpas_hip_mesh m, const double *__restrict__ u, double *__restrict__ ux,
                              double *__restrict__ uy, double *__restrict__ uz, double *__restrict__ uzonal,
                              double *__restrict__ umerid) {
    const int L = m.nLevels, nC = m.nCells, ME = m.maxEdges;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (nC + 1) * L) return;
    int c = i / L, k = i - c * L;
    if (c == nC) {
        ux[i] = uy[i] = uz[i] = uzonal[i] = umerid[i] = 0.0;
        return;
    }
    int n = m.nEdgesOnCell[c];
    double x = 0.0, y = 0.0, z = 0.0;
    for (int j = 0; j < n; ++j) {
        int e = m.edgesOnCell[c * ME + j];
        const double *cf = m.coeffs_reconstruct + ((size_t)c * ME + j) * 3;
        double ue = u[e * L + k];
        double cx = cf[0] * ue, cy = cf[1] * ue, cz = cf[2] * ue;
        x = (j == 0) ? cx : x + cx;
        y = (j == 0) ? cy : y + cy;
        z = (j == 0) ? cz : z + cz;
    }
    double clat = cos(m.latCell[c]), slat = sin(m.latCell[c]);
    double clon = cos(m.lonCell[c]), slon = sin(m.lonCell[c]);
    ux[i] = x;
    uy[i] = y;
    uz[i] = z;
    uzonal[i] = -x * slon + y * clon;
    umerid[i] = -(x * clon + y * slon) * slat + z * clat;
}

void launch_reconstruct(const mpas_hip_mesh &m, const double *u, double *ux, double *uy, double *uz,
                        double *uzonal, double *umerid, hipStream_t s) {
    MPAS_LAUNCH(k_reconstruct, dim3(blocks_for((long long)(m.nCells + 1) * m.nLevels)), dim3(kBlock), 0,
                       s, m, u, ux, uy, uz, uzonal, umerid);
}

}  // namespace mpas
