// Batched tridiagonal factorization and solve with their adjoints and
// tangents: the stand-alone operators behind TridiagonalFactorize
// (autodiff/tridiagonal_factor.py:145-277) and TridiagonalSolve
// (autodiff/tridiagonal_solve.py:177-352).  The large step has its own copies
// of these recurrences fused into the vertical-implicit and acoustic kernels;
// these entry points serve callers of the operators themselves.
//
// Layout: a, b, c, alpha, gamma are (nColumns, nLevels) row-major; rhs and the
// solution are (nColumns, nLevels + 1), the last level being the boundary
// value the back substitution starts from.  Every column is an independent
// serial recurrence over its levels, so one lane runs one column: the lane
// walks its column's contiguous row, every 128-byte line is fetched once and
// consumed over 16 consecutive levels from the vector L1.  Entries that the
// recurrence leaves undefined (alpha, gamma and every derivative at level 0)
// are written as zero, as the reference's zero-initialised outputs are.
#include "common.hpp"

namespace mpas {
namespace tri {

__device__ __forceinline__ double at(const double *p, long long i) { return p ? p[i] : 0.0; }

// alpha(k) = 1 / (b(k) - a(k) gamma(k-1)), gamma(k) = c(k) alpha(k), k = 1..L-1
__global__ void k_tri_factor(long long n, int L, const double *__restrict__ a, const double *__restrict__ b,
                             const double *__restrict__ c, double *__restrict__ alpha, double *__restrict__ gamma) {
    const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n) return;
    const long long o = col * L;
    alpha[o] = 0.0;
    gamma[o] = 0.0;
    double gp = 0.0;
    for (int k = 1; k < L; ++k) {
        const double al = 1.0 / (b[o + k] - a[o + k] * gp);
        gp = c[o + k] * al;
        alpha[o + k] = al;
        gamma[o + k] = gp;
    }
}

// reverse sweep (tridiagonal_factor.py:183-230): gamma(k-1) feeds alpha(k), so
// its adjoint is carried down the column
__global__ void k_tri_factor_vjp(long long n, int L, const double *__restrict__ a, const double *__restrict__ c,
                                 const double *__restrict__ alpha, const double *__restrict__ gamma,
                                 const double *__restrict__ g_alpha, const double *__restrict__ g_gamma,
                                 double *__restrict__ g_a, double *__restrict__ g_b, double *__restrict__ g_c) {
    const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n) return;
    const long long o = col * L;
    double ag = at(g_gamma, o + L - 1);  // adjoint of gamma(k), k = L-1 first
    for (int k = L - 1; k >= 1; --k) {
        const double al = alpha[o + k];
        g_c[o + k] = ag * al;
        const double total = at(g_alpha, o + k) + ag * c[o + k];
        const double dd = -total * (al * al);  // adjoint of the pivot b - a gamma(k-1)
        g_b[o + k] = dd;
        g_a[o + k] = -dd * gamma[o + k - 1];
        ag = at(g_gamma, o + k - 1) - dd * a[o + k];
    }
    g_a[o] = 0.0;
    g_b[o] = 0.0;
    g_c[o] = 0.0;
}

// tangent sweep (tridiagonal_factor.py:233-277)
__global__ void k_tri_factor_jvp(long long n, int L, const double *__restrict__ a, const double *__restrict__ c,
                                 const double *__restrict__ alpha, const double *__restrict__ gamma,
                                 const double *__restrict__ da, const double *__restrict__ db,
                                 const double *__restrict__ dc, double *__restrict__ d_alpha,
                                 double *__restrict__ d_gamma) {
    const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n) return;
    const long long o = col * L;
    d_alpha[o] = 0.0;
    d_gamma[o] = 0.0;
    double dgp = 0.0;
    for (int k = 1; k < L; ++k) {
        const double al = alpha[o + k];
        const double term = at(db, o + k) - at(da, o + k) * gamma[o + k - 1] - a[o + k] * dgp;
        const double dal = -(al * al) * term;
        dgp = at(dc, o + k) * al + c[o + k] * dal;
        d_alpha[o + k] = dal;
        d_gamma[o + k] = dgp;
    }
}

// u(0) = r(0), u(k) = (r(k) - a(k) u(k-1)) alpha(k) for k = 1..L-1, u(L) = r(L);
// x(L) = u(L), x(k) = u(k) - gamma(k) x(k+1) for k = L-1..0
__global__ void k_tri_solve(long long n, int L, const double *__restrict__ a, const double *__restrict__ alpha,
                            const double *__restrict__ gamma, const double *__restrict__ rhs,
                            double *__restrict__ out) {
    const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n) return;
    const long long o = col * L, ow = col * (L + 1);
    double u = rhs[ow];
    out[ow] = u;
    for (int k = 1; k < L; ++k) {
        u = (rhs[ow + k] - a[o + k] * u) * alpha[o + k];
        out[ow + k] = u;
    }
    double x = rhs[ow + L];
    out[ow + L] = x;
    for (int k = L - 1; k >= 0; --k) {
        x = out[ow + k] - gamma[o + k] * x;
        out[ow + k] = x;
    }
}

// reverse sweeps (tridiagonal_solve.py:232-291): first through the back
// substitution (adjoint of u into g_rhs, gamma's adjoint), then through the
// elimination, whose intermediates u are recomputed into g_a.  Each buffer
// slot is read before the sweep overwrites it with its final value.
__global__ void k_tri_solve_vjp(long long n, int L, const double *__restrict__ a, const double *__restrict__ alpha,
                                const double *__restrict__ gamma, const double *__restrict__ rhs,
                                const double *__restrict__ out, const double *__restrict__ g_out,
                                double *__restrict__ g_a, double *__restrict__ g_alpha, double *__restrict__ g_gamma,
                                double *__restrict__ g_rhs) {
    const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n) return;
    const long long o = col * L, ow = col * (L + 1);
    // the elimination's u(0..L-1), kept in g_a
    double u = rhs[ow];
    g_a[o] = u;
    for (int k = 1; k < L; ++k) {
        u = (rhs[ow + k] - a[o + k] * u) * alpha[o + k];
        g_a[o + k] = u;
    }
    // back substitution, reversed: adjoint of u(k) into g_rhs(k)
    double carry = g_out[ow];
    for (int k = 0; k < L; ++k) {
        g_rhs[ow + k] = carry;
        g_gamma[o + k] = -carry * out[ow + k + 1];
        carry = g_out[ow + k + 1] - gamma[o + k] * carry;
    }
    g_rhs[ow + L] = carry;
    // elimination, reversed
    for (int k = L - 1; k >= 1; --k) {
        const double gu = g_rhs[ow + k];
        const double up = g_a[o + k - 1];
        const double t = rhs[ow + k] - a[o + k] * up;
        g_alpha[o + k] = gu * t;
        const double gt = gu * alpha[o + k];
        g_rhs[ow + k] = gt;
        g_a[o + k] = -gt * up;
        g_rhs[ow + k - 1] = g_rhs[ow + k - 1] - gt * a[o + k];
    }
    g_a[o] = 0.0;
    g_alpha[o] = 0.0;
}

// tangent sweeps (tridiagonal_solve.py:294-352)
__global__ void k_tri_solve_jvp(long long n, int L, const double *__restrict__ a, const double *__restrict__ alpha,
                                const double *__restrict__ gamma, const double *__restrict__ rhs,
                                const double *__restrict__ out, const double *__restrict__ da,
                                const double *__restrict__ dalpha, const double *__restrict__ dgamma,
                                const double *__restrict__ drhs, double *__restrict__ dout) {
    const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n) return;
    const long long o = col * L, ow = col * (L + 1);
    double u = rhs[ow], du = at(drhs, ow);
    dout[ow] = du;
    for (int k = 1; k < L; ++k) {
        const double t = rhs[ow + k] - a[o + k] * u;
        const double dt = at(drhs, ow + k) - at(da, o + k) * u - a[o + k] * du;
        du = dt * alpha[o + k] + t * at(dalpha, o + k);
        u = t * alpha[o + k];
        dout[ow + k] = du;
    }
    double dx = at(drhs, ow + L);
    dout[ow + L] = dx;
    for (int k = L - 1; k >= 0; --k) {
        dx = dout[ow + k] - at(dgamma, o + k) * out[ow + k + 1] - gamma[o + k] * dx;
        dout[ow + k] = dx;
    }
}

inline int checked(long long n, int L, const char *what) {
    if (n < 0 || L < 1) {
        set_error(std::string(what) + ": need nColumns >= 0 and nLevels >= 1");
        return 1;
    }
    return 0;
}

}  // namespace tri
}  // namespace mpas

using namespace mpas;
using namespace mpas::tri;

#define TRI_LAUNCH(kernel, n, stream, ...)                                                               \
    do {                                                                                                 \
        if ((n) > 0) MPAS_LAUNCH(kernel, dim3(blocks_for(n)), dim3(kBlock), 0, stream, n, __VA_ARGS__); \
    } while (0)

extern "C" {

int mpas_hip_tridiagonal_factorize(int64_t nColumns, int32_t nLevels, const double *a, const double *b,
                                   const double *c, double *alpha, double *gamma, void *stream) {
    if (checked(nColumns, nLevels, "tridiagonal_factorize")) return 1;
    TRI_LAUNCH(k_tri_factor, (long long)nColumns, (hipStream_t)stream, nLevels, a, b, c, alpha, gamma);
    return check_launch("tridiagonal_factorize");
}

int mpas_hip_tridiagonal_factorize_vjp(int64_t nColumns, int32_t nLevels, const double *a, const double *c,
                                       const double *alpha, const double *gamma, const double *grad_alpha,
                                       const double *grad_gamma, double *grad_a, double *grad_b, double *grad_c,
                                       void *stream) {
    if (checked(nColumns, nLevels, "tridiagonal_factorize_vjp")) return 1;
    TRI_LAUNCH(k_tri_factor_vjp, (long long)nColumns, (hipStream_t)stream, nLevels, a, c, alpha, gamma, grad_alpha,
               grad_gamma, grad_a, grad_b, grad_c);
    return check_launch("tridiagonal_factorize_vjp");
}

int mpas_hip_tridiagonal_factorize_jvp(int64_t nColumns, int32_t nLevels, const double *a, const double *c,
                                       const double *alpha, const double *gamma, const double *da, const double *db,
                                       const double *dc, double *d_alpha, double *d_gamma, void *stream) {
    if (checked(nColumns, nLevels, "tridiagonal_factorize_jvp")) return 1;
    TRI_LAUNCH(k_tri_factor_jvp, (long long)nColumns, (hipStream_t)stream, nLevels, a, c, alpha, gamma, da, db, dc,
               d_alpha, d_gamma);
    return check_launch("tridiagonal_factorize_jvp");
}

int mpas_hip_tridiagonal_solve(int64_t nColumns, int32_t nLevels, const double *a, const double *alpha,
                               const double *gamma, const double *rhs, double *out, void *stream) {
    if (checked(nColumns, nLevels, "tridiagonal_solve")) return 1;
    TRI_LAUNCH(k_tri_solve, (long long)nColumns, (hipStream_t)stream, nLevels, a, alpha, gamma, rhs, out);
    return check_launch("tridiagonal_solve");
}

int mpas_hip_tridiagonal_solve_vjp(int64_t nColumns, int32_t nLevels, const double *a, const double *alpha,
                                   const double *gamma, const double *rhs, const double *out, const double *grad_out,
                                   double *grad_a, double *grad_alpha, double *grad_gamma, double *grad_rhs,
                                   void *stream) {
    if (checked(nColumns, nLevels, "tridiagonal_solve_vjp")) return 1;
    TRI_LAUNCH(k_tri_solve_vjp, (long long)nColumns, (hipStream_t)stream, nLevels, a, alpha, gamma, rhs, out,
               grad_out, grad_a, grad_alpha, grad_gamma, grad_rhs);
    return check_launch("tridiagonal_solve_vjp");
}

int mpas_hip_tridiagonal_solve_jvp(int64_t nColumns, int32_t nLevels, const double *a, const double *alpha,
                                   const double *gamma, const double *rhs, const double *out, const double *da,
                                   const double *d_alpha, const double *d_gamma, const double *d_rhs, double *d_out,
                                   void *stream) {
    if (checked(nColumns, nLevels, "tridiagonal_solve_jvp")) return 1;
    TRI_LAUNCH(k_tri_solve_jvp, (long long)nColumns, (hipStream_t)stream, nLevels, a, alpha, gamma, rhs, out, da,
               d_alpha, d_gamma, d_rhs, d_out);
    return check_launch("tridiagonal_solve_jvp");
}

}  // extern "C"
