// Physical constants and the scalar statement functions shared by the
// forward kernels and their adjoints (mpas_pytorch/constants.py:47-133).
// Only the __device__ / __forceinline__ keywords tie this file to HIP.
#pragma once

#include <cmath>

#define MPAS_GRAVITY 9.80616
#define MPAS_RGAS 287.0
#define MPAS_CP 1004.5
#define MPAS_P0 100000.0

namespace mpas {

__device__ __forceinline__ double sign_pm(double x) { return (x >= 0.0) ? 1.0 : -1.0; }
__device__ __forceinline__ double max0(double x) { return (x < 0.0) ? 0.0 : x; }
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (b > a) ? b : a; }

// flux4/flux3 statement functions (dyn_tend.py:839-890, advance_scalars.py:128-162)
__device__ __forceinline__ double flux4(double qm2, double qm1, double q0, double qp1, double ua) {
    return ua * (7.0 * (q0 + qm1) - (qp1 + qm2)) / 12.0;
}
__device__ __forceinline__ double flux3(double qm2, double qm1, double q0, double qp1, double ua,
                                        double coef3) {
    return flux4(qm2, qm1, q0, qp1, ua) + coef3 * fabs(ua) * ((qp1 - qm2) - 3.0 * (q0 - qm1)) / 12.0;
}

}  // namespace mpas
