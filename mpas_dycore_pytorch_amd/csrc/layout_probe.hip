// Measurement hook behind the choice of the level pitch of the device state:
// column streams at a given row pitch, timed with HIP events.
//
// The device state keeps each (entity, level) array with its levels innermost
// at pitch nLevels (440-byte columns at 55 levels).  A pitch of 56 makes every
// column a whole number of 64-byte sectors and 64 a whole number of 128-byte
// lines; this probe measures what that alignment buys for the access patterns
// of the large step's kernels: one wavefront per column, lane = level.
#include "common.hpp"

namespace mpas {
namespace probe {

// read three columns, write two (the shape of a cell or edge update)
__global__ __launch_bounds__(256) void k_probe_stream(long long nb, long long n, int L, int pitch, int lanes,
                                                      const double *__restrict__ a, const double *__restrict__ b,
                                                      const double *__restrict__ c, double *__restrict__ x,
                                                      double *__restrict__ y) {
    const long long col = xcd_block(nb) * 4 + (threadIdx.x >> 6);
    const int k = threadIdx.x & 63;
    if (col >= n || k >= lanes) return;
    const long long i = col * pitch + k;
    const bool live = k < L;
    const double va = live ? a[i] : 0.0, vb = live ? b[i] : 0.0, vc = live ? c[i] : 0.0;
    x[i] = va + vb * vc;
    y[i] = va - vc;
}

// read the own column and six neighbour columns of one array, write one
// (the shape of a stencil gather over a curve-ordered mesh)
__global__ __launch_bounds__(256) void k_probe_gather(long long nb, long long n, int L, int pitch, int lanes,
                                                      const double *__restrict__ a, double *__restrict__ x) {
    const long long col = xcd_block(nb) * 4 + (threadIdx.x >> 6);
    const int k = threadIdx.x & 63;
    if (col >= n || k >= lanes) return;
    double s = 0.0;
    if (k < L) {
        for (int d = -3; d <= 3; ++d) {
            long long cc = col + (long long)d * 97;
            cc = cc < 0 ? cc + n : (cc >= n ? cc - n : cc);
            s += a[cc * pitch + k];
        }
    }
    x[col * pitch + k] = s;
}

}  // namespace probe
}  // namespace mpas

using namespace mpas;
using namespace mpas::probe;

extern "C" int mpas_hip_layout_probe(int64_t nColumns, int32_t nLevels, int32_t pitch, int32_t variant,
                                     int32_t reps, double *ms) {
    if (nColumns < 1000 || nLevels < 1 || nLevels > 63 || pitch < nLevels || pitch > 64 || reps < 1 ||
        (variant != 0 && variant != 1) || !ms || nColumns < 4 * 97) {
        set_error("layout_probe: need nColumns >= 1000, 1 <= nLevels <= pitch <= 64, variant 0 or 1, reps >= 1");
        return 1;
    }
    const size_t bytes = (size_t)nColumns * pitch * sizeof(double);
    double *buf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int rc = 0;
    for (int j = 0; j < 5 && rc == 0; ++j)
        if (hipMalloc(&buf[j], bytes) != hipSuccess) rc = 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == 0 && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = 1;
    if (rc == 0) {
        for (int j = 0; j < 3; ++j) fill_zero(buf[j], (size_t)nColumns * pitch, nullptr);
        // at pitch nLevels + 1 the spare lane writes the column's padding element, so
        // every 64-byte sector of a column is written whole
        const int lanes = (pitch == nLevels + 1) ? pitch : nLevels;
        const long long nb = (nColumns + 3) / 4;
        auto launch = [&]() {
            if (variant == 0)
                MPAS_LAUNCH(k_probe_stream, dim3(xcd_grid(nb)), dim3(256), 0, nullptr, nb, (long long)nColumns,
                            (int)nLevels, (int)pitch, lanes, buf[0], buf[1], buf[2], buf[3], buf[4]);
            else
                MPAS_LAUNCH(k_probe_gather, dim3(xcd_grid(nb)), dim3(256), 0, nullptr, nb, (long long)nColumns,
                            (int)nLevels, (int)pitch, lanes, buf[0], buf[3]);
        };
        for (int r = 0; r < 3; ++r) launch();
        hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps; ++r) launch();
        hipEventRecord(e1, nullptr);
        float t = 0.0f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
            set_error("layout_probe: event timing failed");
            rc = 1;
        }
        *ms = (double)t / reps;
        if (rc == 0) rc = check_launch("layout_probe");
    } else {
        set_error("layout_probe: allocation failed");
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    for (int j = 0; j < 5; ++j)
        if (buf[j]) hipFree(buf[j]);
    return rc;
}
