// C ABI and the orchestration of one split-explicit large step
// (atm_srk3 / _atm_srk3_impl, mpas_pytorch/dynamics/time_integration.py:
// 3420-4894, stage body execute_rk_step :5000-9154).  Host code only
// enqueues kernels on the caller's stream: no allocation, no host sync.
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#if defined(MPAS_MIXED) && MPAS_MIXED
#include "../../include/mpas_hip_mixed.h"
#endif

namespace mpas {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

// ---------------------------------------------------------------------------
// per-kernel event timer (host state; used only by measurement code)
// ---------------------------------------------------------------------------
struct KernelTimer {
    std::string name;
    std::vector<hipEvent_t> start, stop;
    int used = 0;
    bool armed = false;
};
static KernelTimer g_timer;

bool timer_armed_for(const char *kernel_name) {
    return g_timer.armed && g_timer.used < (int)g_timer.start.size() && g_timer.name == kernel_name;
}
void timer_before(hipStream_t s) { hipEventRecord(g_timer.start[g_timer.used], s); }
void timer_after(hipStream_t s) {
    hipEventRecord(g_timer.stop[g_timer.used], s);
    ++g_timer.used;
}

static void timer_release() {
    for (auto e : g_timer.start) hipEventDestroy(e);
    for (auto e : g_timer.stop) hipEventDestroy(e);
    g_timer.start.clear();
    g_timer.stop.clear();
    g_timer.used = 0;
    g_timer.armed = false;
}

// ---------------------------------------------------------------------------
// launch-shape table written by the autotuner: wavefronts per block of the
// wavefront-per-entity kernels (fixed capacity, so lookups never see a
// reallocation)
// ---------------------------------------------------------------------------
struct TuneEntry {
    char name[64];
    int waves;  // 0: the default
    int order;  // 0: chunks per XCD (the default), 1: linear
};
constexpr int kTuneCapacity = 128;
static TuneEntry g_tune[kTuneCapacity];
static int g_tune_n = 0;

int tuned_waves(const char *kernel_name) {
    for (int i = 0; i < g_tune_n; ++i)
        if (g_tune[i].waves > 0 && std::strcmp(g_tune[i].name, kernel_name) == 0) return g_tune[i].waves;
    return kWaves;
}

int tuned_order(const char *kernel_name) {
    for (int i = 0; i < g_tune_n; ++i)
        if (std::strcmp(g_tune[i].name, kernel_name) == 0) return g_tune[i].order;
    return 0;
}

// the table entry of a kernel, added if absent (nullptr when the table is full)
static TuneEntry *tune_entry(const char *kernel_name) {
    for (int i = 0; i < g_tune_n; ++i)
        if (std::strcmp(g_tune[i].name, kernel_name) == 0) return &g_tune[i];
    if (g_tune_n == kTuneCapacity) return nullptr;
    TuneEntry &t = g_tune[g_tune_n];
    std::strcpy(t.name, kernel_name);
    t.waves = 0;
    t.order = 0;
    ++g_tune_n;
    return &t;
}

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return 1;
    }
    return 0;
}

__global__ void k_fill_zero(double *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void k_copy(double *__restrict__ dst, const double *__restrict__ src, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_add(double *__restrict__ dst, const double *__restrict__ src, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = dst[i] + src[i];
}

__global__ void k_scale(double *__restrict__ dst, const double *__restrict__ src, double f, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] * f;
}

void fill_zero(double *p, size_t n, hipStream_t s) {
    if (n == 0) return;
    MPAS_LAUNCH(k_fill_zero, dim3(blocks_for((long long)n)), dim3(kBlock), 0, s, p, n);
}

static void copy_d(double *dst, const double *src, size_t n, hipStream_t s) {
    MPAS_LAUNCH(k_copy, dim3(blocks_for((long long)n)), dim3(kBlock), 0, s, dst, src, n);
}
static void add_d(double *dst, const double *src, size_t n, hipStream_t s) {
    MPAS_LAUNCH(k_add, dim3(blocks_for((long long)n)), dim3(kBlock), 0, s, dst, src, n);
}
static void scale_d(double *dst, const double *src, double f, size_t n, hipStream_t s) {
    MPAS_LAUNCH(k_scale, dim3(blocks_for((long long)n)), dim3(kBlock), 0, s, dst, src, f, n);
}

#if defined(MPAS_MIXED) && MPAS_MIXED
// binary32 forms for the edge/vertex fields of the mixed-precision step
__global__ void k_fill_zero32(float *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}
__global__ void k_copy32(float *__restrict__ dst, const float *__restrict__ src, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}
__global__ void k_add32(float *__restrict__ dst, const float *__restrict__ src, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (float)((double)dst[i] + (double)src[i]);
}
__global__ void k_scale32(float *__restrict__ dst, const float *__restrict__ src, double f, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (float)((double)src[i] * f);
}
void fill_zero(float *p, size_t n, hipStream_t s) {
    if (n == 0) return;
    MPAS_LAUNCH(k_fill_zero32, dim3(blocks_for((long long)n)), dim3(kBlock), 0, s, p, n);
}
static void copy_d(float *dst, const float *src, size_t n, hipStream_t s) {
    MPAS_LAUNCH(k_copy32, dim3(blocks_for((long long)n)), dim3(kBlock), 0, s, dst, src, n);
}
static void add_d(float *dst, const float *src, size_t n, hipStream_t s) {
    MPAS_LAUNCH(k_add32, dim3(blocks_for((long long)n)), dim3(kBlock), 0, s, dst, src, n);
}
static void scale_d(float *dst, const float *src, double f, size_t n, hipStream_t s) {
    MPAS_LAUNCH(k_scale32, dim3(blocks_for((long long)n)), dim3(kBlock), 0, s, dst, src, f, n);
}
#endif

// ---------------------------------------------------------------------------
// workspace layout
// ---------------------------------------------------------------------------
struct Work {
    double *ru_save;
    double *rw_save, *rtheta_p_save, *rho_p_save, *theta_m_init, *rho_zz_old_split;
    double *scalars_old, *scalars_tmp;
    double *ruAvg_split;
    double *wwAvg_split;
    double *cqw;
    hreal *cqu;
    double *qtot;
    double *cofrz, *cofwr, *cofwz, *coftz, *cofwt, *a_tri, *alpha_tri, *gamma_tri;
    double *tend_rho;
    hreal *tend_u_euler;
    double *tend_w_euler, *tend_theta_euler, *kdiff;
    hreal *tend_u;
    double *tend_w, *tend_theta, *h_divergence, *rthdynten;
    double *ru_p, *ruAvg;
    double *rho_pp, *rtheta_pp, *rtheta_pp_old, *rw_p, *wwAvg;
    double *rho_zz_int;
    hreal *horiz_flux;
    double *scratch;
    size_t total;  // doubles
};

static size_t kernel_scratch(const mpas_hip_mesh &m) {
    size_t a = dyn_tend_scratch_doubles(m), b = diagnostics_scratch_doubles(m), c = mono_scratch_doubles(m);
    size_t r = a > b ? a : b;
    return r > c ? r : c;
}

static Work carve_work(const mpas_hip_mesh &m, int S, double *base) {
    const size_t L = m.nLevels, C = (size_t)(m.nCells + 1) * L, E = (size_t)(m.nEdges + 1) * L,
                 W = (size_t)(m.nCells + 1) * (L + 1);
    Work w;
    Carver k{reinterpret_cast<char *>(base)};  // 256-byte aligned slices
    w.ru_save = k.take<double>(E);
    w.rw_save = k.take<double>(W);
    w.rtheta_p_save = k.take<double>(C);
    w.rho_p_save = k.take<double>(C);
    w.theta_m_init = k.take<double>(C);
    w.rho_zz_old_split = k.take<double>(C);
    w.scalars_old = k.take<double>(C * S);
    w.scalars_tmp = k.take<double>(C * S);
    w.ruAvg_split = k.take<double>(E);
    w.wwAvg_split = k.take<double>(W);
    w.cqw = k.take<double>(W);
    w.cqu = k.take<hreal>(E);
    w.qtot = k.take<double>(C);
    w.cofrz = k.take<double>(L);
    w.cofwr = k.take<double>(C);
    w.cofwz = k.take<double>(C);
    w.coftz = k.take<double>(W);
    w.cofwt = k.take<double>(C);
    w.a_tri = k.take<double>(C);
    w.alpha_tri = k.take<double>(C);
    w.gamma_tri = k.take<double>(C);
    w.tend_rho = k.take<double>(C);
    w.tend_u_euler = k.take<hreal>(E);
    w.tend_w_euler = k.take<double>(W);
    w.tend_theta_euler = k.take<double>(C);
    w.kdiff = k.take<double>(C);
    w.tend_u = k.take<hreal>(E);
    w.tend_w = k.take<double>(W);
    w.tend_theta = k.take<double>(C);
    w.h_divergence = k.take<double>(C);
    w.rthdynten = k.take<double>(C);
    w.ru_p = k.take<double>(E);
    w.ruAvg = k.take<double>(E);
    w.rho_pp = k.take<double>(C);
    w.rtheta_pp = k.take<double>(C);
    w.rtheta_pp_old = k.take<double>(C);
    w.rw_p = k.take<double>(W);
    w.wwAvg = k.take<double>(W);
    w.rho_zz_int = k.take<double>(C);
    w.horiz_flux = k.take<hreal>(E * (S > 0 ? S : 1));
    w.scratch = k.take<double>(kernel_scratch(m));
    w.total = (size_t)(k.p - reinterpret_cast<char *>(base)) / sizeof(double);
    return w;
}

// The state as the step sees it: the edge/vertex fields in hreal.
struct StepState {
    double *u;
    double *w, *theta_m, *rho_zz, *scalars, *rt_diabatic_tend;
    double *rho_p, *rtheta_p, *exner, *pressure_p;
    double *ru;
    double *rw;
    hreal *v;
    double *divergence;
    hreal *vorticity;
    double *ke;
    hreal *pv_edge, *rho_edge;
    double *ur_cell, *vr_cell;
    double *uReconstructX, *uReconstructY, *uReconstructZ;
    double *uReconstructZonal, *uReconstructMeridional;
    double *ruAvg;
    double *wwAvg;
};
template <typename St>
static StepState step_state(St *p) {
    return StepState{p->u, p->w, p->theta_m, p->rho_zz, p->scalars, p->rt_diabatic_tend, p->rho_p, p->rtheta_p,
                     p->exner, p->pressure_p, p->ru, p->rw, p->v, p->divergence, p->vorticity, p->ke, p->pv_edge,
                     p->rho_edge, p->ur_cell, p->vr_cell, p->uReconstructX, p->uReconstructY, p->uReconstructZ,
                     p->uReconstructZonal, p->uReconstructMeridional, p->ruAvg, p->wwAvg};
}

// The curvature terms read the mesh's Coriolis factors and the cell winds.
static int validate_curvature(const mpas_hip_mesh *m, int curvature, const void *ur_cell, const void *vr_cell) {
    if (curvature && (!m->curvEdge || !m->curvCell || !ur_cell || !vr_cell)) {
        set_error("curvature terms need the mesh's curvEdge/curvCell and the state's ur_cell/vr_cell");
        return 1;
    }
    return 0;
}

static int validate_mesh(const mpas_hip_mesh *m) {
    if (!m) {
        set_error("null mesh");
        return 1;
    }
    if (m->nLevels < 3 || m->nLevels > 63) {
        set_error("nLevels must be in [3, 63] (one wavefront per column)");
        return 1;
    }
    if (m->maxEdges < 1 || m->maxEdges > kMaxEdges || m->maxEdges2 < 1 || m->maxEdges2 > kMaxEdges2 ||
        m->maxAdv < 2 || m->maxAdv > kMaxAdv || m->vertexDegree != 3) {
        set_error("mesh slot widths exceed the compiled caps (maxEdges <= 8, maxEdges2 <= 16, "
                  "maxAdv <= 16, vertexDegree == 3)");
        return 1;
    }
    long long big = (long long)(m->nEdges + 1) * (m->nLevels + 1) * 16;
    if (big >= (1LL << 31)) {
        set_error("mesh too large for 32-bit flat indexing");
        return 1;
    }
    return 0;
}

#if !(defined(MPAS_MIXED) && MPAS_MIXED)
// ---------------------------------------------------------------------------
// Workspace of mpas_hip_srk3_step_noalloc: a file-scope device array, so the
// call needs no allocation.  Every GPU the library runs on holds its own copy
// (128 MiB) from the module's load until the process ends; its device address
// is looked up once per device (hipGetSymbolAddress) and kept.  All calls on
// one GPU share the array: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr size_t kStaticWorkDoubles = (size_t)16 << 20;  // 128 MiB of doubles
static __device__ __align__(256) double step_static_work[kStaticWorkDoubles];

constexpr int kStaticWorkDevices = 64;
static double *g_static_work_addr[kStaticWorkDevices];
static std::mutex g_static_work_mu;

static double *static_work_address() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kStaticWorkDevices) {
        set_error("srk3_step_noalloc: no current HIP device (or device index beyond 63)");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_static_work_mu);
    if (!g_static_work_addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(step_static_work));
        if (e != hipSuccess || !p) {
            set_error(std::string("srk3_step_noalloc: hipGetSymbolAddress of the static workspace failed: ") +
                      hipGetErrorString(e));
            return nullptr;
        }
        g_static_work_addr[dev] = static_cast<double *>(p);
    }
    return g_static_work_addr[dev];
}
#endif

static int validate_moist(int S, int ms, int me) {
    if (S < 0 || ms < 1 || me > S) {
        set_error("moist species range must satisfy 1 <= moist_start and moist_end <= nScalars");
        return 1;
    }
    return 0;
}

}  // namespace mpas

using namespace mpas;

#define STREAM(s) ((hipStream_t)(s))

extern "C" {

int mpas_hip_tuning_set(const char *kernel_name, int32_t waves_per_block) {
    if (!kernel_name || std::strlen(kernel_name) >= sizeof(g_tune[0].name)) {
        set_error("tuning_set: kernel name missing or too long");
        return 1;
    }
    if (waves_per_block != 0 && waves_per_block != 1 && waves_per_block != 2 && waves_per_block != 4 &&
        waves_per_block != 8 && waves_per_block != 16) {
        set_error("tuning_set: waves per block must be 1, 2, 4, 8 or 16 (0 restores the default)");
        return 1;
    }
    TuneEntry *t = tune_entry(kernel_name);
    if (!t) {
        set_error("tuning_set: table full");
        return 1;
    }
    t->waves = waves_per_block;
    return 0;
}

int32_t mpas_hip_tuning_get(const char *kernel_name) { return kernel_name ? tuned_waves(kernel_name) : kWaves; }

int mpas_hip_tuning_set_order(const char *kernel_name, int32_t order) {
    if (!kernel_name || std::strlen(kernel_name) >= sizeof(g_tune[0].name)) {
        set_error("tuning_set_order: kernel name missing or too long");
        return 1;
    }
    if (order != 0 && order != 1) {
        set_error("tuning_set_order: the block order is 0 (chunks per XCD) or 1 (linear)");
        return 1;
    }
    TuneEntry *t = tune_entry(kernel_name);
    if (!t) {
        set_error("tuning_set_order: table full");
        return 1;
    }
    t->order = order;
    return 0;
}

int32_t mpas_hip_tuning_get_order(const char *kernel_name) { return kernel_name ? tuned_order(kernel_name) : 0; }

const char *mpas_hip_last_error(void) {
    return g_last_error.c_str();
}

int mpas_hip_timer_arm(const char *kernel_name, int32_t capacity) {
    timer_release();
    if (!kernel_name || capacity < 1) {
        set_error("timer_arm: need a kernel name and capacity >= 1");
        return 1;
    }
    g_timer.name = kernel_name;
    g_timer.start.resize(capacity);
    g_timer.stop.resize(capacity);
    for (int i = 0; i < capacity; ++i) {
        if (hipEventCreate(&g_timer.start[i]) != hipSuccess || hipEventCreate(&g_timer.stop[i]) != hipSuccess) {
            set_error("timer_arm: hipEventCreate failed");
            return 1;
        }
    }
    g_timer.used = 0;
    g_timer.armed = true;
    return 0;
}

int mpas_hip_timer_read_each(double *ms_each, int32_t capacity, int32_t *launches) {
    if (!ms_each || !launches || capacity < g_timer.used) {
        set_error("timer_read_each: need an output array of at least the recorded launch count");
        return 1;
    }
    for (int i = 0; i < g_timer.used; ++i) {
        float ms = 0.0f;
        if (hipEventSynchronize(g_timer.stop[i]) != hipSuccess ||
            hipEventElapsedTime(&ms, g_timer.start[i], g_timer.stop[i]) != hipSuccess) {
            set_error("timer_read_each: event query failed");
            return 1;
        }
        ms_each[i] = ms;
    }
    *launches = g_timer.used;
    timer_release();
    return 0;
}

int mpas_hip_timer_read(double *total_ms, int32_t *launches) {
    double t = 0.0;
    for (int i = 0; i < g_timer.used; ++i) {
        if (hipEventSynchronize(g_timer.stop[i]) != hipSuccess) {
            set_error("timer_read: hipEventSynchronize failed");
            return 1;
        }
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, g_timer.start[i], g_timer.stop[i]) != hipSuccess) {
            set_error("timer_read: hipEventElapsedTime failed");
            return 1;
        }
        t += ms;
    }
    *total_ms = t;
    *launches = g_timer.used;
    timer_release();
    return 0;
}

#if defined(MPAS_MIXED) && MPAS_MIXED
int32_t mpas_hip_mixed_state_size(void) { return (int32_t)sizeof(mpas_hip_mixed_state); }

int64_t mpas_hip_mixed_workspace_doubles(int32_t nCells, int32_t nEdges, int32_t nVertices, int32_t nLevels,
                                         int32_t nScalars) {
    mpas_hip_mesh m;
    std::memset(&m, 0, sizeof(m));
    m.nCells = nCells;
    m.nEdges = nEdges;
    m.nVertices = nVertices;
    m.nLevels = nLevels;
    Work w = carve_work(m, nScalars, nullptr);
    return (int64_t)w.total;
}
#else
int mpas_hip_struct_sizes(int32_t *mesh, int32_t *state, int32_t *workspace, int32_t *config) {
    *mesh = (int32_t)sizeof(mpas_hip_mesh);
    *state = (int32_t)sizeof(mpas_hip_state);
    *workspace = (int32_t)sizeof(mpas_hip_workspace);
    *config = (int32_t)sizeof(mpas_hip_config);
    return 0;
}

int64_t mpas_hip_kernel_scratch_doubles(int32_t nCells, int32_t nEdges, int32_t nVertices, int32_t nLevels) {
    mpas_hip_mesh m;
    std::memset(&m, 0, sizeof(m));
    m.nCells = nCells;
    m.nEdges = nEdges;
    m.nVertices = nVertices;
    m.nLevels = nLevels;
    return (int64_t)kernel_scratch(m);
}

int64_t mpas_hip_workspace_doubles(int32_t nCells, int32_t nEdges, int32_t nVertices, int32_t nLevels,
                                   int32_t nScalars) {
    mpas_hip_mesh m;
    std::memset(&m, 0, sizeof(m));
    m.nCells = nCells;
    m.nEdges = nEdges;
    m.nVertices = nVertices;
    m.nLevels = nLevels;
    Work w = carve_work(m, nScalars, nullptr);
    return (int64_t)w.total;
}

#if !(defined(MPAS_MIXED) && MPAS_MIXED)
int mpas_hip_moist_coefficients(const mpas_hip_mesh *m, const double *scalars, int32_t nScalars, int32_t moist_start,
                                int32_t moist_end, double *cqw, double *cqu, double *qtot, void *stream) {
    if (validate_mesh(m) || validate_moist(nScalars, moist_start, moist_end)) return 1;
    launch_moist(*m, scalars, nScalars, moist_start, moist_end, cqw, cqu, qtot, STREAM(stream));
    return check_launch("moist_coefficients");
}

int mpas_hip_vert_imp_coefs(const mpas_hip_mesh *m, double dts, double epssm, const double *cqw, const double *p,
                            const double *t, const double *rt, const double *qtot, double *cofrz, double *cofwr,
                            double *cofwz, double *coftz, double *cofwt, double *a_tri, double *alpha_tri,
                            double *gamma_tri, void *stream) {
    if (validate_mesh(m)) return 1;
    launch_vimp(*m, dts, epssm, cqw, p, t, rt, qtot, cofrz, cofwr, cofwz, coftz, cofwt, a_tri, alpha_tri, gamma_tri,
                STREAM(stream));
    return check_launch("vert_imp_coefs");
}

int mpas_hip_solve_diagnostics(const mpas_hip_mesh *m, double dt, double apvm, int32_t rk_step, const double *u,
                               const double *h, const double *v_in, double *h_edge, double *v, double *divergence,
                               double *vorticity, double *ke, double *pv_edge, double *scratch, void *stream) {
    if (validate_mesh(m)) return 1;
    launch_diagnostics(*m, dt, apvm, rk_step, u, h, v_in, h_edge, v, divergence, vorticity, ke, pv_edge, scratch,
                       STREAM(stream));
    return check_launch("solve_diagnostics");
}

int mpas_hip_dyn_tend(const mpas_hip_mesh *m, const mpas_hip_dyn_params *p, const double *u, const double *v,
                      const double *w, const double *ru, const double *rw, const double *ru_save, const double *rw_save,
                      const double *theta_m, const double *theta_m_save, const double *rho_zz, const double *rho_edge,
                      const double *divergence, const double *vorticity, const double *ke, const double *pv_edge,
                      const double *pp, const double *rr_save, const double *cqw, const double *cqu,
                      const double *rt_diabatic_tend, const double *qtot, const double *ur_cell,
                      const double *vr_cell, double *tend_rho, double *tend_u_euler,
                      double *tend_w_euler, double *tend_theta_euler, double *kdiff, double *tend_u, double *tend_w,
                      double *tend_theta, double *h_divergence, double *rthdynten, double *scratch, void *stream) {
    if (validate_mesh(m) || validate_curvature(m, p->curvature, ur_cell, vr_cell)) return 1;
    DynInputs in{u, v, w, ru, rw, ru_save, rw_save, theta_m, theta_m_save, rho_zz, rho_edge, divergence, vorticity,
                 ke, pv_edge, pp, rr_save, cqw, cqu, rt_diabatic_tend, qtot, ur_cell, vr_cell};
    DynCarried io{tend_rho, tend_u_euler, tend_w_euler, tend_theta_euler, kdiff};
    DynOutputs out{tend_u, tend_w, tend_theta, h_divergence, rthdynten};
    launch_dyn_tend(*m, *p, in, io, out, scratch, STREAM(stream));
    return check_launch("dyn_tend");
}

int mpas_hip_smlstep_pert(const mpas_hip_mesh *m, const double *u_tend, double *w_tend, void *stream) {
    if (validate_mesh(m)) return 1;
    launch_smlstep(*m, u_tend, w_tend, STREAM(stream));
    return check_launch("smlstep_pert");
}

int mpas_hip_acoustic_step(const mpas_hip_mesh *m, const mpas_hip_acoustic_params *p, const double *cofrz,
                           const double *cofwr, const double *cofwz, const double *coftz, const double *cofwt,
                           const double *a_tri, const double *alpha_tri, const double *gamma_tri, const double *rho_zz,
                           const double *theta_m, const double *exner, const double *cqu, const double *tend_ru,
                           const double *tend_rho, const double *tend_rt, const double *tend_rw, const double *w,
                           const double *rw, const double *rw_save, double *ru_p, double *ruAvg, double *rho_pp,
                           double *rtheta_pp, double *rtheta_pp_old, double *rw_p, double *wwAvg, void *stream) {
    if (validate_mesh(m)) return 1;
    AcousticInputs in{cofrz, cofwr, cofwz, coftz, cofwt, a_tri, alpha_tri, gamma_tri, rho_zz, theta_m,
                      exner, cqu,   tend_ru, tend_rho, tend_rt, tend_rw, w, rw, rw_save};
    AcousticState io{ru_p, ruAvg, rho_pp, rtheta_pp, rtheta_pp_old, rw_p, wwAvg};
    launch_acoustic(*m, *p, in, io, STREAM(stream), DampFold{0, 0.0, nullptr, nullptr, nullptr});
    return check_launch("acoustic_step");
}

int mpas_hip_divergence_damping(const mpas_hip_mesh *m, double dts, double smdiv, double len_disp,
                                const double *theta_m, const double *rtheta_pp, const double *rtheta_pp_old,
                                double *ru_p, void *stream) {
    if (validate_mesh(m)) return 1;
    launch_divdamp(*m, dts, smdiv, len_disp, theta_m, rtheta_pp, rtheta_pp_old, ru_p, STREAM(stream));
    return check_launch("divergence_damping");
}

int mpas_hip_recover_large_step(const mpas_hip_mesh *m, int32_t rk_step, double dt, int32_t ns,
                                const double *rho_p_save, const double *rho_pp, const double *rtheta_p_save,
                                const double *rtheta_pp, const double *rt_diabatic_tend, const double *rw_save,
                                const double *rw_p, const double *ru_save, const double *ru_p, double *wwAvg,
                                double *ruAvg, double *rho_p, double *rho_zz, double *rtheta_p, double *theta_m,
                                double *exner, double *pressure_p, double *rw, double *w, double *ru, double *u,
                                void *stream) {
    if (validate_mesh(m)) return 1;
    RecoverInputs in{rho_p_save, rho_pp, rtheta_p_save, rtheta_pp, rt_diabatic_tend, rw_save, rw_p, ru_save, ru_p};
    RecoverOutputs out{wwAvg, ruAvg, rho_p, rho_zz, rtheta_p, theta_m, exner, pressure_p, rw, w, ru, u};
    launch_recover(*m, rk_step, dt, ns, in, out, STREAM(stream), DampFold{0, 0.0, nullptr, nullptr, nullptr});
    return check_launch("recover_large_step");
}

int mpas_hip_advance_scalars(const mpas_hip_mesh *m, int32_t nScalars, double dt, double coef3, double weight_old,
                             double weight_new, const double *scalar_old, const double *scalar_new_in,
                             const double *rho_zz_old, const double *rho_zz_new, const double *uhAvg,
                             const double *wwAvg, double *scalar_new_out, double *horiz_flux, void *stream) {
    if (validate_mesh(m)) return 1;
    if (scalar_new_out == scalar_new_in) {
        set_error("advance_scalars: output must not alias scalar_new_in");
        return 1;
    }
    launch_advance_scalars(*m, nScalars, dt, coef3, weight_old, weight_new, scalar_old, scalar_new_in, rho_zz_old,
                           rho_zz_new, uhAvg, wwAvg, scalar_new_out, horiz_flux, STREAM(stream));
    return check_launch("advance_scalars");
}

int mpas_hip_advance_scalars_mono(const mpas_hip_mesh *m, int32_t nScalars, double dt, double coef3,
                                  const double *scalars_old, const double *scalars_new_in, const double *rho_zz_old,
                                  const double *uhAvg, const double *wwAvg, double *scalars_new_out,
                                  double *rho_zz_int, double *scratch, void *stream) {
    if (validate_mesh(m)) return 1;
    launch_advance_scalars_mono(*m, nScalars, dt, coef3, scalars_old, scalars_new_in, rho_zz_old, uhAvg, wwAvg,
                                scalars_new_out, rho_zz_int, scratch, STREAM(stream));
    return check_launch("advance_scalars_mono");
}

int mpas_hip_init_coupled_diagnostics(const mpas_hip_mesh *m, const double *theta, const double *rho,
                                      const double *qv, const double *u, const double *w, const double *theta_base,
                                      double *theta_m, double *rho_zz, double *ru, double *rw, double *rho_p,
                                      double *rtheta_base, double *rtheta_p, double *exner, double *exner_base,
                                      double *pressure_p, double *pressure_base, void *stream) {
    if (validate_mesh(m)) return 1;
    InitCoupledOutputs o{theta_m, rho_zz,     ru,         rw,         rho_p,       rtheta_base,
                         rtheta_p, exner, exner_base, pressure_p, pressure_base};
    launch_init_coupled(*m, theta, rho, qv, u, w, theta_base, o, STREAM(stream));
    return check_launch("init_coupled_diagnostics");
}

int mpas_hip_reconstruct(const mpas_hip_mesh *m, const double *u, double *ux, double *uy, double *uz, double *uzonal,
                         double *umerid, void *stream) {
    if (validate_mesh(m)) return 1;
    launch_reconstruct(*m, u, ux, uy, uz, uzonal, umerid, STREAM(stream));
    return check_launch("reconstruct");
}

#endif  // plain build exports

int mpas_hip_history_fields(int32_t nCells, int32_t nLevels, int32_t nScalars, const double *theta_m,
                            const double *scalars, const double *rho_zz, const double *zz, const double *pressure_p,
                            const double *pressure_base, double *theta, double *rho, double *pressure, void *stream) {
    if (nCells < 0 || nLevels < 1 || nScalars < 0) {
        set_error("history_fields: bad dimensions");
        return 1;
    }
    if ((theta && !theta_m) || (rho && (!rho_zz || !zz)) || (pressure && (!pressure_p || !pressure_base))) {
        set_error("history_fields: an output was requested without its inputs");
        return 1;
    }
    launch_history_fields(nCells, nLevels, nScalars, theta_m, scalars, rho_zz, zz, pressure_p, pressure_base, theta,
                          rho, pressure, STREAM(stream));
    return check_launch("history_fields");
}

#endif  // per-kernel entry points (plain build)

// ---------------------------------------------------------------------------
// one large step
// ---------------------------------------------------------------------------
static int srk3_step_impl(const mpas_hip_mesh *mesh, StepState *st, const mpas_hip_workspace *work,
                          const mpas_hip_config *cfg, hipStream_t s) {
    if (validate_mesh(mesh) || validate_curvature(mesh, cfg->curvature, st->ur_cell, st->vr_cell)) return 1;
    const mpas_hip_mesh &m = *mesh;
    const int S = cfg->nScalars;
    if (validate_moist(S, cfg->moist_start, cfg->moist_end)) return 1;
    Work w = carve_work(m, S, work->base);
    if ((int64_t)w.total > work->n_doubles) {
        set_error("workspace too small");
        return 1;
    }
    if (cfg->split < 1 || cfg->ns < 1) {
        set_error("split and ns must be >= 1");
        return 1;
    }
    if (cfg->order != 0 && cfg->order != 2 && cfg->order != 3) {
        set_error("time integration order must be 2 or 3");
        return 1;
    }
    const int L = m.nLevels, nC = m.nCells;
    const size_t C = (size_t)(nC + 1) * L, E = (size_t)(m.nEdges + 1) * L, Wd = (size_t)(nC + 1) * (L + 1);
    const double dt = cfg->dt;

    // atm_rk_integration_setup (time_integration.py:2497-2785)
    fill_zero(st->theta_m + (size_t)nC * L, L, s);
    copy_d(w.theta_m_init, st->theta_m, C, s);
    copy_d(w.rho_zz_old_split, st->rho_zz, C, s);
    copy_d(w.scalars_old, st->scalars, C * S, s);
    copy_d(w.ru_save, st->ru, E, s);
    copy_d(w.rw_save, st->rw, Wd, s);
    copy_d(w.rtheta_p_save, st->rtheta_p, C, s);
    copy_d(w.rho_p_save, st->rho_p, C, s);

    // compute_rk_weights (time_integration.py:2787-2856): three stages for
    // both orders; RK2 changes the stage lengths and substep counts
    const bool rk2 = cfg->order == 2;
    const double dt_dyn = dt / (double)cfg->split;
    const int half_ns = cfg->ns / 2 > 1 ? cfg->ns / 2 : 1;
    const double dts_ns = dt_dyn / (double)cfg->ns;
    const double rk_timestep[3] = {rk2 ? dt_dyn / 2.0 : dt_dyn / 3.0, dt_dyn / 2.0, dt_dyn};
    const double rk_sub[3] = {rk2 ? dts_ns : dt_dyn / 3.0, dts_ns, dts_ns};
    const int nsub[3] = {rk2 ? half_ns : 1, half_ns, cfg->ns};
    const double resm = (1.0 - cfg->epssm) / (1.0 + cfg->epssm);
    const double rcv = MPAS_RGAS / (MPAS_CP - MPAS_RGAS);
    const double c2 = MPAS_CP * rcv;

    // Dynamics-tendency scalars, evaluated as the reference evaluates them.
    mpas_hip_dyn_params P;
    P.dt = dt;
    P.coef_3rd_order = cfg->coef3;
    {
        const double inv_dt = 1.0 / dt;
        const double cl = cfg->c_s * cfg->len_disp;
        P.kdiff_coef = cl * cl;
        P.kdiff_cap = (0.01 * (cfg->len_disp * cfg->len_disp)) * inv_dt;
        if (cfg->mix_fixed) {
            P.h_mom_eddy_visc4 = cfg->h_mom_eddy_visc4;
            P.h_theta_eddy_visc4 = cfg->h_theta_eddy_visc4;
        } else {
            P.h_mom_eddy_visc4 = cfg->visc4_2dsmag * std::pow(cfg->len_disp, 3.0);
            P.h_theta_eddy_visc4 = P.h_mom_eddy_visc4;
        }
        P.del4u_div_factor = cfg->del4u_div_factor;
        P.prandtl_inv = 1.0 / 1.0;
        P.rayleigh_levels = cfg->rayleigh_levels;
        P.rayleigh_coef_inverse = cfg->rayleigh_coef_inverse;
        P.mix_fixed = cfg->mix_fixed ? 1 : 0;
        P.kdiff_fixed = cfg->h_theta_eddy_visc2;
        P.cam_levels = cfg->cam_levels;
        P.cam_visc2 = cfg->cam_visc2;
        P.curvature = cfg->curvature ? 1 : 0;
        P.v_mom_eddy_visc2 = cfg->v_mom_eddy_visc2;
        P.v_theta_eddy_visc2 = cfg->v_theta_eddy_visc2;
        P.mix_full = cfg->mix_full ? 1 : 0;
    }

    for (int sub = 1; sub <= cfg->split; ++sub) {
        for (int rk = 1; rk <= 3; ++rk) {
            if (sub == 1 && rk == 1) {
                launch_moist(m, st->scalars, S, cfg->moist_start, cfg->moist_end, w.cqw, w.cqu, w.qtot, s);
                if (!cfg->diagnostics_current) {
                    // pre-dynamics-tendency diagnostic refresh (time_integration.py:5687-5900)
                    launch_diagnostics(m, dt, cfg->apvm, -1, st->u, st->rho_zz, st->v, st->rho_edge, st->v,
                                       st->divergence, st->vorticity, st->ke, st->pv_edge, w.scratch, s);
                }
            }
            // the coefficients depend on dts: RK3 recomputes them for its
            // second stage, RK2 keeps the first stage's (time_integration.py:5558-5569)
            if (rk == 1 || (rk == 2 && !rk2))
                launch_vimp(m, rk_sub[rk - 1], cfg->epssm, w.cqw, st->exner, st->theta_m, st->rtheta_p, w.qtot,
                            w.cofrz, w.cofwr, w.cofwz, w.coftz, w.cofwt, w.a_tri, w.alpha_tri, w.gamma_tri, s);
            P.rk_step = rk;
            DynInputs din{st->u,          st->v,          st->w,          st->ru,      st->rw,
                          w.ru_save,      w.rw_save,      st->theta_m,    w.theta_m_init, st->rho_zz,
                          st->rho_edge,   st->divergence, st->vorticity,  st->ke,      st->pv_edge,
                          st->pressure_p, w.rho_p_save,   w.cqw,          w.cqu,       st->rt_diabatic_tend,
                          w.qtot,         st->ur_cell,    st->vr_cell};
            DynCarried dio{w.tend_rho, w.tend_u_euler, w.tend_w_euler, w.tend_theta_euler, w.kdiff};
            DynOutputs dout{w.tend_u, w.tend_w, w.tend_theta, w.h_divergence, w.rthdynten};
            launch_dyn_tend(m, P, din, dio, dout, w.scratch, s);
            launch_smlstep(m, w.tend_u, w.tend_w, s);
            AcousticInputs ain{w.cofrz,  w.cofwr,        w.cofwz,    w.coftz,        w.cofwt, w.a_tri, w.alpha_tri,
                               w.gamma_tri, st->rho_zz,  w.theta_m_init, st->exner, w.cqu,   w.tend_u, w.tend_rho,
                               w.tend_theta, w.tend_w,   st->w,      st->rw,         w.rw_save};
            AcousticState aio{w.ru_p, w.ruAvg, w.rho_pp, w.rtheta_pp, w.rtheta_pp_old, w.rw_p, w.wwAvg};
            // The divergence damping that closes each acoustic substep is folded
            // into the next substep's edge pass, and the last one into the
            // recovery's edge pass (divergence_damping.py:187-209).
            const DampFold fold{1, 2.0 * cfg->smdiv * cfg->len_disp / rk_sub[rk - 1], w.theta_m_init, w.rtheta_pp,
                                w.rtheta_pp_old};
            for (int small = 1; small <= nsub[rk - 1]; ++small) {
                mpas_hip_acoustic_params AP;
                AP.small_step = small;
                AP.dts = rk_sub[rk - 1];
                AP.epssm = cfg->epssm;
                AP.resm = resm;
                AP.c2 = c2;
                DampFold f = fold;
                f.on = small > 1 ? 1 : 0;
                launch_acoustic(m, AP, ain, aio, s, f);
            }
            RecoverInputs rin{w.rho_p_save, w.rho_pp, w.rtheta_p_save, w.rtheta_pp, st->rt_diabatic_tend,
                              w.rw_save,    w.rw_p,   w.ru_save,       w.ru_p};
            RecoverOutputs rout{w.wwAvg,      w.ruAvg,       st->rho_p, st->rho_zz, st->rtheta_p, st->theta_m,
                                st->exner,    st->pressure_p, st->rw,   st->w,      st->ru,       st->u};
            launch_recover(m, rk, rk_timestep[rk - 1], nsub[rk - 1], rin, rout, s, fold);
            launch_diagnostics(m, dt, cfg->apvm, rk, st->u, st->rho_zz, st->v, st->rho_edge, st->v, st->divergence,
                               st->vorticity, st->ke, st->pv_edge, w.scratch, s);
        }
        // split-flux accumulation and substep reset (time_integration.py:3820-3942)
        if (sub == 1) {
            copy_d(w.ruAvg_split, w.ruAvg, E, s);
            copy_d(w.wwAvg_split, w.wwAvg, Wd, s);
        } else {
            add_d(w.ruAvg_split, w.ruAvg, E, s);
            add_d(w.wwAvg_split, w.wwAvg, Wd, s);
        }
        if (sub < cfg->split) {
            fill_zero(st->theta_m + (size_t)nC * L, L, s);
            copy_d(w.ru_save, st->ru, E, s);
            copy_d(w.rw_save, st->rw, Wd, s);
            copy_d(w.rtheta_p_save, st->rtheta_p, C, s);
            copy_d(w.rho_p_save, st->rho_p, C, s);
            copy_d(w.theta_m_init, st->theta_m, C, s);
        }
    }
    const double inv_split = 1.0 / (double)cfg->split;
    scale_d(st->ruAvg, w.ruAvg_split, inv_split, E, s);
    scale_d(st->wwAvg, w.wwAvg_split, inv_split, Wd, s);
    // split transport RK3 (time_integration.py:3976-4675); rho_zz_new is the
    // post-dynamics rho_zz already in the state.  Without scalar advection the
    // scalars keep their values (:3975)
    if (S > 0 && cfg->scalar_advection) {
        // transport stage lengths and density weights (time_integration.py:3367-3371,
        // advance_scalars.py:271-286)
        const double dts[3] = {rk2 ? dt / 2.0 : dt / 3.0, dt / 2.0, dt};
        const double wn1 = rk2 ? 1.0 / 2.0 : 1.0 / 3.0;
        // stage 1: state.scalars -> tmp; stage 2: tmp -> state.scalars
        launch_advance_scalars(m, S, dts[0], cfg->coef3, 1.0 - wn1, wn1, w.scalars_old, st->scalars,
                               w.rho_zz_old_split, st->rho_zz, st->ruAvg, st->wwAvg, w.scalars_tmp, w.horiz_flux, s);
        launch_advance_scalars(m, S, dts[1], cfg->coef3, 1.0 - 1.0 / 2.0, 1.0 / 2.0, w.scalars_old, w.scalars_tmp,
                               w.rho_zz_old_split, st->rho_zz, st->ruAvg, st->wwAvg, st->scalars, w.horiz_flux, s);
        if (cfg->monotonic) {
            launch_advance_scalars_mono(m, S, dts[2], cfg->coef3, w.scalars_old, st->scalars, w.rho_zz_old_split,
                                        st->ruAvg, st->wwAvg, st->scalars, w.rho_zz_int, w.scratch, s);
        } else {
            launch_advance_scalars(m, S, dts[2], cfg->coef3, 0.0, 1.0, w.scalars_old, st->scalars,
                                   w.rho_zz_old_split, st->rho_zz, st->ruAvg, st->wwAvg, w.scalars_tmp,
                                   w.horiz_flux, s);
            copy_d(st->scalars, w.scalars_tmp, C * S, s);
        }
    }
    launch_reconstruct(m, st->u, st->uReconstructX, st->uReconstructY, st->uReconstructZ, st->uReconstructZonal,
                       st->uReconstructMeridional, s);
    copy_d(st->ur_cell, st->uReconstructZonal, C, s);
    copy_d(st->vr_cell, st->uReconstructMeridional, C, s);
    return check_launch("srk3_step");
}

#if defined(MPAS_MIXED) && MPAS_MIXED
int mpas_hip_mixed_srk3_step(const mpas_hip_mesh *mesh, mpas_hip_mixed_state *state, const mpas_hip_workspace *work,
                             const mpas_hip_config *cfg, void *stream) {
    StepState st = step_state(state);
    return srk3_step_impl(mesh, &st, work, cfg, STREAM(stream));
}
#else
int mpas_hip_srk3_step(const mpas_hip_mesh *mesh, mpas_hip_state *state, const mpas_hip_workspace *work,
                       const mpas_hip_config *cfg, void *stream) {
    StepState st = step_state(state);
    return srk3_step_impl(mesh, &st, work, cfg, STREAM(stream));
}

int64_t mpas_hip_srk3_step_noalloc_capacity(void) { return (int64_t)kStaticWorkDoubles; }

int mpas_hip_srk3_step_noalloc(const mpas_hip_mesh *mesh, mpas_hip_state *state, const mpas_hip_config *cfg,
                               void *stream) {
    if (validate_mesh(mesh)) return 1;
    if (!state || !cfg) {
        set_error("srk3_step_noalloc: null state or config");
        return 1;
    }
    const Work need = carve_work(*mesh, cfg->nScalars, nullptr);
    if (need.total > kStaticWorkDoubles) {
        set_error("srk3_step_noalloc: this mesh needs " + std::to_string(need.total * 8 / 1024) +
                  " KiB of workspace, beyond the " + std::to_string(kStaticWorkDoubles * 8 / 1024) +
                  " KiB the library holds per GPU; call mpas_hip_srk3_step with a workspace instead");
        return 1;
    }
    double *base = static_work_address();
    if (!base) return 1;
    hipStream_t s = STREAM(stream);
    // the array keeps the previous call's contents: clear the carved part, as a
    // freshly allocated workspace is (a fill kernel, so the call stays capturable)
    fill_zero(base, need.total, s);
    const mpas_hip_workspace ws{base, (int64_t)kStaticWorkDoubles};
    StepState st = step_state(state);
    return srk3_step_impl(mesh, &st, &ws, cfg, s);
}
#endif

}  // extern "C"
