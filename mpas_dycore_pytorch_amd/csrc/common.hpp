// Shared definitions for the gfx950 split-explicit dynamics kernels.
//
// Arithmetic is IEEE binary64 with round-to-nearest and no contraction
// (built with -ffp-contract=off): every expression keeps the association of
// the reference statement it restates, so results match the reference CPU
// float64 path bit for bit apart from libm calls (pow, sin, cos).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mpas_hip.h"
#include "scalar_math.hpp"

namespace mpas {

// Storage type of the edge- and vertex-located fields whose rounding stays
// below the mixed build's accuracy bar (v, pv_edge, rho_edge, vorticity, the
// tendencies of u, cqu, the tracer fluxes, and the edge/vertex scratch).
// binary64 in the plain build; binary32 in the mixed-precision build
// (-DMPAS_MIXED=1), which halves their bytes.  u and the mass fluxes (ru,
// ruAvg, the acoustic ru_p), and every cell and interface field (the
// vertical-implicit coefficients and column solve, rho, theta, exner,
// pressure, w), stay binary64 in both builds.
#if defined(MPAS_MIXED) && MPAS_MIXED
using hreal = float;
#else
using hreal = double;
#endif

// Arithmetic type of the horizontal tendency, diagnostic and transport
// gathers: binary64 in the plain build, binary32 in the mixed build.  The
// column work (vertical-implicit coefficients, the tridiagonal solve, exner
// and pressure recovery, the base state) is binary64 in both.
#if defined(MPAS_MIXED) && MPAS_MIXED
using freal = float;
#else
using freal = double;
#endif

// A slot-ordered sum.  The plain build forms the reference's left-to-right
// sum (the first slot assigned, each later one added).  The mixed build keeps
// a Neumaier compensation term beside the binary32 running sum (TwoSum of the
// sum and the new term, the rounding error accumulated separately) and adds
// it in once at the end, so a flux divergence that nearly cancels keeps the
// accuracy of its binary32 terms instead of losing it to the running sum.
struct SlotSum {
    freal s;
#if defined(MPAS_MIXED) && MPAS_MIXED
    freal c;
#endif
    __device__ __forceinline__ void first(freal x) {
        s = x;
#if defined(MPAS_MIXED) && MPAS_MIXED
        c = 0.0f;
#endif
    }
    __device__ __forceinline__ void add(freal x) {
#if defined(MPAS_MIXED) && MPAS_MIXED
        const freal t = s + x;
        c = c + ((fabsf(s) >= fabsf(x)) ? ((s - t) + x) : ((x - t) + s));
        s = t;
#else
        s = s + x;
#endif
    }
    // slot j of a row with n live slots: assigned at j == 0, added below n
    __device__ __forceinline__ void slot(int j, int n, freal x) {
        if (j < n) {
            if (j == 0) first(x);
            else add(x);
        }
    }
    __device__ __forceinline__ freal value() const {
#if defined(MPAS_MIXED) && MPAS_MIXED
        return s + c;
#else
        return s;
#endif
    }
};

// Host-evaluated kernel scalars (layouts fixed by the C ABI header).
using DynParams = mpas_hip_dyn_params;
using AcousticParams = mpas_hip_acoustic_params;

// Earth radius of the curvature terms (the reference's constants.py:84)
constexpr double kEarthRadius = 6371229.0;

// Whether a dynamics tendency runs with the curvature terms or the vertical
// eddy viscosities: the passes that carry them are built twice, and the plain
// build of each (without their code, so with fewer registers) runs otherwise.
inline bool dyn_options(const DynParams &P) {
    return P.curvature || P.v_mom_eddy_visc2 > 0.0 || P.v_theta_eddy_visc2 > 0.0;
}

// Heights of the vertical mixing stencil at level k (1..L-2): zm, z0, zp are the
// midpoints of interfaces (k-1, k), (k, k+1), (k+1, k+2) of a cell column
// (dyn_tend.py:5073-5080), or of the mean of an edge's two cell columns
// (:2750-2762).  The mixing is ((f(k+1) - f(k)) / (zp - z0) - (f(k) - f(k-1)) /
// (z0 - zm)) / (0.5 (zp - zm)) times its coefficient.
struct VmixZ {
    double zm, z0, zp;
};
__device__ __forceinline__ VmixZ vmix_z_cell(const double *zg, int k) {
    return {0.5 * (zg[k - 1] + zg[k]), 0.5 * (zg[k] + zg[k + 1]), 0.5 * (zg[k + 1] + zg[k + 2])};
}
__device__ __forceinline__ VmixZ vmix_z_edge(const double *g1, const double *g2, int k) {
    const double z1 = 0.5 * (g1[k - 1] + g2[k - 1]), z2 = 0.5 * (g1[k] + g2[k]);
    const double z3 = 0.5 * (g1[k + 1] + g2[k + 1]), z4 = 0.5 * (g1[k + 2] + g2[k + 2]);
    return {0.5 * (z1 + z2), 0.5 * (z2 + z3), 0.5 * (z3 + z4)};
}
// the mixed u of level k: u itself, or its departure from the reference profile
__device__ __forceinline__ double vmix_u(const mpas_hip_mesh &m, const DynParams &P, double u, int e, int k) {
    return P.mix_full ? u : (u - m.u_init[k] * m.cosAngleEdge[e]) - m.v_init[k] * m.sinAngleEdge[e];
}

// CAM damping floor of kdiff at level k (one of the top P.cam_levels levels):
// cam_visc2 (1 - (nLevels - (k + 1)) / cam_levels), dyn_tend.py:3442-3446
__device__ __forceinline__ double cam_kdiff(const DynParams &P, int L, int k) {
    return P.cam_visc2 * (1.0 - (double)(L - (k + 1)) / (double)P.cam_levels);
}

constexpr int kBlock = 256;

// Compile-time caps on the slot counts, so slot loops unroll fully: every
// load of a loop is issued before the first accumulation (the sums still run
// in slot order).  Slots past the live count read a clamped, in-bounds slot
// and are not accumulated.  Meshes beyond the caps are rejected on entry.
constexpr int kMaxEdges = 8;    // edgesOnCell / cellsOnCell / verticesOnCell
constexpr int kMaxEdges2 = 16;  // edgesOnEdge
constexpr int kMaxAdv = 16;     // advCellsForEdge
__device__ __forceinline__ int clamp_slot(int j, int n) { return j < n ? j : n - 1; }

// Slot rows across a wavefront: lane j loads slot clamp(j, n) of a
// connectivity or coefficient row with one vector load, and each slot is then
// broadcast to the whole wavefront with readlane (slot index uniform).  This
// puts every gather address of a slot loop in registers before the first
// gather is issued.  All 64 lanes must execute the row load.
__device__ __forceinline__ int bcast(int v, int j) { return __builtin_amdgcn_readlane(v, j); }
__device__ __forceinline__ double bcast(double v, int j) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), j);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);
    return __hiloint2double(hi, lo);
}

inline unsigned blocks_for(long long n, int block = kBlock) {
    return (unsigned)((n + block - 1) / block);
}

// XCD-aware block order.  The dispatcher hands consecutive workgroups to the
// 8 XCDs round-robin; remapping block b to (b % 8) * per + b / 8 gives each
// XCD (and its L2) one contiguous range of the space-filling-curve ordered
// entities.  The grid is padded to a multiple of 8; blocks mapped past the
// end exit.
constexpr int kXcds = 8;
// A launch whose block count carries kLinearOrder keeps the hardware's
// round-robin block order instead (block b takes the b-th group of entities):
// a per-kernel choice of the autotuner (tuned_order).
constexpr long long kLinearOrder = 1LL << 62;
inline unsigned xcd_grid(long long nblocks) {
    nblocks &= ~kLinearOrder;
    return (unsigned)(((nblocks + kXcds - 1) / kXcds) * kXcds);
}
__device__ __forceinline__ long long xcd_block(long long nblocks) {
    const long long b = blockIdx.x;
    if (nblocks & kLinearOrder) return b;
    const long long per = (nblocks + kXcds - 1) / kXcds;
    return (b % kXcds) * per + b / kXcds;
}

// Wavefront-per-entity mapping: a 256-thread block holds 4 wavefronts, each
// wavefront one entity (cell, edge or vertex) with lane = vertical level.
// The entity index is wave-uniform, so connectivity and mesh metrics are read
// with scalar loads and every neighbour gather is one contiguous column.
// Each wavefront walks kEntPerWave consecutive entities (1 measured fastest
// at level 7 x 55 on MI355X: walking 4 did not raise L1 reuse enough to
// offset the lost parallelism).
// The number of wavefronts per block (1, 2, 4, 8 or 16; the derivative
// kernels at most 4) is a launch-time choice per kernel: kWaves unless the
// autotuner set another (tuned_waves); the kernels read it back from
// blockDim.  Larger blocks keep more consecutive (curve-ordered) entities on
// one CU, so neighbour columns they share are found in its L1.
constexpr int kWaves = 4;
constexpr int kMaxWaveBlock = 1024;  // the forward wavefront kernels take up to 16 wavefronts per block
constexpr int kEntPerWave = 1;
int tuned_waves(const char *kernel_name);
int tuned_order(const char *kernel_name);  // 0: blocks dealt to the XCDs in chunks, 1: linear
inline long long wave_blocks(long long n_entities, int waves = kWaves) {
    return (n_entities + waves * kEntPerWave - 1) / (waves * kEntPerWave);
}
__device__ __forceinline__ int wave_entity_base(long long nblocks) {
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int waves = (int)(blockDim.x >> 6);
    return ((int)(xcd_block(nblocks) * waves) + w) * kEntPerWave;
}
__device__ __forceinline__ int wave_lane() { return (int)(threadIdx.x & 63); }

#define MPAS_WAVE_LAUNCH(kernel, n_entities, stream, ...)                                          \
    do {                                                                                           \
        const int wpb_ = ::mpas::tuned_waves(#kernel);                                             \
        const long long nb_ = ::mpas::wave_blocks(n_entities, wpb_);                               \
        MPAS_LAUNCH(kernel, dim3(::mpas::xcd_grid(nb_)), dim3(64 * wpb_), 0, stream,               \
                    nb_ | (::mpas::tuned_order(#kernel) ? ::mpas::kLinearOrder : 0LL), __VA_ARGS__);   \
    } while (0)

// Zero-fill by kernel rather than hipMemsetAsync so launch sequences stay
// safe to capture in a HIP graph.
void fill_zero(double *p, size_t n, hipStream_t s);
#if defined(MPAS_MIXED) && MPAS_MIXED
void fill_zero(float *p, size_t n, hipStream_t s);
#endif

// Bump allocator over a caller-owned scratch block: slices of any element
// type, each starting on a 256-byte boundary (size with slice_doubles).
struct Carver {
    char *p;
    template <typename T>
    T *take(size_t n) {
        T *r = reinterpret_cast<T *>(p);
        p += (n * sizeof(T) + 255) & ~(size_t)255;
        return r;
    }
};
template <typename T>
inline size_t slice_doubles(size_t n) {
    return ((n * sizeof(T) + 255) & ~(size_t)255) / sizeof(double);
}

// Error plumbing: every launcher records the first HIP error.
void set_error(const std::string &msg);
int check_launch(const char *what);

// Per-kernel event timer (measurement hook): when armed for a kernel name,
// each launch of that kernel is bracketed by HIP events on its stream.
bool timer_armed_for(const char *kernel_name);
void timer_before(hipStream_t s);
void timer_after(hipStream_t s);

}  // namespace mpas

#define MPAS_LAUNCH(kernel, grid, block, shm, stream, ...)                           \
    do {                                                                             \
        const bool mpas_timed_ = ::mpas::timer_armed_for(#kernel);                   \
        if (mpas_timed_) ::mpas::timer_before(stream);                               \
        hipLaunchKernelGGL(kernel, grid, block, shm, stream, __VA_ARGS__);           \
        if (mpas_timed_) ::mpas::timer_after(stream);                                \
    } while (0)

// Internal launchers (one per kernel family).
namespace mpas {
void launch_moist(const mpas_hip_mesh &m, const double *scalars, int S, int ms, int me, double *cqw,
                  hreal *cqu, double *qtot, hipStream_t s);
void launch_vimp(const mpas_hip_mesh &m, double dts, double epssm, const double *cqw, const double *p,
                 const double *t, const double *rt, const double *qtot, double *cofrz, double *cofwr,
                 double *cofwz, double *coftz, double *cofwt, double *a_tri, double *alpha_tri,
                 double *gamma_tri, hipStream_t s);
void launch_diagnostics(const mpas_hip_mesh &m, double dt, double apvm, int rk_step, const double *u,
                        const double *h, const hreal *v_in, hreal *h_edge, hreal *v, double *div,
                        hreal *vort, double *ke, hreal *pv_edge, double *scratch, hipStream_t s);
size_t diagnostics_scratch_doubles(const mpas_hip_mesh &m);

struct DynInputs {
    const double *u;
    const hreal *v;
    const double *w;
    const double *ru;
    const double *rw;
    const double *ru_save;
    const double *rw_save, *theta_m, *theta_m_save, *rho_zz;
    const hreal *rho_edge;
    const double *divergence;
    const hreal *vorticity;
    const double *ke;
    const hreal *pv_edge;
    const double *pp, *rr_save, *cqw;
    const hreal *cqu;
    const double *rt_diabatic_tend, *qtot;
    const double *ur_cell, *vr_cell;  // read only with the curvature terms on
};
struct DynCarried {
    double *tend_rho;
    hreal *tend_u_euler;
    double *tend_w_euler, *tend_theta_euler, *kdiff;
};
struct DynOutputs {
    hreal *tend_u;
    double *tend_w, *tend_theta, *h_divergence, *rthdynten;
};
void launch_dyn_tend(const mpas_hip_mesh &m, const DynParams &P, const DynInputs &in,
                     const DynCarried &io, const DynOutputs &out, double *scratch, hipStream_t s);
size_t dyn_tend_scratch_doubles(const mpas_hip_mesh &m);

void launch_smlstep(const mpas_hip_mesh &m, const hreal *u_tend, double *w_tend, hipStream_t s);

struct AcousticInputs {
    const double *cofrz, *cofwr, *cofwz, *coftz, *cofwt, *a_tri, *alpha_tri, *gamma_tri;
    const double *rho_zz, *theta_m, *exner;
    const hreal *cqu, *tend_ru;
    const double *tend_rho, *tend_rt, *tend_rw, *w, *rw, *rw_save;
};
struct AcousticState {
    double *ru_p, *ruAvg;
    double *rho_pp, *rtheta_pp, *rtheta_pp_old, *rw_p, *wwAvg;
};
// Divergence damping carried into a later edge pass (on = 0: none).
struct DampFold {
    int on;
    double coef;
    const double *theta_m, *rtheta_pp, *rtheta_pp_old;
};
void launch_acoustic(const mpas_hip_mesh &m, const AcousticParams &P, const AcousticInputs &in,
                     const AcousticState &io, hipStream_t s, const DampFold &fold);
void launch_divdamp(const mpas_hip_mesh &m, double dts, double smdiv, double len_disp,
                    const double *theta_m, const double *rtheta_pp, const double *rtheta_pp_old,
                    double *ru_p, hipStream_t s);

struct RecoverInputs {
    const double *rho_p_save, *rho_pp, *rtheta_p_save, *rtheta_pp, *rt_diabatic_tend, *rw_save, *rw_p;
    const double *ru_save, *ru_p;
};
struct RecoverOutputs {
    double *wwAvg;
    double *ruAvg;
    double *rho_p, *rho_zz, *rtheta_p, *theta_m, *exner, *pressure_p, *rw, *w;
    double *ru, *u;
};
void launch_recover(const mpas_hip_mesh &m, int rk_step, double dt, int ns, const RecoverInputs &in,
                    const RecoverOutputs &out, hipStream_t s, const DampFold &fold);

void launch_advance_scalars(const mpas_hip_mesh &m, int S, double dt, double coef3, double w_old,
                            double w_new, const double *scalar_old, const double *scalar_new_in,
                            const double *rho_zz_old, const double *rho_zz_new, const double *uhAvg,
                            const double *wwAvg, double *scalar_new_out, hreal *horiz_flux,
                            hipStream_t s);
void launch_advance_scalars_mono(const mpas_hip_mesh &m, int S, double dt, double coef3,
                                 const double *scalars_old, const double *scalars_new_in,
                                 const double *rho_zz_old, const double *uhAvg, const double *wwAvg,
                                 double *scalars_new_out, double *rho_zz_int, double *scratch,
                                 hipStream_t s);
size_t mono_scratch_doubles(const mpas_hip_mesh &m);

void launch_reconstruct(const mpas_hip_mesh &m, const double *u, double *ux, double *uy, double *uz,
                        double *uzonal, double *umerid, hipStream_t s);

struct InitCoupledOutputs {
    double *theta_m, *rho_zz;
    double *ru;
    double *rw, *rho_p, *rtheta_base, *rtheta_p, *exner, *exner_base, *pressure_p, *pressure_base;
};
void launch_init_coupled(const mpas_hip_mesh &m, const double *theta, const double *rho, const double *qv,
                         const double *u, const double *w, const double *theta_base, const InitCoupledOutputs &o,
                         hipStream_t s);
void launch_history_fields(int nC, int L, int S, const double *theta_m, const double *scalars, const double *rho_zz,
                           const double *zz, const double *pressure_p, const double *pressure_base, double *theta,
                           double *rho, double *pressure, hipStream_t s);
}  // namespace mpas
