// Launch plumbing shared by the adjoint (kernels_adjoint.hip) and tangent
// (kernels_tangent.hip) kernels: the slot arrays, the per-launch scalar
// arguments, the wavefront-per-entity kernel and launch macros, and the
// layout of the scratch buffer (sized by mpas_hip_vjp_scratch_doubles).
#pragma once
#include "common.hpp"

#include "adjoint_bodies.hpp"
#include "adjoint_dyn.hpp"
#include "adjoint_transport.hpp"
#include "tangent_bodies.hpp"
#include "tangent_dyn.hpp"
#include "tangent_transport.hpp"

namespace mpas {
namespace adj {

constexpr int kMaxSlots = 40;
struct Slots {
    const double *f[kMaxSlots];
    double *b[kMaxSlots];
};

inline Slots make_slots(int n, const double *const *fwd, double *const *bar) {
    Slots s{};
    for (int i = 0; i < n; ++i) {
        s.f[i] = fwd[i];
        s.b[i] = bar[i];
    }
    return s;
}

#define ADJ_F (S_.f)
#define ADJ_B (S_.b)

// (entity, level) kernels: one wavefront per entity, lane = level (or interface)
#define ADJ_WAVE_KERNEL(name, call)                                                                 \
    __global__ __launch_bounds__(256) void name(long long nb, long long n_ent, int n_lev, mpas_hip_mesh m, \
                                                Slots S_, KArgs A) {                                \
        const int ent = wave_entity_base(nb);                                                       \
        const int k = wave_lane();                                                                  \
        if (ent >= n_ent || k >= n_lev) return;                                                     \
        call;                                                                                       \
    }

// (entity, level) kernels whose lanes exchange values with shuffles: one
// wavefront per entity, every lane present (the only early return is
// wave-uniform); the call guards the lanes past the column itself
#define ADJ_COLUMN_KERNEL(name, call)                                                               \
    __global__ __launch_bounds__(256) void name(long long nb, long long n_ent, int n_lev, mpas_hip_mesh m, \
                                                Slots S_, KArgs A) {                                \
        const int ent = wave_entity_base(nb);                                                       \
        if (ent >= n_ent) return;                                                                   \
        const int k = wave_lane();                                                                  \
        (void)n_lev;                                                                                \
        call;                                                                                       \
    }

// Scalar arguments shared by the adjoint and tangent kernels (unused fields ignored).
struct KArgs {
    int rk, S, q0, q1, s;
    double dt, coef, epssm, apvm, coef3, w_old, w_new;
    mpas_hip_dyn_params dp;
    mpas_hip_acoustic_params ap;
    DiagScratch ds;
    DiagDot dd;
    DynAdjScratch dy;
    ScalAdjScratch sc;
    MonoScratch mo;
    double *col;
    mpas_hip_inverse_tables inv;
};

// ---------------------------------------------------------------- scratch
inline size_t dyn_adj_doubles(const mpas_hip_mesh &m) {
    const size_t L = m.nLevels, C = (m.nCells + 1) * L, E = (m.nEdges + 1) * L, V = (m.nVertices + 1) * L,
                 W = (m.nCells + 1) * (L + 1);
    return C + E + C + V + C + W + C + W + C + C + V + E + 3 * E + 9 * E + 2 * C;
}
inline DynAdjScratch carve_dyn(const mpas_hip_mesh &m, double *p) {
    const size_t L = m.nLevels, C = (m.nCells + 1) * L, E = (m.nEdges + 1) * L, V = (m.nVertices + 1) * L,
                 W = (m.nCells + 1) * (L + 1);
    DynAdjScratch S;
    S.dpdz = p; p += C;
    S.delsq_u = p; p += E;
    S.delsq_div = p; p += C;
    S.delsq_vort = p; p += V;
    S.g_thadv = p; p += C;
    S.g_hadv = p; p += W;
    S.g_dpdz = p; p += C;
    S.g_delsq_w = p; p += W;
    S.g_delsq_theta = p; p += C;
    S.g_delsq_div = p; p += C;
    S.g_delsq_vort = p; p += V;
    S.g_delsq_u = p; p += E;
    S.g_tf = p; p += E;
    S.g_wf = p; p += E;
    S.g_q = p; p += E;
    for (int j = 0; j < 9; ++j) {
        S.x[j] = p;
        p += E;
    }
    S.g_dd = p; p += C;
    S.g_dof = p; p += C;
    return S;
}
inline size_t mono_adj_doubles(const mpas_hip_mesh &m) {
    const size_t L = m.nLevels, C = (m.nCells + 1) * L, E = (m.nEdges + 1) * L, W = (m.nCells + 1) * (L + 1);
    return 16 * C + 6 * W + 6 * E + (4 + kMaxEdgesAdj) * C + 2 * E;
}
inline MonoScratch carve_mono(const mpas_hip_mesh &m, double *p) {
    const size_t L = m.nLevels, C = (m.nCells + 1) * L, E = (m.nEdges + 1) * L, W = (m.nCells + 1) * (L + 1);
    MonoScratch X;
    X.smax = p; p += C; X.smin = p; p += C; X.wdtn = p; p += W; X.wup = p; p += W; X.wmid = p; p += W;
    X.srp = p; p += C; X.sinr = p; p += C; X.soutr = p; p += C; X.sinl = p; p += C; X.soutl = p; p += C;
    X.fho = p; p += E; X.fup = p; p += E; X.ftmp = p; p += E;
    X.g_srf = p; p += C; X.g_sinl = p; p += C; X.g_soutl = p; p += C; X.g_sinr = p; p += C; X.g_soutr = p; p += C;
    X.g_wmid = p; p += W; X.g_wup = p; p += W; X.g_fup = p; p += E; X.g_ftmp = p; p += E;
    X.g_smax = p; p += C; X.g_smin = p; p += C; X.g_wdtn = p; p += W;
    X.g_sr = p; p += C;
    for (int j = 0; j < 3 + kMaxEdgesAdj; ++j) {
        X.r5[j] = p;
        p += C;
    }
    X.g_hv = p; p += E; X.g_up = p; p += E;
    return X;
}

}  // namespace adj
}  // namespace mpas

#define ADJ_LAUNCH(kernel, n_ent, n_lev, stream)                                                      \
    do {                                                                                              \
        const int wpb_ = ::mpas::tuned_waves(#kernel) > 4 ? 4 : ::mpas::tuned_waves(#kernel);         \
        const long long nb_ = ::mpas::wave_blocks(n_ent, wpb_);                                       \
        MPAS_LAUNCH(kernel, dim3(::mpas::xcd_grid(nb_)), dim3(64 * wpb_), 0, stream,                  \
                    nb_ | (::mpas::tuned_order(#kernel) ? ::mpas::kLinearOrder : 0LL), (long long)(n_ent), \
                    (int)(n_lev), m, S_, A);                                                          \
    } while (0)


#define STREAM(s) ((hipStream_t)(s))
