// Adjoint kernels of the large step: launch wrappers over the bodies in
// adjoint_bodies.hpp / adjoint_dyn.hpp / adjoint_transport.hpp and the
// extern "C" entry points mpas_hip_*_vjp.
//
// Every entry point takes fwd[] (the forward inputs, then the forward
// outputs, of one call) and bar[] (their adjoints, same slots; input
// adjoints are accumulated into, output adjoints are read and may be used as
// accumulators).  Both are host arrays of device pointers.  Cross-entity
// contributions are added with hardware FP64 atomics (global_atomic_add_f64),
// so results are reproducible to rounding, not bit for bit.
#include "common.hpp"

namespace mpas {
namespace adj {
__device__ __forceinline__ void acc(double *p, double v) { unsafeAtomicAdd(p, v); }
}  // namespace adj
}  // namespace mpas

#include "adjoint_bodies.hpp"
#include "adjoint_dyn.hpp"
#include "adjoint_transport.hpp"

namespace mpas {
namespace adj {

constexpr int kMaxSlots = 40;
struct Slots {
    const double *f[kMaxSlots];
    double *b[kMaxSlots];
};

static Slots make_slots(int n, const double *const *fwd, double *const *bar) {
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

// Scalar arguments shared by the adjoint kernels (unused fields ignored).
struct KArgs {
    int rk, S, q0, q1, s;
    double dt, coef, epssm, apvm, coef3, w_old, w_new;
    mpas_hip_dyn_params dp;
    mpas_hip_acoustic_params ap;
    DiagScratch ds;
    DynAdjScratch dy;
    ScalAdjScratch sc;
    MonoScratch mo;
    double *col;
};

ADJ_WAVE_KERNEL(k_adj_smlstep, smlstep_vjp(m, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_divdamp, divdamp_vjp(m, A.coef, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_moist_cell, moist_cell_vjp(m, A.S, A.q0, A.q1, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_moist_edge, moist_edge_vjp(m, A.S, A.q0, A.q1, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_reconstruct, reconstruct_vjp(m, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_recover_w, recover_w_vjp(m, A.coef, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_recover_edge, recover_edge_vjp(m, A.coef, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_recover_cell, recover_cell_vjp(m, A.rk, A.dt, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_diag_edge2, diag_edge2_vjp(m, A.dt, A.apvm, A.rk, ADJ_F, ADJ_B, A.ds, ent, k))
ADJ_WAVE_KERNEL(k_adj_diag_cell, diag_cell_vjp(m, ADJ_F, ADJ_B, A.ds, ent, k))
ADJ_WAVE_KERNEL(k_adj_diag_vertex, diag_vertex_vjp(m, ADJ_F, ADJ_B, A.ds, ent, k))
ADJ_WAVE_KERNEL(k_adj_diag_edge1, diag_edge1_vjp(m, ADJ_F, ADJ_B, A.ds, ent, k))
ADJ_WAVE_KERNEL(k_adj_acoustic_edge, acoustic_edge_vjp(m, A.ap, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_fwd_cells, dyn_fwd_cells(m, ADJ_F, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_fwd_edges, dyn_fwd_edges(m, ADJ_F, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_fwd_del4, dyn_fwd_del4(m, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_a1, dyn_a1(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_a2, dyn_a2(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_a3, dyn_a3(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_a4, dyn_a4(m, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_a5, dyn_a5(m, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_dyn_a6, dyn_a6(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_adj_scal_cell, scal_cell_vjp(m, A.S, A.dt, A.coef3, A.w_old, A.w_new, ADJ_F, ADJ_B, A.sc, ent, k))
ADJ_WAVE_KERNEL(k_adj_scal_edge, scal_edge_vjp(m, A.S, ADJ_F, ADJ_B, A.sc, ent, k))
ADJ_WAVE_KERNEL(k_adj_scal_vert, scal_vert_vjp(m, A.S, A.dt, A.coef3, A.w_old, A.w_new, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_f1, mono_f1(m, A.S, A.s, A.coef3, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_f2, mono_f2(m, A.S, A.s, A.dt, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_f3, mono_f3(m, A.S, A.s, A.dt, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_f4, mono_f4(m, A.S, A.s, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r1, mono_r1(m, A.S, A.s, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r2, mono_r2(m, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r3, mono_r3(m, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r4, mono_r4(m, A.S, A.s, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r5, mono_r5(m, A.S, A.s, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r6, mono_r6(m, A.S, A.s, A.dt, A.coef3, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r7, mono_r7(m, A.S, A.s, A.dt, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_adj_mono_r8, mono_r8(m, A.dt, ADJ_F, ADJ_B, ent, k))

// column kernels: one thread per column
__global__ __launch_bounds__(256) void k_adj_vimp_col(mpas_hip_mesh m, Slots S_, KArgs A) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nCells) return;
    vimp_col_vjp(m, A.dt, A.epssm, ADJ_F, ADJ_B, c);
}
__global__ __launch_bounds__(256) void k_adj_acoustic_col(mpas_hip_mesh m, Slots S_, KArgs A) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nCells) return;
    acoustic_col_vjp(m, A.ap, ADJ_F, ADJ_B, A.col, c);
}

#define ADJ_LAUNCH(kernel, n_ent, n_lev, stream)                                                      \
    do {                                                                                              \
        const long long nb_ = ::mpas::wave_blocks(n_ent);                                             \
        MPAS_LAUNCH(kernel, dim3(::mpas::xcd_grid(nb_)), dim3(64 * ::mpas::kWaves), 0, stream, nb_,   \
                    (long long)(n_ent), (int)(n_lev), m, S_, A);                                      \
    } while (0)

// ---------------------------------------------------------------- scratch
static size_t dyn_adj_doubles(const mpas_hip_mesh &m) {
    const size_t L = m.nLevels, C = (m.nCells + 1) * L, E = (m.nEdges + 1) * L, V = (m.nVertices + 1) * L,
                 W = (m.nCells + 1) * (L + 1);
    return C + E + C + V + C + W + C + W + C + C + V + E;
}
static DynAdjScratch carve_dyn(const mpas_hip_mesh &m, double *p) {
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
    return S;
}
static size_t mono_adj_doubles(const mpas_hip_mesh &m) {
    const size_t L = m.nLevels, C = (m.nCells + 1) * L, E = (m.nEdges + 1) * L, W = (m.nCells + 1) * (L + 1);
    return 16 * C + 6 * W + 6 * E;
}
static MonoScratch carve_mono(const mpas_hip_mesh &m, double *p) {
    const size_t L = m.nLevels, C = (m.nCells + 1) * L, E = (m.nEdges + 1) * L, W = (m.nCells + 1) * (L + 1);
    MonoScratch X;
    X.smax = p; p += C; X.smin = p; p += C; X.wdtn = p; p += W; X.wup = p; p += W; X.wmid = p; p += W;
    X.srp = p; p += C; X.sinr = p; p += C; X.soutr = p; p += C; X.sinl = p; p += C; X.soutl = p; p += C;
    X.fho = p; p += E; X.fup = p; p += E; X.ftmp = p; p += E;
    X.g_srf = p; p += C; X.g_sinl = p; p += C; X.g_soutl = p; p += C; X.g_sinr = p; p += C; X.g_soutr = p; p += C;
    X.g_wmid = p; p += W; X.g_wup = p; p += W; X.g_fup = p; p += E; X.g_ftmp = p; p += E;
    X.g_smax = p; p += C; X.g_smin = p; p += C; X.g_wdtn = p; p += W;
    return X;
}

size_t vjp_scratch_doubles(const mpas_hip_mesh &m, int S) {
    const size_t L = m.nLevels, C = (m.nCells + 1) * L, E = (m.nEdges + 1) * L, V = (m.nVertices + 1) * L;
    size_t n = dyn_adj_doubles(m);
    n = std::max(n, 2 * V + C + E);                                          // diagnostics
    n = std::max(n, (size_t)acoustic::kScratchArrays * (L + 1) * m.nCells);  // acoustic columns
    n = std::max(n, C * (size_t)(S > 0 ? S : 1));                            // explicit transport
    n = std::max(n, mono_adj_doubles(m));                                    // monotonic transport
    return n;
}

}  // namespace adj
}  // namespace mpas

using namespace mpas;
using namespace mpas::adj;

#define STREAM(s) ((hipStream_t)(s))

static int vjp_check(const mpas_hip_mesh *m, const char *what) {
    if (!m) {
        set_error(std::string(what) + ": null mesh");
        return 1;
    }
    if (m->nLevels < 3 || m->nLevels > 63 || m->maxEdges > kMaxEdges || m->maxEdges2 > kMaxEdges2 ||
        m->maxAdv > kMaxAdv) {
        set_error(std::string(what) + ": mesh outside the kernels' limits");
        return 1;
    }
    return 0;
}
static int need(double *const *bar, int slot, const char *what) {
    if (!bar[slot]) {
        set_error(std::string(what) + ": this output adjoint must be provided (it is used as an accumulator)");
        return 1;
    }
    return 0;
}

extern "C" {

int64_t mpas_hip_vjp_scratch_doubles(int32_t nCells, int32_t nEdges, int32_t nVertices, int32_t nLevels,
                                     int32_t nScalars) {
    mpas_hip_mesh m{};
    m.nCells = nCells;
    m.nEdges = nEdges;
    m.nVertices = nVertices;
    m.nLevels = nLevels;
    return (int64_t)vjp_scratch_doubles(m, nScalars);
}

int mpas_hip_smlstep_pert_vjp(const mpas_hip_mesh *mesh, const double *const *fwd, double *const *bar, void *stream) {
    if (vjp_check(mesh, "smlstep_pert_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(smlstep::N, fwd, bar);
    KArgs A{};
    ADJ_LAUNCH(k_adj_smlstep, m.nCells + 1, m.nLevels + 1, STREAM(stream));
    return check_launch("smlstep_pert_vjp");
}

int mpas_hip_divergence_damping_vjp(const mpas_hip_mesh *mesh, double dts, double smdiv, double len_disp,
                                    const double *const *fwd, double *const *bar, void *stream) {
    if (vjp_check(mesh, "divergence_damping_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(divdamp::N, fwd, bar);
    KArgs A{};
    A.coef = 2.0 * smdiv * len_disp / dts;
    ADJ_LAUNCH(k_adj_divdamp, m.nEdges + 1, m.nLevels, STREAM(stream));
    return check_launch("divergence_damping_vjp");
}

int mpas_hip_moist_coefficients_vjp(const mpas_hip_mesh *mesh, int32_t nScalars, int32_t moist_start,
                                    int32_t moist_end, const double *const *fwd, double *const *bar, void *stream) {
    if (vjp_check(mesh, "moist_coefficients_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(moist::N, fwd, bar);
    KArgs A{};
    A.S = nScalars;
    A.q0 = moist_start - 1;
    A.q1 = moist_end;
    ADJ_LAUNCH(k_adj_moist_cell, m.nCells, m.nLevels, STREAM(stream));
    ADJ_LAUNCH(k_adj_moist_edge, m.nEdges, m.nLevels, STREAM(stream));
    return check_launch("moist_coefficients_vjp");
}

int mpas_hip_reconstruct_vjp(const mpas_hip_mesh *mesh, const double *const *fwd, double *const *bar, void *stream) {
    if (vjp_check(mesh, "reconstruct_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(recon::N, fwd, bar);
    KArgs A{};
    ADJ_LAUNCH(k_adj_reconstruct, m.nCells, m.nLevels, STREAM(stream));
    return check_launch("reconstruct_vjp");
}

int mpas_hip_recover_large_step_vjp(const mpas_hip_mesh *mesh, int32_t rk_step, double dt, int32_t ns,
                                    const double *const *fwd, double *const *bar, void *stream) {
    if (vjp_check(mesh, "recover_large_step_vjp")) return 1;
    if (need(bar, recover::rho_zz, "recover_large_step_vjp") || need(bar, recover::ru, "recover_large_step_vjp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(recover::N, fwd, bar);
    KArgs A{};
    A.rk = rk_step;
    A.dt = dt;
    A.coef = 1.0 / (double)ns;
    ADJ_LAUNCH(k_adj_recover_w, m.nCells, m.nLevels + 1, STREAM(stream));
    ADJ_LAUNCH(k_adj_recover_edge, m.nEdges, m.nLevels, STREAM(stream));
    ADJ_LAUNCH(k_adj_recover_cell, m.nCells, m.nLevels, STREAM(stream));
    return check_launch("recover_large_step_vjp");
}

int mpas_hip_solve_diagnostics_vjp(const mpas_hip_mesh *mesh, double dt, double apvm, int32_t rk_step,
                                   const double *const *fwd, double *const *bar, double *scratch, void *stream) {
    if (vjp_check(mesh, "solve_diagnostics_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    const size_t L = m.nLevels, C = (m.nCells + 1) * L, E = (m.nEdges + 1) * L, V = (m.nVertices + 1) * L;
    Slots S_ = make_slots(diag::N, fwd, bar);
    KArgs A{};
    A.dt = dt;
    A.apvm = apvm;
    A.rk = rk_step;
    A.ds = DiagScratch{scratch, scratch + V, scratch + 2 * V, scratch + 2 * V + C};
    fill_zero(scratch, 2 * V + C + E, s);
    ADJ_LAUNCH(k_adj_diag_edge2, m.nEdges, m.nLevels, s);
    ADJ_LAUNCH(k_adj_diag_cell, m.nCells, m.nLevels, s);
    ADJ_LAUNCH(k_adj_diag_vertex, m.nVertices, m.nLevels, s);
    ADJ_LAUNCH(k_adj_diag_edge1, m.nEdges, m.nLevels, s);
    return check_launch("solve_diagnostics_vjp");
}

int mpas_hip_vert_imp_coefs_vjp(const mpas_hip_mesh *mesh, double dts, double epssm, const double *const *fwd,
                                double *const *bar, void *stream) {
    if (vjp_check(mesh, "vert_imp_coefs_vjp")) return 1;
    if (need(bar, vimp::cofwz, "vert_imp_coefs_vjp") || need(bar, vimp::coftz, "vert_imp_coefs_vjp") ||
        need(bar, vimp::cofwt, "vert_imp_coefs_vjp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(vimp::N, fwd, bar);
    KArgs A{};
    A.dt = dts;
    A.epssm = epssm;
    MPAS_LAUNCH(k_adj_vimp_col, dim3(blocks_for(m.nCells)), dim3(kBlock), 0, STREAM(stream), m, S_, A);
    return check_launch("vert_imp_coefs_vjp");
}

int mpas_hip_acoustic_step_vjp(const mpas_hip_mesh *mesh, const mpas_hip_acoustic_params *p, const double *const *fwd,
                               double *const *bar, double *scratch, void *stream) {
    if (vjp_check(mesh, "acoustic_step_vjp")) return 1;
    if (need(bar, acoustic::ru_p, "acoustic_step_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(acoustic::N, fwd, bar);
    KArgs A{};
    A.ap = *p;
    A.col = scratch;
    MPAS_LAUNCH(k_adj_acoustic_col, dim3(blocks_for(m.nCells)), dim3(kBlock), 0, s, m, S_, A);
    ADJ_LAUNCH(k_adj_acoustic_edge, m.nEdges, m.nLevels, s);
    return check_launch("acoustic_step_vjp");
}

int mpas_hip_dyn_tend_vjp(const mpas_hip_mesh *mesh, const mpas_hip_dyn_params *p, const double *const *fwd,
                          double *const *bar, double *scratch, void *stream) {
    if (vjp_check(mesh, "dyn_tend_vjp")) return 1;
    if (need(bar, dyn::h_divergence, "dyn_tend_vjp") || need(bar, dyn::kdiff, "dyn_tend_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(dyn::N, fwd, bar);
    KArgs A{};
    A.dp = *p;
    A.dy = carve_dyn(m, scratch);
    const int L = m.nLevels, nC = m.nCells, nE = m.nEdges, nV = m.nVertices;
    const bool rk1 = (p->rk_step == 1);
    fill_zero(scratch, dyn_adj_doubles(m), s);
    const int ncv = nC > nV ? nC : nV;
    if (rk1) {
        ADJ_LAUNCH(k_adj_dyn_fwd_cells, nC, L, s);
        ADJ_LAUNCH(k_adj_dyn_fwd_edges, nE, L, s);
        ADJ_LAUNCH(k_adj_dyn_fwd_del4, ncv, L, s);
    }
    ADJ_LAUNCH(k_adj_dyn_a1, nC, L + 1, s);
    if (rk1) ADJ_LAUNCH(k_adj_dyn_a2, nC, L + 1, s);
    ADJ_LAUNCH(k_adj_dyn_a3, nE, L, s);
    if (rk1 && p->h_mom_eddy_visc4 > 0.0) ADJ_LAUNCH(k_adj_dyn_a4, ncv, L, s);
    if (rk1) ADJ_LAUNCH(k_adj_dyn_a5, nE, L, s);
    ADJ_LAUNCH(k_adj_dyn_a6, nC, L, s);
    return check_launch("dyn_tend_vjp");
}

int mpas_hip_advance_scalars_vjp(const mpas_hip_mesh *mesh, int32_t nScalars, double dt, double coef3,
                                 double weight_old, double weight_new, const double *const *fwd, double *const *bar,
                                 double *scratch, void *stream) {
    if (vjp_check(mesh, "advance_scalars_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(scal::N, fwd, bar);
    KArgs A{};
    A.S = nScalars;
    A.dt = dt;
    A.coef3 = coef3;
    A.w_old = weight_old;
    A.w_new = weight_new;
    A.sc = ScalAdjScratch{scratch};
    ADJ_LAUNCH(k_adj_scal_cell, m.nCells, m.nLevels, s);
    ADJ_LAUNCH(k_adj_scal_edge, m.nEdges, m.nLevels, s);
    ADJ_LAUNCH(k_adj_scal_vert, m.nCells, m.nLevels + 1, s);
    return check_launch("advance_scalars_vjp");
}

int mpas_hip_advance_scalars_mono_vjp(const mpas_hip_mesh *mesh, int32_t nScalars, double dt, double coef3,
                                      const double *const *fwd, double *const *bar, double *scratch, void *stream) {
    if (vjp_check(mesh, "advance_scalars_mono_vjp")) return 1;
    if (need(bar, mono::rho_zz_int, "advance_scalars_mono_vjp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(mono::N, fwd, bar);
    KArgs A{};
    A.S = nScalars;
    A.dt = dt;
    A.coef3 = coef3;
    A.mo = carve_mono(m, scratch);
    const int L = m.nLevels, nC = m.nCells, nE = m.nEdges;
    for (int sc = 0; sc < nScalars; ++sc) {
        A.s = sc;
        fill_zero(scratch, mono_adj_doubles(m), s);
        ADJ_LAUNCH(k_adj_mono_f1, nC, L + 1, s);
        ADJ_LAUNCH(k_adj_mono_f2, nE, L, s);
        ADJ_LAUNCH(k_adj_mono_f3, nC, L + 1, s);
        ADJ_LAUNCH(k_adj_mono_f4, nC, L, s);
        ADJ_LAUNCH(k_adj_mono_r1, nC, L, s);
        ADJ_LAUNCH(k_adj_mono_r2, nC, L + 1, s);
        ADJ_LAUNCH(k_adj_mono_r3, nE, L, s);
        ADJ_LAUNCH(k_adj_mono_r4, nC, L, s);
        ADJ_LAUNCH(k_adj_mono_r5, nC, L, s);
        ADJ_LAUNCH(k_adj_mono_r6, nC, L + 1, s);
        ADJ_LAUNCH(k_adj_mono_r7, nE, L, s);
    }
    ADJ_LAUNCH(k_adj_mono_r8, nC, L, s);
    return check_launch("advance_scalars_mono_vjp");
}

}  // extern "C"
