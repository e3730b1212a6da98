// Tangent-linear kernels of the large step: launch wrappers over the bodies
// in tangent_bodies.hpp / tangent_dyn.hpp / tangent_transport.hpp and the
// extern "C" entry points mpas_hip_*_jvp.
//
// Every entry point takes fwd[] (the forward inputs, then the forward
// outputs, of one call) and dot[] (their tangents, same slots; input
// tangents are read, a null input meaning a zero tangent; output tangents
// are written and must all be given).  Both are host arrays of device
// pointers.  Every thread gathers the tangents of the elements it owns, so
// there are no atomics and results are deterministic.
#include "adjoint_launch.hpp"

namespace mpas {
namespace adj {

ADJ_WAVE_KERNEL(k_tan_smlstep, smlstep_jvp(m, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_divdamp, divdamp_jvp(m, A.coef, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_moist_cell, moist_cell_jvp(m, A.S, A.q0, A.q1, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_moist_edge, moist_edge_jvp(m, A.S, A.q0, A.q1, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_reconstruct, reconstruct_jvp(m, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_recover_cell, recover_cell_jvp(m, A.rk, A.dt, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_recover_edge, recover_edge_jvp(m, A.coef, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_recover_w, recover_w_jvp(m, A.coef, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_diag_vertex, diag_vertex_jvp(m, ADJ_F, ADJ_B, A.dd, ent, k))
ADJ_WAVE_KERNEL(k_tan_diag_cell, diag_cell_jvp(m, ADJ_F, ADJ_B, A.dd, ent, k))
ADJ_WAVE_KERNEL(k_tan_diag_edge, diag_edge_jvp(m, A.dt, A.apvm, A.rk, ADJ_F, ADJ_B, A.dd, ent, k))
ADJ_WAVE_KERNEL(k_tan_acoustic_edge, acoustic_edge_jvp(m, A.ap, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_j1, dyn_j1(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_j2, dyn_j2(m, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_j3, dyn_j3(m, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_j4, dyn_j4(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_j5, dyn_j5(m, A.dp, ADJ_F, ADJ_B, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_fwd_cells, dyn_fwd_cells(m, ADJ_F, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_fwd_edges, dyn_fwd_edges(m, ADJ_F, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_dyn_fwd_del4, dyn_fwd_del4(m, A.dy, ent, k))
ADJ_WAVE_KERNEL(k_tan_scal_edge, scal_edge_jvp(m, A.S, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_scal_cell, scal_cell_jvp(m, A.S, A.dt, A.coef3, A.w_old, A.w_new, ADJ_F, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_rho, mono_rho_jvp(m, A.dt, ADJ_B, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_f1, mono_f1(m, A.S, A.s, A.coef3, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_f2, mono_f2(m, A.S, A.s, A.dt, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_f3, mono_f3(m, A.S, A.s, A.dt, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_f4, mono_f4(m, A.S, A.s, ADJ_F, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_m1, mono_m1(m, A.S, A.s, A.coef3, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_m2, mono_m2(m, A.S, A.s, A.dt, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_m3, mono_m3(m, A.S, A.s, A.dt, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_m4, mono_m4(m, A.S, A.s, ADJ_F, ADJ_B, A.mo, ent, k))
ADJ_WAVE_KERNEL(k_tan_mono_m5, mono_m5(m, A.S, A.s, ADJ_F, ADJ_B, A.mo, ent, k))

// column kernels: one thread per column, the padding column included
__global__ __launch_bounds__(256) void k_tan_vimp_col(mpas_hip_mesh m, Slots S_, KArgs A) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > m.nCells) return;
    vimp_col_jvp(m, A.dt, A.epssm, ADJ_F, ADJ_B, c);
}
__global__ __launch_bounds__(256) void k_tan_acoustic_col(mpas_hip_mesh m, Slots S_, KArgs A) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > m.nCells) return;
    acoustic_col_jvp(m, A.ap, ADJ_F, ADJ_B, A.col, c);
}

}  // namespace adj
}  // namespace mpas

using namespace mpas;
using namespace mpas::adj;

static int jvp_check(const mpas_hip_mesh *m, const char *what) {
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
// every output tangent in slots [first, n) must be given
static int need_outputs(double *const *dot, int first, int n, const char *what) {
    for (int i = first; i < n; ++i)
        if (!dot[i]) {
            set_error(std::string(what) + ": every output tangent must be provided (slot " + std::to_string(i) + ")");
            return 1;
        }
    return 0;
}
static dim3 column_grid(int n) { return dim3((unsigned)((n + 255) / 256)); }

extern "C" {

int mpas_hip_smlstep_pert_jvp(const mpas_hip_mesh *mesh, const double *const *fwd, double *const *dot, void *stream) {
    if (jvp_check(mesh, "smlstep_pert_jvp") || need_outputs(dot, smlstep::w_tend_out, smlstep::N, "smlstep_pert_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(smlstep::N, fwd, dot);
    KArgs A{};
    ADJ_LAUNCH(k_tan_smlstep, m.nCells + 1, m.nLevels + 1, STREAM(stream));
    return check_launch("smlstep_pert_jvp");
}

int mpas_hip_divergence_damping_jvp(const mpas_hip_mesh *mesh, double dts, double smdiv, double len_disp,
                                    const double *const *fwd, double *const *dot, void *stream) {
    if (jvp_check(mesh, "divergence_damping_jvp") ||
        need_outputs(dot, divdamp::ru_p_out, divdamp::N, "divergence_damping_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(divdamp::N, fwd, dot);
    KArgs A{};
    A.coef = 2.0 * smdiv * len_disp / dts;
    ADJ_LAUNCH(k_tan_divdamp, m.nEdges + 1, m.nLevels, STREAM(stream));
    return check_launch("divergence_damping_jvp");
}

int mpas_hip_moist_coefficients_jvp(const mpas_hip_mesh *mesh, int32_t nScalars, int32_t moist_start,
                                    int32_t moist_end, const double *const *fwd, double *const *dot, void *stream) {
    if (jvp_check(mesh, "moist_coefficients_jvp") ||
        need_outputs(dot, moist::cqw, moist::N, "moist_coefficients_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(moist::N, fwd, dot);
    KArgs A{};
    A.S = nScalars;
    A.q0 = moist_start - 1;
    A.q1 = moist_end;
    ADJ_LAUNCH(k_tan_moist_cell, m.nCells + 1, m.nLevels + 1, STREAM(stream));
    ADJ_LAUNCH(k_tan_moist_edge, m.nEdges + 1, m.nLevels, STREAM(stream));
    return check_launch("moist_coefficients_jvp");
}

int mpas_hip_reconstruct_jvp(const mpas_hip_mesh *mesh, const double *const *fwd, double *const *dot, void *stream) {
    if (jvp_check(mesh, "reconstruct_jvp") || need_outputs(dot, recon::ux, recon::N, "reconstruct_jvp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(recon::N, fwd, dot);
    KArgs A{};
    ADJ_LAUNCH(k_tan_reconstruct, m.nCells + 1, m.nLevels, STREAM(stream));
    return check_launch("reconstruct_jvp");
}

int mpas_hip_recover_large_step_jvp(const mpas_hip_mesh *mesh, int32_t rk_step, double dt, int32_t ns,
                                    const double *const *fwd, double *const *dot, void *stream) {
    if (jvp_check(mesh, "recover_large_step_jvp") ||
        need_outputs(dot, recover::wwAvg, recover::N, "recover_large_step_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(recover::N, fwd, dot);
    KArgs A{};
    A.rk = rk_step;
    A.dt = dt;
    A.coef = 1.0 / (double)ns;
    ADJ_LAUNCH(k_tan_recover_cell, m.nCells + 1, m.nLevels, STREAM(stream));
    ADJ_LAUNCH(k_tan_recover_edge, m.nEdges + 1, m.nLevels, STREAM(stream));
    ADJ_LAUNCH(k_tan_recover_w, m.nCells + 1, m.nLevels + 1, STREAM(stream));
    return check_launch("recover_large_step_jvp");
}

int mpas_hip_solve_diagnostics_jvp(const mpas_hip_mesh *mesh, double dt, double apvm, int32_t rk_step,
                                   const double *const *fwd, double *const *dot, double *scratch, void *stream) {
    if (jvp_check(mesh, "solve_diagnostics_jvp") ||
        need_outputs(dot, diag::h_edge, diag::N, "solve_diagnostics_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    const size_t L = m.nLevels, V = (m.nVertices + 1) * L;
    Slots S_ = make_slots(diag::N, fwd, dot);
    KArgs A{};
    A.dt = dt;
    A.apvm = apvm;
    A.rk = rk_step;
    A.dd = DiagDot{scratch, scratch + V, scratch + 2 * V};
    ADJ_LAUNCH(k_tan_diag_vertex, m.nVertices + 1, m.nLevels, s);
    ADJ_LAUNCH(k_tan_diag_cell, m.nCells + 1, m.nLevels, s);
    ADJ_LAUNCH(k_tan_diag_edge, m.nEdges + 1, m.nLevels, s);
    return check_launch("solve_diagnostics_jvp");
}

int mpas_hip_vert_imp_coefs_jvp(const mpas_hip_mesh *mesh, double dts, double epssm, const double *const *fwd,
                                double *const *dot, void *stream) {
    if (jvp_check(mesh, "vert_imp_coefs_jvp") || need_outputs(dot, vimp::cofrz, vimp::N, "vert_imp_coefs_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    Slots S_ = make_slots(vimp::N, fwd, dot);
    KArgs A{};
    A.dt = dts;
    A.epssm = epssm;
    MPAS_LAUNCH(k_tan_vimp_col, column_grid(m.nCells + 1), dim3(256), 0, STREAM(stream), m, S_, A);
    return check_launch("vert_imp_coefs_jvp");
}

int mpas_hip_acoustic_step_jvp(const mpas_hip_mesh *mesh, const mpas_hip_acoustic_params *p, const double *const *fwd,
                               double *const *dot, double *scratch, void *stream) {
    if (jvp_check(mesh, "acoustic_step_jvp") || need_outputs(dot, acoustic::ru_p, acoustic::N, "acoustic_step_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(acoustic::N, fwd, dot);
    KArgs A{};
    A.ap = *p;
    A.col = scratch;
    ADJ_LAUNCH(k_tan_acoustic_edge, m.nEdges + 1, m.nLevels, s);
    MPAS_LAUNCH(k_tan_acoustic_col, column_grid(m.nCells + 1), dim3(256), 0, s, m, S_, A);
    return check_launch("acoustic_step_jvp");
}

int mpas_hip_dyn_tend_jvp(const mpas_hip_mesh *mesh, const mpas_hip_dyn_params *p, const double *const *fwd,
                          double *const *dot, double *scratch, void *stream) {
    if (jvp_check(mesh, "dyn_tend_jvp") || need_outputs(dot, dyn::tend_rho, dyn::N, "dyn_tend_jvp")) return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(dyn::N, fwd, dot);
    KArgs A{};
    A.dp = *p;
    A.dy = carve_dyn(m, scratch);
    const int L = m.nLevels, nC = m.nCells, nE = m.nEdges, nV = m.nVertices;
    const bool rk1 = (p->rk_step == 1);
    const int ncv = nC > nV ? nC : nV;
    if (rk1) {
        ADJ_LAUNCH(k_tan_dyn_fwd_cells, nC, L, s);
        ADJ_LAUNCH(k_tan_dyn_fwd_edges, nE, L, s);
        ADJ_LAUNCH(k_tan_dyn_fwd_del4, ncv, L, s);
    }
    ADJ_LAUNCH(k_tan_dyn_j1, nC + 1, L + 1, s);
    if (rk1) ADJ_LAUNCH(k_tan_dyn_j2, nE + 1, L, s);
    if (rk1 && p->h_mom_eddy_visc4 > 0.0) ADJ_LAUNCH(k_tan_dyn_j3, ncv, L, s);
    ADJ_LAUNCH(k_tan_dyn_j4, nE + 1, L, s);
    ADJ_LAUNCH(k_tan_dyn_j5, nC + 1, L + 1, s);
    return check_launch("dyn_tend_jvp");
}

int mpas_hip_advance_scalars_jvp(const mpas_hip_mesh *mesh, int32_t nScalars, double dt, double coef3,
                                 double weight_old, double weight_new, const double *const *fwd, double *const *dot,
                                 void *stream) {
    if (jvp_check(mesh, "advance_scalars_jvp") ||
        need_outputs(dot, scal::scalar_new_out, scal::N, "advance_scalars_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(scal::N, fwd, dot);
    KArgs A{};
    A.S = nScalars;
    A.dt = dt;
    A.coef3 = coef3;
    A.w_old = weight_old;
    A.w_new = weight_new;
    ADJ_LAUNCH(k_tan_scal_edge, m.nEdges + 1, m.nLevels, s);
    ADJ_LAUNCH(k_tan_scal_cell, m.nCells + 1, m.nLevels, s);
    return check_launch("advance_scalars_jvp");
}

int mpas_hip_advance_scalars_mono_jvp(const mpas_hip_mesh *mesh, int32_t nScalars, double dt, double coef3,
                                      const double *const *fwd, double *const *dot, double *scratch, void *stream) {
    if (jvp_check(mesh, "advance_scalars_mono_jvp") ||
        need_outputs(dot, mono::scalars_new_out, mono::N, "advance_scalars_mono_jvp"))
        return 1;
    const mpas_hip_mesh &m = *mesh;
    hipStream_t s = STREAM(stream);
    Slots S_ = make_slots(mono::N, fwd, dot);
    KArgs A{};
    A.S = nScalars;
    A.dt = dt;
    A.coef3 = coef3;
    A.mo = carve_mono(m, scratch);
    const int L = m.nLevels, nC = m.nCells, nE = m.nEdges;
    ADJ_LAUNCH(k_tan_mono_rho, nC + 1, L, s);
    for (int sc = 0; sc < nScalars; ++sc) {
        A.s = sc;
        ADJ_LAUNCH(k_tan_mono_f1, nC, L + 1, s);
        ADJ_LAUNCH(k_tan_mono_f2, nE, L, s);
        ADJ_LAUNCH(k_tan_mono_f3, nC, L + 1, s);
        ADJ_LAUNCH(k_tan_mono_f4, nC, L, s);
        ADJ_LAUNCH(k_tan_mono_m1, nC, L + 1, s);
        ADJ_LAUNCH(k_tan_mono_m2, nE, L, s);
        ADJ_LAUNCH(k_tan_mono_m3, nC, L + 1, s);
        ADJ_LAUNCH(k_tan_mono_m4, nC, L, s);
        ADJ_LAUNCH(k_tan_mono_m5, nC + 1, L, s);
    }
    return check_launch("advance_scalars_mono_jvp");
}

}  // extern "C"
