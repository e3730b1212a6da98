"""Algorithmic HBM bytes of the split-explicit step (roofline denominator).

Definition: every distinct array a function (or kernel) reads is counted
once and every array it writes is counted once per invocation, with perfect
on-chip reuse of stencil neighbours.  Connectivity and mesh metrics are
counted once per invocation at their device width (int32 indices, float64
metrics).  Padding rows are ignored.

Two inventories:

* ``FUNCTIONS`` - per reference function of the large step
  (``atm_srk3``, ``mpas_pytorch/dynamics/time_integration.py:3420-4894``)
  with its calls per step; ``step_bytes`` sums them.  A fused or split GPU
  implementation is charged the same per-function total.
* ``KERNELS`` - per HIP kernel launch of this library (the arrays that
  kernel must touch), used for the per-kernel roofline fraction.

Classes: C = nCells*L, W = nCells*(L+1), E = nEdges*L, V = nVertices*L,
CS = C*nScalars, ES = E*nScalars, L1 = L, and mesh tables ``cell_e``
(nEdgesOnCell + edgesOnCell + edgesOnCell_sign + invAreaCell),
``edge_c`` (cellsOnEdge + dcEdge/dvEdge/invDcEdge), ``edge_v``
(verticesOnEdge + invDvEdge), ``vert_e`` (edgesOnVertex + sign +
invAreaTriangle), ``edge_e`` (nEdgesOnEdge + edgesOnEdge + weightsOnEdge),
``edge_adv`` (nAdvCellsForEdge + advCellsForEdge + adv_coefs +
adv_coefs_3rd), ``cell_v`` (verticesOnCell + kiteForCell),
``vert_kite`` (kiteAreasOnVertex + fVertex), ``cell_c`` (cellsOnCell),
``cell_rec`` (coeffs_reconstruct + latCell + lonCell), ``cell_def``
(defc_a + defc_b).
"""

from __future__ import annotations

from dataclasses import dataclass

ME = 6      # maxEdges
ME2 = 12    # maxEdges2 (edgesOnEdge width of these meshes)
MADV = 15   # advCellsForEdge width


def class_bytes(cls: str, nC: int, nE: int, nV: int, L: int, S: int) -> int:
    """Bytes of one array of class ``cls``."""
    f8, i4 = 8, 4
    table = {
        "C": nC * L * f8, "W": nC * (L + 1) * f8, "E": nE * L * f8, "V": nV * L * f8,
        "CS": nC * L * S * f8, "ES": nE * L * S * f8, "L1": L * f8,
        "cell_e": nC * (i4 + ME * i4 + ME * f8 + f8),
        "edge_c": nE * (2 * i4 + 3 * f8),
        "edge_v": nE * (2 * i4 + f8),
        "vert_e": nV * (3 * i4 + 3 * f8 + f8),
        "edge_e": nE * (i4 + ME2 * i4 + ME2 * f8),
        "edge_adv": nE * (i4 + MADV * i4 + 2 * MADV * f8),
        "cell_v": nC * (ME * i4 + ME * i4),
        "vert_kite": nV * (3 * f8 + f8),
        "cell_c": nC * ME * i4,
        "cell_rec": nC * (ME * 3 * f8 + 2 * f8),
        "cell_def": nC * 2 * ME * f8,
        "edge_1": nE * f8,
    }
    return table[cls]


@dataclass(frozen=True)
class Inventory:
    calls: int
    reads: tuple[str, ...]
    writes: tuple[str, ...]

    def bytes(self, nC, nE, nV, L, S) -> int:
        return sum(class_bytes(c, nC, nE, nV, L, S) for c in self.reads + self.writes)


# ---------------------------------------------------------------------------
# reference functions, per call (RK3, split 3, 2 acoustic substeps: the
# per-stage counts are 1, 1, 2 substeps -> 4 per split step, 12 per step, of
# which 9 are the first substep of their stage)
# ---------------------------------------------------------------------------
FUNCTIONS: dict[str, Inventory] = {
    # moist_coefficients.py:124: scalars -> cqw, cqu, qtot
    "atm_compute_moist_coefficients": Inventory(1, ("CS", "edge_c"), ("W", "E", "C")),
    # vert_imp_coefs.py:188: zz, cqw, p, t, rb, rtb, pb, rt, qtot -> cofrz, cofwr, cofwz, coftz,
    # cofwt, a, alpha, gamma
    "atm_compute_vert_imp_coefs": Inventory(6, ("C", "W", "C", "C", "C", "C", "C", "C", "C", "L1"),
                                            ("L1", "C", "C", "W", "C", "C", "C", "C")),
    # dyn_tend.py:10836 at rk_step 1: u, v, ru, rho_edge, pv_edge, cqu, zxu (E); divergence, ke,
    # theta_m, theta_m_save, rho_zz, pp, rr_save, qtot, rt_diabatic, rho_base, zz (C); vorticity (V);
    # w, rw, rw_save, cqw (W) -> tend_u, tend_u_euler (E), tend_w, tend_w_euler (W), tend_theta,
    # tend_theta_euler, tend_rho, kdiff, h_divergence, rthdynten (C)
    "atm_compute_dyn_tend[rk1]": Inventory(
        3, ("E",) * 7 + ("C",) * 10 + ("V",) + ("W",) * 4
        + ("cell_e", "edge_c", "edge_v", "vert_e", "edge_e", "edge_adv", "cell_def", "edge_1", "edge_1"),
        ("E", "E", "W", "W", "C", "C", "C", "C", "C", "C")),
    # rk_step 2/3: u, ru, ru_save, rho_edge, pv_edge, tend_u_euler (E); ke, theta_m, theta_m_save,
    # rho_zz, rt_diabatic, tend_theta_euler, tend_rho (C); w, rw, rw_save, tend_w_euler (W)
    # -> tend_u (E), tend_w (W), tend_theta, h_divergence, rthdynten (C)
    "atm_compute_dyn_tend[rk2,3]": Inventory(
        6, ("E",) * 6 + ("C",) * 7 + ("W",) * 4 + ("cell_e", "edge_c", "edge_e", "edge_adv"),
        ("E", "W", "C", "C", "C")),
    # smlstep_pert.py:373: u_tend, w_tend, zz (+ flat-terrain zb/zb3 skipped) -> w_tend
    "atm_set_smlstep_pert_variables": Inventory(9, ("E", "W", "C", "cell_e"), ("W",)),
    # acoustic_step.py:1297, a later substep (small_step > 1; 3 per step): ru_p, ruAvg, tend_ru, cqu,
    # zxu (E); rho_pp, rtheta_pp, tend_rho, tend_rt, cofwr, cofwz, cofwt, a, alpha, gamma, zz, rho_zz,
    # theta_m, exner, dss (C); rw_p, wwAvg, tend_rw, coftz, rw_save, rw, w (W) -> ru_p, ruAvg (E),
    # rho_pp, rtheta_pp, rtheta_pp_old (C), rw_p, wwAvg (W)
    "atm_advance_acoustic_step[later substeps]": Inventory(
        3, ("E",) * 5 + ("C",) * 15 + ("W",) * 7 + ("cell_e", "edge_c", "L1"),
        ("E", "E", "C", "C", "C", "W", "W")),
    # the first substep of a stage (small_step == 1; 9 per step, acoustic_step.py:1702-1761): the
    # perturbations start from zero, so no ru_p, ruAvg, cqu, zxu, rho_pp, rtheta_pp, exner, rw_p or
    # wwAvg is read: tend_ru (E); tend_rho, tend_rt, cofwr, cofwz, cofwt, a, alpha, gamma, zz,
    # rho_zz, theta_m, dss (C); tend_rw, coftz, rw_save, rw, w (W) -> the same writes
    "atm_advance_acoustic_step[first substep]": Inventory(
        9, ("E",) + ("C",) * 12 + ("W",) * 5 + ("cell_e", "edge_c", "L1"),
        ("E", "E", "C", "C", "C", "W", "W")),
    # divergence_damping.py:299: ru_p, theta_m, rtheta_pp, rtheta_pp_old -> ru_p
    "atm_divergence_damping_3d": Inventory(12, ("E", "C", "C", "C", "edge_c"), ("E",)),
    # recover_large_step.py:1090: rho_p_save, rho_pp, rtheta_p_save, rtheta_pp, rt_diabatic, rho_base,
    # rtheta_base, zz (C); rw_save, rw_p, wwAvg (W); ru_save, ru_p, ruAvg (E) -> rho_p, rho_zz, rtheta_p,
    # theta_m (C), rw, w, wwAvg (W), ruAvg, ru, u (E); exner, pressure_p at rk 3 (amortised: 1/3 of
    # the calls read exner_base and write exner, pressure_p)
    "atm_recover_large_step_variables": Inventory(
        9, ("C",) * 8 + ("W",) * 3 + ("E",) * 3 + ("cell_e", "edge_c"),
        ("C",) * 4 + ("W",) * 3 + ("E",) * 3),
    "atm_recover_large_step_variables[rk3 exner]": Inventory(3, ("C",), ("C", "C")),
    # solve_diagnostics.py:1052: u, h, (v_in) -> h_edge, v, pv_edge (E), divergence, ke (C), vorticity (V)
    "atm_compute_solve_diagnostics": Inventory(
        10, ("E", "C", "cell_e", "edge_c", "edge_v", "vert_e", "cell_v", "vert_kite"),
        ("E", "E", "E", "C", "C", "V")),
    "atm_compute_solve_diagnostics[v recon]": Inventory(4, ("edge_e",), ()),
    # advance_scalars.py:1272 (2 explicit stages): scalar_old, scalar_new (CS), rho_zz_old, rho_zz_new
    # (C), uhAvg (E), wwAvg (W) -> scalar_new (CS), horiz_flux (ES)
    "atm_advance_scalars": Inventory(2, ("CS", "CS", "C", "C", "E", "W", "cell_e", "edge_adv", "L1"),
                                     ("CS", "ES")),
    # advance_scalars_mono.py:2878: scalars_old, scalars_new (CS), rho_zz_old (C), uhAvg (E), wwAvg (W)
    # -> scalars_new (CS), rho_zz_int (C)
    "atm_advance_scalars_mono": Inventory(1, ("CS", "CS", "C", "E", "W", "cell_e", "cell_c", "edge_c",
                                              "edge_adv", "L1"), ("CS", "C")),
    # vector_reconstruction.py:311: u -> 5 cell vectors (+ ur_cell, vr_cell copies)
    "mpas_reconstruct_2d": Inventory(1, ("E", "cell_e", "cell_rec"), ("C",) * 7),
    # atm_srk3 set-up copies and split averaging (time_integration.py:2584-2654, :3820-3942):
    # 6 state copies, 2 flux accumulations, 2 averages, rho_zz_old/scalars_old snapshots
    "atm_srk3 copies": Inventory(1, ("E", "W", "C", "C", "C", "C", "CS", "E", "W", "E", "W"),
                                 ("E", "W", "C", "C", "C", "C", "CS", "E", "W", "E", "W")),
}


def step_bytes(nC: int, nE: int, nV: int, L: int, S: int) -> int:
    """Algorithmic bytes of one large step (sum over the reference functions)."""
    return sum(inv.calls * inv.bytes(nC, nE, nV, L, S) for inv in FUNCTIONS.values())


def step_bytes_breakdown(nC: int, nE: int, nV: int, L: int, S: int) -> dict[str, int]:
    return {k: inv.calls * inv.bytes(nC, nE, nV, L, S) for k, inv in FUNCTIONS.items()}


# ---------------------------------------------------------------------------
# HIP kernels of this library, per launch, per variant.  A kernel whose inputs
# depend on the RK stage or the acoustic substep has one inventory per variant;
# ``launch_schedule`` gives the variant of every launch of one large step, in
# launch order, so a timed set of launches is charged launch by launch.
# ---------------------------------------------------------------------------
KERNEL_VARIANTS: dict[str, dict[str, Inventory]] = {
    # column Thomas acoustic cell pass.  later substeps: ru_p (E); theta_m, tend_rho, tend_rt,
    # rho_pp, rtheta_pp, cofwz, cofwr, cofwt, zz, a, alpha, gamma, rho_zz, dss (C); rw_p, wwAvg,
    # tend_rw, coftz, rw_save, rw, w (W); cofrz, rdzw, fzm, fzp (L) -> rtheta_pp_old, rho_pp,
    # rtheta_pp (C), rw_p, wwAvg (W).  first substep: no rho_pp, rtheta_pp, rw_p, wwAvg read
    "k_acoustic_cells": {
        "later": Inventory(1, ("E",) + ("C",) * 14 + ("W",) * 7 + ("L1",) * 4 + ("cell_e", "edge_c"),
                           ("C", "C", "C", "W", "W")),
        "first": Inventory(1, ("E",) + ("C",) * 12 + ("W",) * 5 + ("L1",) * 4 + ("cell_e", "edge_c"),
                           ("C", "C", "C", "W", "W")),
    },
    # edge pass.  later substeps (with the previous substep's divergence damping folded in):
    # rtheta_pp, rho_pp, exner, zz, theta_m, rtheta_pp_old (C); cqu, tend_ru, ru_p, ruAvg, zxu (E)
    # -> ru_p, ruAvg (E).  first substep: tend_ru (E) -> ru_p, ruAvg (E)
    "k_acoustic_edges": {
        "later": Inventory(1, ("C",) * 6 + ("E",) * 5 + ("edge_c",), ("E", "E")),
        "first": Inventory(1, ("E",), ("E", "E")),
    },
    # stand-alone divergence damping (per-function entry point only; the step folds it)
    "k_divdamp": {"any": Inventory(1, ("C", "C", "C", "E", "edge_c"), ("E",))},
    # Smagorinsky kdiff (first stage only): u, v (E) -> kdiff (C)
    "k_dyn_kdiff": {"rk1": Inventory(1, ("E", "E", "cell_e", "cell_def"), ("C",))},
    # mass-flux divergence.  rk1: ru, rho_edge (E); rw, w (W); rr_save, qtot, rho_base, theta_m,
    # kdiff (C) -> h_divergence, tend_rho, dpdz, delsq_theta, tend_theta_euler (C), delsq_w,
    # tend_w_euler (W).  rk2/3: ru (E) -> h_divergence (C)
    "k_dyn_cells_a": {
        "rk1": Inventory(1, ("E", "E", "W", "W") + ("C",) * 5 + ("cell_e", "edge_c", "edge_1", "L1"),
                         ("C",) * 5 + ("W", "W")),
        "rk23": Inventory(1, ("E", "cell_e", "edge_c"), ("C",)),
    },
    # u tendency.  rk1 (del4 follows): u, pv_edge, rho_edge, cqu, zxu (E); rw (W); ke, h_div, pp,
    # dpdz, zz, divergence, kdiff (C); vorticity (V) -> tend_u, tend_u_euler, delsq_u (E).
    # rk2/3: u, pv_edge, rho_edge, tend_u_euler (E); rw (W); ke, h_div (C) -> tend_u (E)
    "k_dyn_edges_a": {
        "rk1": Inventory(1, ("E",) * 5 + ("W",) + ("C",) * 7 + ("V",)
                         + ("edge_c", "edge_e", "edge_v", "edge_1", "L1"), ("E", "E", "E")),
        "rk23": Inventory(1, ("E",) * 4 + ("W",) + ("C",) * 2 + ("edge_c", "edge_e", "L1"), ("E",)),
    },
    # del4 first half (first stage): delsq_u (E) -> delsq_vort (V), delsq_div (C)
    "k_dyn_del4_vc": {"rk1": Inventory(1, ("E", "vert_e", "cell_e"), ("V", "C"))},
    # del4 second half (first stage): rho_edge, tend_u_euler, tend_u (E); delsq_div (C);
    # delsq_vort (V) -> tend_u_euler, tend_u (E)
    "k_dyn_edges_b": {"rk1": Inventory(1, ("E", "E", "E", "C", "V", "edge_c", "edge_v", "edge_1"), ("E", "E"))},
    # advective fluxes through each edge.  rk1: ru (E), w (W), theta_m (C) -> w and theta fluxes (E).
    # rk2/3: ru, ru_save (E), w (W), theta_m, theta_m_save (C) -> w, theta and theta-perturbation fluxes
    "k_dyn_edge_adv": {
        "rk1": Inventory(1, ("E", "W", "C", "edge_adv", "L1"), ("E", "E")),
        "rk23": Inventory(1, ("E", "E", "W", "C", "C", "edge_adv", "edge_c", "L1"), ("E", "E", "E")),
    },
    # w / theta tendencies.  rk1: wflux, tflux (E); w, rw, rw_save, tend_w_euler, cqw, delsq_w (W);
    # theta_m, theta_m_save, rho_zz, pp, rt_diab, dpdz, delsq_theta, tend_rho, tend_theta_euler (C)
    # -> tend_w_euler, tend_w (W), tend_theta_euler, tend_theta, rthdynten (C).  rk2/3: wflux, tflux,
    # tpert (E); w, rw, rw_save, tend_w_euler (W); theta_m, theta_m_save, rho_zz, rt_diab, tend_rho,
    # tend_theta_euler (C) -> tend_w (W), tend_theta, rthdynten (C)
    "k_dyn_cells_b": {
        "rk1": Inventory(1, ("E",) * 2 + ("W",) * 6 + ("C",) * 9 + ("cell_e", "edge_c", "edge_1", "L1"),
                         ("W", "W", "C", "C", "C")),
        "rk23": Inventory(1, ("E",) * 3 + ("W",) * 4 + ("C",) * 6 + ("cell_e", "L1"), ("W", "C", "C")),
    },
    "k_recover_cells": {"any": Inventory(1, ("C",) * 8 + ("W",) * 3, ("C",) * 4 + ("W",) * 2)},
    "k_recover_edges": {"any": Inventory(1, ("E", "E", "E", "C", "edge_c"), ("E", "E", "E"))},
    "k_recover_w": {"any": Inventory(1, ("E", "W", "C", "C", "cell_e"), ("W",))},
    "k_smlstep": {"any": Inventory(1, ("W", "C"), ("W",))},
    "k_vimp": {"any": Inventory(1, ("C",) * 8 + ("W",), ("C",) * 6 + ("W",))},
    "k_diag_vertices": {"any": Inventory(1, ("E", "E", "vert_e"), ("V", "V", "V"))},
    "k_diag_cells": {"any": Inventory(1, ("E", "E", "V", "V", "cell_e", "cell_v", "vert_kite"), ("C", "C", "C"))},
    "k_diag_edges2": {"any": Inventory(1, ("E", "V", "C", "C", "edge_c", "edge_v", "edge_e"), ("E", "E", "E"))},
}
# one inventory per kernel, for callers that do not split by variant: the general (later / rk2,3) one
KERNELS: dict[str, Inventory] = {
    name: v.get("later") or v.get("rk23") or v.get("any") or v.get("rk1") for name, v in KERNEL_VARIANTS.items()}


def substep_counts(order: int = 3, ns: int = 2) -> list[int]:
    """Acoustic substeps of the three RK stages (compute_rk_weights, time_integration.py:2787-2856)."""
    half = max(1, ns // 2)
    return [half, half, ns] if order == 2 else [1, half, ns]


def launch_schedule(name: str, order: int = 3, ns: int = 2, split: int = 3) -> list[str]:
    """The variant of every launch of kernel ``name`` in one large step, in launch order."""
    if name in ("k_acoustic_cells", "k_acoustic_edges"):
        per_split = [("first" if s == 1 else "later") for n in substep_counts(order, ns) for s in range(1, n + 1)]
        return per_split * split
    if name in ("k_dyn_cells_a", "k_dyn_edges_a", "k_dyn_edge_adv", "k_dyn_cells_b"):
        return ["rk1", "rk23", "rk23"] * split
    if name in ("k_dyn_kdiff", "k_dyn_del4_vc", "k_dyn_edges_b"):
        return ["rk1"] * split
    if name in ("k_recover_cells", "k_recover_edges", "k_recover_w", "k_smlstep"):
        return ["any"] * (3 * split)
    if name == "k_vimp":
        return ["any"] * ((2 if order == 3 else 1) * split)
    return ["any"]


def kernel_bytes(name: str, nC: int, nE: int, nV: int, L: int, S: int, variant: str | None = None) -> int:
    """Algorithmic bytes of one launch of ``name`` (the given variant, or the general one)."""
    if variant is None:
        return KERNELS[name].bytes(nC, nE, nV, L, S)
    return KERNEL_VARIANTS[name][variant].bytes(nC, nE, nV, L, S)


def schedule_bytes(name: str, variants: list[str], nC: int, nE: int, nV: int, L: int, S: int) -> int:
    """Algorithmic bytes of a sequence of launches of ``name`` (variant by variant)."""
    return sum(kernel_bytes(name, nC, nE, nV, L, S, v) for v in variants)
