"""Replays the recorded reference kernel calls through an implementation.

``run_call(api, mesh, variant, inputs)`` calls the entry point of ``api``
(the CPU oracle ``oracle.oracle`` or the GPU ``mpas_dycore_pytorch_amd.kernels``;
both expose the same function names and argument meaning) that matches the
recorded reference variant, and returns ``[(name, got, expected_key)]``
pairs for comparison with the recorded outputs.
"""

from __future__ import annotations

import numpy as np

ACTIVE = "_vertical_implicit_coefficients_%s_active"


def _full(a):
    """Re-pad a compact (active-rows) vertical coefficient to the padded layout."""
    a = np.asarray(a)
    full = np.zeros((a.shape[0] + 1,) + a.shape[1:])
    full[: a.shape[0]] = a
    return full


def _np(x):
    try:
        import torch

        if isinstance(x, torch.Tensor):
            return x.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return np.asarray(x)


def run_call(api, mesh, variant: str, I: dict, out_keys) -> list[tuple[str, np.ndarray, str]]:
    pairs: list[tuple[str, object, str]] = []
    if variant.startswith("mpas.moist"):
        r = api.moist_coefficients(mesh, I["scalars"], I["moist_start"], I["moist_end"])
        pairs = [(k, r[k], k) for k in ("cqw", "cqu", "qtot")]
    elif variant.startswith("mpas.vertical_implicit"):
        r = api.vert_imp_coefs(mesh, I["dts"], I["epssm"], I["cqw"], I["p"], I["t"], I["rt"], I["qtot"])
        pairs = [("cofrz", r["cofrz"], "cofrz")]
        pairs += [(k, r[k], ACTIVE % k) for k in ("cofwr", "cofwz", "coftz", "cofwt", "a_tri", "alpha_tri",
                                                   "gamma_tri")]
    elif "diagnostic_solve" in variant:
        r = api.solve_diagnostics(mesh, I["dt"], I["config_apvm_upwinding"], I["rk_step"], I["u"], I["h"],
                                  I.get("v_in"))
        pairs = [(k, r[k], k) for k in r if k in out_keys]
    elif "dyn_tend" in variant:
        p = api.dyn_params(I["rk_step"], I["dt"], I["config_len_disp"])
        carried = None
        if I["rk_step"] != 1:
            carried = {k: I[k + "_in"] for k in ("tend_rho", "tend_u_euler", "tend_w_euler", "tend_theta_euler",
                                                 "kdiff")}
        r = api.dyn_tend(mesh, p, I, carried)
        pairs = [(k, r[k], k) for k in ("tend_u", "tend_rho", "tend_theta", "tend_w", "tend_u_euler",
                                         "tend_w_euler", "tend_theta_euler", "h_divergence", "kdiff",
                                         "rthdynten")]
    elif "smlstep" in variant:
        pairs = [("w_tend", api.smlstep_pert(mesh, I["u_tend"], I["w_tend"]), "w_tend")]
    elif "acoustic" in variant:
        f = dict(I)
        for k in ("cofwr", "cofwz", "coftz", "cofwt", "a_tri", "alpha_tri", "gamma_tri"):
            f[k] = _full(I[ACTIVE % k])
        p = api.acoustic_params(I["small_step"], I["dts"], I["epssm"])
        r = api.acoustic_step(mesh, p, f)
        pairs = [(k, r[k], k) for k in r if k in out_keys]
    elif "divergence_damping" in variant:
        r = api.divergence_damping(mesh, I["dts"], I["config_smdiv"], I["config_len_disp"], I["theta_m"],
                                   I["rtheta_pp"], I["rtheta_pp_old"], I["ru_p"])
        pairs = [("ru_p", r, "ru_p")]
    elif "recovery" in variant:
        r = api.recover_large_step(mesh, I["rk_step"], I["dt"], I["ns"], I)
        pairs = [(k, r[k], k) for k in r if k in out_keys]
    elif "explicit_scalar" in variant:
        wn = {1: 1.0 / 3.0, 2: 0.5}.get(int(I["rk_step"]), 1.0)
        r, hf = api.advance_scalars(mesh, I["dt"], I["coef_3rd_order"], 1.0 - wn, wn, I["scalar_old"],
                                    I["scalar_new"], I["rho_zz_old"], I["rho_zz_new"], I["uhAvg"], I["wwAvg"])
        pairs = [("scalar_new", r, "scalar_new"), ("horiz_flux", hf, "horiz_flux_arr")]
    elif "monotonic" in variant:
        r, rzi = api.advance_scalars_mono(mesh, I["dt"], I["coef_3rd_order"], I["scalars_old"], I["scalars_new"],
                                          I["rho_zz_old"], I["uhAvg"], I["wwAvg"])
        pairs = [("scalars_new", r, "scalars_new"), ("rho_zz_int", rzi, "rho_zz_int")]
    elif "vector_reconstruction" in variant:
        r = api.reconstruct(mesh, I["u"])
        pairs = [(k, r[k], k) for k in r if k in out_keys]
    else:
        raise KeyError(variant)
    return [(n, _np(g), e) for n, g, e in pairs]


# outputs computed with libm transcendental functions (pow / sin / cos):
# equal to the CPU result within a few ulp, not bit for bit
LIBM_OUTPUTS = {"exner", "pressure_p", "uReconstructZonal", "uReconstructMeridional", "ur_cell", "vr_cell"}
