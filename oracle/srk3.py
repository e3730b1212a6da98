"""Oracle orchestration of one split-explicit large step (test infrastructure only).

Restates ``atm_srk3`` / ``_atm_srk3_impl``
(``mpas_pytorch/dynamics/time_integration.py:3420-4894``) over the C oracle
kernels of ``mpas_oracle.c``:

* ``atm_rk_integration_setup`` (:2497-2785): time-level-1 copies
  (``theta_m_init``, ``ru_save``, ``rw_save``, ``rtheta_p_save``,
  ``rho_p_save``, ``rho_zz_old_split``), ``theta_m_save``/``rr_save``;
* per dynamics substep and RK stage (``execute_rk_step`` :5000-9154):
  moist coefficients (stage 1 of substep 1), vertical implicit coefficients
  (stages 1 and 2), dynamics tendency, small-step perturbation, acoustic
  substeps with divergence damping, large-step recovery, diagnostic solve;
* split-flux averaging and substep state reset (:3820-3942);
* split scalar transport (:3976-4675): explicit stages 1-2, monotonic
  stage 3, stage dt = [dt/3, dt/2, dt] (RK2: [dt/2, dt/2, dt]);
* cell-vector reconstruction (:4693-4754).

Stage weights follow ``compute_rk_weights`` (:2787-2856) for
``config_time_integration_order`` 3 or 2 (the reference's default).
"""

from __future__ import annotations

import numpy as np

from . import oracle as O  # noqa: F401  (package-relative import when used as oracle.srk3)


def _weights(dt_dyn: float, ns: int, order: int = 3):
    """``compute_rk_weights`` (time_integration.py:2787-2856): stage lengths,
    acoustic substep lengths and substep counts of the three stages."""
    half = max(1, ns // 2)
    if order == 2:
        return [dt_dyn / 2.0, dt_dyn / 2.0, dt_dyn], [dt_dyn / float(ns)] * 3, [half, half, ns]
    if order != 3:
        raise ValueError(f"time integration order {order} (2 or 3)")
    return [dt_dyn / 3.0, dt_dyn / 2.0, dt_dyn], [dt_dyn / 3.0, dt_dyn / float(ns), dt_dyn / float(ns)], [1, half, ns]


def _explicit_weights(rk_step: int, order: int = 3) -> tuple[float, float]:
    """``_explicit_scalar_transport_density_time_weights`` (advance_scalars.py:271-286)
    with advance_density = 1 (split transport)."""
    if rk_step == 1:
        wn = 1.0 / 2.0 if order == 2 else 1.0 / 3.0
    elif rk_step == 2:
        wn = 1.0 / 2.0
    else:
        wn = 1.0
    return 1.0 - wn, wn


DIAG_KEYS = ("h_edge", "v", "divergence", "vorticity", "ke", "pv_edge")


def srk3(state: dict, mesh: "O.Mesh", dt: float, *, len_disp: float, ns: int = 2, split: int = 3,
         epssm: float = 0.1, smdiv: float = 0.1, apvm: float = 0.5, coef3: float = 0.25,
         monotonic: bool = True, diagnostics_current: bool = False, order: int = 3,
         rayleigh_levels: int = 0, rayleigh_timescale_days: float = 5.0, mixing: dict | None = None,
         curvature: bool = False, advect_scalars: bool = True) -> dict:
    """One large step on NumPy arrays in the reference layout; returns a new dict.
    ``rayleigh_levels`` > 0 damps u on that many top levels (config_rayleigh_damp_u);
    ``mixing`` passes the horizontal-mixing and CAM keywords of ``oracle.dyn_params``;
    ``curvature`` adds the curvature terms (config_curvature_u), which read the input
    state's ``ur_cell`` / ``vr_cell`` all step long.  ``advect_scalars=False``
    (config_scalar_advection) leaves the scalars as they are."""
    st = {k: (np.array(v, dtype=np.float64, copy=True) if isinstance(v, np.ndarray) and v.dtype.kind == "f" else v)
          for k, v in state.items()}
    nc, ne, L = mesh.nc, mesh.ne, mesh.L
    S = st["scalars"].shape[2]
    st["theta_m"][nc] = 0.0
    zeros_c = lambda: np.zeros((nc + 1, L))  # noqa: E731
    rt_diab = st.get("rt_diabatic_tend", zeros_c())
    st["rt_diabatic_tend"] = rt_diab

    # atm_rk_integration_setup
    theta_m_init = st["theta_m"].copy()
    rho_zz_old_split = st["rho_zz"].copy()
    scalars_old = st["scalars"].copy()
    scalars_new = st["scalars"].copy()
    ru_save, rw_save = st["ru"].copy(), st["rw"].copy()
    rtheta_p_save, rho_p_save = st["rtheta_p"].copy(), st["rho_p"].copy()
    theta_m_save, rr_save = st["theta_m"].copy(), st["rho_p"].copy()
    ruAvg_split = np.zeros((ne + 1, L))
    wwAvg_split = np.zeros((nc + 1, L + 1))
    acoustic = {"ru_p": np.zeros((ne + 1, L)), "ruAvg": np.zeros((ne + 1, L)),
                "rho_pp": zeros_c(), "rtheta_pp": zeros_c(), "rtheta_pp_old": zeros_c(),
                "rw_p": np.zeros((nc + 1, L + 1)), "wwAvg": np.zeros((nc + 1, L + 1))}

    dt_dyn = dt / split
    rk_timestep, rk_sub, nsub = _weights(dt_dyn, ns, order)
    moist = None
    vcoef = None
    carried = None
    for sub in range(1, split + 1):
        for rk in (1, 2, 3):
            if rk == 1 and sub == 1:
                moist = O.moist_coefficients(mesh, st["scalars"], int(st.get("moist_start", 1)),
                                             int(st.get("moist_end", S)))
                if not diagnostics_current:
                    d = O.solve_diagnostics(mesh, dt, apvm, None, st["u"], st["rho_zz"])
                    st.update({k: d[k] for k in DIAG_KEYS})
                    st["rho_edge"] = d["h_edge"]
            # RK3 recomputes the coefficients for stage 2; RK2 keeps stage 1's
            # (time_integration.py:5558-5569)
            if rk == 1 or (rk == 2 and order == 3):
                vcoef = O.vert_imp_coefs(mesh, rk_sub[rk - 1], epssm, moist["cqw"], st["exner"],
                                         st["theta_m"], st["rtheta_p"], moist["qtot"])
            fields = {
                "u": st["u"], "v": st["v"], "w": st["w"], "ru": st["ru"], "rw": st["rw"],
                "ru_save": ru_save, "rw_save": rw_save, "theta_m": st["theta_m"],
                "theta_m_save": theta_m_save, "rho_zz": st["rho_zz"], "rho_edge": st["rho_edge"],
                "divergence": st["divergence"], "vorticity": st["vorticity"], "ke": st["ke"],
                "pv_edge": st["pv_edge"], "pp": st["pressure_p"], "rr_save": rr_save,
                "cqw": moist["cqw"], "cqu": moist["cqu"], "rt_diabatic_tend": rt_diab,
                "qtot": moist["qtot"], "ur_cell": st.get("ur_cell", zeros_c()),
                "vr_cell": st.get("vr_cell", zeros_c()),
            }
            params = O.dyn_params(rk, dt, len_disp, coef_3rd_order=coef3, rayleigh_levels=rayleigh_levels,
                                  rayleigh_timescale_days=rayleigh_timescale_days, curvature=curvature,
                                  **(mixing or {}))
            tend = O.dyn_tend(mesh, params, fields, carried if rk > 1 else None)
            carried = {k: tend[k] for k in ("tend_rho", "tend_u_euler", "tend_w_euler", "tend_theta_euler", "kdiff")}
            tend_rw = O.smlstep_pert(mesh, tend["tend_u"], tend["tend_w"])
            afields = dict(vcoef)
            afields.update({"rho_zz": st["rho_zz"], "theta_m": theta_m_init, "exner": st["exner"],
                            "cqu": moist["cqu"], "tend_ru": tend["tend_u"], "tend_rho": tend["tend_rho"],
                            "tend_rt": tend["tend_theta"], "tend_rw": tend_rw, "w": st["w"],
                            "rw": st["rw"], "rw_save": rw_save})
            for small in range(1, nsub[rk - 1] + 1):
                afields.update(acoustic)
                ap = O.acoustic_params(small, rk_sub[rk - 1], epssm)
                acoustic = O.acoustic_step(mesh, ap, afields)
                acoustic["ru_p"] = O.divergence_damping(mesh, rk_sub[rk - 1], smdiv, len_disp, theta_m_init,
                                                        acoustic["rtheta_pp"], acoustic["rtheta_pp_old"],
                                                        acoustic["ru_p"])
            rfields = {"rho_p_save": rho_p_save, "rho_pp": acoustic["rho_pp"], "rtheta_p_save": rtheta_p_save,
                       "rtheta_pp": acoustic["rtheta_pp"], "rt_diabatic_tend": rt_diab, "rw_save": rw_save,
                       "rw_p": acoustic["rw_p"], "ru_save": ru_save, "ru_p": acoustic["ru_p"],
                       "wwAvg": acoustic["wwAvg"], "ruAvg": acoustic["ruAvg"]}
            rec = O.recover_large_step(mesh, rk, rk_timestep[rk - 1], nsub[rk - 1], rfields,
                                       exner_in=st["exner"], pressure_p_in=st["pressure_p"])
            acoustic["wwAvg"], acoustic["ruAvg"] = rec["wwAvg"], rec["ruAvg"]
            for k in ("rho_p", "rho_zz", "rtheta_p", "theta_m", "exner", "pressure_p", "rw", "w", "ru", "u"):
                st[k] = rec[k]
            d = O.solve_diagnostics(mesh, dt, apvm, rk, st["u"], st["rho_zz"], st["v"])
            st.update({k: d[k] for k in DIAG_KEYS})
            st["rho_edge"] = d["h_edge"]
        # split-flux accumulation and substep reset (time_integration.py:3820-3942)
        if sub == 1:
            ruAvg_split = acoustic["ruAvg"].copy()
            wwAvg_split = acoustic["wwAvg"].copy()
        else:
            ruAvg_split = ruAvg_split + acoustic["ruAvg"]
            wwAvg_split = wwAvg_split + acoustic["wwAvg"]
        if sub < split:
            st["theta_m"][nc] = 0.0
            ru_save, rw_save = st["ru"].copy(), st["rw"].copy()
            rtheta_p_save, rho_p_save = st["rtheta_p"].copy(), st["rho_p"].copy()
            theta_m_init = st["theta_m"].copy()
            rr_save = rho_p_save
            theta_m_save = st["theta_m"].copy()
    inv_split = 1.0 / float(split)
    ruAvg = ruAvg_split * inv_split
    wwAvg = wwAvg_split * inv_split
    rho_zz_new = st["rho_zz"].copy()
    # split transport (time_integration.py:3976-4675), skipped without scalar advection (:3975)
    for rk_t in ((1, 2, 3) if advect_scalars else ()):
        # stage lengths (time_integration.py:3367-3371)
        dt_t = [dt / 2.0 if order == 2 else dt / 3.0, dt / 2.0, dt][rk_t - 1]
        if rk_t == 3 and monotonic:
            scalars_new, _ = O.advance_scalars_mono(mesh, dt_t, coef3, scalars_old, scalars_new,
                                                    rho_zz_old_split, ruAvg, wwAvg)
        else:
            wo, wn = _explicit_weights(rk_t, order)
            scalars_new, _ = O.advance_scalars(mesh, dt_t, coef3, wo, wn, scalars_old, scalars_new,
                                               rho_zz_old_split, rho_zz_new, ruAvg, wwAvg)
    st["scalars"] = scalars_new
    st["rho_zz"] = rho_zz_new
    st["ruAvg"], st["wwAvg"] = ruAvg, wwAvg
    st.update(O.reconstruct(mesh, st["u"]))
    return st
