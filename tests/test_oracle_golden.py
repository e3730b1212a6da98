"""The CPU oracle against the reference's own outputs (golden vectors).

Every fixture under ``tests/golden`` was written by running the reference
(``tests/golden/make_golden.py``); these tests pin the oracle to it:

* each of the 20 kernel-call variants that one reference large step makes
  (level-1 mesh, 10 levels, one tracer) must be reproduced bit for bit;
* whole large steps (level 2, dry and with a tracer, 2-3 steps) within
  1e-14 relative per field;
* a level-4, 26-level case (the 2-step CPU configuration) through field
  norms and seeded samples, from a state prepared by this package's own
  cold-start chain.
"""

from __future__ import annotations

import json

import numpy as np
import pytest

from conftest import golden_path, kernel_calls, rel_err, srk3_case
from kernel_cases import run_call

from oracle import oracle as O
from oracle import srk3 as OS

CALLS_MERGED, CALLS = kernel_calls()


@pytest.fixture(scope="module")
def kmesh():
    return O.Mesh(CALLS_MERGED)


@pytest.mark.parametrize("idx", range(len(CALLS)), ids=[c[0] for c in CALLS])
def test_kernel_call_bit_exact(kmesh, idx):
    variant, I, expected = CALLS[idx]
    for name, got, key in run_call(O, kmesh, variant, I, expected.keys()):
        ref = np.asarray(expected[key])
        assert np.array_equal(got[: ref.shape[0]], ref), f"{variant}:{name} rel err {rel_err(got, ref):.3e}"


def test_every_recorded_variant_is_covered():
    kinds = {v.split(":")[0] for v, _, _ in CALLS}
    assert len(CALLS) == 20
    assert any("dyn_tend" in k for k in kinds) and any("monotonic" in k for k in kinds)


FIELDS = ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "pressure_p", "rtheta_p", "rho_p", "ru", "rw",
          "v", "divergence", "vorticity", "ke", "pv_edge", "uReconstructZonal", "uReconstructMeridional")


@pytest.mark.parametrize("case", ["l2_L26_dry", "l2_L20_tracer", "l2_L20_terrain"])
def test_srk3_whole_steps(case):
    st, outs, meta = srk3_case(case)
    mesh = O.Mesh(st)
    cur = st
    for step, ref in enumerate(outs):
        cur = OS.srk3(cur, mesh, meta["dt"], len_disp=float(st["nominalMinDc"]),
                      diagnostics_current=(step > 0))
        for k in FIELDS:
            if k in ref:
                err = rel_err(cur[k], ref[k])
                assert err <= 1e-14, f"{case} step {step + 1} {k}: rel err {err:.3e}"


def test_srk3_whole_steps_rk2_reference_defaults():
    """The reference's default configuration (RK2, config_len_disp 0 resolved from nominalMinDc):
    the oracle's RK2 orchestration matches the reference's two steps."""
    st, outs, meta = srk3_case("l2_L20_defaults")
    mesh = O.Mesh(st)
    cur = st
    for step, ref in enumerate(outs):
        cur = OS.srk3(cur, mesh, meta["dt"], len_disp=float(st["nominalMinDc"]), order=2,
                      diagnostics_current=(step > 0))
        for k in FIELDS:
            if k in ref:
                err = rel_err(cur[k], ref[k])
                assert err <= 1e-14, f"rk2 step {step + 1} {k}: rel err {err:.3e}"
    # a plausible slip (RK3 weights) is caught by the same bar
    wrong = OS.srk3(st, mesh, meta["dt"], len_disp=float(st["nominalMinDc"]), order=3)
    assert rel_err(wrong["theta_m"], outs[0]["theta_m"]) > 1e-12


def test_srk3_whole_steps_rayleigh_damping_of_u():
    """Rayleigh damping of u on the top 5 levels (config_rayleigh_damp_u, timescale 0.05 days)
    over two RK3 steps: the oracle against the reference's recording; without the damping
    the oracle misses it, so the option is exercised."""
    st, outs, meta = srk3_case("l2_L20_rayleigh")
    extra = meta["config_extra"]
    assert extra["config_rayleigh_damp_u"] and extra["config_number_rayleigh_damp_u_levels"] == 5
    mesh = O.Mesh(st)
    kw = {"rayleigh_levels": extra["config_number_rayleigh_damp_u_levels"],
          "rayleigh_timescale_days": extra["config_rayleigh_damp_u_timescale_days"]}
    cur = st
    for step, ref in enumerate(outs):
        cur = OS.srk3(cur, mesh, meta["dt"], len_disp=float(st["nominalMinDc"]), diagnostics_current=(step > 0),
                      **kw)
        for k in FIELDS:
            if k in ref:
                err = rel_err(cur[k], ref[k])
                assert err <= 1e-14, f"rayleigh step {step + 1} {k}: rel err {err:.3e}"
    undamped = OS.srk3(st, mesh, meta["dt"], len_disp=float(st["nominalMinDc"]))
    assert rel_err(undamped["u"], outs[0]["u"]) > 1e-6


def _mixing_kwargs(extra: dict) -> dict:
    """oracle.dyn_params keywords of a recorded configuration's mixing keys."""
    kw = {}
    if "config_horiz_mixing" in extra:
        kw["horiz_mixing"] = extra["config_horiz_mixing"]
    for key, name in (("config_h_theta_eddy_visc2", "h_theta_eddy_visc2"),
                      ("config_h_mom_eddy_visc4", "h_mom_eddy_visc4"),
                      ("config_h_theta_eddy_visc4", "h_theta_eddy_visc4"),
                      ("config_mpas_cam_coef", "cam_coef"), ("config_number_cam_damping_levels", "cam_levels"),
                      ("config_v_mom_eddy_visc2", "v_mom_eddy_visc2"),
                      ("config_v_theta_eddy_visc2", "v_theta_eddy_visc2"), ("config_mix_full", "mix_full")):
        if key in extra:
            kw[name] = extra[key]
    return kw


@pytest.mark.parametrize("case", ["l2_L20_fixedmix", "l2_L20_cam", "l2_L20_curvature", "l2_L20_vmix",
                                  "l2_L20_vmix_dep"])
def test_srk3_whole_steps_mixing_options(case):
    """"2d_fixed" horizontal mixing (constant kdiff, separate del4 coefficients for
    momentum and theta), the CAM damping of kdiff on the top 6 levels, the
    curvature terms of u and w, and the vertical eddy viscosities on the full
    fields and on their departures from the reference profiles, over two RK3
    steps: the oracle against the reference's recordings at 1e-14.  Without the option the oracle misses the
    recording, so each option is exercised."""
    st, outs, meta = srk3_case(case)
    extra = meta["config_extra"]
    kw = _mixing_kwargs(extra)
    curvature = bool(extra.get("config_curvature_u", False))
    assert kw or curvature, extra
    mesh = O.Mesh(st)
    cur = st
    for step, ref in enumerate(outs):
        cur = OS.srk3(cur, mesh, meta["dt"], len_disp=float(st["nominalMinDc"]), diagnostics_current=(step > 0),
                      mixing=kw, curvature=curvature)
        for k in FIELDS:
            if k in ref:
                err = rel_err(cur[k], ref[k])
                assert err <= 1e-14, f"{case} step {step + 1} {k}: rel err {err:.3e}"
    plain = OS.srk3(st, mesh, meta["dt"], len_disp=float(st["nominalMinDc"]))
    assert rel_err(plain["u"], outs[0]["u"]) > 1e-9
    assert rel_err(plain["theta_m"], outs[0]["theta_m"]) > 1e-12


@pytest.mark.parametrize("case,kw", [("l2_L20_nonmono", {"monotonic": False}),
                                     ("l2_L20_posdef", {"monotonic": True}),
                                     ("l2_L20_noadv", {"advect_scalars": False})])
def test_srk3_whole_steps_transport_switches(case, kw):
    """The transport switches over two RK3 steps against the reference's recordings:
    config_monotonic=False (an explicit final stage), config_positive_definite=True
    with config_monotonic=False (the reference runs the limited final stage, so the
    oracle's monotonic step must match it) and config_scalar_advection=False (the
    scalars stay as they are)."""
    st, outs, meta = srk3_case(case)
    mesh = O.Mesh(st)
    cur = st
    for step, ref in enumerate(outs):
        cur = OS.srk3(cur, mesh, meta["dt"], len_disp=float(st["nominalMinDc"]), diagnostics_current=(step > 0), **kw)
        for k in FIELDS:
            if k in ref:
                err = rel_err(cur[k], ref[k])
                assert err <= 1e-14, f"{case} step {step + 1} {k}: rel err {err:.3e}"
    if case == "l2_L20_noadv":
        assert np.array_equal(np.asarray(outs[-1]["scalars"]), np.asarray(st["scalars"]))
    else:
        other = OS.srk3(st, mesh, meta["dt"], len_disp=float(st["nominalMinDc"]), monotonic=not kw["monotonic"])
        assert rel_err(other["scalars"], outs[0]["scalars"]) > 1e-12


def test_one_dynamics_substep_transports_in_stage():
    """With config_dynamics_split_steps=1 the reference skips its split transport loop
    (time_integration.py:3953) and transports the scalars inside every RK stage.  The
    recording pins both halves of that: the dynamics of one substep match the
    oracle's at 1e-14, and its scalars differ from the split transport's, which is
    why the GPU step rejects the configuration instead of running it."""
    st, outs, meta = srk3_case("l2_L20_split1")
    mesh = O.Mesh(st)
    cur = OS.srk3(st, mesh, meta["dt"], len_disp=float(st["nominalMinDc"]), split=1)
    for k in FIELDS:
        if k in outs[0] and k != "scalars":
            assert rel_err(cur[k], outs[0][k]) <= 1e-14, k
    assert rel_err(cur["scalars"], outs[0]["scalars"]) > 1e-9


@pytest.mark.slow
def test_level4_summary_from_own_cold_start():
    from mpas_dycore_pytorch_amd import initial_state as IS
    from mpas_dycore_pytorch_amd import mesh as M

    with open(golden_path("summary_cfg1_l4_L26.json")) as f:
        summ = json.load(f)
    m = M.build_mesh(summ["level"], summ["nVertLevels"], stretch=summ["stretch"])
    raw = IS.raw_state(m, seed=summ["seed"], n_tracers=summ["n_tracers"])
    om = O.Mesh({**m, **IS.init_coupled_diagnostics(m, raw), **raw})
    dt = summ["dt"]
    st = IS.prepare_state(m, raw,
                          diagnostics=lambda u, h: O.solve_diagnostics(om, dt, 0.5, None, u, h),
                          reconstruct=lambda u: O.reconstruct(om, u))
    om = O.Mesh(st)
    cur = st
    for step, rec in enumerate(summ["steps_out"]):
        cur = OS.srk3(cur, om, dt, len_disp=float(m["nominalMinDc"]), diagnostics_current=(step > 0))
        for k, r in rec.items():
            a = np.asarray(cur[k]).reshape(cur[k].shape[0], -1)[:-1]
            scale = max(r["linf"], 1e-300)
            vals = a.reshape(-1)[np.asarray(r["idx"])]
            err = float(np.max(np.abs(vals - np.asarray(r["val"])))) / scale
            assert err <= 1e-11, f"step {step + 1} {k}: sample rel err {err:.3e}"
            assert abs(float(np.sqrt(np.sum(a * a))) - r["l2"]) <= 1e-11 * max(r["l2"], 1e-300), k
            assert abs(float(np.max(np.abs(a))) - r["linf"]) <= 1e-11 * scale, k
