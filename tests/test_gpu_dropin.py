"""The drop-in contract on the GPU, under the reference's own defaults.

* The reference's default configuration (``config_defaults.py``: RK2, the
  default order, and ``config_len_disp = 0`` resolved from ``nominalMinDc``)
  through ``atm_srk3``, the resident stepper and ``run_simulation``, against
  the reference's two steps (``srk3_l2_L20_defaults.npz``) and the oracle;
* the RK2 step's derivatives: the taped forward is the fused step bit for bit,
  JVP and VJP are transposes, and the JVP matches oracle central differences;
* a differentiable run from raw fields differentiates the GPU cold start
  (``atmosphere_model.py:1589-1597``): JVP/VJP identity and central
  differences of the GPU forward;
* step-boundary receipts and the nonfinite guards in the driver;
* ``h_edge`` is ``rho_edge`` for both derivative modes.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import perturbed_state, rel_err, srk3_case

pytestmark = pytest.mark.gpu

FIELDS = ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "pressure_p", "rtheta_p", "rho_p", "ru", "rw",
          "v", "divergence", "vorticity", "ke", "pv_edge", "uReconstructZonal", "uReconstructMeridional")


def _dev_state(st, dev):
    return {k: (torch.as_tensor(np.asarray(v, dtype=np.float64), device=dev)
                if isinstance(v, np.ndarray) and v.dtype.kind == "f" else v) for k, v in st.items()}


def test_atm_srk3_with_reference_default_config():
    from oracle import oracle as O, srk3 as OS
    from mpas_dycore_pytorch_amd.time_integration import atm_srk3

    print("[dropin] atm_srk3, reference defaults (RK2), l2_L20_defaults", flush=True)
    st, refs, meta = srk3_case("l2_L20_defaults")
    dev = torch.device("cuda")
    cur = _dev_state(st, dev)
    ora = st
    om = O.Mesh(st)
    for step, ref in enumerate(refs):
        cur = atm_srk3(cur, st, {}, meta["dt"])  # the reference's default configuration
        ora = OS.srk3(ora, om, meta["dt"], len_disp=float(st["nominalMinDc"]), order=2,
                      diagnostics_current=(step > 0))
        torch.cuda.synchronize()
        for k in FIELDS:
            if k in ref:
                got = cur[k].cpu().numpy()
                e_ref, e_ora = rel_err(got, ref[k]), rel_err(got, ora[k])
                assert e_ref <= 1e-11 and e_ora <= 1e-11, f"step {step + 1} {k}: {e_ref:.2e} {e_ora:.2e}"


def test_run_simulation_with_reference_default_config():
    from mpas_dycore_pytorch_amd import run_simulation, run_timestep
    from mpas_dycore_pytorch_amd.boundary import BOUNDARY_KEY

    print("[dropin] run_simulation / run_timestep, reference defaults", flush=True)
    st, refs, meta = srk3_case("l2_L20_defaults")
    sim = run_simulation(st, st, {}, meta["dt"], nsteps=2, device="cuda")
    assert isinstance(sim[BOUNDARY_KEY], dict)
    for k in FIELDS:
        if k in refs[1]:
            err = rel_err(sim[k].cpu().numpy(), refs[1][k])
            assert err <= 1e-11, f"{k}: {err:.2e}"
    # the receipt lets the second call continue without a new cold start: bit for bit
    one = run_timestep(st, st, {}, meta["dt"], device="cuda")
    two = run_timestep(one, st, {}, meta["dt"], device="cuda")
    for k in ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "pressure_p", "ke", "pv_edge"):
        assert torch.equal(two[k], sim[k]), k


def test_boundary_receipt_and_guards_in_the_driver():
    from mpas_dycore_pytorch_amd import run_timestep
    from mpas_dycore_pytorch_amd.errors import MpasConfigError

    print("[dropin] boundary receipt and nonfinite guards", flush=True)
    st, _, meta = srk3_case("l2_L20_defaults")
    one = run_timestep(st, st, {}, meta["dt"], device="cuda")
    with pytest.raises(MpasConfigError, match="identity"):
        run_timestep(one, st, {}, meta["dt"] / 2, device="cuda")
    with pytest.raises(MpasConfigError, match="identity"):
        run_timestep(one, st, {"config_smdiv": 0.05}, meta["dt"], device="cuda")
    edited = dict(one)
    edited["ke"] = one["ke"] * 1.5
    with pytest.raises(MpasConfigError, match="edited"):
        run_timestep(edited, st, {}, meta["dt"], device="cuda")
    # a prognostic edit is allowed: the diagnostics are derived again from it
    moved = dict(one)
    moved["theta_m"] = one["theta_m"].clone()
    moved["theta_m"][: int(st["nCells"])] += 0.01
    out = run_timestep(moved, st, {}, meta["dt"], device="cuda")
    assert bool(torch.isfinite(out["theta_m"]).all())
    bad = dict(one)
    bad["w"] = one["w"].clone()
    bad["w"][0, 3] = float("nan")
    with pytest.raises(MpasConfigError, match="nonfinite values in w"):
        run_timestep(bad, st, {}, meta["dt"], device="cuda")


def test_rk2_step_derivatives():
    from oracle import oracle as O, srk3 as OS
    from mpas_dycore_pytorch_amd.adjoint import STEP_INPUTS, Tape, step_jvp, step_vjp, taped_step
    from mpas_dycore_pytorch_amd.time_integration import Srk3Stepper, resolve_config

    print("[dropin] RK2 step: taped = fused, JVP/VJP identity, oracle FD", flush=True)
    st0, _, meta = srk3_case("l2_L20_defaults")
    st = perturbed_state(st0, 21)
    dev = torch.device("cuda")
    cfg = resolve_config({}, st, st, meta["dt"], int(np.asarray(st["scalars"]).shape[2]))
    assert cfg.order == 2
    stepper = Srk3Stepper(st, st, {}, meta["dt"], device=dev)
    ckpt = {k: stepper.t[k].clone() for k in STEP_INPUTS}
    dm = stepper.mesh
    stepper.step(1)
    fused = stepper.state_dict()
    taped = taped_step(dm, ckpt, cfg, meta["dt"], Tape(dm))
    for k in ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "pressure_p", "pv_edge"):
        assert torch.equal(taped[k], fused[k]), k
    gen = torch.Generator().manual_seed(3)
    keys = ("u", "w", "theta_m")
    pad = {dm.nc + 1, dm.ne + 1, dm.nv + 1}

    def rand(like):
        r = torch.randn(like.shape, generator=gen, dtype=torch.float64).to(dev) * float(like.std() or 1.0)
        if like.shape[0] in pad:
            r[-1] = 0.0
        return r

    t = {k: rand(ckpt[k]) for k in keys}
    t["w"][:, 0] = 0.0
    t["w"][:, -1] = 0.0
    outk = ("u", "w", "theta_m", "scalars")
    ybar = {k: rand(fused[k]) / float(fused[k].std()) ** 2 for k in outk}
    _, jt = step_jvp(dm, ckpt, cfg, meta["dt"], t)
    _, g = step_vjp(dm, ckpt, cfg, meta["dt"], ybar)
    jv = sum(float(torch.sum(ybar[k] * jt[k])) for k in outk)
    vj = sum(float(torch.sum(g[k] * t[k])) for k in keys)
    rel = abs(jv - vj) / max(abs(jv), abs(vj))
    om = O.Mesh(st)
    tn = {k: t[k].cpu().numpy() for k in keys}
    # central differences on the dynamics outputs only: the monotonic limiter of the transport
    # has switching points that a finite step can cross, the dynamics of one step do not read it
    fdk = ("u", "w", "theta_m")
    yb = {k: ybar[k].cpu().numpy() for k in fdk}
    jv_dyn = sum(float(torch.sum(ybar[k] * jt[k])) for k in fdk)

    def f(sign, h):
        s2 = dict(st)
        s2.update({k: np.asarray(st[k], dtype=np.float64) + sign * h * tn[k] for k in keys})
        y = OS.srk3(s2, om, meta["dt"], len_disp=float(st["nominalMinDc"]), order=2)
        return sum(float(np.sum(yb[k] * np.asarray(y[k]))) for k in fdk)

    fds = [(f(1, h) - f(-1, h)) / (2 * h) for h in (1e-6, 1e-7)]
    rf = min(abs(fd - jv_dyn) / max(abs(fd), abs(jv_dyn)) for fd in fds)
    print(f"[dropin] RK2 jvp {jv:.15e} vjp {vj:.15e} rel {rel:.1e}; dynamics jvp {jv_dyn:.15e} fd {fds} "
          f"rel {rf:.1e}", flush=True)
    assert rel <= 1e-11
    assert rf < 1e-5


def _raw_level2():
    from mpas_dycore_pytorch_amd import initial_state as IS
    from mpas_dycore_pytorch_amd import mesh as M

    m = M.build_mesh(2, 20, stretch="quadratic")
    raw = IS.raw_state(m, seed=77, n_tracers=1)
    # off the switching surfaces of the hard primal (w = 0 ties every vertical upwind choice)
    rng = np.random.default_rng(78)
    nc, ne = int(m["nCells"]), int(m["nEdges"])
    raw["u"][:ne] += 0.3 * rng.standard_normal(raw["u"][:ne].shape)
    raw["w"][:nc, 1:-1] += 1e-3 * rng.standard_normal(raw["w"][:nc, 1:-1].shape)
    raw["qv"][:nc] += 1e-4 * rng.random(raw["qv"][:nc].shape)
    raw["scalars"][:nc] += 1e-4 * rng.random(raw["scalars"][:nc].shape)
    return m, raw


def test_vjp_through_the_cold_start_from_raw_fields():
    """run_simulation.vjp / .jvp from a raw state: the gradient reaches theta, rho, qv, u and w
    through the GPU cold start and one large step."""
    from mpas_dycore_pytorch_amd import run_simulation

    print("[dropin] differentiable cold start from raw fields (level 2 x 20, 1 step)", flush=True)
    m, raw = _raw_level2()
    dev = torch.device("cuda")
    cfg = {"config_time_integration_order": 3}
    dt = 900.0
    state = _dev_state(raw, dev)
    nc = int(m["nCells"])
    area = torch.as_tensor(1.0 / np.asarray(m["invAreaCell"], dtype=np.float64)[:nc], device=dev)
    keys = ("theta", "rho", "qv", "u", "w")
    gen = torch.Generator().manual_seed(12)
    pad = {nc + 1, int(m["nEdges"]) + 1}
    seed_w = torch.randn((nc, int(m["nVertLevels"])), generator=gen, dtype=torch.float64).to(dev)
    th_ref = state["theta"][:nc].clone()  # a constant offset: keeps the loss small against its rounding

    def objective(out):
        return (torch.sum(area[:, None] * seed_w * (out["theta_m"][:nc] - th_ref))
                + 1e-2 * torch.sum(out["u"][:-1] ** 2))

    grads = run_simulation.vjp(state, m, objective, track_keys=keys, config=cfg, dt=dt, nsteps=1, device=dev)
    for k in keys:
        assert k in grads and bool(torch.isfinite(grads[k]).all()), k
    direction = {}
    for k in keys:
        v = state[k]
        d = torch.randn(v.shape, generator=gen, dtype=torch.float64).to(dev) * float(v[:-1].std() or 1.0) * 1e-3
        if v.shape[0] in pad:
            d[-1] = 0.0
        direction[k] = d
    direction["w"][:, 0] = 0.0
    direction["w"][:, -1] = 0.0
    tan = run_simulation.jvp(state, m, direction, config=cfg, dt=dt, nsteps=1, device=dev)
    with torch.no_grad():
        u_out = run_simulation(state, m, cfg, dt, nsteps=1, device="cuda")["u"]
    jv = float(torch.sum(area[:, None] * seed_w * tan["theta_m"][:nc])
               + 2e-2 * torch.sum(u_out[:-1] * tan["u"][:-1]))
    vj = sum(float(torch.sum(grads[k] * direction[k])) for k in keys)
    rel = abs(jv - vj) / max(abs(jv), abs(vj))

    def loss(sign, h):
        s2 = {k: (v + sign * h * direction[k] if k in direction else v) for k, v in state.items()}
        with torch.no_grad():
            out = run_simulation(s2, m, cfg, dt, nsteps=1, device="cuda")
            return float(objective(out))

    # steps small enough that no vertical upwind choice (the sign of w) flips between the two sides
    fds = [(loss(1, h) - loss(-1, h)) / (2 * h) for h in (1e-4, 1e-5)]
    rf = min(abs(fd - vj) / max(abs(fd), abs(vj)) for fd in fds)
    print(f"[dropin] cold start: jvp {jv:.15e} vjp {vj:.15e} rel {rel:.1e}; fd {fds} rel {rf:.1e}", flush=True)
    assert rel <= 1e-10
    assert rf < 1e-5


def test_h_edge_is_rho_edge_for_both_modes():
    from mpas_dycore_pytorch_amd import run_simulation
    from mpas_dycore_pytorch_amd.boundary import mark_complete_step_boundary

    print("[dropin] h_edge cotangent and tangent", flush=True)
    st0, _, meta = srk3_case("l2_L20_tracer")
    st = perturbed_state(st0, 4)
    dev = torch.device("cuda")
    cfg = {"config_time_integration_order": 3, "config_len_disp": float(st["nominalMinDc"])}
    state = _dev_state(st, dev)
    mark_complete_step_boundary(state, mesh=st, config=cfg, dt=meta["dt"])
    ne = int(st["nEdges"])
    wgt = torch.randn((ne, int(st["nVertLevels"])), generator=torch.Generator().manual_seed(1),
                      dtype=torch.float64).to(dev)
    kw = {"config": cfg, "dt": meta["dt"], "nsteps": 1, "device": dev}
    keys = ("u", "theta_m")
    ga = run_simulation.vjp(dict(state), st, lambda o: torch.sum(wgt * o["h_edge"][:ne]), track_keys=keys, **kw)
    gb = run_simulation.vjp(dict(state), st, lambda o: torch.sum(wgt * o["rho_edge"][:ne]), track_keys=keys, **kw)
    for k in keys:
        scale = float(gb[k].abs().max())
        assert scale > 0.0, k
        # equal up to the order of the adjoint kernels' remaining atomic additions
        assert float((ga[k] - gb[k]).abs().max()) <= 1e-12 * scale, k
    d = {"u": torch.randn(state["u"].shape, generator=torch.Generator().manual_seed(2),
                          dtype=torch.float64).to(dev)}
    d["u"][-1] = 0.0
    tan = run_simulation.jvp(dict(state), st, d, **kw)
    assert float(tan["rho_edge"].abs().max()) > 0.0
    assert torch.equal(tan["h_edge"], tan["rho_edge"])


@pytest.mark.parametrize("derived", [False, True])
def test_init_coupled_derivatives_against_central_differences(derived):
    """The cold-start kernels alone (atm_init_coupled_diagnostics, raw or derived inputs, over
    terrain): the JVP against central differences of the GPU forward, and the VJP as its transpose."""
    from mpas_dycore_pytorch_amd import kernels as K
    from mpas_dycore_pytorch_amd.adjoint import ICD_DIFF_OUT, Tape, taped_cold_start, vjp_scratch
    from mpas_dycore_pytorch_amd.device import DeviceMesh

    print(f"[dropin] init_coupled derivatives, derived={derived}, l2_L20_terrain", flush=True)
    st, _, meta = srk3_case("l2_L20_terrain")
    dev = torch.device("cuda")
    dm = DeviceMesh(st, dev)
    assert dm.s.flatTerrain == 0
    keys = ("theta_m", "rho_zz", "u", "w") if derived else ("theta", "rho", "qv", "u", "w")
    X = {k: torch.as_tensor(np.asarray(st[k], dtype=np.float64), device=dev) for k in keys}
    if "qv" in X:
        X["qv"] = X["qv"] + 1e-4
    tb = torch.as_tensor(np.asarray(st["theta_base"], dtype=np.float64), device=dev)
    gen = torch.Generator().manual_seed(5)
    pad = {dm.nc + 1, dm.ne + 1}
    d = {}
    for k, v in X.items():
        r = torch.randn(v.shape, generator=gen, dtype=torch.float64).to(dev) * v.abs().clamp_min(1e-3) * 1e-3
        if v.shape[0] in pad:
            r[-1] = 0.0
        d[k] = r
    d["w"][:, 0] = 0.0
    d["w"][:, -1] = 0.0
    tape = Tape(dm)
    out = taped_cold_start(dm, X, tb, meta["dt"], 0.5, tape)
    scratch = vjp_scratch(dm, 0)
    dots = tape.forward_tangents({id(X[k]): d[k] for k in keys}, scratch, {id(out[k]) for k in ICD_DIFF_OUT})

    def fwd(sign, h):
        Y = {k: X[k] + sign * h * d[k] for k in keys}
        if derived:
            return K.init_coupled_diagnostics(dm, None, None, None, Y["u"], Y["w"], tb, Y["theta_m"], Y["rho_zz"])
        return K.init_coupled_diagnostics(dm, Y["theta"], Y["rho"], Y["qv"], Y["u"], Y["w"], tb)

    h = 1e-3
    yp, ym = fwd(1, h), fwd(-1, h)
    for k in ICD_DIFF_OUT:
        fd = (yp[k] - ym[k]) / (2 * h)
        jt = dots[id(out[k])]
        err = float((fd - jt).norm() / max(float(jt.norm()), 1e-300))
        print(f"[dropin] init_coupled d{k}: rel l2 {err:.1e}", flush=True)
        # rw sums terrain terms that nearly cancel, so its differences carry more rounding
        assert err < (1e-6 if k == "rw" else 1e-8), k
    ybar = {k: torch.randn(out[k].shape, generator=gen, dtype=torch.float64).to(dev) for k in ICD_DIFF_OUT}
    for k in ybar:
        ybar[k][-1] = 0.0
    bars = tape.backward({id(out[k]): ybar[k] for k in ICD_DIFF_OUT}, scratch)
    jv = sum(float(torch.sum(ybar[k] * dots[id(out[k])])) for k in ICD_DIFF_OUT)
    vj = sum(float(torch.sum(bars[id(X[k])] * d[k])) for k in keys if id(X[k]) in bars)
    rel = abs(jv - vj) / max(abs(jv), abs(vj))
    print(f"[dropin] init_coupled identity: {jv:.15e} {vj:.15e} rel {rel:.1e}", flush=True)
    assert rel <= 1e-12
