"""Non-default dynamics-tendency options on the GPU.

Rayleigh damping of u (``config_rayleigh_damp_u``: the top 5 levels, timescale
0.05 days; reference ``dyn_tend.py:2942-2996``, coefficient ``:10066-10072``):

* two RK3 steps against the reference's own recording of the same case
  (``tests/golden/srk3_l2_L20_rayleigh.npz``), per field at the whole-step bar,
  and the undamped step missing it;
* the step's derivatives: the taped step equals the fused one bit for bit, the
  JVP/VJP identity, and the JVP against central differences of the oracle with
  the damping on.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch
from conftest import perturbed_state, rel_err, srk3_case

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-11
FIELDS = ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "pressure_p", "rtheta_p", "rho_p", "ru", "rw",
          "v", "divergence", "vorticity", "ke", "pv_edge", "uReconstructZonal", "uReconstructMeridional")


def _config(st, meta, damp=True):
    cfg = {"config_time_integration_order": 3, "config_len_disp": float(st["nominalMinDc"])}
    if damp:
        cfg.update(meta["config_extra"])
    return cfg


def test_rayleigh_damping_whole_steps_match_reference():
    from mpas_dycore_pytorch_amd.time_integration import Srk3Stepper

    st, refs, meta = srk3_case("l2_L20_rayleigh")
    print(f"[options] Rayleigh damping of u: {meta['config_extra']}", flush=True)
    stepper = Srk3Stepper(st, st, _config(st, meta), meta["dt"], device=torch.device("cuda"))
    assert stepper.cfg.rayleigh_levels == 5
    worst = 0.0
    for step, ref in enumerate(refs):
        stepper.step(1)
        got = {k: v.cpu().numpy() for k, v in stepper.state_dict().items()}
        for k in FIELDS:
            if k in ref:
                err = rel_err(got[k], ref[k])
                worst = max(worst, err)
                assert err <= STEP_TOL, f"rayleigh step {step + 1} {k}: rel err {err:.3e}"
    plain = Srk3Stepper(st, st, _config(st, meta, damp=False), meta["dt"], device=torch.device("cuda"))
    plain.step(1)
    miss = rel_err(plain.state_dict()["u"].cpu().numpy(), refs[0]["u"])
    print(f"[options] Rayleigh: worst field error {worst:.2e}; undamped u misses the reference by {miss:.2e}",
          flush=True)
    assert miss > 1e-6


def test_rayleigh_damping_step_derivatives():
    from oracle import oracle as O, srk3 as OS
    from mpas_dycore_pytorch_amd.adjoint import STEP_INPUTS, Tape, step_jvp, step_vjp, taped_step
    from mpas_dycore_pytorch_amd.time_integration import Srk3Stepper, resolve_config

    st0, _, meta = srk3_case("l2_L20_rayleigh")
    st = perturbed_state(st0, 31)
    dev = torch.device("cuda")
    config = _config(st, meta)
    cfg = resolve_config(config, st, st, meta["dt"], int(np.asarray(st["scalars"]).shape[2]))
    stepper = Srk3Stepper(st, st, config, meta["dt"], device=dev)
    ckpt = {k: stepper.t[k].clone() for k in STEP_INPUTS}
    dm = stepper.mesh
    stepper.step(1)
    fused = stepper.state_dict()
    taped = taped_step(dm, ckpt, cfg, meta["dt"], Tape(dm))
    for k in ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "pressure_p", "pv_edge"):
        assert torch.equal(taped[k], fused[k]), k
    gen = torch.Generator().manual_seed(5)
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
    outk = ("u", "w", "theta_m")
    ybar = {k: rand(fused[k]) / float(fused[k].std()) ** 2 for k in outk}
    _, jt = step_jvp(dm, ckpt, cfg, meta["dt"], t)
    _, g = step_vjp(dm, ckpt, cfg, meta["dt"], ybar)
    jv = sum(float(torch.sum(ybar[k] * jt[k])) for k in outk)
    vj = sum(float(torch.sum(g[k] * t[k])) for k in keys)
    rel = abs(jv - vj) / max(abs(jv), abs(vj))
    om = O.Mesh(st)
    tn = {k: t[k].cpu().numpy() for k in keys}
    yb = {k: ybar[k].cpu().numpy() for k in outk}
    extra = meta["config_extra"]
    kw = {"rayleigh_levels": extra["config_number_rayleigh_damp_u_levels"],
          "rayleigh_timescale_days": extra["config_rayleigh_damp_u_timescale_days"]}

    def f(sign, h):
        s2 = dict(st)
        s2.update({k: np.asarray(st[k], dtype=np.float64) + sign * h * tn[k] for k in keys})
        y = OS.srk3(s2, om, meta["dt"], len_disp=float(st["nominalMinDc"]), **kw)
        return sum(float(np.sum(yb[k] * np.asarray(y[k]))) for k in outk)

    fds = [(f(1, h) - f(-1, h)) / (2 * h) for h in (1e-6, 1e-7)]
    rf = min(abs(fd - jv) / max(abs(fd), abs(jv)) for fd in fds)
    print(f"[options] Rayleigh step: jvp {jv:.15e} vjp {vj:.15e} rel {rel:.1e}; fd {fds} rel {rf:.1e}", flush=True)
    assert rel <= 1e-11
    assert rf < 1e-5
