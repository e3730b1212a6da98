"""The mixed-precision workload at full size: level 8 (655362 cells, 1966080 edges) x 55 levels,
dt 150 s, 48 large steps, 2 passive tracers.

The mixed-precision step (horizontal tendencies, diagnostics and transport in binary32 with
compensated flux sums; vertical-implicit coefficients, the tridiagonal solve, exner recovery,
the mass fluxes and the base state in binary64) against the FP64 GPU step (itself within 1e-11
of the reference float64 path) from the same state: per-field relative L2 and Linf error at
steps 12, 24 and 48, every field of the returned state at or below 1e-5.

The mixed-precision reverse mode (checkpoints from the mixed forward, the adjoint sweep in
binary64 around that trajectory) is compared with the FP64 adjoint on a level-6 x 55 run.
"""

from __future__ import annotations

import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

FIELDS = ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "pressure_p", "rtheta_p", "rho_p", "ru", "rw",
          "v", "divergence", "vorticity", "ke", "pv_edge", "rho_edge", "uReconstructZonal",
          "uReconstructMeridional", "ruAvg", "wwAvg")
BAR = 1e-5


def _prepared(level, n_levels, n_tracers, dt, seed):
    from mpas_dycore_pytorch_amd import initial_state as IS
    from mpas_dycore_pytorch_amd import mesh as M
    from mpas_dycore_pytorch_amd import run_simulation

    m = M.build_mesh(level, n_levels, stretch="quadratic")
    raw = IS.raw_state(m, seed=seed, n_tracers=n_tracers)
    st = run_simulation(raw, m, {"config_time_integration_order": 3}, dt, nsteps=0, device="cuda")
    return m, st


def _errors(a: torch.Tensor, b: torch.Tensor) -> tuple[float, float]:
    a = a[:-1].double()
    b = b[:-1].double()
    d = a - b
    l2 = float(torch.linalg.vector_norm(d) / max(float(torch.linalg.vector_norm(a)), 1e-300))
    linf = float(d.abs().max() / max(float(a.abs().max()), 1e-300))
    return l2, linf


def test_mixed_precision_level8_48_steps_every_field():
    from mpas_dycore_pytorch_amd.time_integration import Srk3Stepper

    dt = 150.0
    t0 = time.time()
    print("[config5] building level 8 x 55 with 2 tracers", flush=True)
    m, st = _prepared(8, 55, 2, dt, 8055)
    cfg = {"config_time_integration_order": 3}
    ref = Srk3Stepper(st, m, cfg, dt, reorder=True)
    mix = Srk3Stepper(st, m, {**cfg, "config_precision": "mixed"}, dt, reorder=True)
    del st
    assert mix.mixed and mix.t["pv_edge"].dtype == torch.float32 and mix.t["ru"].dtype == torch.float64
    print(f"[config5] prepared in {time.time() - t0:.1f} s", flush=True)
    done, worst = 0, {}
    for n in (12, 24, 48):
        ref.step(n - done)
        mix.step(n - done)
        done = n
        torch.cuda.synchronize()
        a, b = ref.state_dict(), mix.state_dict()
        for k in FIELDS:
            l2, linf = _errors(a[k], b[k])
            assert np.isfinite(l2) and np.isfinite(linf), k
            print(f"[config5] step {n:2d} {k:22s} rel L2 {l2:.2e}  rel Linf {linf:.2e}", flush=True)
            worst[k] = max(worst.get(k, 0.0), l2, linf)
        del a, b
    bad = {k: v for k, v in worst.items() if not v <= BAR}
    assert not bad, bad


def test_mixed_precision_vjp_against_fp64_adjoint():
    from mpas_dycore_pytorch_amd import run_simulation
    from mpas_dycore_pytorch_amd.boundary import mark_complete_step_boundary

    dt = 450.0
    m, st = _prepared(6, 55, 1, dt, 606)
    state = {k: (torch.as_tensor(v, device="cuda") if isinstance(v, np.ndarray) and v.dtype.kind == "f" else v)
             for k, v in st.items()}
    nc = int(m["nCells"])
    area = torch.as_tensor(1.0 / np.asarray(m["invAreaCell"], dtype=np.float64)[:nc], device="cuda")
    th0 = state["theta_m"][:nc].clone()

    def objective(out):
        d = out["theta_m"][:nc] - th0
        return torch.sum(area[:, None] * d * d)

    keys = ("u", "w", "theta_m", "rho_zz")
    grads = {}
    for precision in ("fp64", "mixed"):
        cfg = {"config_time_integration_order": 3, "config_precision": precision}
        s = dict(state)
        mark_complete_step_boundary(s, mesh=m, config=cfg, dt=dt)
        grads[precision] = run_simulation.vjp(s, m, objective, track_keys=keys, config=cfg, dt=dt, nsteps=6,
                                              device="cuda")
    for k in keys:
        l2, linf = _errors(grads["fp64"][k], grads["mixed"][k])
        print(f"[config5] mixed-precision VJP (6 steps, level 6 x 55) d/d{k}: rel L2 {l2:.2e} rel Linf {linf:.2e}",
              flush=True)
        assert l2 <= 1e-3, k
