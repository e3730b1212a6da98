"""GPU cold start and the driver API (``run_timestep`` / ``run_simulation``).

* ``atm_init_coupled_diagnostics`` on the GPU against the host restatement
  (itself within 1 ulp of the reference): bit for bit except exner (pow,
  1 ulp), with pressure recomputed exactly from the GPU's exner;
* the full cold start + one step through ``run_timestep`` against the
  reference's outputs for the golden case (1e-11);
* ``run_simulation`` over 3 steps equals 3 successive ``run_timestep`` calls
  bit for bit, and ``on_step`` sees every step.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import rel_err, srk3_case

pytestmark = pytest.mark.gpu


def _raw_case():
    from mpas_dycore_pytorch_amd import initial_state as IS
    from mpas_dycore_pytorch_amd import mesh as M

    m = M.build_mesh(3, 30, stretch="quadratic")
    raw = IS.raw_state(m, seed=5, n_tracers=1)
    return m, raw


def test_init_coupled_on_gpu(gpu_available):
    from mpas_dycore_pytorch_amd import initial_state as IS
    from mpas_dycore_pytorch_amd import kernels as K
    from mpas_dycore_pytorch_amd.device import DeviceMesh

    print("init_coupled level 3", flush=True)
    m, raw = _raw_case()
    host = IS.init_coupled_diagnostics(m, raw)
    dm = DeviceMesh({**m, **raw}, torch.device("cuda", 0))
    got = K.init_coupled_diagnostics(dm, raw["theta"], raw["rho"], raw["qv"], raw["u"], raw["w"], raw["theta_base"])
    got = {k: v.cpu().numpy() for k, v in got.items()}
    nc = m["nCells"]
    for k in ("theta_m", "rho_zz", "ru", "rw", "rho_p", "rtheta_base", "rtheta_p"):
        assert np.array_equal(got[k], host[k]), k
    for k in ("exner", "exner_base"):
        ulps = np.abs(got[k].view(np.int64) - host[k].view(np.int64))
        assert ulps.max() <= 1, k
    zz = m["zz"][:nc]
    ex, exb, rtp, rtb = got["exner"][:nc], got["exner_base"][:nc], got["rtheta_p"][:nc], got["rtheta_base"][:nc]
    assert np.array_equal(got["pressure_p"][:nc], zz * 287.0 * (ex * rtp + rtb * (ex - exb)))
    assert np.array_equal(got["pressure_base"][:nc], zz * 287.0 * exb * rtb)
    assert got["rho_zz"][nc].tolist() == [1.0] * m["nVertLevels"]


def test_init_coupled_on_gpu_over_terrain(gpu_available):
    """Cold start over synthetic terrain (the zb/zb3 terms of rw) against the
    reference's own prepared fields: bit for bit except exner (pow, 1 ulp)."""
    from mpas_dycore_pytorch_amd import kernels as K
    from mpas_dycore_pytorch_amd.device import DeviceMesh

    print("init_coupled terrain l2_L20_terrain", flush=True)
    st, _, _ = srk3_case("l2_L20_terrain")
    dm = DeviceMesh(st, torch.device("cuda", 0))
    assert dm.s.flatTerrain == 0
    got = K.init_coupled_diagnostics(dm, st["theta"], st["rho"], st["qv"], st["u"], st["w"], st["theta_base"])
    got = {k: v.cpu().numpy() for k, v in got.items()}
    nc, ne = int(st["nCells"]), int(st["nEdges"])
    rows = {"ru": ne}
    for k in ("theta_m", "rho_zz", "ru", "rw", "rho_p", "rtheta_p"):
        n = rows.get(k, nc)
        ref = np.asarray(st[k])[:n]
        assert np.array_equal(got[k][:n], ref), (k, rel_err(got[k][:n], ref))
    ulps = np.abs(got["exner"][:nc].view(np.int64) - np.asarray(st["exner"])[:nc].view(np.int64))
    assert ulps.max() <= 1


def test_run_timestep_from_raw_matches_reference(gpu_available):
    from mpas_dycore_pytorch_amd import run_timestep

    print("run_timestep golden l2_L20_tracer", flush=True)
    st, refs, meta = srk3_case("l2_L20_tracer")
    cfg = {"config_time_integration_order": 3, "config_len_disp": float(st["nominalMinDc"])}
    out = run_timestep(st, st, cfg, meta["dt"], device="cpu")
    for k in ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "pressure_p", "v", "ke", "pv_edge"):
        err = rel_err(out[k].numpy(), refs[0][k])
        assert err <= 1e-11, f"{k}: {err:.3e}"


def test_run_simulation_equals_repeated_run_timestep(gpu_available):
    from mpas_dycore_pytorch_amd import run_simulation, run_timestep

    print("run_simulation level 3", flush=True)
    m, raw = _raw_case()
    st = {**raw}
    cfg = {"config_time_integration_order": 3, "config_len_disp": float(m["nominalMinDc"])}
    seen = []
    sim = run_simulation(st, m, cfg, 600.0, nsteps=3, device="cuda", on_step=lambda i, s: seen.append(i))
    assert seen == [1, 2, 3]
    cur = st
    for _ in range(3):
        cur = run_timestep(cur, m, cfg, 600.0, device="cuda")
    for k in ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "ke"):
        assert torch.equal(sim[k], cur[k]), k
