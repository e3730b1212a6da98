"""Boundary formats with a device state: GPU history diagnostics and exact restart continuation."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import srk3_case

pytestmark = pytest.mark.gpu


def test_history_fields_on_gpu_match_host(tmp_path):
    from mpas_dycore_pytorch_amd import io as IO

    st, _, _ = srk3_case("l2_L20_tracer")
    st = dict(st)
    nc, L = int(st["nCells"]), int(st["nVertLevels"])
    st.setdefault("pressure_base", np.linspace(1e5, 1e3, (nc + 1) * L).reshape(nc + 1, L))
    host = IO.history_fields(st, st)
    dev = {k: (torch.as_tensor(v, device="cuda") if isinstance(v, np.ndarray) and v.dtype.kind == "f" else v)
           for k, v in st.items()}
    gpu = IO.history_fields(dev, st)
    torch.cuda.synchronize()
    assert gpu["theta"].is_cuda
    for k in host:
        assert torch.equal(gpu[k].cpu(), host[k]), k
    IO.write_history(tmp_path / "h.mpb", dev, st, xtime="2000-01-01_00:00:00")
    h = IO.read_history(tmp_path / "h.mpb")
    assert np.array_equal(h["fields"]["theta"][0], host["theta"].numpy())


def test_restart_continues_bit_for_bit(tmp_path):
    """3 straight steps equal 2 steps, write_restart, read_restart, 1 step."""
    from mpas_dycore_pytorch_amd import io as IO
    from mpas_dycore_pytorch_amd.time_integration import Srk3Stepper

    st, _, meta = srk3_case("l2_L20_tracer")
    cfg = {"config_time_integration_order": 3, "config_len_disp": float(st["nominalMinDc"])}
    dt = meta["dt"]
    a = Srk3Stepper(st, st, cfg, dt)
    a.step(3)
    b = Srk3Stepper(st, st, cfg, dt)
    b.step(2)
    torch.cuda.synchronize()
    p = tmp_path / "restart.mpb"
    IO.write_restart(b.state_dict(), st, p, "2000-01-01_00:30:00", config=cfg, dt=dt,
                     model_time_seconds=2 * dt)
    back, _ = IO.read_restart(p, mesh=st, config=cfg, dt=dt)
    merged = dict(st)
    merged.update({k: v for k, v in back.items() if isinstance(v, torch.Tensor)})
    c = Srk3Stepper(merged, st, cfg, dt)
    c.step(1)
    torch.cuda.synchronize()
    sa, sc = a.state_dict(), c.state_dict()
    for k in ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "pressure_p", "rtheta_p", "rho_p", "ru", "rw",
              "pv_edge", "ke", "divergence", "vorticity", "ruAvg", "wwAvg"):
        assert torch.equal(sa[k][:-1], sc[k][:-1]), k
