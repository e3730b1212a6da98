"""Host tests of the drop-in contract pieces that need no GPU.

* step-boundary receipts (the reference's ``mark_complete_step_boundary`` /
  ``require_boundary_continuation_identity`` / ``require_unedited_boundary_support``,
  ``state.py:154-253``): a clone is still a boundary, a prognostic edit asks for
  a new cold start, an edited diagnostic or a different mesh, configuration or
  dt is an error;
* nonfinite guards on entry and exit (``atmosphere_model.py:77-95``);
* the base-state binding of a cached device mesh follows the bound source
  (a new array, or an in-place change of a tensor, is uploaded again).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import srk3_case

from mpas_dycore_pytorch_amd.atmosphere_model import require_finite
from mpas_dycore_pytorch_amd.boundary import (BOUNDARY_KEY, boundary_status, mark_complete_step_boundary,
                                              mesh_digest)
from mpas_dycore_pytorch_amd.device import DeviceMesh
from mpas_dycore_pytorch_amd.errors import MpasConfigError


def _state():
    st, _, meta = srk3_case("l2_L20_tracer")
    state = {k: (torch.as_tensor(v) if isinstance(v, np.ndarray) and v.dtype.kind == "f" else v)
             for k, v in st.items()}
    return st, state, meta


def test_receipt_accepts_a_clone_and_rederives_after_a_prognostic_edit():
    st, state, meta = _state()
    cfg = {"config_len_disp": float(st["nominalMinDc"])}
    mark_complete_step_boundary(state, mesh=st, config=cfg, dt=meta["dt"])
    assert boundary_status(state, mesh=st, config=cfg, dt=meta["dt"]) is True
    clone = {k: (v.clone() if isinstance(v, torch.Tensor) else v) for k, v in state.items()}
    assert boundary_status(clone, mesh=st, config=cfg, dt=meta["dt"]) is True
    # an edit of one value of a prognostic: diagnostics must be derived again
    clone["theta_m"][3, 4] += 1e-9
    assert boundary_status(clone, mesh=st, config=cfg, dt=meta["dt"]) is False
    # a moved value (same sum, other position) is caught by the weighted checksum
    swapped = {k: (v.clone() if isinstance(v, torch.Tensor) else v) for k, v in state.items()}
    a, b = swapped["theta_m"][5, 1].item(), swapped["theta_m"][5, 9].item()
    assert a != b
    swapped["theta_m"][5, 1], swapped["theta_m"][5, 9] = b, a
    assert boundary_status(swapped, mesh=st, config=cfg, dt=meta["dt"]) is False
    # no receipt at all: not a boundary
    assert boundary_status({k: v for k, v in state.items() if k != BOUNDARY_KEY}, mesh=st, config=cfg,
                           dt=meta["dt"]) is False


def test_receipt_rejects_edited_support_and_another_identity():
    st, state, meta = _state()
    cfg = {"config_len_disp": float(st["nominalMinDc"])}
    mark_complete_step_boundary(state, mesh=st, config=cfg, dt=meta["dt"])
    edited = dict(state)
    edited["pv_edge"] = state["pv_edge"].clone()
    edited["pv_edge"][0, 0] += 1.0
    with pytest.raises(MpasConfigError, match="edited"):
        boundary_status(edited, mesh=st, config=cfg, dt=meta["dt"])
    with pytest.raises(MpasConfigError, match="identity"):
        boundary_status(state, mesh=st, config=cfg, dt=meta["dt"] * 2)
    with pytest.raises(MpasConfigError, match="identity"):
        boundary_status(state, mesh=st, config={**cfg, "config_smdiv": 0.2}, dt=meta["dt"])
    other_mesh = dict(st)
    other_mesh["dcEdge"] = st["dcEdge"] * 1.0000001
    with pytest.raises(MpasConfigError, match="identity"):
        boundary_status(state, mesh=other_mesh, config=cfg, dt=meta["dt"])
    # aliases and defaults resolve to the same identity
    assert boundary_status(state, mesh=st, config={"len_disp": float(st["nominalMinDc"]),
                                                   "config_smdiv": 0.1}, dt=meta["dt"]) is True
    assert mesh_digest(st) == mesh_digest(dict(st))


def test_nonfinite_guards():
    st, state, _ = _state()
    require_finite(state, "timestep input")
    for name in ("u", "theta_m", "scalars", "pressure_p"):
        bad = dict(state)
        bad[name] = state[name].clone()
        bad[name].view(-1)[7] = float("nan") if name != "u" else float("inf")
        with pytest.raises(MpasConfigError, match=f"timestep input: nonfinite values in {name}"):
            require_finite(bad, "timestep input")
    # fields outside the reference's list are not inspected
    odd = dict(state)
    odd["pv_edge"] = torch.full_like(state["pv_edge"], float("nan"))
    require_finite(odd, "timestep result")


def test_base_state_binding_follows_its_source():
    st0, _, _ = _state()
    st = {k: v for k, v in st0.items() if k != "rtheta_base"}  # the state supplies it (the mesh is looked up first)
    dm = DeviceMesh.get(st, torch.device("cpu"), st0)
    nc, L = dm.nc, dm.L
    first = torch.full((nc + 1, L), 2.0, dtype=torch.float64)
    second = torch.full((nc + 1, L), 3.0, dtype=torch.float64)
    s1 = dict(st, rtheta_base=first)
    dm = DeviceMesh.get(st, torch.device("cpu"), s1)
    assert float(dm.t["rtheta_base"][0, 0]) == 2.0
    dm = DeviceMesh.get(st, torch.device("cpu"), dict(st, rtheta_base=second))
    assert float(dm.t["rtheta_base"][0, 0]) == 3.0
    # the same tensor changed in place is uploaded again; an unchanged one is not
    second.fill_(4.0)
    dm = DeviceMesh.get(st, torch.device("cpu"), dict(st, rtheta_base=second))
    assert float(dm.t["rtheta_base"][0, 0]) == 4.0
    bound = dm.t["rtheta_base"]
    dm = DeviceMesh.get(st, torch.device("cpu"), dict(st, rtheta_base=second))
    assert dm.t["rtheta_base"] is bound
    # NumPy sources carry no version: always uploaded again
    arr = np.full((nc + 1, L), 5.0)
    dm = DeviceMesh.get(st, torch.device("cpu"), dict(st, rtheta_base=arr))
    arr[...] = 6.0
    dm = DeviceMesh.get(st, torch.device("cpu"), dict(st, rtheta_base=arr))
    assert float(dm.t["rtheta_base"][0, 0]) == 6.0
