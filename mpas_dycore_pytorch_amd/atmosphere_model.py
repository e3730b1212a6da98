"""Public driver API over the GPU step: ``run_timestep`` and ``run_simulation``.

Mirrors the reference driver (``mpas_pytorch/atmosphere_model.py``):

* ``run_timestep(state, mesh, config=None, dt=None, device=..., differentiable=False)``
  (:900) prepares the state unless it is a completed step boundary, then
  advances one large step (:738-780);
* ``run_simulation(initial_state, mesh, config=None, dt=None, nsteps=1, device=...,
  on_step=None, ..., on_phase=None, on_initial=None)`` (:1474) prepares once and
  advances ``nsteps`` steps, calling ``on_step(step, state)`` after each.

The cold-start preparation (``_prepare_timestep_state`` :782-897:
``atm_init_coupled_diagnostics``, the padding sentinels, the first
diagnostic solve and the vector reconstruction) runs on the GPU.  Outputs are
returned on ``device`` (the reference's default is the CPU; computation is
always on the current ROCm GPU).  Results carry the key
``STEP_BOUNDARY_KEY`` so that a later call does not re-derive diagnostics,
as the reference's step-boundary receipt does (:776-779).
"""

from __future__ import annotations

from typing import Any, Callable, Mapping

import numpy as np
import torch

from . import kernels as K
from .device import DeviceMesh, require_gpu, to_numpy
from .errors import MpasConfigError, MpasMissingFieldError, MpasUnsupportedCapabilityError
from .adjoint import STEP_INPUTS
from .time_integration import Srk3Stepper, _atm_srk3_differentiable, atm_srk3

# keys run_simulation.vjp differentiates by default (atmosphere_model.py:157-168)
INTER_STEP_PROGNOSTIC_KEYS = ("u", "w", "theta_m", "rho_zz", "scalars", "rt_diabatic_tend")

STEP_BOUNDARY_KEY = "_mpas_hip_step_boundary"
RAW_KEYS = ("theta", "rho", "qv", "u", "w", "theta_base")


def _dt(config: Mapping[str, Any] | None, dt: float | None) -> float:
    if dt is None:
        dt = float((config or {}).get("config_dt", 720.0))
    dt = float(dt)
    if not np.isfinite(dt) or dt <= 0.0:
        raise MpasConfigError("dt must be finite and positive")
    return dt


def cold_start(state: Mapping[str, Any], mesh: Mapping[str, Any], config: Mapping[str, Any] | None, dt: float,
               device: torch.device) -> dict[str, Any]:
    """GPU cold start of raw fields; returns the prepared mapping (device tensors)."""
    missing = [k for k in RAW_KEYS if k not in state and k not in mesh]
    if missing:
        raise MpasMissingFieldError(fields=missing, source="state")
    merged = {**mesh, **state}
    dm = DeviceMesh(merged, device, None)
    icd = K.init_coupled_diagnostics(dm, merged["theta"], merged["rho"], merged["qv"], merged["u"], merged["w"],
                                     merged["theta_base"])
    dm.bind_base("rtheta_base", icd["rtheta_base"])
    dm.bind_base("exner_base", icd["exner_base"])
    apvm = float((config or {}).get("config_apvm_upwinding", 0.5))
    d = K.solve_diagnostics(dm, dt, apvm, None, merged["u"], icd["rho_zz"])
    r = K.reconstruct(dm, merged["u"])
    out = dict(state)
    out.update(icd)
    out.update({k: d[k] for k in ("h_edge", "v", "divergence", "vorticity", "ke", "pv_edge")})
    out["rho_edge"] = d["h_edge"]
    out.update(r)
    out.setdefault("coef_3rd_order", float((config or {}).get("config_coef_3rd_order", 0.25)))
    out.setdefault("c_s", float((config or {}).get("config_smagorinsky_coef", 0.125)))
    if "scalars" in out and "num_scalars" not in out:
        out["num_scalars"] = int(out["scalars"].shape[-1])
    return out


def _prepare(state, mesh, config, dt, gpu):
    if state.get(STEP_BOUNDARY_KEY, False):
        return dict(state)
    return cold_start(state, mesh, config, dt, gpu)


def _finish(stepper: Srk3Stepper, template: Mapping[str, Any], device: torch.device) -> dict[str, Any]:
    out = dict(template)
    for k, v in stepper.state_dict().items():
        out[k] = v if device.type == "cuda" else v.to(device)
    out[STEP_BOUNDARY_KEY] = True
    return out


def _reorder(config: Mapping[str, Any] | None) -> bool:
    return bool((config or {}).get("config_mesh_reordering_sfc", True))


def run_timestep(state: dict[str, Any], mesh: dict[str, Any], config: dict | None = None, dt: float | None = None,
                 device: torch.device | str = "cuda", differentiable: bool = False) -> dict[str, Any]:
    """Advance the state by one large step (reference ``run_timestep``)."""
    if differentiable:
        return _run_differentiable(state, mesh, config, _dt(config, dt), 1, torch.device(device), None)
    dt = _dt(config, dt)
    gpu = require_gpu(None)
    out_dev = torch.device(device)
    prepared = _prepare(dict(state), dict(mesh), config, dt, gpu)
    stepper = Srk3Stepper(prepared, mesh, config, dt, device=gpu, reorder=_reorder(config))
    stepper.step(1)
    return _finish(stepper, prepared, out_dev)


def run_simulation(initial_state: dict[str, Any], mesh: dict[str, Any], config: dict | None = None,
                   dt: float | None = None, nsteps: int = 1, device: torch.device | str = "cuda",
                   on_step: Callable[[int, dict[str, Any]], None] | None = None, differentiable: bool = False,
                   checkpoint_timesteps: bool | None = None, checkpoint_use_reentrant: bool = False,
                   checkpoint_materialize_final: bool = True, on_phase: Callable[[str], None] | None = None,
                   on_initial: Callable[[dict[str, Any]], None] | None = None) -> dict[str, Any]:
    """Run ``nsteps`` large steps (reference ``run_simulation``).

    With ``differentiable=True`` every large step is one autograd node that
    keeps its input state as the checkpoint and recomputes the step, acoustic
    substeps included, in the backward (the reference's default outer
    timestep checkpointing, ``autodiff/checkpoint.py:106``).  The last step
    keeps the kernel tape of its forward instead, since its backward comes
    first; ``checkpoint_timesteps=False`` keeps every step's tape (no
    recompute, memory grows with ``nsteps``; ``atmosphere_model.py:1504-1509``)."""
    if isinstance(nsteps, bool) or not isinstance(nsteps, int) or nsteps < 0:
        raise MpasConfigError("nsteps must be a non-negative integer")
    if differentiable:
        return _run_differentiable(initial_state, mesh, config, _dt(config, dt), nsteps, torch.device(device),
                                   on_step, keep_all=checkpoint_timesteps is False)
    dt = _dt(config, dt)
    gpu = require_gpu(None)
    out_dev = torch.device(device)
    prepared = _prepare(dict(initial_state), dict(mesh), config, dt, gpu)
    if on_initial is not None:
        on_initial(prepared)
    stepper = Srk3Stepper(prepared, mesh, config, dt, device=gpu, reorder=_reorder(config))
    if on_phase is not None:
        on_phase("forecast_integration_begin")
    for step in range(1, nsteps + 1):
        stepper.step(1)
        if on_step is not None:
            on_step(step, _finish(stepper, prepared, out_dev))
    if on_phase is not None:
        torch.cuda.synchronize(gpu)
        on_phase("forecast_integration_end")
    if nsteps == 0:
        out = dict(prepared)
        out[STEP_BOUNDARY_KEY] = True
        return out
    return _finish(stepper, prepared, out_dev)


def _run_differentiable(state: Mapping[str, Any], mesh: Mapping[str, Any], config: Mapping[str, Any] | None,
                        dt: float, nsteps: int, device: torch.device,
                        on_step: Callable[[int, dict[str, Any]], None] | None,
                        keep_all: bool = False) -> dict[str, Any]:
    if not state.get(STEP_BOUNDARY_KEY, False):
        raise MpasUnsupportedCapabilityError(
            "differentiable runs start from a prepared step-boundary state (run_simulation(..., nsteps=0) "
            "returns one); the cold start itself is not differentiated", capability="differentiable_cold_start",
            requested=True)
    cur = dict(state)
    for step in range(1, nsteps + 1):
        # a tape is worth keeping only if a backward can reach this step
        graph = torch.is_grad_enabled() and any(
            isinstance(v, torch.Tensor) and v.requires_grad for k, v in cur.items() if k in STEP_INPUTS)
        keep = graph and (keep_all or step == nsteps)
        cur = _atm_srk3_differentiable(cur, mesh, config, dt, strict=True, keep_tape=keep)
        cur[STEP_BOUNDARY_KEY] = True
        if on_step is not None:
            on_step(step, cur)
    if device.type != "cuda":
        cur = {k: (v.to(device) if isinstance(v, torch.Tensor) else v) for k, v in cur.items()}
    return cur


def _run_simulation_vjp(state: dict[str, Any], mesh: dict[str, Any], objective: Callable[[dict[str, Any]], torch.Tensor],
                        track_keys=None, **kwargs: Any) -> dict[str, Any]:
    """Adjoint model: d objective(run_simulation(state)) / d state (reference
    ``run_simulation.vjp``, atmosphere_model.py:2101-2155), through the
    hand-written adjoint kernels with one checkpoint per large step."""
    kwargs.setdefault("differentiable", True)
    clone_state = dict(state)
    keys = tuple(k for k in INTER_STEP_PROGNOSTIC_KEYS if k in clone_state) if track_keys is None else tuple(track_keys)
    tracked: dict[str, torch.Tensor] = {}
    for key in keys:
        if key not in clone_state:
            raise KeyError(f"Tracked VJP key '{key}' is missing from state")
        v = clone_state[key]
        if not isinstance(v, torch.Tensor):
            raise TypeError(f"Tracked VJP key '{key}' is not a tensor")
        leaf = v.detach().clone().to(torch.float64).requires_grad_(True)
        clone_state[key] = leaf
        tracked[key] = leaf
    out = run_simulation(clone_state, mesh, **kwargs)
    loss = objective(out)
    if not isinstance(loss, torch.Tensor):
        raise TypeError(f"objective must return a tensor, got {type(loss).__name__}")
    if loss.ndim != 0:
        raise MpasConfigError("objective must return a scalar tensor")
    grads = torch.autograd.grad(loss, list(tracked.values()), allow_unused=True)
    return {k: g for k, g in zip(tracked, grads) if g is not None}


def _run_simulation_jvp(state: dict[str, Any], mesh: dict[str, Any], perturbation: dict[str, Any],
                        **kwargs: Any) -> dict[str, Any]:
    """Tangent-linear model (reference ``run_simulation.jvp``,
    atmosphere_model.py:2041-2098): forward-mode dual numbers through
    ``run_simulation(differentiable=True)``, whose every large step takes its
    tangent from the hand-written tangent-linear kernels.

    ``perturbation`` maps state keys to tangents of the same shape, dtype and
    device; other entries are constants.  Returns the tangents of the output
    state's tensors."""
    from torch.autograd import forward_ad

    kwargs.setdefault("differentiable", True)
    with forward_ad.dual_level():
        dual_state = dict(state)
        for key, tangent in perturbation.items():
            if not isinstance(tangent, torch.Tensor):
                raise TypeError(f"Perturbation '{key}' must be a tensor, got {type(tangent).__name__}")
            if key not in dual_state:
                raise KeyError(f"Perturbation key '{key}' is missing from state")
            value = dual_state[key]
            if not isinstance(value, torch.Tensor):
                raise TypeError(f"State entry '{key}' is not a tensor and cannot receive a tangent")
            if tangent.shape != value.shape:
                raise MpasConfigError(f"Perturbation '{key}' has shape {tuple(tangent.shape)}; "
                                      f"expected {tuple(value.shape)}")
            if tangent.dtype != value.dtype:
                raise MpasConfigError(f"Perturbation '{key}' has dtype {tangent.dtype}; expected {value.dtype}")
            if tangent.device != value.device:
                raise MpasConfigError(f"Perturbation '{key}' is on {tangent.device}; expected {value.device}")
            if key not in STEP_INPUTS:
                raise MpasUnsupportedCapabilityError(
                    f"run_simulation.jvp: '{key}' is not an input of the large step; tangents are taken "
                    f"with respect to {', '.join(STEP_INPUTS)}", capability="jvp_input", requested=key)
            dual_state[key] = forward_ad.make_dual(value, tangent)
        out = run_simulation(dual_state, mesh, **kwargs)
        result: dict[str, Any] = {}
        for key, val in out.items():
            if isinstance(val, torch.Tensor):
                _, t = forward_ad.unpack_dual(val)
                if t is not None:
                    result[key] = t
        return result


run_simulation.vjp = _run_simulation_vjp  # type: ignore[attr-defined]
run_simulation.jvp = _run_simulation_jvp  # type: ignore[attr-defined]
