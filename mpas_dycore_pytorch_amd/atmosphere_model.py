"""Public driver API over the GPU step: ``run_timestep`` and ``run_simulation``.

Mirrors the reference driver (``mpas_pytorch/atmosphere_model.py``):

* ``run_timestep(state, mesh, config=None, dt=None, device=..., differentiable=False)``
  (:900) prepares the state unless it is a completed step boundary, then
  advances one large step (:738-780);
* ``run_simulation(initial_state, mesh, config=None, dt=None, nsteps=1, device=...,
  on_step=None, ..., on_phase=None, on_initial=None)`` (:1474) prepares once and
  advances ``nsteps`` steps, calling ``on_step(step, state)`` after each;
* ``run_simulation.vjp`` / ``run_simulation.jvp`` (:2041-2155), the adjoint
  and tangent-linear models, through the hand-written derivative kernels.

The cold-start preparation (``_prepare_timestep_state`` :782-897:
``atm_init_coupled_diagnostics``, the padding sentinels, the first
diagnostic solve and the vector reconstruction) runs on the GPU, and in a
differentiable run it is differentiated too, as in the reference.  Inputs and
results are checked for nonfinite values (:77-95).  A non-differentiable
result carries a step-boundary receipt bound to its values, mesh,
configuration and dt (``boundary.py``; the reference's
``mark_complete_step_boundary``, ``state.py:194-214``).  Outputs are returned
on ``device`` (the reference's default is the CPU; computation is always on
the current ROCm GPU).
"""

from __future__ import annotations

from typing import Any, Callable, Iterable, Mapping

import numpy as np
import torch

from . import kernels as K
from .adjoint import COLD_START_INPUTS, STEP_INPUTS, Tape, taped_cold_start, vjp_scratch
from .boundary import (BOUNDARY_KEY, boundary_status, mark_complete_step_boundary, strip_boundary)
from .device import DeviceMesh, require_gpu
from .errors import MpasConfigError, MpasMissingFieldError, MpasUnsupportedCapabilityError
from .time_integration import Srk3Stepper, _atm_srk3_differentiable, _cfg

# keys run_simulation.vjp differentiates by default (atmosphere_model.py:157-168)
INTER_STEP_PROGNOSTIC_KEYS = ("u", "w", "theta_m", "rho_zz", "scalars", "rt_diabatic_tend")

STEP_BOUNDARY_KEY = BOUNDARY_KEY
RAW_KEYS = ("theta", "rho", "qv")
# fields whose values must be finite on entry and on exit (atmosphere_model.py:77-95)
FINITE_FIELDS = ("u", "w", "theta", "rho", "theta_m", "rho_zz", "scalars", "qv", "rho_base", "theta_base", "zz",
                 "pressure_p", "pressure_base")


def require_finite(state: Mapping[str, Any], context: str) -> None:
    """Raise ``MpasConfigError`` naming the first listed field that holds a NaN or an infinity."""
    bad = [name for name in FINITE_FIELDS
           if isinstance(state.get(name), torch.Tensor) and not bool(torch.isfinite(state[name].detach()).all())]
    if bad:
        raise MpasConfigError(f"{context}: nonfinite values in {bad[0]}"
                              + (f" (also {', '.join(bad[1:])})" if len(bad) > 1 else ""))


def _dt(config: Mapping[str, Any] | None, dt: float | None) -> float:
    if dt is None:
        dt = float((config or {}).get("config_dt", 720.0))
    dt = float(dt)
    if not np.isfinite(dt) or dt <= 0.0:
        raise MpasConfigError("dt must be finite and positive")
    return dt


def _theta_base_of(merged: Mapping[str, Any]) -> Any:
    if "theta_base" not in merged:
        raise MpasMissingFieldError(fields=["theta_base"], source="state")
    return merged["theta_base"]


def _cold_start_fields(merged: Mapping[str, Any]) -> dict[str, Any]:
    """u, w and the thermodynamic inputs of the cold start: theta_m/rho_zz when
    the state has them (the reference's derived path), else raw theta/rho/qv."""
    missing = [k for k in ("u", "w") if k not in merged]
    if "theta_m" in merged and "rho_zz" in merged:
        thermo = ("theta_m", "rho_zz")
    else:
        thermo = RAW_KEYS
        missing += [k for k in RAW_KEYS if k not in merged]
    if missing:
        raise MpasMissingFieldError(fields=missing + ["(theta/rho/qv or theta_m/rho_zz)"], source="state")
    return {k: merged[k] for k in ("u", "w") + thermo}


def _finish_prepared(state: Mapping[str, Any], prepared: Mapping[str, Any],
                     config: Mapping[str, Any] | None) -> dict[str, Any]:
    out = dict(state)
    out.update(prepared)
    out["h_edge"] = out["rho_edge"]
    out.setdefault("coef_3rd_order", float((config or {}).get("config_coef_3rd_order", 0.25)))
    out.setdefault("c_s", float((config or {}).get("config_smagorinsky_coef", 0.125)))
    if "scalars" in out and "num_scalars" not in out:
        out["num_scalars"] = int(out["scalars"].shape[-1])
    return out


def cold_start(state: Mapping[str, Any], mesh: Mapping[str, Any], config: Mapping[str, Any] | None, dt: float,
               device: torch.device) -> dict[str, Any]:
    """GPU cold start; returns the prepared mapping (device tensors)."""
    merged = {**mesh, **state}
    fields = _cold_start_fields(merged)
    dm = DeviceMesh.get(mesh, device, state)
    tb = _theta_base_of(merged)
    icd = K.init_coupled_diagnostics(dm, fields.get("theta"), fields.get("rho"), fields.get("qv"), fields["u"],
                                     fields["w"], tb, fields.get("theta_m"), fields.get("rho_zz"))
    dm.bind_base("rtheta_base", icd["rtheta_base"])
    dm.bind_base("exner_base", icd["exner_base"])
    apvm = float(_cfg(config)["config_apvm_upwinding"])
    d = K.solve_diagnostics(dm, dt, apvm, None, fields["u"], icd["rho_zz"])
    prepared = dict(icd)
    prepared.update({k: d[k] for k in ("h_edge", "v", "divergence", "vorticity", "ke", "pv_edge")})
    prepared["rho_edge"] = d["h_edge"]
    prepared.update(K.reconstruct(dm, fields["u"]))
    return _finish_prepared(state, prepared, config)


class _ColdStart(torch.autograd.Function):
    """The cold start as one autograd node: the forward records the GPU
    kernels on a tape (kept: it holds the cold start's own fields only); the
    backward and the jvp sweep that tape through the hand-written adjoint and
    tangent kernels of ``atm_init_coupled_diagnostics``, the diagnostic solve
    and the reconstruction."""

    @staticmethod
    def forward(box: dict, *inputs: torch.Tensor):
        dm, tb, dt, apvm = box["dm"], box["theta_base"], box["dt"], box["apvm"]
        fields = {k: v.detach() for k, v in zip(box["in_keys"], inputs)}
        tape = Tape(dm)
        with torch.cuda.device(dm.device):
            out = taped_cold_start(dm, fields, tb, dt, apvm, tape)
        keys = [k for k in sorted(out) if k not in ("rtheta_base", "exner_base", "pressure_base")]
        box["out_keys"] = keys
        box["constants"] = {k: out[k] for k in ("rtheta_base", "exner_base", "pressure_base")}
        box["_ctx"] = (tape, fields, out, list(box["in_keys"]), keys)
        return tuple(out[k] if k != "h_edge" else out[k].clone() for k in keys)

    @staticmethod
    def setup_context(ctx, inputs, output):
        box = inputs[0]
        ctx.tape, ctx.fields, ctx.out, ctx.in_keys, ctx.out_keys = box.pop("_ctx")

    @staticmethod
    def backward(ctx, *grads: torch.Tensor):
        tape, dm = ctx.tape, ctx.tape.dm
        seeds: dict[int, torch.Tensor] = {}
        for k, g in zip(ctx.out_keys, grads):
            if g is None:
                continue
            t = ctx.out[k]
            g = g.to(device=dm.device, dtype=torch.float64).contiguous()
            seeds[id(t)] = seeds[id(t)] + g if id(t) in seeds else g
        with torch.cuda.device(dm.device):
            scratch = vjp_scratch(dm, 0)
            bars = tape.backward(seeds, scratch, tape.depending_on(ctx.fields.values()))
        res = []
        for k in ctx.in_keys:
            b = bars.get(id(ctx.fields[k]))
            res.append(b if b is not None else torch.zeros_like(ctx.fields[k]))
        return (None,) + tuple(res)

    @staticmethod
    def jvp(ctx, _box_tangent, *tangents: torch.Tensor):
        tape, dm = ctx.tape, ctx.tape.dm
        seeds = {id(ctx.fields[k]): t.to(device=dm.device, dtype=torch.float64).contiguous()
                 for k, t in zip(ctx.in_keys, tangents) if t is not None}
        keep = {id(v) for v in ctx.out.values()}
        with torch.cuda.device(dm.device):
            dots = tape.forward_tangents(seeds, vjp_scratch(dm, 0), keep)
        res = []
        for k in ctx.out_keys:
            t = ctx.out[k]
            d = dots.get(id(t))
            res.append(d if d is not None else torch.zeros_like(t))
        return tuple(res)


def _differentiable_cold_start(state: Mapping[str, Any], mesh: Mapping[str, Any],
                               config: Mapping[str, Any] | None, dt: float, device: torch.device
                               ) -> dict[str, Any]:
    merged = {**mesh, **state}
    fields = _cold_start_fields(merged)
    dm = DeviceMesh.get(mesh, device, state)
    tb = K._d(_theta_base_of(merged), dm).detach()
    in_keys = list(fields)
    tensors = []
    for k in in_keys:
        v = fields[k]
        tensors.append(v if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v, dtype=np.float64)))
    box = {"dm": dm, "theta_base": tb, "dt": dt, "apvm": float(_cfg(config)["config_apvm_upwinding"]),
           "in_keys": in_keys}
    outs = _ColdStart.apply(box, *[t.to(device=device, dtype=torch.float64) for t in tensors])
    prepared = dict(zip(box["out_keys"], outs))
    prepared.update(box["constants"])
    dm.bind_base("rtheta_base", box["constants"]["rtheta_base"])
    dm.bind_base("exner_base", box["constants"]["exner_base"])
    out = _finish_prepared(state, prepared, config)
    out["h_edge"] = prepared["h_edge"]
    for k in ("u", "w"):
        out[k] = state[k] if k in state else mesh[k]
    return out


def _prepare(state: Mapping[str, Any], mesh: Mapping[str, Any], config: Mapping[str, Any] | None, dt: float,
             gpu: torch.device, differentiable: bool = False) -> dict[str, Any]:
    """The state the first step reads: the input itself when it is a complete,
    unedited step boundary of this (mesh, configuration, dt), otherwise the
    cold start of it."""
    require_finite(state, "timestep input")
    if boundary_status(state, mesh=mesh, config=config, dt=dt):
        return strip_boundary(state)
    plain = strip_boundary(state)
    if differentiable:
        return _differentiable_cold_start(plain, mesh, config, dt, gpu)
    return cold_start(plain, mesh, config, dt, gpu)


def _finish(stepper: Srk3Stepper, template: Mapping[str, Any], device: torch.device) -> dict[str, Any]:
    out = dict(template)
    for k, v in stepper.state_dict().items():
        out[k] = v if device.type == "cuda" else v.to(device)
    return out


def _bound(result: dict[str, Any], mesh: Mapping[str, Any], config: Mapping[str, Any] | None,
           dt: float) -> dict[str, Any]:
    mark_complete_step_boundary(result, mesh=mesh, config=config, dt=dt)
    return result


def _reorder(config: Mapping[str, Any] | None) -> bool:
    return bool((config or {}).get("config_mesh_reordering_sfc", True))


def run_timestep(state: dict[str, Any], mesh: dict[str, Any], config: dict | None = None, dt: float | None = None,
                 device: torch.device | str = "cuda", differentiable: bool = False) -> dict[str, Any]:
    """Advance the state by one large step (reference ``run_timestep``)."""
    dt = _dt(config, dt)
    if differentiable:
        result = _run_differentiable(state, mesh, config, dt, 1, torch.device(device), None)
        require_finite(result, "timestep result")
        return result
    gpu = require_gpu(None)
    prepared = _prepare(state, mesh, config, dt, gpu)
    stepper = Srk3Stepper(prepared, mesh, config, dt, device=gpu, reorder=_reorder(config))
    stepper.step(1)
    result = _finish(stepper, prepared, torch.device(device))
    require_finite(result, "timestep result")
    return _bound(result, mesh, config, dt)


def run_simulation(initial_state: dict[str, Any], mesh: dict[str, Any], config: dict | None = None,
                   dt: float | None = None, nsteps: int = 1, device: torch.device | str = "cuda",
                   on_step: Callable[[int, dict[str, Any]], None] | None = None, differentiable: bool = False,
                   checkpoint_timesteps: bool | None = None, checkpoint_use_reentrant: bool = False,
                   checkpoint_materialize_final: bool = True, on_phase: Callable[[str], None] | None = None,
                   on_initial: Callable[[dict[str, Any]], None] | None = None) -> dict[str, Any]:
    """Run ``nsteps`` large steps (reference ``run_simulation``).

    With ``differentiable=True`` the cold start (when the input is not a
    complete step boundary) and every large step are autograd nodes.  Each
    step keeps its input state as the checkpoint and recomputes the step,
    acoustic substeps included, in the backward (the reference's default
    outer timestep checkpointing, ``autodiff/checkpoint.py:106``).  The last
    step keeps the kernel tape of its forward instead, since its backward
    comes first; ``checkpoint_timesteps=False`` keeps every step's tape (no
    recompute, memory grows with ``nsteps``; ``atmosphere_model.py:1504-1509``).
    A non-differentiable run with ``nsteps > 0`` returns a bound step
    boundary (``atmosphere_model.py:2025-2033``)."""
    if isinstance(nsteps, bool) or not isinstance(nsteps, int) or nsteps < 0:
        raise MpasConfigError("nsteps must be a non-negative integer")
    dt = _dt(config, dt)
    out_dev = torch.device(device)
    if differentiable:
        result = _run_differentiable(initial_state, mesh, config, dt, nsteps, out_dev, on_step,
                                     keep_all=checkpoint_timesteps is False, on_initial=on_initial,
                                     on_phase=on_phase)
        require_finite(result, "forecast result")
        return result
    gpu = require_gpu(None)
    prepared = _prepare(initial_state, mesh, config, dt, gpu)
    if on_initial is not None:
        on_initial(prepared)
    if nsteps == 0:
        require_finite(prepared, "forecast result")
        return prepared
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
    result = _finish(stepper, prepared, out_dev)
    require_finite(result, "forecast result")
    return _bound(result, mesh, config, dt)


def _run_differentiable(state: Mapping[str, Any], mesh: Mapping[str, Any], config: Mapping[str, Any] | None,
                        dt: float, nsteps: int, device: torch.device,
                        on_step: Callable[[int, dict[str, Any]], None] | None,
                        keep_all: bool = False, on_initial=None, on_phase=None) -> dict[str, Any]:
    gpu = require_gpu(None)
    cur = _prepare(state, mesh, config, dt, gpu, differentiable=True)
    if on_initial is not None:
        on_initial(cur)
    if on_phase is not None:
        on_phase("forecast_integration_begin")
    for step in range(1, nsteps + 1):
        # a tape is worth keeping only if a backward can reach this step
        graph = torch.is_grad_enabled() and any(
            isinstance(v, torch.Tensor) and v.requires_grad for k, v in cur.items() if k in STEP_INPUTS)
        keep = graph and (keep_all or step == nsteps)
        # the base state is bound by the first step; later steps of the run reuse it
        cur = _atm_srk3_differentiable(cur, mesh, config, dt, strict=True, keep_tape=keep,
                                       rebind_base=step == 1)
        if on_step is not None:
            on_step(step, cur)
    if on_phase is not None:
        on_phase("forecast_integration_end")
    if device.type != "cuda":
        cur = {k: (v.to(device) if isinstance(v, torch.Tensor) else v) for k, v in cur.items()}
    return cur


# ---------------------------------------------------------------------------
# derivative drivers (reference atmosphere_model.py:2041-2155)
# ---------------------------------------------------------------------------
def _differentiable_keys(state: Mapping[str, Any], track_keys: Iterable[str] | None) -> tuple[str, ...]:
    if track_keys is not None:
        return tuple(track_keys)
    return tuple(k for k in INTER_STEP_PROGNOSTIC_KEYS if k in state)


def _leaf_copy(state: Mapping[str, Any], key: str) -> torch.Tensor:
    """A fresh leaf holding the value of ``state[key]`` that autograd tracks."""
    if key not in state:
        raise KeyError(f"run_simulation.vjp: no state entry {key!r} to differentiate with respect to")
    value = state[key]
    if not isinstance(value, torch.Tensor):
        raise TypeError(f"run_simulation.vjp: state entry {key!r} is a {type(value).__name__}, not a tensor")
    return value.detach().clone().requires_grad_(True)


def _checked_loss(loss: Any) -> torch.Tensor:
    if not isinstance(loss, torch.Tensor):
        raise TypeError(f"run_simulation.vjp: the objective returned a {type(loss).__name__}; it must return "
                        "a tensor")
    if loss.ndim != 0:
        raise MpasConfigError(f"run_simulation.vjp: the objective returned shape {tuple(loss.shape)}; it must "
                              "return a 0-d tensor")
    return loss


def _run_simulation_vjp(state: dict[str, Any], mesh: dict[str, Any], objective: Callable[[dict[str, Any]], torch.Tensor],
                        track_keys: Iterable[str] | None = None, **kwargs: Any) -> dict[str, Any]:
    """Gradient of ``objective(run_simulation(state, mesh, **kwargs))`` with
    respect to the tracked state entries (reference ``run_simulation.vjp``).

    Each tracked entry (by default the inter-step prognostics present in
    ``state``) is replaced by a leaf copy, the differentiable run is made,
    and one reverse sweep returns d(objective)/d(entry); entries the
    objective does not depend on are left out of the result.  The sweep
    recomputes each large step from its checkpoint and applies the
    hand-written adjoint kernels."""
    kwargs.setdefault("differentiable", True)
    leaves = {key: _leaf_copy(state, key) for key in _differentiable_keys(state, track_keys)}
    start = dict(state)
    start.update(leaves)
    loss = _checked_loss(objective(run_simulation(start, mesh, **kwargs)))
    found = torch.autograd.grad(loss, tuple(leaves.values()), allow_unused=True)
    return {key: g for key, g in zip(leaves, found) if g is not None}


def _checked_tangent(state: Mapping[str, Any], key: str, tangent: Any) -> None:
    """Reject a perturbation the way the reference does (type, presence,
    target type, shape, dtype, device, in that order), plus one that names
    neither a large-step nor a cold-start input."""
    if not isinstance(tangent, torch.Tensor):
        raise TypeError(f"run_simulation.jvp: the perturbation of {key!r} is a {type(tangent).__name__}, "
                        "not a tensor")
    if key not in state:
        raise KeyError(f"run_simulation.jvp: no state entry {key!r} to perturb")
    target = state[key]
    if not isinstance(target, torch.Tensor):
        raise TypeError(f"run_simulation.jvp: state entry {key!r} is a {type(target).__name__}; only tensors "
                        "take a tangent")
    for what, got, want in (("shape", tuple(tangent.shape), tuple(target.shape)),
                            ("dtype", tangent.dtype, target.dtype), ("device", tangent.device, target.device)):
        if got != want:
            raise MpasConfigError(f"run_simulation.jvp: the perturbation of {key!r} has {what} {got}, the state "
                                  f"entry {want}")
    if key not in STEP_INPUTS and key not in COLD_START_INPUTS:
        raise MpasUnsupportedCapabilityError(
            f"run_simulation.jvp: {key!r} is neither a large-step input ({', '.join(STEP_INPUTS)}) nor a "
            f"cold-start input ({', '.join(COLD_START_INPUTS)})", capability="jvp_input", requested=key)


def _run_simulation_jvp(state: dict[str, Any], mesh: dict[str, Any], perturbation: dict[str, Any],
                        **kwargs: Any) -> dict[str, Any]:
    """Directional derivative of ``run_simulation(state, mesh, **kwargs)``
    along ``perturbation`` (reference ``run_simulation.jvp``).

    ``perturbation`` maps state entries to tangents of the same shape, dtype
    and device; every other entry is held constant.  The run is made on
    forward-mode dual numbers; each large step (and the cold start) takes
    its tangent from the tangent-linear kernels.  Returns the tangent of
    every output tensor a perturbation reaches."""
    from torch.autograd import forward_ad

    kwargs.setdefault("differentiable", True)
    for key, tangent in perturbation.items():
        _checked_tangent(state, key, tangent)
    with forward_ad.dual_level():
        start = dict(state)
        for key, tangent in perturbation.items():
            start[key] = forward_ad.make_dual(state[key], tangent)
        final = run_simulation(start, mesh, **kwargs)
        tangents = {}
        for key, value in final.items():
            if isinstance(value, torch.Tensor):
                t = forward_ad.unpack_dual(value).tangent
                if t is not None:
                    tangents[key] = t
        return tangents


run_simulation.vjp = _run_simulation_vjp  # type: ignore[attr-defined]
run_simulation.jvp = _run_simulation_jvp  # type: ignore[attr-defined]
