"""Complete-step boundaries: the receipt a finished large step leaves on its state.

A state returned by ``run_timestep`` / ``run_simulation`` already carries the
diagnostics the next step reads, so a later call must not derive them again
from the prognostics (that changes the trajectory by rounding).  The
reference marks such a state with a receipt bound to the state's values and
to its continuation identity (mesh, configuration, dt)
(``mpas_pytorch/state.py:194-236``) and, on the next call,

* rejects a continuation under a different mesh, configuration or dt
  (``require_boundary_continuation_identity``, ``state.py:217-229``);
* rejects edited diagnostics or support fields
  (``require_unedited_boundary_support``, ``state.py:154-191``);
* derives the diagnostics again when only the prognostics were edited
  (``has_complete_step_boundary``, ``state.py:232-253``).

The reference keeps a byte copy of every field on the host.  Here the
receipt holds a 128-bit checksum per field, formed where the tensor lives
(on the GPU for device states): the sum of the field's 64-bit words and a
position-weighted sum of them, both modulo 2**64.  A clone has the same
checksum, so a deep-copied boundary is still recognised; an in-place edit or
a replaced tensor changes it.  The mesh is identified by a SHA-256 over its
arrays (as ``io_schema.mesh_digest`` does), cached per mesh object.
"""

from __future__ import annotations

import hashlib
import json
from typing import Any, Mapping

import numpy as np
import torch

from .errors import MpasConfigError

BOUNDARY_KEY = "_mpas_hip_step_boundary"
BINDING = "mpas-hip-boundary-checksum.v1"
PROGNOSTICS = ("u", "w", "rho_zz", "theta_m", "scalars")
# 61 odd weights; word i of a field is weighted by _WEIGHTS[i % 61]
_NW = 61
_WEIGHTS_NP = (np.arange(_NW, dtype=np.int64) * 0x9E3779B1 + 0x632BE5AB) | 1


def _as_tensor(v: Any) -> torch.Tensor | None:
    if isinstance(v, torch.Tensor):
        return v.detach()
    if isinstance(v, np.ndarray) and v.dtype != object:
        return torch.from_numpy(np.ascontiguousarray(v))
    return None


def _words(t: torch.Tensor) -> torch.Tensor:
    """The field's bytes as int64 words (zero-padded to a multiple of 8 bytes)."""
    t = t.contiguous().reshape(-1)
    if t.dtype == torch.bool:
        t = t.to(torch.uint8)
    b = t.view(torch.uint8) if t.element_size() != 8 else None
    if b is None:
        return t.view(torch.int64)
    pad = (-b.numel()) % 8
    if pad:
        b = torch.cat([b, torch.zeros(pad, dtype=torch.uint8, device=b.device)])
    return b.view(torch.int64)


def _checksums(fields: Mapping[str, torch.Tensor]) -> dict[str, tuple]:
    """Per field: (dtype, shape, plain word sum, weighted word sum)."""
    sums = []
    names = []
    for name, t in fields.items():
        w = _words(t)
        n = w.numel()
        full = (n // _NW) * _NW
        wt = torch.as_tensor(_WEIGHTS_NP, device=w.device)
        s2 = (w[:full].view(-1, _NW) * wt).sum() if full else torch.zeros((), dtype=torch.int64, device=w.device)
        if n > full:
            s2 = s2 + (w[full:] * wt[: n - full]).sum()
        s1 = w.sum() if n else torch.zeros((), dtype=torch.int64, device=w.device)
        sums.append(torch.stack([s1, s2]).cpu())
        names.append(name)
    out = {}
    for name, s in zip(names, sums):
        t = fields[name]
        out[name] = (str(t.dtype), tuple(t.shape), int(s[0]), int(s[1]))
    return out


def _snapshot(state: Mapping[str, Any], support_only: bool) -> dict[str, tuple]:
    fields: dict[str, torch.Tensor] = {}
    scalars: dict[str, tuple] = {}
    for key, value in state.items():
        if key.startswith("_"):
            continue
        if support_only and key in PROGNOSTICS:
            continue
        t = _as_tensor(value)
        if t is not None:
            fields[key] = t
        elif isinstance(value, (bool, int, float, str, np.generic)):
            scalars[key] = ("scalar", json.dumps(value, sort_keys=True, default=str))
    snap = _checksums(fields)
    snap.update(scalars)
    return snap


_MESH_DIGESTS: dict[int, tuple[Any, str]] = {}


def mesh_digest(mesh: Mapping[str, Any]) -> str:
    """SHA-256 over the mesh's named arrays and scalars, cached per mesh object."""
    hit = _MESH_DIGESTS.get(id(mesh))
    if hit is not None and hit[0] is mesh:
        return hit[1]
    h = hashlib.sha256()
    n = 0
    for name, value in sorted(mesh.items()):
        if name.startswith("_"):
            continue
        if isinstance(value, torch.Tensor):
            a = value.detach().cpu().contiguous().numpy()
        elif isinstance(value, np.ndarray):
            a = np.ascontiguousarray(value)
        elif isinstance(value, (bool, int, float, str, np.generic)):
            h.update(name.encode() + b"\0=" + json.dumps(value, default=str).encode() + b"\0")
            n += 1
            continue
        else:
            continue
        h.update(name.encode() + b"\0" + str(a.dtype).encode() + str(a.shape).encode() + b"\0")
        h.update(a.tobytes())
        n += 1
    if n == 0:
        raise MpasConfigError("a mesh needs at least one array or scalar to identify it")
    d = h.hexdigest()
    if len(_MESH_DIGESTS) > 8:
        _MESH_DIGESTS.clear()
    _MESH_DIGESTS[id(mesh)] = (mesh, d)
    return d


def config_digest(config: Mapping[str, Any] | None) -> str:
    """Identity of the scientific configuration: the reference defaults filled
    in and short aliases resolved, private keys and the checkpoint policy left
    out (``io_schema.continuation_configuration_digest``)."""
    from .time_integration import _cfg

    items = {k: v for k, v in _cfg(config).items() if not str(k).startswith("_") and k != "gradient_checkpointing"}
    return hashlib.sha256(json.dumps(items, sort_keys=True, default=str).encode()).hexdigest()


def mark_complete_step_boundary(state: dict[str, Any], *, mesh: Mapping[str, Any] | None = None,
                                config: Mapping[str, Any] | None = None, dt: float | None = None) -> None:
    """Bind ``state`` (in place) as a complete step boundary for (mesh, config, dt)."""
    for name in PROGNOSTICS:
        if not isinstance(state.get(name), torch.Tensor):
            raise MpasConfigError(f"a complete step boundary needs the tensor {name!r}")
    snap = _snapshot(state, support_only=False)
    state[BOUNDARY_KEY] = {
        "binding": BINDING,
        "prognostics": {k: snap[k] for k in PROGNOSTICS},
        "support": {k: v for k, v in snap.items() if k not in PROGNOSTICS},
        "mesh": mesh_digest(mesh) if mesh is not None else None,
        "config": config_digest(config),
        "dt": float(dt) if dt is not None else None,
    }


def boundary_status(state: Mapping[str, Any], *, mesh: Mapping[str, Any], config: Mapping[str, Any] | None,
                    dt: float) -> bool:
    """True when ``state`` is a complete boundary whose prognostics are unedited.

    Raises ``MpasConfigError`` when the receipt belongs to another mesh,
    configuration or dt, or when a diagnostic or support field was edited
    after the step.  Returns False (the caller derives the diagnostics
    again) when there is no receipt or only prognostics changed."""
    marker = state.get(BOUNDARY_KEY)
    if not isinstance(marker, dict):
        return False
    if marker.get("binding") != BINDING:
        raise MpasConfigError("step boundary receipt of an unknown binding; prepare the state again")
    if ((marker.get("mesh") is not None and marker["mesh"] != mesh_digest(mesh))
            or (marker.get("config") is not None and marker["config"] != config_digest(config))
            or (marker.get("dt") is not None and marker["dt"] != float(dt))):
        raise MpasConfigError("complete boundary continuation identity mismatch (mesh, configuration or dt "
                              "differs from the step that produced this state)")
    snap = _snapshot(state, support_only=False)
    support = {k: v for k, v in snap.items() if k not in PROGNOSTICS}
    if support != marker["support"]:
        changed = sorted(k for k in set(support) | set(marker["support"])
                         if support.get(k) != marker["support"].get(k))
        raise MpasConfigError("complete boundary diagnostics or support inputs were edited: "
                              + ", ".join(changed[:8]))
    return all(snap.get(k) == marker["prognostics"].get(k) for k in PROGNOSTICS)


def strip_boundary(state: Mapping[str, Any]) -> dict[str, Any]:
    """A copy of ``state`` without its receipt."""
    return {k: v for k, v in state.items() if k != BOUNDARY_KEY}
