"""On-disk formats at the step's boundaries: mesh / initial state, restart, history.

The reference writes NetCDF (``restart.py:383-690``, ``output_writer.py:329``)
through ``netCDF4``, which this stack does not carry.  The same contents go
into one self-describing binary container instead, with the reference's
variable and dimension names so that a NetCDF converter is a loop:

    8 bytes   magic ``MPASHIPB``
    8 bytes   little-endian length H of the header
    H bytes   UTF-8 JSON header: schema, version, attributes, dimensions and
              one entry per variable (name, dims, dtype, shape, offset, nbytes)
    ...       the arrays, C order, little endian, each 64-byte aligned

The header carries ``payload_sha256``, the SHA-256 of the canonical header
(without that key) followed by every array's bytes in variable order; readers
recompute it and reject any mismatch (the reference's
``restart_payload_sha256``, ``restart.py:349``).  Files are written to a
temporary name and renamed into place, so a reader never sees a partial file.

* :func:`write_mesh` / :func:`read_mesh` and :func:`write_state` /
  :func:`read_state`: the mesh dict (``ic_loader.py:1243``) and a prepared
  state, every array in the reference's padded layout;
  :func:`mesh_digest` is the mesh identity both carry;
* :func:`write_restart` / :func:`read_restart`: the inter-step state (active
  rows) with its mesh, configuration and timestep identities; reading it back
  and stepping continues bit for bit;
* :func:`write_history` / :func:`read_history`: the twelve diagnostic fields of
  ``docs/API.md:179-201``; theta, rho and pressure are derived on the GPU
  (``mpas_hip_history_fields``) when the state lives there.
"""

from __future__ import annotations

import ctypes as C
import datetime as _dt
import hashlib
import json
import os
from pathlib import Path
from typing import Any, Mapping

import numpy as np
import torch

from .errors import MpasConfigError, MpasDeviceError, MpasMissingFieldError

MAGIC = b"MPASHIPB"
VERSION = 1
ALIGN = 64
RVORD = 1.6083623693379792  # rv / rd (constants.py:96)

# ---------------------------------------------------------------------------
# container
# ---------------------------------------------------------------------------


def _canonical(header: Mapping[str, Any]) -> bytes:
    h = {k: v for k, v in header.items() if k != "payload_sha256"}
    return json.dumps(h, sort_keys=True, separators=(",", ":"), allow_nan=False).encode()


def write_container(path: str | Path, schema: str, arrays: Mapping[str, tuple[np.ndarray, tuple[str, ...]]],
                    dimensions: Mapping[str, int], attributes: Mapping[str, Any] | None = None,
                    clobber: bool = False) -> str:
    """Write ``arrays`` (name -> (array, dimension names)); returns the payload SHA-256."""
    path = Path(path)
    if path.exists() and not clobber:
        raise FileExistsError(f"{path} already exists")
    entries, blobs, offset = [], [], 0
    for name, (a, dims) in arrays.items():
        a = np.ascontiguousarray(a)
        if a.dtype.kind == "f" and not np.isfinite(a).all():
            raise MpasConfigError(f"field {name!r} has nonfinite values")
        if a.dtype.byteorder == ">":
            a = a.astype(a.dtype.newbyteorder("<"))
        dims = tuple(dims)
        if len(dims) != a.ndim or any(int(dimensions[d]) != n for d, n in zip(dims, a.shape)):
            raise MpasConfigError(f"field {name!r}: shape {a.shape} does not match dimensions {dims}")
        entries.append({"name": name, "dims": list(dims), "dtype": a.dtype.str, "shape": list(a.shape),
                        "offset": offset, "nbytes": int(a.nbytes)})
        blobs.append(a)
        offset += (a.nbytes + ALIGN - 1) // ALIGN * ALIGN
    header = {"schema": schema, "version": VERSION, "attributes": dict(attributes or {}),
              "dimensions": {k: int(v) for k, v in dimensions.items()}, "variables": entries}
    digest = hashlib.sha256(_canonical(header))
    for a in blobs:
        digest.update(a.tobytes())
    header["payload_sha256"] = digest.hexdigest()
    text = json.dumps(header, sort_keys=True, allow_nan=False).encode()
    start = (16 + len(text) + ALIGN - 1) // ALIGN * ALIGN
    tmp = path.with_name(path.name + f".tmp{os.getpid()}")
    with open(tmp, "wb") as f:
        f.write(MAGIC)
        f.write(len(text).to_bytes(8, "little"))
        f.write(text)
        f.write(b"\0" * (start - 16 - len(text)))
        pos = 0
        for e, a in zip(entries, blobs):
            f.write(b"\0" * (e["offset"] - pos))
            f.write(a.tobytes())
            pos = e["offset"] + e["nbytes"]
    os.replace(tmp, path)
    return header["payload_sha256"]


def read_container(path: str | Path, validate: bool = True) -> tuple[dict[str, np.ndarray], dict[str, Any]]:
    """(arrays, header) of a container; ``validate`` checks the payload SHA-256."""
    data = Path(path).read_bytes()
    if data[:8] != MAGIC:
        raise MpasConfigError(f"{path}: not an MPASHIPB container")
    n = int.from_bytes(data[8:16], "little")
    try:
        header = json.loads(data[16:16 + n].decode())
    except (UnicodeDecodeError, json.JSONDecodeError) as exc:
        raise MpasConfigError(f"{path}: malformed header") from exc
    if header.get("version") != VERSION:
        raise MpasConfigError(f"{path}: container version {header.get('version')} is not {VERSION}")
    start = (16 + n + ALIGN - 1) // ALIGN * ALIGN
    arrays = {}
    digest = hashlib.sha256(_canonical(header))
    for e in header["variables"]:
        lo = start + int(e["offset"])
        raw = data[lo:lo + int(e["nbytes"])]
        if len(raw) != int(e["nbytes"]):
            raise MpasConfigError(f"{path}: field {e['name']!r} is truncated")
        digest.update(raw)
        arrays[e["name"]] = np.frombuffer(raw, dtype=np.dtype(e["dtype"])).reshape(e["shape"]).copy()
    if validate and digest.hexdigest() != header.get("payload_sha256"):
        raise MpasConfigError(f"{path}: payload SHA-256 mismatch (file altered or truncated)")
    return arrays, header


# ---------------------------------------------------------------------------
# identities
# ---------------------------------------------------------------------------


def _np(v: Any) -> np.ndarray:
    if isinstance(v, torch.Tensor):
        return v.detach().cpu().numpy()
    return np.asarray(v)


def mesh_digest(mesh: Mapping[str, Any]) -> str:
    """SHA-256 of every array and scalar of a mesh mapping (sorted by name)."""
    h = hashlib.sha256()
    for k in sorted(mesh):
        v = mesh[k]
        if isinstance(v, (np.ndarray, torch.Tensor, int, float, np.generic)) and not isinstance(v, bool):
            a = np.ascontiguousarray(_np(v))
            if a.dtype.kind not in "biuf":
                continue
            h.update(json.dumps([k, a.dtype.str, list(a.shape)]).encode())
            h.update(a.tobytes())
    return h.hexdigest()


def configuration_digest(config: Mapping[str, Any] | None) -> str:
    items = {k: v for k, v in (config or {}).items() if isinstance(v, (bool, int, float, str))}
    return hashlib.sha256(json.dumps(items, sort_keys=True, allow_nan=False).encode()).hexdigest()


def _mpas_time(s: str) -> str:
    try:
        t = _dt.datetime.strptime(s.strip(), "%Y-%m-%d_%H:%M:%S")
    except ValueError as exc:
        raise MpasConfigError(f"time {s!r} is not YYYY-MM-DD_hh:mm:ss") from exc
    return t.strftime("%Y-%m-%d_%H:%M:%S")


# ---------------------------------------------------------------------------
# mesh and prepared state
# ---------------------------------------------------------------------------


def _generic_arrays(mapping: Mapping[str, Any]) -> tuple[dict, dict]:
    arrays, dims = {}, {}
    for k in sorted(mapping):
        v = mapping[k]
        if isinstance(v, bool) or not isinstance(v, (np.ndarray, torch.Tensor, int, float, np.generic)):
            continue
        a = np.ascontiguousarray(_np(v))
        if a.dtype.kind not in "biuf":
            continue
        names = tuple(f"{k}__d{i}" for i in range(a.ndim))
        dims.update(zip(names, a.shape))
        arrays[k] = (a, names)
    return arrays, dims


def write_mesh(path: str | Path, mesh: Mapping[str, Any], clobber: bool = False) -> str:
    """Write a mesh dict (reference padded layout); returns its mesh digest."""
    arrays, dims = _generic_arrays(mesh)
    digest = mesh_digest(mesh)
    write_container(path, "mpas_hip.mesh", arrays, dims, {"mesh_sha256": digest}, clobber)
    return digest


def read_mesh(path: str | Path) -> dict[str, Any]:
    arrays, header = read_container(path)
    if header["schema"] != "mpas_hip.mesh":
        raise MpasConfigError(f"{path}: schema {header['schema']!r} is not a mesh")
    mesh = {k: (a.item() if a.ndim == 0 else a) for k, a in arrays.items()}
    if mesh_digest(mesh) != header["attributes"]["mesh_sha256"]:
        raise MpasConfigError(f"{path}: mesh identity mismatch")
    return mesh


def write_state(path: str | Path, state: Mapping[str, Any], mesh: Mapping[str, Any] | None = None,
                clobber: bool = False) -> None:
    """Write a (prepared or initial) state dict; with ``mesh`` the file carries its identity."""
    arrays, dims = _generic_arrays(state)
    attrs = {"mesh_sha256": mesh_digest(mesh)} if mesh is not None else {}
    write_container(path, "mpas_hip.state", arrays, dims, attrs, clobber)


def read_state(path: str | Path, mesh: Mapping[str, Any] | None = None) -> dict[str, Any]:
    arrays, header = read_container(path)
    if header["schema"] != "mpas_hip.state":
        raise MpasConfigError(f"{path}: schema {header['schema']!r} is not a state")
    want = header["attributes"].get("mesh_sha256")
    if mesh is not None and want is not None and want != mesh_digest(mesh):
        raise MpasConfigError(f"{path}: written for another mesh")
    return {k: (a.item() if a.ndim == 0 else torch.from_numpy(a) if a.dtype.kind == "f" else a)
            for k, a in arrays.items()}


# ---------------------------------------------------------------------------
# restart
# ---------------------------------------------------------------------------
# restart fields by location (restart.py:43-145): the step's inter-step state
RESTART_REQUIRED = ("u", "w", "rho_zz", "theta_m", "scalars", "pressure_p", "exner", "ru", "rw", "rho_p",
                    "rtheta_p")
RESTART_OPTIONAL = ("pressure_base", "rtheta_base", "exner_base", "rho_base", "theta_base", "rt_diabatic_tend",
                    "v", "vorticity", "divergence", "ke", "pv_edge", "rho_edge", "ur_cell", "vr_cell",
                    "uReconstructX", "uReconstructY", "uReconstructZ", "uReconstructZonal",
                    "uReconstructMeridional", "ruAvg", "wwAvg")
EDGE_FIELDS = frozenset({"u", "ru", "v", "rho_edge", "pv_edge", "ruAvg"})
VERTEX_FIELDS = frozenset({"vorticity"})
W_FIELDS = frozenset({"w", "rw", "wwAvg"})
SCALAR_FIELDS = frozenset({"scalars"})


def _counts(mesh: Mapping[str, Any], state: Mapping[str, Any]) -> dict[str, int]:
    get = lambda k: int(_np(mesh[k] if k in mesh else state[k]))  # noqa: E731
    return {"nCells": get("nCells"), "nEdges": get("nEdges"), "nVertices": get("nVertices"),
            "nVertLevels": get("nVertLevels")}


def _location(name: str) -> tuple[str, tuple[str, ...]]:
    if name in EDGE_FIELDS:
        return "nEdges", ("nEdges", "nVertLevels")
    if name in VERTEX_FIELDS:
        return "nVertices", ("nVertices", "nVertLevels")
    if name in W_FIELDS:
        return "nCells", ("nCells", "nVertLevelsP1")
    if name in SCALAR_FIELDS:
        return "nCells", ("nCells", "nVertLevels", "nScalars")
    return "nCells", ("nCells", "nVertLevels")


def write_restart(state: Mapping[str, Any], mesh: Mapping[str, Any], path: str | Path, xtime: str, *,
                  initial_time: str | None = None, config: Mapping[str, Any] | None = None,
                  dt: float | None = None, model_time_seconds: float = 0.0,
                  source_provenance: str = "caller-provided in-memory state", clobber: bool = False) -> None:
    """Restart of the inter-step state (reference ``write_restart``, restart.py:383): active rows
    only, with the mesh, configuration and timestep identities of an exact continuation."""
    missing = [k for k in RESTART_REQUIRED if not isinstance(state.get(k), (torch.Tensor, np.ndarray))]
    if missing:
        raise MpasMissingFieldError(fields=missing, source="restart state")
    if not np.isfinite(model_time_seconds) or model_time_seconds < 0.0:
        raise MpasConfigError("model_time_seconds must be finite and nonnegative")
    timestep = float(dt if dt is not None else (config or {}).get("config_dt", 0.0))
    if not np.isfinite(timestep) or timestep < 0.0:
        raise MpasConfigError("restart timestep must be finite and nonnegative")
    n = _counts(mesh, state)
    S = int(_np(state["scalars"]).shape[-1])
    dims = dict(n, nVertLevelsP1=n["nVertLevels"] + 1, nScalars=S)
    arrays = {}
    for k in RESTART_REQUIRED + RESTART_OPTIONAL:
        v = state.get(k, mesh.get(k))
        if not isinstance(v, (torch.Tensor, np.ndarray)):
            continue
        rows, d = _location(k)
        a = np.asarray(_np(v), dtype=np.float64)[: n[rows]]
        arrays[k] = (a, d)
    attrs = {"xtime": _mpas_time(xtime), "initial_time": _mpas_time(initial_time or xtime),
             "model_time_seconds": float(model_time_seconds), "timestep_seconds": timestep,
             "mesh_sha256": mesh_digest(mesh), "configuration_sha256": configuration_digest(config),
             "source_provenance": source_provenance, "runtime_dtype": "float64"}
    write_container(path, "mpas_hip.restart", arrays, dims, attrs, clobber)


def read_restart(path: str | Path, *, mesh: Mapping[str, Any] | None = None,
                 config: Mapping[str, Any] | None = None, dt: float | None = None,
                 device: str | torch.device = "cpu") -> tuple[dict[str, Any], str]:
    """(state, valid time) of a restart (reference ``read_restart``); padding rows restored.

    With ``mesh`` / ``config`` / ``dt`` the file's identities must match them."""
    if torch.device(device).type != "cpu":
        raise MpasDeviceError(f"read_restart loads to the CPU, not {device}", requested=str(device))
    arrays, header = read_container(path)
    if header["schema"] != "mpas_hip.restart":
        raise MpasConfigError(f"{path}: schema {header['schema']!r} is not a restart")
    a = header["attributes"]
    if mesh is not None and a["mesh_sha256"] != mesh_digest(mesh):
        raise MpasConfigError(f"{path}: restart written for another mesh")
    if config is not None and a["configuration_sha256"] != configuration_digest(config):
        raise MpasConfigError(f"{path}: restart written with another configuration")
    if dt is not None and float(dt) != float(a["timestep_seconds"]):
        raise MpasConfigError(f"{path}: restart timestep {a['timestep_seconds']} differs from dt={dt}")
    d = header["dimensions"]
    state: dict[str, Any] = {}
    for k, v in arrays.items():
        pad = np.zeros((1,) + v.shape[1:], dtype=v.dtype)  # zero padding row (restart.py:612-650)
        state[k] = torch.from_numpy(np.concatenate([v, pad], axis=0))
    for k in ("nCells", "nEdges", "nVertices", "nVertLevels"):
        state[k] = int(d[k])
    state["_restart_model_time_seconds"] = float(a["model_time_seconds"])
    return state, a["xtime"]


# ---------------------------------------------------------------------------
# history
# ---------------------------------------------------------------------------
HISTORY_FIELDS = {
    "u": (("Time", "nEdges", "nVertLevels"), "m s-1"),
    "w": (("Time", "nCells", "nVertLevelsP1"), "m s-1"),
    "theta": (("Time", "nCells", "nVertLevels"), "K"),
    "rho": (("Time", "nCells", "nVertLevels"), "kg m-3"),
    "zz": (("Time", "nCells", "nVertLevels"), "1"),
    "qv": (("Time", "nCells", "nVertLevels"), "kg kg-1"),
    "pressure": (("Time", "nCells", "nVertLevels"), "Pa"),
    "divergence": (("Time", "nCells", "nVertLevels"), "s-1"),
    "vorticity": (("Time", "nVertices", "nVertLevels"), "s-1"),
    "ke": (("Time", "nCells", "nVertLevels"), "m2 s-2"),
    "uReconstructZonal": (("Time", "nCells", "nVertLevels"), "m s-1"),
    "uReconstructMeridional": (("Time", "nCells", "nVertLevels"), "m s-1"),
}


def history_fields(state: Mapping[str, Any], mesh: Mapping[str, Any],
                   fields: list[str] | None = None) -> dict[str, torch.Tensor]:
    """The history fields (active rows) of a state (output_writer.py:209-330).

    theta = theta_m / (1 + rvord qv), rho = rho_zz zz and pressure =
    pressure_p + pressure_base are derived on the GPU by
    ``mpas_hip_history_fields`` when ``theta_m`` is a device tensor."""
    combined = {**mesh, **state}
    n = _counts(mesh, state)
    nc, ne, nv, L = n["nCells"], n["nEdges"], n["nVertices"], n["nVertLevels"]
    want = list(HISTORY_FIELDS) if fields is None else list(fields)
    unknown = [f for f in want if f not in HISTORY_FIELDS]
    if unknown:
        raise MpasConfigError(f"Unsupported history fields: {unknown}")
    required = {"u": ("u",), "w": ("w",), "theta": ("theta_m", "scalars"), "rho": ("rho_zz", "zz"),
                "zz": ("zz",), "qv": ("scalars",)}
    missing = sorted({k for f in want for k in required.get(f, ()) if k not in combined})
    if missing:
        raise MpasMissingFieldError(fields=missing, source="history state")
    t = lambda k: torch.as_tensor(_np(combined[k])) if not isinstance(combined[k], torch.Tensor) else combined[k]  # noqa: E731
    out: dict[str, torch.Tensor] = {}
    theta_m = t("theta_m") if "theta_m" in combined else None
    on_gpu = theta_m is not None and theta_m.is_cuda
    derived = {f for f in ("theta", "rho", "pressure") if f in want}
    if "pressure" in derived and not ("pressure_p" in combined and "pressure_base" in combined):
        derived.discard("pressure")
        if fields is not None:
            raise MpasMissingFieldError(fields=["pressure_p", "pressure_base"], source="history state")
    if derived and on_gpu:
        out.update(_derive_on_gpu(combined, t, nc, L, derived))
    elif derived:
        qv = t("scalars")[:nc, :L, 0].double()
        if "theta" in derived:
            out["theta"] = t("theta_m")[:nc, :L].double() / (1.0 + RVORD * qv)
        if "rho" in derived:
            out["rho"] = t("rho_zz")[:nc, :L].double() * t("zz")[:nc, :L].double()
        if "pressure" in derived:
            out["pressure"] = t("pressure_p")[:nc, :L].double() + t("pressure_base")[:nc, :L].double()
    rows = {"u": ne, "vorticity": nv}
    aliases = {"uReconstructZonal": "ur_cell", "uReconstructMeridional": "vr_cell"}
    for f in want:
        if f in out or f in derived:
            continue
        if f == "qv":
            out[f] = t("scalars")[:nc, :L, 0].double()
            continue
        key = f if f in combined else aliases.get(f)
        if key is None or key not in combined:
            if fields is not None:
                raise MpasMissingFieldError(fields=[f], source="history state")
            continue
        width = L + 1 if f == "w" else L
        out[f] = t(key)[: rows.get(f, nc), :width].double()
    return out


def _derive_on_gpu(combined, t, nc: int, L: int, derived: set[str]) -> dict[str, torch.Tensor]:
    from . import _native as N
    from .device import stream_handle

    lib = N.load()
    theta_m = t("theta_m").double().contiguous()
    dev = theta_m.device
    g = lambda k: t(k).to(device=dev, dtype=torch.float64).contiguous() if k in combined else None  # noqa: E731
    scalars, rho_zz, zz = g("scalars"), g("rho_zz"), g("zz")
    pp, pb = g("pressure_p"), g("pressure_base")
    S = int(scalars.shape[-1]) if scalars is not None else 0
    outs = {f: torch.empty((nc, L), dtype=torch.float64, device=dev) for f in ("theta", "rho", "pressure")}
    ptr = lambda x: C.c_void_p(x.data_ptr() if x is not None else 0)  # noqa: E731
    with torch.cuda.device(dev):
        N.check(lib.mpas_hip_history_fields(nc, L, S, ptr(theta_m), ptr(scalars), ptr(rho_zz), ptr(zz), ptr(pp),
                                            ptr(pb), ptr(outs["theta"]), ptr(outs["rho"]),
                                            ptr(outs["pressure"] if pp is not None and pb is not None else None),
                                            C.c_void_p(stream_handle(dev))), "history_fields")
    return {f: outs[f] for f in derived}


def write_history(filepath: str | Path, state: Mapping[str, Any], mesh: Mapping[str, Any], *, xtime: str = "",
                  global_attrs: Mapping[str, Any] | None = None, fields: list[str] | None = None,
                  config: Mapping[str, Any] | None = None, dt: float | None = None,
                  model_time_seconds: float = 0.0, initial_time: str | None = None,
                  source_provenance: str = "caller-provided in-memory state", clobber: bool = False) -> None:
    """One history record of the selected fields (reference ``write_history``, output_writer.py:329)."""
    vals = history_fields(state, mesh, fields)
    n = _counts(mesh, state)
    dims = dict(n, nVertLevelsP1=n["nVertLevels"] + 1, Time=1)
    valid = _mpas_time(xtime or "1970-01-01_00:00:00")
    if not np.isfinite(model_time_seconds) or model_time_seconds < 0.0:
        raise MpasConfigError("model_time_seconds must be finite and nonnegative")
    arrays = {f: (v.detach().cpu().numpy()[None], HISTORY_FIELDS[f][0]) for f, v in vals.items()}
    start = _dt.datetime(1970, 1, 1)
    t_valid = _dt.datetime.strptime(valid, "%Y-%m-%d_%H:%M:%S")
    arrays["time"] = (np.array([(t_valid - start).total_seconds()], dtype=np.float64), ("Time",))
    attrs = dict(global_attrs or {})
    attrs.update({"xtime": valid, "initial_time": _mpas_time(initial_time or valid),
                  "model_time_seconds": float(model_time_seconds), "time_units": "seconds since 1970-01-01 00:00:00",
                  "calendar": "gregorian", "units": {f: HISTORY_FIELDS[f][1] for f in vals},
                  "mesh_sha256": mesh_digest(mesh), "configuration_sha256": configuration_digest(config),
                  "timestep_seconds": float(dt) if dt is not None else None,
                  "source_provenance": source_provenance})
    write_container(filepath, "mpas_hip.history", arrays, dims, attrs, clobber)


def read_history(filepath: str | Path, *, validate: bool = True) -> dict[str, Any]:
    """``{"fields", "dimensions", "attributes"}`` of a history file (reference ``read_history``)."""
    arrays, header = read_container(filepath, validate)
    if header["schema"] != "mpas_hip.history":
        raise MpasConfigError(f"{filepath}: schema {header['schema']!r} is not a history")
    dims = dict(header["dimensions"])
    dims["Time"] = None  # the record dimension
    return {"fields": arrays, "dimensions": dims, "attributes": header["attributes"]}


def _main(argv: list[str] | None = None) -> int:
    """``python -m mpas_dycore_pytorch_amd.io mesh --level 5 --levels 55 -o mesh.mpb``: write a
    synthetic icosahedral mesh (the reusable generator) to a container."""
    import argparse

    from .mesh import build_mesh

    ap = argparse.ArgumentParser(description=_main.__doc__)
    sub = ap.add_subparsers(dest="cmd", required=True)
    pm = sub.add_parser("mesh")
    pm.add_argument("--level", type=int, required=True)
    pm.add_argument("--levels", type=int, default=55)
    pm.add_argument("--stretch", default="quadratic", choices=("uniform", "quadratic"))
    pm.add_argument("-o", "--output", required=True)
    pm.add_argument("--clobber", action="store_true")
    args = ap.parse_args(argv)
    m = build_mesh(args.level, args.levels, stretch=args.stretch)
    digest = write_mesh(args.output, m, clobber=args.clobber)
    print(json.dumps({"path": args.output, "nCells": int(m["nCells"]), "mesh_sha256": digest}))
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
