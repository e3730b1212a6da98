"""Batched ensembles: many states, possibly on meshes of different sizes, one launch per kernel.

Members that share a mesh and a time step form a group.  A group's members
are laid side by side as one disjoint-union mesh: cells, edges and vertices of
member ``m`` follow those of member ``m - 1``, connectivity is offset
accordingly, and the union keeps one trailing padding row per entity.  Every
kernel of the large step then advances all members of the group in one launch
(one persistent grid per stage over the union), and a group's large step can
be captured once in a HIP graph.  Groups run on their own streams, so members
of different sizes and time steps (an uneven ensemble) overlap on the GPU.

Each member is put in space-filling-curve order before the union is formed, so
that the neighbours of a cell stay close in memory; results are returned in
the member's own order.  A member's results are bitwise those of stepping it
alone: the union only relabels entities and never mixes members.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Mapping, Sequence

import numpy as np
import torch

from .device import to_numpy
from .errors import MpasConfigError
from .time_integration import Srk3Stepper

# 1-based connectivity tables and the entity their values index
CONNECTIVITY = {
    "cellsOnEdge": "cell", "verticesOnEdge": "vertex", "edgesOnCell": "edge", "cellsOnCell": "cell",
    "verticesOnCell": "vertex", "edgesOnVertex": "edge", "edgesOnEdge": "edge", "advCellsForEdge": "cell",
    "cellsOnVertex": "cell",
}
COUNT_KEYS = {"nCells": "cell", "nEdges": "edge", "nVertices": "vertex"}


def _dims(mesh: Mapping[str, Any]) -> dict[str, int]:
    return {"cell": int(to_numpy(mesh["nCells"])), "edge": int(to_numpy(mesh["nEdges"])),
            "vertex": int(to_numpy(mesh["nVertices"]))}


def _row_kind(a: np.ndarray, dims: Mapping[str, int]) -> str | None:
    if a.ndim == 0:
        return None
    for kind, n in dims.items():
        if a.shape[0] == n + 1:
            return kind
    return None


def union(mappings: Sequence[Mapping[str, Any]], meshes: Sequence[Mapping[str, Any]]) -> dict[str, Any]:
    """Disjoint union of per-member mappings (mesh or state) over their meshes.

    Arrays with one row per cell, edge or vertex (plus the padding row) are
    concatenated over their live rows and closed by the first member's padding
    row; connectivity values are offset; entity counts are summed; every other
    entry (vertical metrics, scalars) must agree across members and is shared."""
    dims = [_dims(m) for m in meshes]
    off = {k: np.cumsum([0] + [d[k] for d in dims[:-1]]) for k in ("cell", "edge", "vertex")}
    out: dict[str, Any] = {}
    keys = list(mappings[0].keys())
    for key in keys:
        vals = [mp[key] for mp in mappings]
        if key in COUNT_KEYS:
            out[key] = int(sum(d[COUNT_KEYS[key]] for d in dims))
            continue
        if isinstance(vals[0], torch.Tensor) or isinstance(vals[0], np.ndarray):
            arrs = [to_numpy(v) for v in vals]
            kinds = {_row_kind(a, d) for a, d in zip(arrs, dims)}
            if len(kinds) == 1 and None not in kinds:
                kind = kinds.pop()
                parts = []
                for i, (a, d) in enumerate(zip(arrs, dims)):
                    live = np.array(a[: d[kind]], copy=True)
                    if key in CONNECTIVITY:
                        live = np.where(live > 0, live + off[CONNECTIVITY[key]][i], live)
                    parts.append(live)
                parts.append(arrs[0][dims[0][kind]:])
                out[key] = np.concatenate(parts, axis=0)
                continue
            if not all(a.shape == arrs[0].shape and np.array_equal(a, arrs[0]) for a in arrs[1:]):
                raise MpasConfigError(f"ensemble members disagree on the shared array {key!r}")
            out[key] = arrs[0]
            continue
        if all(_same_scalar(v, vals[0]) for v in vals[1:]):
            out[key] = vals[0]
        else:
            raise MpasConfigError(f"ensemble members disagree on {key!r}")
    return out


def _same_scalar(a: Any, b: Any) -> bool:
    try:
        return bool(np.all(np.asarray(a) == np.asarray(b)))
    except Exception:  # pragma: no cover - non-comparable objects
        return a is b


def split(values: Mapping[str, torch.Tensor], meshes: Sequence[Mapping[str, Any]]) -> list[dict[str, torch.Tensor]]:
    """Per-member views (padded like the member's own arrays) of union-mesh arrays."""
    dims = [_dims(m) for m in meshes]
    total = {k: sum(d[k] for d in dims) for k in ("cell", "edge", "vertex")}
    outs: list[dict[str, torch.Tensor]] = [dict() for _ in meshes]
    for key, t in values.items():
        kind = None
        for k, n in total.items():
            if t.ndim >= 1 and t.shape[0] == n + 1:
                kind = k
        if kind is None:
            for o in outs:
                o[key] = t
            continue
        start = 0
        pad = t[total[kind]:]
        for i, d in enumerate(dims):
            n = d[kind]
            outs[i][key] = torch.cat([t[start:start + n], pad], dim=0)
            start += n
    return outs


@dataclass
class _Group:
    members: list[int]
    meshes: list[Any]
    perms: list[Any]
    stepper: Srk3Stepper
    stream: torch.cuda.Stream
    dt: float


class EnsembleStepper:
    """Advance an ensemble; members sharing (mesh, dt) run as one batched grid.

    ``members`` is a sequence of ``(state, mesh, dt)`` or ``(state, mesh, dt,
    config)``; a member without its own config uses ``config``.  Members that
    share the mesh object, dt and config form one batched group.  ``step(n)``
    advances every member ``n`` large steps (each with its own dt);
    ``states()`` returns the members' state dicts (device tensors) in the
    caller's member and entity order."""

    def __init__(self, members: Sequence[tuple], config: Mapping[str, Any] | None = None,
                 device: torch.device | str | None = None, reorder: bool = True):
        from .reorder import permute_mapping
        from .time_integration import _reordered_mesh

        self.n_members = len(members)
        groups: dict[tuple[int, float, int], list[int]] = {}
        cfgs: dict[int, Mapping[str, Any] | None] = {}
        for i, mem in enumerate(members):
            mesh, dt = mem[1], mem[2]
            cfg = mem[3] if len(mem) > 3 else config
            cfgs[id(cfg)] = cfg
            groups.setdefault((id(mesh), float(dt), id(cfg)), []).append(i)
        self.groups: list[_Group] = []
        for (_, dt, cfg_id), idx in groups.items():
            mesh = members[idx[0]][1]
            config_g = cfgs[cfg_id]
            if reorder:
                perm, mesh_p = _reordered_mesh(mesh)
                dims = tuple(int(to_numpy(mesh[k])) for k in ("nCells", "nEdges", "nVertices"))
                states = [permute_mapping(members[i][0], perm, dims) for i in idx]
            else:
                perm, mesh_p = None, mesh
                states = [members[i][0] for i in idx]
            meshes = [mesh_p] * len(idx)
            umesh = union(meshes, meshes)
            ustate = union(states, meshes)
            stepper = Srk3Stepper(ustate, umesh, config_g, dt, device=device, reorder=False)
            stream = torch.cuda.Stream(stepper.device)
            self.groups.append(_Group(idx, meshes, [perm] * len(idx), stepper, stream, dt))
        self.device = self.groups[0].stepper.device

    @property
    def cell_levels(self) -> int:
        """Cell-levels advanced per large step of every member."""
        return sum(g.stepper.mesh.nc * g.stepper.mesh.L for g in self.groups)

    def capture_graphs(self) -> None:
        for g in self.groups:
            with torch.cuda.stream(g.stream):
                g.stepper.capture_graph()

    def step(self, n: int = 1) -> None:
        """Advance all members ``n`` large steps; groups overlap on their streams."""
        cur = torch.cuda.current_stream(self.device)
        for g in self.groups:
            g.stream.wait_stream(cur)
            with torch.cuda.stream(g.stream):
                g.stepper.step(n)
        for g in self.groups:
            cur.wait_stream(g.stream)

    def states(self) -> list[dict[str, torch.Tensor]]:
        out: list[dict[str, torch.Tensor] | None] = [None] * self.n_members
        for g in self.groups:
            parts = split(g.stepper.state_dict(), g.meshes)
            rows = None
            perm = g.perms[0]
            if perm is not None:
                d = _dims(g.meshes[0])
                rows = {
                    d["cell"] + 1: torch.as_tensor(np.concatenate([perm.cell_old_to_new, [d["cell"]]]),
                                                   device=self.device),
                    d["edge"] + 1: torch.as_tensor(np.concatenate([perm.edge_old_to_new, [d["edge"]]]),
                                                   device=self.device),
                    d["vertex"] + 1: torch.as_tensor(np.concatenate([perm.vertex_old_to_new, [d["vertex"]]]),
                                                     device=self.device),
                }
            for i, st in zip(g.members, parts):
                if rows is not None:
                    st = {k: (v.index_select(0, rows[v.shape[0]]) if v.ndim >= 1 and v.shape[0] in rows else v)
                          for k, v in st.items()}
                out[i] = st
        return out  # type: ignore[return-value]
