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

``ensemble_vjp`` differentiates a sum of per-member objectives through the
same unions: every large step of a batch is one differentiable step over the
union mesh (one launch per kernel for all its members, one checkpoint per
step), and one reverse sweep returns every member's gradient.  Members are
disjoint, so the gradient of the sum with respect to member m is the gradient
of member m's own objective.
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


def _rows_kind(n_rows: int, dims: Mapping[str, int]) -> str | None:
    """The entity kind whose count plus the padding row is ``n_rows``, if any."""
    for kind, n in dims.items():
        if n_rows == n + 1:
            return kind
    return None


def _row_kind(a: np.ndarray, dims: Mapping[str, int]) -> str | None:
    return _rows_kind(int(a.shape[0]), dims) if a.ndim else None


def union(mappings: Sequence[Mapping[str, Any]], meshes: Sequence[Mapping[str, Any]]) -> dict[str, Any]:
    """Disjoint union of per-member mappings (mesh or state) over their meshes.

    Arrays with one row per cell, edge or vertex (plus the padding row) are
    concatenated over their live rows and closed by the first member's padding
    row; connectivity values are offset; entity counts are summed; every other
    entry (vertical metrics, scalars) must agree across members and is shared."""
    from .boundary import BOUNDARY_KEY

    dims = [_dims(m) for m in meshes]
    off = {k: np.cumsum([0] + [d[k] for d in dims[:-1]]) for k in ("cell", "edge", "vertex")}
    out: dict[str, Any] = {}
    keys = [k for k in mappings[0].keys() if k != BOUNDARY_KEY]  # receipts are per member
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


def _union_state(states: Sequence[Mapping[str, Any]], dims: Sequence[Mapping[str, int]],
                 device: torch.device) -> dict[str, Any]:
    """Disjoint union of member states on the device: every floating-point
    array with one row per cell, edge or vertex (plus the padding row) is
    concatenated over its live rows and closed by the first member's padding
    row; arrays without such rows and scalars are shared (taken from the first
    member); integer arrays (connectivity, which the union mesh carries) and
    the members' boundary receipts are left out."""
    from .boundary import BOUNDARY_KEY

    out: dict[str, Any] = {}
    for key, v0 in states[0].items():
        if key == BOUNDARY_KEY:
            continue
        if isinstance(v0, (torch.Tensor, np.ndarray)):
            if not (v0.dtype.is_floating_point if isinstance(v0, torch.Tensor) else v0.dtype.kind == "f"):
                continue
            kind = _rows_kind(int(v0.shape[0]), dims[0]) if v0.ndim else None
            if kind is None:
                out[key] = v0
                continue
            ts = [torch.as_tensor(st[key], device=device) for st in states]
            for t, d in zip(ts, dims):
                if t.shape[0] != d[kind] + 1:
                    raise MpasConfigError(f"ensemble member field {key!r} has {t.shape[0]} rows, expected {d[kind] + 1}")
            out[key] = torch.cat([t[: d[kind]] for t, d in zip(ts, dims)] + [ts[0][dims[0][kind]:]], dim=0)
            continue
        out[key] = v0
    return out


_UNION_MESHES: dict[tuple[int, int], tuple[Any, dict[str, Any]]] = {}


def _union_mesh(mesh: Mapping[str, Any], count: int) -> dict[str, Any]:
    """The disjoint union of ``count`` copies of ``mesh`` (cached per mesh object and count)."""
    hit = _UNION_MESHES.get((id(mesh), count))
    if hit is not None and hit[0] is mesh:
        return hit[1]
    um = union([mesh] * count, [mesh] * count)
    if len(_UNION_MESHES) > 8:
        _UNION_MESHES.clear()
    _UNION_MESHES[(id(mesh), count)] = (mesh, um)
    return um


def ensemble_vjp(members: Sequence[tuple], objective, nsteps: int = 1,
                 track_keys: Sequence[str] = ("u", "w", "theta_m", "rho_zz"),
                 config: Mapping[str, Any] | None = None, device: torch.device | str | None = None,
                 max_batch_cells: int | None = None) -> list[dict[str, torch.Tensor]]:
    """Per-member gradients of ``sum_m objective(m, state_m)``, ``state_m`` the
    member's state after ``nsteps`` large steps.

    ``members`` is a sequence of ``(state, mesh, dt)`` or ``(state, mesh, dt,
    config)`` as for :class:`EnsembleStepper`; each state is a complete step
    boundary (a state ``run_simulation`` returned).  Members sharing the mesh
    object, dt and config are batched: each batch (at most ``max_batch_cells``
    cells, by default 200000, which bounds the one recorded step the reverse
    sweep holds) advances as one differentiable run over the union mesh, and
    one reverse sweep gives all its members' gradients.  ``objective(m,
    state)`` receives member ``m``'s output state in the member's own layout
    and returns a scalar tensor.  Returns, per member, the gradient with
    respect to each entry of ``track_keys`` (zero where the objective does not
    depend on it).  A member's gradient equals the one ``run_simulation.vjp``
    gives for the member alone bit for bit (the union keeps each member's
    entity order, so every gather sums in the same order)."""
    from .device import require_gpu
    from .time_integration import _atm_srk3_differentiable

    gpu = require_gpu(device)
    budget = int(max_batch_cells) if max_batch_cells else 200_000
    groups: dict[tuple[int, float, int], list[int]] = {}
    cfgs: dict[int, Mapping[str, Any] | None] = {}
    for i, mem in enumerate(members):
        cfg = mem[3] if len(mem) > 3 else config
        cfgs[id(cfg)] = cfg
        groups.setdefault((id(mem[1]), float(mem[2]), id(cfg)), []).append(i)
    grads: list[dict[str, torch.Tensor] | None] = [None] * len(members)
    for (_, dt, cfg_id), idx in groups.items():
        mesh = members[idx[0]][1]
        d = _dims(mesh)
        per = max(1, budget // max(d["cell"], 1))
        for b0 in range(0, len(idx), per):
            batch = idx[b0:b0 + per]
            meshes = [mesh] * len(batch)
            umesh = mesh if len(batch) == 1 else _union_mesh(mesh, len(batch))
            ustate = _union_state([members[i][0] for i in batch], [d] * len(batch), gpu)
            leaves = {}
            for k in track_keys:
                if not isinstance(ustate.get(k), torch.Tensor):
                    raise MpasConfigError(f"ensemble_vjp: no tensor state entry {k!r} to differentiate")
                leaves[k] = ustate[k].detach().clone().requires_grad_(True)
            cur = dict(ustate)
            cur.update(leaves)
            for step in range(1, nsteps + 1):
                cur = _atm_srk3_differentiable(cur, umesh, cfgs[cfg_id], dt, strict=True, keep_tape=step == nsteps,
                                               rebind_base=step == 1)
            outs = split({k: v for k, v in cur.items() if isinstance(v, torch.Tensor)}, meshes)
            loss = None
            for i, out in zip(batch, outs):
                term = objective(i, out)
                loss = term if loss is None else loss + term
            found = torch.autograd.grad(loss, tuple(leaves.values()), allow_unused=True)
            gfull = {k: (g if g is not None else torch.zeros_like(leaves[k])) for k, g in zip(leaves, found)}
            for i, part in zip(batch, split(gfull, meshes)):
                grads[i] = {k: v.contiguous() for k, v in part.items()}
            del cur, outs, loss, found, gfull, leaves, ustate
    return grads  # type: ignore[return-value]
