"""Device-resident mesh and state for the gfx950 step.

The public boundary keeps the reference layout (``mpas_pytorch`` docs/API.md
"State and controls"): float64 tensors with one trailing padding row per
entity, levels innermost, 1-based connectivity with 0 for "none".  On the
device the connectivity is int32 and 0-based (-1 for "none"); every other
array keeps the reference shape, so device results compare index for index.

PyTorch owns every device allocation (mesh arrays, state, the per-mesh
workspace); the HIP library only receives pointers.
"""

from __future__ import annotations

import ctypes as C
from typing import Any, Mapping

import numpy as np
import torch

from . import _native as N
from .errors import MpasDeviceError, MpasMeshMismatchError, MpasMissingFieldError

# 1-based connectivity -> 0-based; the entity each index points to
INDEX_TARGET = {
    "cellsOnEdge": "nCells", "verticesOnEdge": "nVertices", "edgesOnCell": "nEdges",
    "cellsOnCell": "nCells", "verticesOnCell": "nVertices", "kiteForCell": 3,
    "edgesOnVertex": "nEdges", "edgesOnEdge": "nEdges", "advCellsForEdge": "nCells",
}
COUNT_FIELDS = ("nEdgesOnCell", "nEdgesOnEdge", "nAdvCellsForEdge")
BASE_FIELDS = ("rho_base", "rtheta_base", "exner_base")
OMEGA = 7.29212e-5  # rad/s, the reference's rotation rate (constants.py:88)


def to_numpy(v: Any) -> np.ndarray:
    if isinstance(v, torch.Tensor):
        return v.detach().cpu().numpy()
    return np.asarray(v)


def require_gpu(device: torch.device | str | None = None) -> torch.device:
    if not torch.cuda.is_available():
        raise MpasDeviceError(requested=str(device or "cuda"))
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    if dev.type != "cuda":
        raise MpasDeviceError(f"the HIP step runs on a ROCm GPU, not {dev}", requested=str(dev))
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


def _get(mesh: Mapping, state: Mapping | None, name: str):
    if name in mesh:
        return mesh[name]
    if state is not None and name in state:
        return state[name]
    raise MpasMissingFieldError(fields=[name], source="mesh")


class DeviceMesh:
    """The mesh of one ``atm_srk3`` call resident on one GPU.

    Built once per (mesh mapping, device) and cached; connectivity is
    validated on the host before anything reaches a kernel so that every
    gather a kernel makes stays inside its array.
    """

    _cache: dict[tuple, "DeviceMesh"] = {}

    def __init__(self, mesh: Mapping, device: torch.device, state: Mapping | None = None):
        self.device = device
        nc, ne, nv = (int(to_numpy(mesh[k])) for k in ("nCells", "nEdges", "nVertices"))
        L = int(to_numpy(_get(mesh, state, "nVertLevels")))
        self.nc, self.ne, self.nv, self.L = nc, ne, nv, L
        if L < 3 or L > 63:
            raise MpasMeshMismatchError(f"nVertLevels={L}: the column kernels take 3..63 levels")
        self.t: dict[str, torch.Tensor] = {}
        s = N.MeshStruct()
        s.nCells, s.nEdges, s.nVertices, s.nLevels = nc, ne, nv, L
        eoc = to_numpy(mesh["edgesOnCell"])
        s.maxEdges = int(eoc.shape[1])
        s.maxEdges2 = int(to_numpy(mesh["edgesOnEdge"]).shape[1])
        s.vertexDegree = 3
        s.maxAdv = int(to_numpy(mesh["advCellsForEdge"]).shape[1])
        dims = {"nCells": nc, "nEdges": ne, "nVertices": nv}
        rows = {"cellsOnEdge": ne, "verticesOnEdge": ne, "edgesOnCell": nc, "cellsOnCell": nc,
                "verticesOnCell": nc, "kiteForCell": nc, "edgesOnVertex": nv, "edgesOnEdge": ne,
                "advCellsForEdge": ne, "nEdgesOnCell": nc, "nEdgesOnEdge": ne, "nAdvCellsForEdge": ne}
        counts = {}
        for name in COUNT_FIELDS:
            a = to_numpy(mesh[name]).astype(np.int64)
            if a.shape[0] < rows[name]:
                raise MpasMeshMismatchError(dim_name=name, init_value=a.shape[0], grid_value=rows[name])
            counts[name] = a[: rows[name]]
        width = {"edgesOnCell": "nEdgesOnCell", "cellsOnCell": "nEdgesOnCell", "verticesOnCell": "nEdgesOnCell",
                 "kiteForCell": "nEdgesOnCell", "edgesOnEdge": "nEdgesOnEdge",
                 "advCellsForEdge": "nAdvCellsForEdge"}
        for name in N.MESH_INDEX_FIELDS:
            a = to_numpy(mesh[name]).astype(np.int64)
            if name in INDEX_TARGET:
                a = a - 1
                tgt = INDEX_TARGET[name]
                hi = tgt if isinstance(tgt, int) else dims[tgt]
                live = a[: rows[name]]
                if live.ndim == 2 and name in width:
                    cnt = counts[width[name]]
                    if cnt.min() < 1 or cnt.max() > live.shape[1]:
                        raise MpasMeshMismatchError(f"{width[name]} outside 1..{live.shape[1]}")
                    used = np.arange(live.shape[1])[None, :] < cnt[:, None]
                    bad = used & ((live < 0) | (live >= hi))
                else:
                    bad = (live < 0) | (live >= hi)
                if bad.any():
                    raise MpasMeshMismatchError(f"{name}: connectivity value outside 1..{hi}")
            t = torch.as_tensor(np.ascontiguousarray(a.astype(np.int32)), device=device)
            self.t[name] = t
            setattr(s, name, t.data_ptr())
        flat = True
        for name in N.MESH_REAL_FIELDS_A:
            if name == "zxu" and "zxu" not in mesh:
                a = np.zeros((ne + 1, L))
            elif name in ("zb_cell", "zb3_cell") and name not in mesh:
                a = np.zeros((nc + 1, s.maxEdges, L + 1))
            else:
                a = to_numpy(_get(mesh, state, name)).astype(np.float64)
            if name in ("zxu", "zb_cell", "zb3_cell") and np.any(a != 0.0):
                flat = False
            t = torch.as_tensor(np.ascontiguousarray(a), device=device)
            self.t[name] = t
            setattr(s, name, t.data_ptr())
        s.flatTerrain = 1 if flat else 0
        for name in ("cf1", "cf2", "cf3"):
            setattr(s, name, float(to_numpy(_get(mesh, state, name))))
        # Coriolis factors of the curvature terms, formed as the reference forms them
        # ((2 omega cos angleEdge) cos latEdge; 2 omega cos latCell: dyn_tend.py:2494-2502,
        # :4258-4265); a mesh without angleEdge/latEdge cannot run with curvature on
        two_omega = 2.0 * OMEGA
        self.has_edge_angles = "angleEdge" in mesh and "latEdge" in mesh
        if self.has_edge_angles:
            ce = two_omega * np.cos(to_numpy(mesh["angleEdge"]).astype(np.float64)) * np.cos(
                to_numpy(mesh["latEdge"]).astype(np.float64))
        else:
            ce = np.zeros(ne + 1)
        cc = two_omega * np.cos(to_numpy(_get(mesh, state, "latCell")).astype(np.float64))
        # vertical mixing: interface heights, and cos/sin of angleEdge for the
        # departures of u from its reference profile
        zg = to_numpy(_get(mesh, state, "zgrid")).astype(np.float64) if ("zgrid" in mesh or (
            state is not None and "zgrid" in state)) else np.zeros((nc + 1, L + 1))
        ang = to_numpy(mesh["angleEdge"]).astype(np.float64) if "angleEdge" in mesh else np.zeros(ne + 1)
        for name, a in (("curvEdge", ce), ("curvCell", cc), ("zgrid", zg), ("cosAngleEdge", np.cos(ang)),
                        ("sinAngleEdge", np.sin(ang))):
            t = torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)
            self.t[name] = t
            setattr(s, name, t.data_ptr())
        self.s = s
        self._base_src: dict[str, Any] = {}
        self.set_base(mesh, state)
        self._work: dict[tuple[int, bool], torch.Tensor] = {}
        self._kscratch: torch.Tensor | None = None

    def set_base(self, mesh: Mapping, state: Mapping | None) -> None:
        """(Re)bind the hydrostatic base state and the reference profiles of the
        vertical mixing (``t_init``, ``u_init``, ``v_init``), looked up in the mesh
        then the state.

        ``rtheta_base`` and ``exner_base`` are outputs of the cold start; until
        it has run they are bound to zeros (``missing_base`` lists them)."""
        self.missing_base = []
        # reference profiles of the vertical mixing departures (zeros where absent)
        for name, shape in (("t_init", (self.nc + 1, self.L)), ("u_init", (self.L,)), ("v_init", (self.L,))):
            src = mesh.get(name, None) if name in mesh else (state or {}).get(name, None)
            a = np.zeros(shape) if src is None else to_numpy(src).astype(np.float64).reshape(shape)
            t = torch.as_tensor(np.ascontiguousarray(a), device=self.device)
            self.t[name] = t
            setattr(self.s, name, t.data_ptr())
        for name in BASE_FIELDS:
            try:
                src = _get(mesh, state, name)
            except MpasMissingFieldError:
                if name == "rho_base":
                    raise
                self.missing_base.append(name)
                if name not in self.t:
                    self.bind_base(name, torch.zeros((self.nc + 1, self.L), dtype=torch.float64, device=self.device))
                continue
            if name in self.t and self._same_source(name, src):
                continue
            if isinstance(src, torch.Tensor):
                t = src.detach().to(device=self.device, dtype=torch.float64).contiguous().clone()
            else:
                t = torch.as_tensor(np.ascontiguousarray(to_numpy(src).astype(np.float64)), device=self.device)
            self.bind_base(name, t)
            self._base_src[name] = self._source_key(src)

    @staticmethod
    def _source_key(src: Any):
        """What identifies a bound base-state source: the object itself (held,
        so its id cannot be reused) and, for a tensor, its version counter,
        which every in-place change advances.  Other sources (NumPy arrays)
        carry no version and are uploaded again at every binding."""
        if isinstance(src, torch.Tensor):
            return (src, src._version)
        return None

    def _same_source(self, name: str, src: Any) -> bool:
        key = self._base_src.get(name)
        return key is not None and key[0] is src and key[1] == src._version

    def bind_base(self, name: str, t: torch.Tensor) -> None:
        """Bind a device tensor as one of the base-state fields."""
        if tuple(t.shape) != (self.nc + 1, self.L):
            raise MpasMeshMismatchError(f"{name} has shape {tuple(t.shape)}, expected {(self.nc + 1, self.L)}")
        t = t.to(device=self.device, dtype=torch.float64).contiguous()
        self.t[name] = t
        self._base_src.pop(name, None)
        setattr(self.s, name, t.data_ptr())

    @classmethod
    def get(cls, mesh: Mapping, device: torch.device, state: Mapping | None = None,
            rebind_base: bool = True) -> "DeviceMesh":
        """The cached device mesh of ``mesh`` on ``device``.  A cached mesh
        binds its base state again from ``mesh``/``state`` unless
        ``rebind_base`` is False, which a caller passes only when it bound
        that same base state earlier in the same run (the large steps after
        the first of a differentiable run): a NumPy base field is uploaded
        again at every binding."""
        key = (id(mesh), str(device))
        dm = cls._cache.get(key)
        if dm is None or dm._mesh_ref is not mesh:
            dm = cls(mesh, device, state)
            dm._mesh_ref = mesh
            if len(cls._cache) > 8:
                cls._cache.clear()
            cls._cache[key] = dm
        elif rebind_base:
            dm.set_base(mesh, state)
        return dm

    _mesh_ref: Any = None

    @property
    def ref(self):
        return C.byref(self.s)

    # shapes of the state arrays
    def shape_of(self, name: str, n_scalars: int = 0) -> tuple[int, ...]:
        nc, ne, nv, L = self.nc, self.ne, self.nv, self.L
        if name in ("u", "ru", "v", "pv_edge", "rho_edge", "ruAvg", "h_edge"):
            return (ne + 1, L)
        if name in ("w", "rw", "wwAvg"):
            return (nc + 1, L + 1)
        if name == "vorticity":
            return (nv + 1, L)
        if name == "scalars":
            return (nc + 1, L, n_scalars)
        return (nc + 1, L)

    def workspace(self, n_scalars: int, mixed: bool = False) -> torch.Tensor:
        """Zero-initialised per-mesh workspace of one large step (kept for reuse);
        ``mixed`` sizes it for the mixed-precision step."""
        w = self._work.get((n_scalars, mixed))
        if w is None:
            if mixed:
                n = N.load_mixed().mpas_hip_mixed_workspace_doubles(self.nc, self.ne, self.nv, self.L, n_scalars)
            else:
                n = N.load().mpas_hip_workspace_doubles(self.nc, self.ne, self.nv, self.L, n_scalars)
            w = torch.zeros(int(n), dtype=torch.float64, device=self.device)
            self._work[(n_scalars, mixed)] = w
        return w

    def inverse_tables(self) -> "N.InverseTablesStruct":
        """CSR inverse connectivity of advCellsForEdge, edgesOnEdge and cellsOnCell (built once).

        For each cell, the (edge, slot) pairs of the advection stencils that
        name it and the (cell, slot) pairs of the neighbour lists that name it;
        for each edge, the (edge, slot) pairs of the PV-flux stencils that name
        it; all ordered by (source row, slot).  The adjoint kernels gather over
        them (each target sums its sources in that order) instead of
        scattering."""
        if self._inv is None:
            t = {}
            for prefix, table, count, n_rows, n_targets in (
                    ("adv", "advCellsForEdge", "nAdvCellsForEdge", self.ne, self.nc),
                    ("eoe", "edgesOnEdge", "nEdgesOnEdge", self.ne, self.ne),
                    ("coc", "cellsOnCell", "nEdgesOnCell", self.nc, self.nc)):
                tab = self.t[table][:n_rows].cpu().numpy()
                cnt = self.t[count][:n_rows].cpu().numpy().astype(np.int64)
                t[prefix] = build_inverse(tab, cnt, n_targets)
            dev = {k: [torch.as_tensor(a, device=self.device) for a in v] for k, v in t.items()}
            s = N.InverseTablesStruct(*[a.data_ptr() for k in ("adv", "eoe", "coc") for a in dev[k]])
            self._inv = (dev, s)
        return self._inv[1]

    _inv = None

    def kernel_scratch(self) -> torch.Tensor:
        if self._kscratch is None:
            n = N.load().mpas_hip_kernel_scratch_doubles(self.nc, self.ne, self.nv, self.L)
            self._kscratch = torch.zeros(int(n), dtype=torch.float64, device=self.device)
        return self._kscratch


def build_inverse(table: np.ndarray, counts: np.ndarray, n_targets: int):
    """CSR transpose of a slot table: (start, source row, slot) ordered by (source row, slot).

    ``table[r, j]`` (0-based, the first ``counts[r]`` slots live) names a
    target; ``start[t]:start[t+1]`` lists every (r, j) with ``table[r, j] == t``;
    entries naming no target in ``0..n_targets-1`` are left out."""
    rows, slots = np.nonzero(np.arange(table.shape[1])[None, :] < counts[:, None])
    targets = table[rows, slots].astype(np.int64)
    # entries outside 0..n_targets-1 (a padding row on a bounded mesh) name no target
    keep = (targets >= 0) & (targets < n_targets)
    rows, slots, targets = rows[keep], slots[keep], targets[keep]
    order = np.argsort(targets, kind="stable")  # (row, slot) order kept within a target
    start = np.zeros(n_targets + 1, dtype=np.int64)
    np.cumsum(np.bincount(targets, minlength=n_targets), out=start[1:])
    return (start.astype(np.int32), rows[order].astype(np.int32), slots[order].astype(np.int32))


def stream_handle(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream
