"""Space-filling-curve ordering of cells, edges and vertices.

Replaces the reference's reverse-Cuthill-McKee reordering
(``mpas_pytorch/mesh_reordering.py:243`` ``apply_rcm_reordering``; permutation
rules ``:59-103``) with an order chosen for the GPU:

* cells follow a 3-D Hilbert curve through their unit-sphere positions
  (Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 2004), so
  the columns a workgroup gathers from are close in memory and a contiguous
  block of cells is a compact patch of the sphere;
* edges are sorted by (min, max) of their two new cell ids, with the old id
  breaking ties, as the reference sorts them (``mesh_reordering.py:82-103``);
* vertices are sorted the same way by their three new cell ids.

Only rows are permuted and ids renumbered; the slot order inside every
connectivity row is kept and ``cellsOnEdge`` keeps its (cell1, cell2)
orientation.  Every slot-ordered sum therefore adds the same numbers in the
same order, and a step on the reordered mesh reproduces the unpermuted step
bit for bit after the inverse permutation.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Mapping

import numpy as np

# connectivity whose values are ids of the named entity (1-based, 0 = none)
VALUE_ENTITY = {
    "cellsOnEdge": "cell", "cellsOnCell": "cell", "advCellsForEdge": "cell", "cellsOnVertex": "cell",
    "edgesOnCell": "edge", "edgesOnEdge": "edge", "edgesOnVertex": "edge",
    "verticesOnCell": "vertex", "verticesOnEdge": "vertex",
}
# per-level 1-D vertical arrays (never permuted even if their length matched an entity count)
VERTICAL_1D = {"rdzw", "rdzu", "fzm", "fzp", "rdnw", "fnm", "fnp", "dzu", "dzw", "zw", "zu", "cofrz"}


def hilbert_index_3d(p: np.ndarray, bits: int = 16) -> np.ndarray:
    """Hilbert-curve index of points ``p`` (n, 3) with coordinates in [-1, 1]."""
    q = np.floor((np.clip(p, -1.0, 1.0) + 1.0) * 0.5 * ((1 << bits) - 1) + 0.5).astype(np.uint64)
    X = [q[:, 0].copy(), q[:, 1].copy(), q[:, 2].copy()]
    M = np.uint64(1 << (bits - 1))
    # inverse undo of excess work
    Q = M
    while Q > 1:
        P = Q - np.uint64(1)
        for i in range(3):
            hit = (X[i] & Q) != 0
            if i == 0:
                X[0] = np.where(hit, X[0] ^ P, X[0])
            else:
                t = (X[0] ^ X[i]) & P
                X0 = np.where(hit, X[0] ^ P, X[0] ^ t)
                X[i] = np.where(hit, X[i], X[i] ^ t)
                X[0] = X0
        Q = Q >> np.uint64(1)
    # Gray encode
    for i in range(1, 3):
        X[i] = X[i] ^ X[i - 1]
    t = np.zeros_like(X[0])
    Q = M
    while Q > 1:
        t = np.where((X[2] & Q) != 0, t ^ (Q - np.uint64(1)), t)
        Q = Q >> np.uint64(1)
    X = [x ^ t for x in X]
    # interleave the transposed bits, most significant first
    h = np.zeros_like(X[0])
    for b in range(bits - 1, -1, -1):
        for i in range(3):
            h = (h << np.uint64(1)) | ((X[i] >> np.uint64(b)) & np.uint64(1))
    return h


@dataclass(frozen=True)
class Permutation:
    """new->old and old->new maps (0-based) for cells, edges and vertices."""

    cell_new_to_old: np.ndarray
    edge_new_to_old: np.ndarray
    vertex_new_to_old: np.ndarray

    @staticmethod
    def _inv(p: np.ndarray) -> np.ndarray:
        inv = np.empty_like(p)
        inv[p] = np.arange(p.size, dtype=p.dtype)
        return inv

    @property
    def cell_old_to_new(self) -> np.ndarray:
        return self._inv(self.cell_new_to_old)

    @property
    def edge_old_to_new(self) -> np.ndarray:
        return self._inv(self.edge_new_to_old)

    @property
    def vertex_old_to_new(self) -> np.ndarray:
        return self._inv(self.vertex_new_to_old)


def _np(v: Any) -> np.ndarray:
    try:
        import torch

        if isinstance(v, torch.Tensor):
            return v.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return np.asarray(v)


def _cell_xyz(mesh: Mapping[str, Any], nc: int) -> np.ndarray:
    if all(k in mesh for k in ("xCell", "yCell", "zCell")):
        p = np.stack([_np(mesh[k])[:nc] for k in ("xCell", "yCell", "zCell")], axis=1).astype(np.float64)
        r = np.linalg.norm(p, axis=1, keepdims=True)
        return p / np.where(r > 0, r, 1.0)
    lat, lon = _np(mesh["latCell"])[:nc], _np(mesh["lonCell"])[:nc]
    return np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], axis=1)


def compute_sfc_permutation(mesh: Mapping[str, Any]) -> Permutation:
    nc, ne, nv = (int(_np(mesh[k])) for k in ("nCells", "nEdges", "nVertices"))
    h = hilbert_index_3d(_cell_xyz(mesh, nc))
    cell_n2o = np.argsort(h, kind="stable").astype(np.int64)
    c_o2n = Permutation._inv(cell_n2o)
    coe = _np(mesh["cellsOnEdge"])[:ne].astype(np.int64) - 1
    a, b = c_o2n[coe[:, 0]], c_o2n[coe[:, 1]]
    edge_n2o = np.lexsort((np.arange(ne), np.maximum(a, b), np.minimum(a, b))).astype(np.int64)
    cov = _np(mesh["cellsOnVertex"])[:nv].astype(np.int64) - 1
    cn = np.sort(c_o2n[np.clip(cov, 0, None)], axis=1)
    vertex_n2o = np.lexsort((np.arange(nv), cn[:, 2], cn[:, 1], cn[:, 0])).astype(np.int64)
    return Permutation(cell_n2o, edge_n2o, vertex_n2o)


def _row_entity(name: str, a: Any, nc: int, ne: int, nv: int) -> str | None:
    """The entity kind (cell / edge / vertex) of an array's rows, from its shape."""
    shape = tuple(a.shape)
    if len(shape) == 0 or (len(shape) == 1 and name in VERTICAL_1D):
        return None
    n = shape[0]
    if n == nc + 1:
        return "cell"
    if n == ne + 1:
        return "edge"
    if n == nv + 1:
        return "vertex"
    return None


def permute_mapping(d: Mapping[str, Any], perm: Permutation, dims: tuple[int, int, int],
                    inverse: bool = False) -> dict[str, Any]:
    """Permute the rows of every per-entity array (padding row stays last) and
    renumber connectivity values.  ``inverse`` maps a permuted mapping back."""
    nc, ne, nv = dims
    if len({nc, ne, nv}) != 3:
        raise ValueError("entity counts must differ to tell array kinds apart")
    n2o = {"cell": perm.cell_new_to_old, "edge": perm.edge_new_to_old, "vertex": perm.vertex_new_to_old}
    o2n = {k: Permutation._inv(v) for k, v in n2o.items()}
    if inverse:
        n2o, o2n = o2n, n2o
    counts = {"cell": nc, "edge": ne, "vertex": nv}
    out: dict[str, Any] = {}
    rows_on: dict[tuple[str, Any], Any] = {}
    for name, v in d.items():
        if isinstance(v, (int, float, str, bool)) or v is None:
            out[name] = v
            continue
        if name not in VALUE_ENTITY and type(v).__module__.startswith("torch") and hasattr(v, "index_select"):
            # a tensor of values: its rows are gathered where it lives (no host round trip)
            ent = _row_entity(name, v, nc, ne, nv)
            if ent is None:
                out[name] = v
                continue
            key = (ent, v.device)
            if key not in rows_on:
                import torch

                rows_on[key] = torch.as_tensor(np.concatenate([n2o[ent], [counts[ent]]]), device=v.device)
            out[name] = v.detach().index_select(0, rows_on[key])
            continue
        a = _np(v)
        ent = _row_entity(name, a, nc, ne, nv)
        if ent is None:
            out[name] = v
            continue
        n = counts[ent]
        rows = np.concatenate([n2o[ent], [n]])
        b = a[rows]
        if name in VALUE_ENTITY:
            tgt = VALUE_ENTITY[name]
            m = o2n[tgt]
            bi = b.astype(np.int64)
            live = (bi >= 1) & (bi <= counts[tgt])
            bi = np.where(live, m[np.clip(bi - 1, 0, counts[tgt] - 1)] + 1, bi)
            b = bi.astype(a.dtype)
        out[name] = np.ascontiguousarray(b)
    return out
