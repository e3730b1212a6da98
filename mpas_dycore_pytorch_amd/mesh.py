"""Synthetic icosahedral-Voronoi MPAS meshes and their static stencil fields.

The reference reads meshes from MPAS grid files (``mpas_pytorch/ic_loader.py``
``load_initial_state`` :1243) and synthesizes several derived stencil fields
when they are absent (``ic_loader.py`` :214-487, :750, :830).  This module
builds the same field set for a subdivided icosahedron so the dynamics step can
run without grid files:

* cells are the points of the level-``n`` subdivided icosahedron
  (``nCells = 10*4**n + 2``), vertices are the circumcentres of its triangles
  (``nVertices = 20*4**n``), edges are the triangle edges (``nEdges = 30*4**n``);
* connectivity is 1-based with 0 as the missing-neighbour sentinel and every
  per-entity array carries one trailing zero padding row, as in the reference
  runtime layout (``docs/API.md`` "State and controls");
* ``edgesOnCell_sign`` / ``edgesOnVertex_sign`` follow
  ``initialization.py:79`` (+1 when the cell is ``cellsOnEdge[e, 0]``; +1 when
  the vertex is ``verticesOnEdge[e, 1]``);
* TRiSK ``edgesOnEdge`` / ``weightsOnEdge`` follow Thuburn et al. (2009) and
  Ringler et al. (2010);
* ``advCellsForEdge`` / ``adv_coefs`` / ``adv_coefs_3rd`` are built from
  ``deriv_two`` with the construction of ``ic_loader.py:358-487``;
* ``deriv_two`` is the second derivative along the edge normal of a
  least-squares quadratic fit over each edge cell and its neighbours
  (Skamarock & Gassmann 2011);
* ``coeffs_reconstruct`` is a least-squares cell-vector reconstruction from
  edge-normal components (the reference uses RBF coefficients from the grid
  file, ``vector_reconstruction.py:311``).

Everything here is host-side set-up (NumPy, float64); nothing is on the
time-stepping path.
"""

from __future__ import annotations

import hashlib
import os
from dataclasses import dataclass

import numpy as np

EARTH_RADIUS_SYNTHETIC = 6.371e6
OMEGA = 7.29212e-5
MAX_EDGES = 6
MAX_EDGES2 = 2 * MAX_EDGES
VERTEX_DEGREE = 3
MAX_ADV = 15  # FIFTEEN in the reference's advCellsForEdge / deriv_two slots


# ---------------------------------------------------------------------------
# geometry helpers
# ---------------------------------------------------------------------------
def _normalize(x: np.ndarray) -> np.ndarray:
    return x / np.linalg.norm(x, axis=-1, keepdims=True)


def _arc(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Great-circle angle between unit vectors (rows)."""
    return np.arctan2(np.linalg.norm(np.cross(a, b), axis=-1), np.sum(a * b, axis=-1))


def _tri_area(a: np.ndarray, b: np.ndarray, c: np.ndarray) -> np.ndarray:
    """Spherical triangle area on the unit sphere (Van Oosterom & Strackee 1983)."""
    num = np.abs(np.sum(a * np.cross(b, c), axis=-1))
    den = 1.0 + np.sum(a * b, axis=-1) + np.sum(b * c, axis=-1) + np.sum(c * a, axis=-1)
    return 2.0 * np.arctan2(num, den)


def _latlon(x: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    lat = np.arcsin(np.clip(x[:, 2], -1.0, 1.0))
    lon = np.arctan2(x[:, 1], x[:, 0])
    lon = np.where(lon < 0.0, lon + 2.0 * np.pi, lon)
    return lat, lon


def _east_north(lat: np.ndarray, lon: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    east = np.stack([-np.sin(lon), np.cos(lon), np.zeros_like(lon)], axis=-1)
    north = np.stack(
        [-np.sin(lat) * np.cos(lon), -np.sin(lat) * np.sin(lon), np.cos(lat)], axis=-1
    )
    return east, north


# ---------------------------------------------------------------------------
# icosahedral triangulation
# ---------------------------------------------------------------------------
def _icosahedron() -> tuple[np.ndarray, np.ndarray]:
    t = (1.0 + 5.0 ** 0.5) / 2.0
    pts = np.array(
        [
            (-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
            (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
            (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1),
        ],
        dtype=np.float64,
    )
    faces = np.array(
        [
            (0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
            (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
            (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
            (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1),
        ],
        dtype=np.int64,
    )
    pts = _normalize(pts)
    a, b, c = pts[faces[:, 0]], pts[faces[:, 1]], pts[faces[:, 2]]
    flip = np.sum(np.cross(b - a, c - a) * a, axis=-1) < 0.0
    faces[flip] = faces[flip][:, [0, 2, 1]]
    return pts, faces


def _subdivide(pts: np.ndarray, faces: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    n_pts = pts.shape[0]
    e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]], axis=0)
    key = np.minimum(e[:, 0], e[:, 1]) * n_pts + np.maximum(e[:, 0], e[:, 1])
    uniq, inv = np.unique(key, return_inverse=True)
    lo, hi = uniq // n_pts, uniq % n_pts
    mids = _normalize(pts[lo] + pts[hi])
    new_pts = np.concatenate([pts, mids], axis=0)
    nf = faces.shape[0]
    ab = n_pts + inv[:nf]
    bc = n_pts + inv[nf : 2 * nf]
    ca = n_pts + inv[2 * nf :]
    a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
    new_faces = np.concatenate(
        [
            np.stack([a, ab, ca], 1),
            np.stack([b, bc, ab], 1),
            np.stack([c, ca, bc], 1),
            np.stack([ab, bc, ca], 1),
        ],
        axis=0,
    )
    return new_pts, new_faces


def icosahedral_triangulation(level: int) -> tuple[np.ndarray, np.ndarray]:
    pts, faces = _icosahedron()
    for _ in range(level):
        pts, faces = _subdivide(pts, faces)
    return pts, faces


# ---------------------------------------------------------------------------
# mesh container
# ---------------------------------------------------------------------------
@dataclass
class HorizontalMesh:
    """Active-entity (unpadded, 0-based) horizontal mesh arrays."""

    level: int
    radius: float
    x_cell: np.ndarray
    x_edge: np.ndarray
    x_vertex: np.ndarray
    cells_on_edge: np.ndarray  # (E, 2) 0-based
    vertices_on_edge: np.ndarray  # (E, 2)
    n_edges_on_cell: np.ndarray  # (C,)
    edges_on_cell: np.ndarray  # (C, 6) 0-based, -1 pad
    cells_on_cell: np.ndarray  # (C, 6)
    vertices_on_cell: np.ndarray  # (C, 6)
    cells_on_vertex: np.ndarray  # (V, 3)
    edges_on_vertex: np.ndarray  # (V, 3)
    kite_areas_on_vertex: np.ndarray  # (V, 3)
    dc_edge: np.ndarray
    dv_edge: np.ndarray
    area_cell: np.ndarray
    area_triangle: np.ndarray
    normal: np.ndarray  # (E, 3) unit edge normals cell1 -> cell2
    tangent: np.ndarray  # (E, 3) k x normal (vertex1 -> vertex2)


def build_horizontal_mesh(level: int, radius: float = EARTH_RADIUS_SYNTHETIC) -> HorizontalMesh:
    pts, faces = icosahedral_triangulation(level)
    n_c = pts.shape[0]
    n_v = faces.shape[0]

    # circumcentres of the counter-clockwise triangles are the Voronoi vertices
    a, b, c = pts[faces[:, 0]], pts[faces[:, 1]], pts[faces[:, 2]]
    x_vertex = _normalize(np.cross(b - a, c - a))

    # edges: unique triangle sides, ordered by (min cell, max cell)
    d0 = np.concatenate([faces[:, 0], faces[:, 1], faces[:, 2]])
    d1 = np.concatenate([faces[:, 1], faces[:, 2], faces[:, 0]])
    dface = np.concatenate([np.arange(n_v)] * 3)
    lo, hi = np.minimum(d0, d1), np.maximum(d0, d1)
    key = lo * n_c + hi
    uniq, inv = np.unique(key, return_inverse=True)
    n_e = uniq.shape[0]
    cells_on_edge = np.stack([uniq // n_c, uniq % n_c], axis=1)
    # directed side lo->hi (counter-clockwise in its triangle) has that
    # triangle on its left, i.e. on the +tangent side: vertex 2.
    vertices_on_edge = np.full((n_e, 2), -1, dtype=np.int64)
    forward = d0 < d1
    vertices_on_edge[inv[forward], 1] = dface[forward]
    vertices_on_edge[inv[~forward], 0] = dface[~forward]
    assert (vertices_on_edge >= 0).all()

    x_edge = _normalize(pts[cells_on_edge[:, 0]] + pts[cells_on_edge[:, 1]])
    chord = pts[cells_on_edge[:, 1]] - pts[cells_on_edge[:, 0]]
    normal = _normalize(chord - np.sum(chord * x_edge, -1, keepdims=True) * x_edge)
    tangent = np.cross(x_edge, normal)
    check = np.sum((x_vertex[vertices_on_edge[:, 1]] - x_vertex[vertices_on_edge[:, 0]]) * tangent, -1)
    assert (check > 0).all(), "vertex ordering does not follow k x n"

    dc_edge = radius * _arc(pts[cells_on_edge[:, 0]], pts[cells_on_edge[:, 1]])
    dv_edge = radius * _arc(x_vertex[vertices_on_edge[:, 0]], x_vertex[vertices_on_edge[:, 1]])

    # per-cell incidence, sorted counter-clockwise about the cell
    inc_cell = np.concatenate([cells_on_edge[:, 0], cells_on_edge[:, 1]])
    inc_edge = np.concatenate([np.arange(n_e), np.arange(n_e)])
    inc_other = np.concatenate([cells_on_edge[:, 1], cells_on_edge[:, 0]])
    lat_c, lon_c = _latlon(pts)
    east_c, north_c = _east_north(lat_c, lon_c)
    rel = pts[inc_other] - pts[inc_cell]
    ang = np.arctan2(np.sum(rel * north_c[inc_cell], -1), np.sum(rel * east_c[inc_cell], -1))
    order = np.lexsort((ang, inc_cell))
    inc_cell, inc_edge, inc_other = inc_cell[order], inc_edge[order], inc_other[order]
    n_edges_on_cell = np.bincount(inc_cell, minlength=n_c)
    assert n_edges_on_cell.max() <= MAX_EDGES
    start = np.concatenate([[0], np.cumsum(n_edges_on_cell)[:-1]])
    slot = np.arange(inc_cell.shape[0]) - start[inc_cell]
    edges_on_cell = np.full((n_c, MAX_EDGES), -1, dtype=np.int64)
    cells_on_cell = np.full((n_c, MAX_EDGES), -1, dtype=np.int64)
    edges_on_cell[inc_cell, slot] = inc_edge
    cells_on_cell[inc_cell, slot] = inc_other

    # verticesOnCell[i]: the vertex shared by edgesOnCell[i] and edgesOnCell[i+1]
    vertices_on_cell = np.full((n_c, MAX_EDGES), -1, dtype=np.int64)
    for i in range(MAX_EDGES):
        has = i < n_edges_on_cell
        cidx = np.nonzero(has)[0]
        e_a = edges_on_cell[cidx, i]
        nxt = (i + 1) % n_edges_on_cell[cidx]
        e_b = edges_on_cell[cidx, nxt]
        va, vb = vertices_on_edge[e_a], vertices_on_edge[e_b]
        shared = np.where(
            (va[:, 0] == vb[:, 0]) | (va[:, 0] == vb[:, 1]), va[:, 0], va[:, 1]
        )
        vertices_on_cell[cidx, i] = shared

    # vertex stencils: counter-clockwise triangle corners; edgesOnVertex[i]
    # joins cellsOnVertex[i] and cellsOnVertex[i+1]
    cells_on_vertex = faces.copy()
    edges_on_vertex = np.stack(
        [inv[:n_v], inv[n_v : 2 * n_v], inv[2 * n_v :]], axis=1
    )

    # kites: vertex, next-side midpoint, cell, previous-side midpoint
    kite = np.zeros((n_v, 3))
    xv = x_vertex
    for i in range(3):
        ci = cells_on_vertex[:, i]
        m_next = x_edge[edges_on_vertex[:, i]]
        m_prev = x_edge[edges_on_vertex[:, (i + 2) % 3]]
        kite[:, i] = _tri_area(xv, m_next, pts[ci]) + _tri_area(xv, pts[ci], m_prev)
    kite *= radius * radius
    area_triangle = kite.sum(axis=1)
    area_cell = np.zeros(n_c)
    np.add.at(area_cell, cells_on_vertex.reshape(-1), kite.reshape(-1))

    return HorizontalMesh(
        level=level,
        radius=radius,
        x_cell=pts,
        x_edge=x_edge,
        x_vertex=x_vertex,
        cells_on_edge=cells_on_edge,
        vertices_on_edge=vertices_on_edge,
        n_edges_on_cell=n_edges_on_cell,
        edges_on_cell=edges_on_cell,
        cells_on_cell=cells_on_cell,
        vertices_on_cell=vertices_on_cell,
        cells_on_vertex=cells_on_vertex,
        edges_on_vertex=edges_on_vertex,
        kite_areas_on_vertex=kite,
        dc_edge=dc_edge,
        dv_edge=dv_edge,
        area_cell=area_cell,
        area_triangle=area_triangle,
        normal=normal,
        tangent=tangent,
    )


# ---------------------------------------------------------------------------
# TRiSK tangential-velocity weights
# ---------------------------------------------------------------------------
def trisk_weights(hm: HorizontalMesh) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """``nEdgesOnEdge``, ``edgesOnEdge`` (0-based, -1 pad), ``weightsOnEdge``.

    For edge e and cell c of e the edges of c are walked counter-clockwise
    starting after e; the kite-area fraction R of c swept so far gives the
    weight (1/2 - R) * dvEdge[e'] / dcEdge[e] with the sign that makes
    ``sum_j w_j u_j`` the velocity component along ``k x n_e``.
    """
    n_e = hm.cells_on_edge.shape[0]
    n_c = hm.x_cell.shape[0]
    eoc, noc, voc = hm.edges_on_cell, hm.n_edges_on_cell, hm.vertices_on_cell
    n_eoe = np.zeros(n_e, dtype=np.int64)
    eoe = np.full((n_e, MAX_EDGES2), -1, dtype=np.int64)
    woe = np.zeros((n_e, MAX_EDGES2))
    # kite area of the vertex shared by slots i and i+1 of a cell, as a
    # fraction of the cell area (verticesOnCell[i] is that vertex)
    vv = np.clip(voc, 0, None)
    corner = np.argmax(hm.cells_on_vertex[vv] == np.arange(n_c)[:, None, None], axis=-1)
    frac = np.take_along_axis(hm.kite_areas_on_vertex[vv], corner[..., None], -1)[..., 0] / hm.area_cell[:, None]
    e_all = np.arange(n_e)
    k0 = np.zeros(n_e, dtype=np.int64)
    for side in (0, 1):
        c = hm.cells_on_edge[:, side]
        n = noc[c]
        j = np.argmax(eoc[c] == e_all[:, None], axis=1)
        s_e = 1.0 if side == 0 else -1.0
        r = np.zeros(n_e)
        for step in range(1, MAX_EDGES):
            live = step < n
            ep = eoc[c, (j + step) % n]
            # walk counter-clockwise; r accumulates the swept kite fraction
            r = np.where(live, r + frac[c, (j + step - 1) % n], r)
            s_ep = np.where(hm.cells_on_edge[ep, 0] == c, 1.0, -1.0)
            w = s_e * s_ep * (0.5 - r) * hm.dv_edge[ep] / hm.dc_edge
            slot = k0 + step - 1
            idx = np.nonzero(live)[0]
            eoe[idx, slot[idx]] = ep[idx]
            woe[idx, slot[idx]] = w[idx]
        k0 = k0 + (n - 1)
    n_eoe[:] = k0
    return n_eoe, eoe, woe


def _trisk_sign_fix(hm: HorizontalMesh, eoe: np.ndarray, woe: np.ndarray) -> np.ndarray:
    """Orient weights so they reconstruct the tangential (k x n) component."""
    lat, lon = _latlon(hm.x_edge)
    east, _ = _east_north(lat, lon)
    # solid-body flow about the polar axis: tangent vector field k x r ~ east*cos(lat)
    vel = east * np.cos(lat)[:, None]
    u = np.sum(vel * hm.normal, -1)
    vt = np.sum(vel * hm.tangent, -1)
    rec = np.sum(np.where(eoe >= 0, woe * u[np.clip(eoe, 0, None)], 0.0), axis=1)
    sgn = np.sign(np.dot(rec, vt))
    return woe * sgn


# ---------------------------------------------------------------------------
# least-squares helpers for deriv_two / reconstruction / deformation
# ---------------------------------------------------------------------------
def _local_plane(hm: HorizontalMesh, center: np.ndarray, pts: np.ndarray, xaxis: np.ndarray):
    """Map points to tangent-plane frames: ``center`` (m, 3), ``pts`` (m, p, 3),
    ``xaxis`` (m, 3); returns x, y of shape (m, p) in metres."""
    k = center
    ex = xaxis - np.sum(xaxis * k, -1, keepdims=True) * k
    ex = ex / np.linalg.norm(ex, axis=-1, keepdims=True)
    ey = np.cross(k, ex)
    kb = np.broadcast_to(k[:, None, :], pts.shape)
    ang = _arc(kb, pts)
    d = pts - np.sum(pts * kb, -1, keepdims=True) * kb
    dn = np.linalg.norm(d, axis=-1)
    dn = np.where(dn > 0.0, dn, 1.0)
    dirx = np.sum(d * ex[:, None, :], -1) / dn
    diry = np.sum(d * ey[:, None, :], -1) / dn
    r = hm.radius * ang
    return r * dirx, r * diry


def deriv_two(hm: HorizontalMesh) -> np.ndarray:
    """(E, 2, MAX_ADV) normal second-derivative weights for each edge cell.

    For edge e and its cell c, a quadratic is least-squares fitted over c and
    its edge neighbours in the tangent plane of c with x along the normal of
    e; the weights are 2x the x^2 row of the fit's pseudo-inverse.  The fit
    runs in coordinates scaled by dcEdge (condition number O(10) instead of
    O(1e12) in metres) and the x^2 row is scaled back by 1/dcEdge^2.
    """
    n_e = hm.cells_on_edge.shape[0]
    out = np.zeros((n_e, 2, MAX_ADV))
    for side in (0, 1):
        c_all = hm.cells_on_edge[:, side]
        n_all = hm.n_edges_on_cell[c_all]
        for n in np.unique(n_all):
            e = np.nonzero(n_all == n)[0]
            c = c_all[e]
            ids = np.concatenate([c[:, None], hm.cells_on_cell[c, :n]], axis=1)
            x, y = _local_plane(hm, hm.x_cell[c], hm.x_cell[ids], hm.normal[e])
            h = hm.dc_edge[e][:, None]
            xs, ys = x / h, y / h
            a = np.stack([np.ones_like(xs), xs, ys, xs * xs, xs * ys, ys * ys], axis=-1)
            pinv = np.linalg.pinv(a)
            out[e, side, : n + 1] = 2.0 * pinv[:, 3, :] / (h * h)
    return out


def advective_stencil(hm: HorizontalMesh, d2: np.ndarray, coef_3rd_order: float):
    """advCellsForEdge construction of ``ic_loader.py:358-471`` (0-based ids).

    The cell list of edge e is c1, c2, then the neighbours of c1 and of c2
    not yet listed, in slot order; the coefficients accumulate the two
    deriv_two rows into those positions in the same order.
    """
    n_e = hm.cells_on_edge.shape[0]
    coc, noc = hm.cells_on_cell, hm.n_edges_on_cell
    c1, c2 = hm.cells_on_edge[:, 0], hm.cells_on_edge[:, 1]
    adv_cells = np.full((n_e, MAX_ADV), -1, dtype=np.int64)
    adv_cells[:, 0], adv_cells[:, 1] = c1, c2
    n_adv = np.full(n_e, 2, dtype=np.int64)
    rows = np.arange(n_e)
    # positions of the neighbour slots of c1 (pos1) and c2 (pos2)
    pos1 = np.zeros((n_e, MAX_EDGES), dtype=np.int64)
    pos2 = np.zeros((n_e, MAX_EDGES), dtype=np.int64)
    for side, (cc, pos) in enumerate(((c1, pos1), (c2, pos2))):
        for i in range(MAX_EDGES):
            live = i < noc[cc]
            nb = coc[cc, i]
            hit = adv_cells == nb[:, None]
            seen = hit.any(axis=1)
            where = np.argmax(hit, axis=1)
            add = live & ~seen
            adv_cells[rows[add], n_adv[add]] = nb[add]
            pos[:, i] = np.where(seen, where, n_adv)
            n_adv = n_adv + add
    adv = np.zeros((n_e, MAX_ADV))
    adv3 = np.zeros((n_e, MAX_ADV))
    adv[rows, 0] += d2[:, 0, 0]
    adv3[rows, 0] += d2[:, 0, 0]
    for i in range(MAX_EDGES):
        live = np.nonzero(i < noc[c1])[0]
        adv[live, pos1[live, i]] += d2[live, 0, i + 1]
        adv3[live, pos1[live, i]] += d2[live, 0, i + 1]
    adv[rows, 1] += d2[:, 1, 0]
    adv3[rows, 1] -= d2[:, 1, 0]
    for i in range(MAX_EDGES):
        live = np.nonzero(i < noc[c2])[0]
        adv[live, pos2[live, i]] += d2[live, 1, i + 1]
        adv3[live, pos2[live, i]] -= d2[live, 1, i + 1]
    used = np.arange(MAX_ADV)[None, :] < n_adv[:, None]
    dcs = (hm.dc_edge ** 2)[:, None]
    adv = np.where(used, -(dcs) * adv / 12.0, 0.0)
    adv3 = np.where(used, -(dcs) * adv3 / 12.0, 0.0)
    adv[rows, 0] += 0.5
    adv[rows, 1] += 0.5
    adv = adv * hm.dv_edge[:, None]
    adv3 = adv3 * hm.dv_edge[:, None]
    adv3 = adv3 * float(coef_3rd_order)
    return n_adv, adv_cells, adv, adv3


def reconstruction_coeffs(hm: HorizontalMesh) -> np.ndarray:
    """(C, MAX_EDGES, 3) least-squares cell vector from edge-normal winds."""
    n_c = hm.x_cell.shape[0]
    out = np.zeros((n_c, MAX_EDGES, 3))
    for nedge in (5, 6):
        cidx = np.nonzero(hm.n_edges_on_cell == nedge)[0]
        if cidx.size == 0:
            continue
        nrm = hm.normal[hm.edges_on_cell[cidx, :nedge]]  # (m, n, 3)
        k = hm.x_cell[cidx]
        # project normals into each cell's tangent plane
        nrm = nrm - np.sum(nrm * k[:, None, :], -1, keepdims=True) * k[:, None, :]
        gram = np.einsum("mni,mnj->mij", nrm, nrm)
        # gram is rank 2 (tangent plane); regularise along k then remove it
        gram = gram + np.einsum("mi,mj->mij", k, k)
        inv = np.linalg.inv(gram)
        coef = np.einsum("mij,mnj->mni", inv, nrm)
        coef = coef - np.sum(coef * k[:, None, :], -1, keepdims=True) * k[:, None, :]
        out[cidx, :nedge, :] = coef
    return out


def deformation_coeffs(hm: HorizontalMesh) -> tuple[np.ndarray, np.ndarray]:
    """(C, MAX_EDGES) defc_a / defc_b for the 2-D deformation of a cell.

    With the edge normal at angle theta from local east, Gauss' theorem over
    the cell gives D11-D22 = sum(a*u - b*v) and D12 = sum(b*u + a*v) with
    a = s*dv*cos(2 theta)/A, b = s*dv*sin(2 theta)/A, s the outward sign.
    """
    n_c = hm.x_cell.shape[0]
    defc_a = np.zeros((n_c, MAX_EDGES))
    defc_b = np.zeros((n_c, MAX_EDGES))
    lat, lon = _latlon(hm.x_cell)
    east, north = _east_north(lat, lon)
    for i in range(MAX_EDGES):
        cidx = np.nonzero(i < hm.n_edges_on_cell)[0]
        e = hm.edges_on_cell[cidx, i]
        nrm = hm.normal[e]
        th = np.arctan2(np.sum(nrm * north[cidx], -1), np.sum(nrm * east[cidx], -1))
        s = np.where(hm.cells_on_edge[e, 0] == cidx, 1.0, -1.0)
        w = s * hm.dv_edge[e] / hm.area_cell[cidx]
        defc_a[cidx, i] = w * np.cos(2.0 * th)
        defc_b[cidx, i] = w * np.sin(2.0 * th)
    return defc_a, defc_b


# ---------------------------------------------------------------------------
# vertical grid
# ---------------------------------------------------------------------------
@dataclass
class VerticalGrid:
    n_levels: int
    z_top: float
    zw: np.ndarray  # (L+1,) interface heights
    dzw: np.ndarray
    rdzw: np.ndarray
    dzu: np.ndarray
    rdzu: np.ndarray
    fzm: np.ndarray
    fzp: np.ndarray
    cf1: float
    cf2: float
    cf3: float


def vertical_grid(n_levels: int, z_top: float = 30000.0, stretch: str = "uniform") -> VerticalGrid:
    """Flat-terrain height coordinate with MPAS interpolation metrics."""
    s = np.arange(n_levels + 1, dtype=np.float64) / n_levels
    if stretch == "uniform":
        zw = z_top * s
    elif stretch == "quadratic":
        zw = z_top * s * s
    else:
        raise ValueError(f"unknown stretch {stretch!r}")
    L = n_levels
    dzw = zw[1:] - zw[:-1]
    rdzw = 1.0 / dzw
    dzu = np.zeros(L)
    rdzu = np.zeros(L)
    fzm = np.zeros(L)
    fzp = np.zeros(L)
    for k in range(1, L):
        dzu[k] = 0.5 * (dzw[k] + dzw[k - 1])
        rdzu[k] = 1.0 / dzu[k]
        fzp[k] = 0.5 * dzw[k] / dzu[k]
        fzm[k] = 0.5 * dzw[k - 1] / dzu[k]
    if L >= 3:
        cof1 = (2.0 * dzu[1] + dzu[2]) / (dzu[1] + dzu[2]) * dzw[0] / dzu[1]
        cof2 = dzu[1] / (dzu[1] + dzu[2]) * dzw[0] / dzu[2]
        cf1 = fzp[1] + cof1
        cf2 = fzm[1] - cof1 - cof2
        cf3 = cof2
    else:
        cf1, cf2, cf3 = 1.0, 0.0, 0.0
    return VerticalGrid(L, z_top, zw, dzw, rdzw, dzu, rdzu, fzm, fzp, cf1, cf2, cf3)


def damping_coefs(zw: np.ndarray, n_cells: int, zd: float = 22000.0, xnutr: float = 0.2) -> np.ndarray:
    """Upper-layer w damping ``dss`` (``initialization.py:333-362``), (C, L)."""
    z_mid = 0.5 * (zw[:-1] + zw[1:])
    zt = zw[-1]
    raw = xnutr * np.sin(0.5 * np.pi * (z_mid - zd) / (zt - zd)) ** 2
    dss = np.where(z_mid > zd, raw, 0.0)
    return np.broadcast_to(dss, (n_cells, zw.shape[0] - 1)).copy()


# ---------------------------------------------------------------------------
# assembled runtime mesh dictionary (reference layout)
# ---------------------------------------------------------------------------
def _pad_rows(a: np.ndarray) -> np.ndarray:
    pad = np.zeros((1,) + a.shape[1:], dtype=a.dtype)
    return np.concatenate([a, pad], axis=0)


def _one_based(a: np.ndarray) -> np.ndarray:
    return np.where(a >= 0, a + 1, 0).astype(np.int64)


def _cache_path(level: int, n_levels: int, stretch: str, radius: float) -> str | None:
    root = os.environ.get("MPAS_AMD_MESH_CACHE")
    if not root:
        return None
    tag = hashlib.sha256(f"v3-{level}-{n_levels}-{stretch}-{radius!r}".encode()).hexdigest()[:16]
    os.makedirs(root, exist_ok=True)
    return os.path.join(root, f"icos_l{level}_L{n_levels}_{stretch}_{tag}.npz")


def build_mesh(
    level: int,
    n_levels: int,
    *,
    z_top: float = 30000.0,
    stretch: str = "uniform",
    radius: float = EARTH_RADIUS_SYNTHETIC,
    coef_3rd_order: float = 0.25,
) -> dict[str, np.ndarray | int | float]:
    """Runtime mesh dict (NumPy) in the reference's padded 1-based layout."""
    cache = _cache_path(level, n_levels, stretch, radius)
    if cache and os.path.exists(cache):
        with np.load(cache, allow_pickle=False) as z:
            return {k: (z[k].item() if z[k].ndim == 0 else z[k]) for k in z.files}

    hm = build_horizontal_mesh(level, radius)
    vg = vertical_grid(n_levels, z_top, stretch)
    n_c, n_e, n_v = hm.x_cell.shape[0], hm.cells_on_edge.shape[0], hm.x_vertex.shape[0]
    L = n_levels

    n_eoe, eoe, woe = trisk_weights(hm)
    woe = _trisk_sign_fix(hm, eoe, woe)
    d2 = deriv_two(hm)
    n_adv, adv_cells, adv, adv3 = advective_stencil(hm, d2, coef_3rd_order)
    coeffs = reconstruction_coeffs(hm)
    defc_a, defc_b = deformation_coeffs(hm)

    lat_c, lon_c = _latlon(hm.x_cell)
    lat_e, lon_e = _latlon(hm.x_edge)
    lat_v, lon_v = _latlon(hm.x_vertex)
    east_e, north_e = _east_north(lat_e, lon_e)
    angle_edge = np.arctan2(np.sum(hm.normal * north_e, -1), np.sum(hm.normal * east_e, -1))

    eoc1 = _one_based(hm.edges_on_cell)
    sign_cell = np.zeros((n_c, MAX_EDGES))
    for i in range(MAX_EDGES):
        valid = hm.edges_on_cell[:, i] >= 0
        e = np.clip(hm.edges_on_cell[:, i], 0, None)
        s = np.where(hm.cells_on_edge[e, 0] == np.arange(n_c), 1.0, -1.0)
        sign_cell[:, i] = np.where(valid, s, 0.0)
    sign_vert = np.zeros((n_v, 3))
    for i in range(3):
        e = hm.edges_on_vertex[:, i]
        sign_vert[:, i] = np.where(hm.vertices_on_edge[e, 1] == np.arange(n_v), 1.0, -1.0)

    kite_for_cell = np.zeros((n_c, MAX_EDGES), dtype=np.int64)
    for i in range(MAX_EDGES):
        cidx = np.nonzero(i < hm.n_edges_on_cell)[0]
        v = hm.vertices_on_cell[cidx, i]
        match = hm.cells_on_vertex[v] == cidx[:, None]
        kite_for_cell[cidx, i] = np.argmax(match, axis=1) + 1

    zgrid = np.broadcast_to(vg.zw, (n_c, L + 1)).copy()
    zz = np.ones((n_c, L))
    dss = damping_coefs(vg.zw, n_c)

    m: dict[str, np.ndarray | int | float] = {
        "nCells": n_c,
        "nEdges": n_e,
        "nVertices": n_v,
        "nVertLevels": L,
        "maxEdges": MAX_EDGES,
        "maxEdges2": MAX_EDGES2,
        "vertexDegree": VERTEX_DEGREE,
        "nAdvCellsMax": MAX_ADV,
        "sphere_radius": radius,
        "xCell": _pad_rows(hm.x_cell[:, 0] * radius),
        "yCell": _pad_rows(hm.x_cell[:, 1] * radius),
        "zCell": _pad_rows(hm.x_cell[:, 2] * radius),
        "latCell": _pad_rows(lat_c),
        "lonCell": _pad_rows(lon_c),
        "xEdge": _pad_rows(hm.x_edge[:, 0] * radius),
        "yEdge": _pad_rows(hm.x_edge[:, 1] * radius),
        "zEdge": _pad_rows(hm.x_edge[:, 2] * radius),
        "latEdge": _pad_rows(lat_e),
        "lonEdge": _pad_rows(lon_e),
        "xVertex": _pad_rows(hm.x_vertex[:, 0] * radius),
        "yVertex": _pad_rows(hm.x_vertex[:, 1] * radius),
        "zVertex": _pad_rows(hm.x_vertex[:, 2] * radius),
        "latVertex": _pad_rows(lat_v),
        "lonVertex": _pad_rows(lon_v),
        "cellsOnEdge": _pad_rows(_one_based(hm.cells_on_edge)),
        "verticesOnEdge": _pad_rows(_one_based(hm.vertices_on_edge)),
        "nEdgesOnCell": _pad_rows(hm.n_edges_on_cell.astype(np.int64)),
        "edgesOnCell": _pad_rows(eoc1),
        "cellsOnCell": _pad_rows(_one_based(hm.cells_on_cell)),
        "verticesOnCell": _pad_rows(_one_based(hm.vertices_on_cell)),
        "cellsOnVertex": _pad_rows(_one_based(hm.cells_on_vertex)),
        "edgesOnVertex": _pad_rows(_one_based(hm.edges_on_vertex)),
        "nEdgesOnEdge": _pad_rows(n_eoe),
        "edgesOnEdge": _pad_rows(_one_based(eoe)),
        "weightsOnEdge": _pad_rows(woe),
        "kiteAreasOnVertex": _pad_rows(hm.kite_areas_on_vertex),
        "kiteForCell": _pad_rows(kite_for_cell),
        "edgesOnCell_sign": _pad_rows(sign_cell),
        "edgesOnVertex_sign": _pad_rows(sign_vert),
        "dcEdge": _pad_rows(hm.dc_edge),
        "dvEdge": _pad_rows(hm.dv_edge),
        "invDcEdge": _pad_rows(1.0 / hm.dc_edge),
        "invDvEdge": _pad_rows(1.0 / hm.dv_edge),
        "areaCell": _pad_rows(hm.area_cell),
        "invAreaCell": _pad_rows(1.0 / hm.area_cell),
        "areaTriangle": _pad_rows(hm.area_triangle),
        "invAreaTriangle": _pad_rows(1.0 / hm.area_triangle),
        "angleEdge": _pad_rows(angle_edge),
        "fEdge": _pad_rows(2.0 * OMEGA * np.sin(lat_e)),
        "fVertex": _pad_rows(2.0 * OMEGA * np.sin(lat_v)),
        "fCell": _pad_rows(2.0 * OMEGA * np.sin(lat_c)),
        "meshDensity": _pad_rows(np.ones(n_c)),
        "meshScalingDel2": _pad_rows(np.ones(n_e)),
        "meshScalingDel4": _pad_rows(np.ones(n_e)),
        "meshScalingRegionalCell": _pad_rows(np.ones(n_c)),
        "meshScalingRegionalEdge": _pad_rows(np.ones(n_e)),
        "bdyMaskCell": _pad_rows(np.zeros(n_c, dtype=np.int64)),
        "bdyMaskEdge": _pad_rows(np.zeros(n_e, dtype=np.int64)),
        "bdyMaskVertex": _pad_rows(np.zeros(n_v, dtype=np.int64)),
        "deriv_two": _pad_rows(d2),
        "nAdvCellsForEdge": _pad_rows(n_adv),
        "advCellsForEdge": _pad_rows(_one_based(adv_cells)),
        "adv_coefs": _pad_rows(adv),
        "adv_coefs_3rd": _pad_rows(adv3),
        "coeffs_reconstruct": _pad_rows(coeffs),
        "defc_a": _pad_rows(defc_a),
        "defc_b": _pad_rows(defc_b),
        # vertical metrics (flat terrain: zeta == z)
        "zgrid": _pad_rows(zgrid),
        "zz": _pad_rows(zz),
        "rdzw": vg.rdzw.copy(),
        "dzw": vg.dzw.copy(),
        "rdzu": vg.rdzu.copy(),
        "dzu": vg.dzu.copy(),
        "fzm": vg.fzm.copy(),
        "fzp": vg.fzp.copy(),
        "rdnw": vg.rdzw.copy(),
        "fnm": vg.fzm.copy(),
        "fnp": vg.fzp.copy(),
        "cf1": float(vg.cf1),
        "cf2": float(vg.cf2),
        "cf3": float(vg.cf3),
        "zb": _pad_rows(np.zeros((n_e, 2, L + 1))),
        "zb3": _pad_rows(np.zeros((n_e, 2, L + 1))),
        "zb_cell": _pad_rows(np.zeros((n_c, MAX_EDGES, L + 1))),
        "zb3_cell": _pad_rows(np.zeros((n_c, MAX_EDGES, L + 1))),
        "zxu": _pad_rows(np.zeros((n_e, L))),
        "dss": _pad_rows(dss),
        "nominalMinDc": float(hm.dc_edge.min()),
    }
    if cache:
        np.savez(cache, **{k: np.asarray(v) for k, v in m.items()})
    return m
