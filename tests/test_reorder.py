"""Space-filling-curve reordering (replaces the reference's RCM reordering,
``mpas_pytorch/mesh_reordering.py:243``).

The permutation only moves rows and renumbers ids, keeping every slot order,
so a large step on the reordered mesh must reproduce the unpermuted step bit
for bit after the inverse permutation.
"""

from __future__ import annotations

import numpy as np

from conftest import srk3_case

from mpas_dycore_pytorch_amd import reorder as R
from oracle import oracle as O
from oracle import srk3 as OS


def test_hilbert_index_is_a_bijection_on_the_grid_and_local():
    bits = 3
    g = np.arange(1 << bits)
    ix, iy, iz = np.meshgrid(g, g, g, indexing="ij")
    q = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    p = q / ((1 << bits) - 1) * 2.0 - 1.0
    h = R.hilbert_index_3d(p, bits=bits)
    assert sorted(h.tolist()) == list(range(1 << (3 * bits)))
    # consecutive curve points are grid neighbours (unit Manhattan step)
    order = np.argsort(h)
    steps = np.abs(np.diff(q[order], axis=0)).sum(axis=1)
    assert np.all(steps == 1)


def test_permutation_round_trip_and_locality():
    st, _, _ = srk3_case("l2_L20_tracer")
    dims = (st["nCells"], st["nEdges"], st["nVertices"])
    perm = R.compute_sfc_permutation(st)
    for p in (perm.cell_new_to_old, perm.edge_new_to_old, perm.vertex_new_to_old):
        assert sorted(p.tolist()) == list(range(p.size))
    ps = R.permute_mapping(st, perm, dims)
    back = R.permute_mapping(ps, perm, dims, inverse=True)
    for k, v in st.items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(back[k], v), k
    nc = st["nCells"]
    d_new = np.abs(ps["cellsOnCell"][:nc] - 1 - np.arange(nc)[:, None]).mean()
    d_old = np.abs(st["cellsOnCell"][:nc] - 1 - np.arange(nc)[:, None]).mean()
    assert d_new < 0.5 * d_old
    # edges sorted by (min, max) of their new cell ids
    coe = ps["cellsOnEdge"][: st["nEdges"]]
    key = np.minimum(coe[:, 0], coe[:, 1]) * (nc + 1) + np.maximum(coe[:, 0], coe[:, 1])
    assert np.all(np.diff(key) >= 0)


def test_reordered_step_is_bitwise_identical():
    st, _, meta = srk3_case("l2_L20_tracer")
    dims = (st["nCells"], st["nEdges"], st["nVertices"])
    perm = R.compute_sfc_permutation(st)
    ps = R.permute_mapping(st, perm, dims)
    ld = float(st["nominalMinDc"])
    a = OS.srk3(st, O.Mesh(st), meta["dt"], len_disp=ld)
    b = R.permute_mapping(OS.srk3(ps, O.Mesh(ps), meta["dt"], len_disp=ld), perm, dims, inverse=True)
    for k in ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "pressure_p", "v", "ke", "pv_edge",
              "vorticity", "divergence", "ru", "rw", "uReconstructZonal", "uReconstructMeridional"):
        assert np.array_equal(a[k], b[k]), k
