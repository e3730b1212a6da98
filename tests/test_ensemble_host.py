"""Disjoint-union ensembles on the CPU: the union relabels entities and mixes nothing.

One oracle large step of the union of three member states equals, bit for
bit, each member stepped alone; ``split`` inverts ``union``.
"""

from __future__ import annotations

import numpy as np
import torch

from mpas_dycore_pytorch_amd import initial_state as IS
from mpas_dycore_pytorch_amd import mesh as M
from mpas_dycore_pytorch_amd.ensemble import split, union


def _members(n=3, level=1, L=10, dt=900.0):
    from oracle import oracle as O

    m = M.build_mesh(level, L, stretch="quadratic")
    states = []
    for seed in range(n):
        raw = IS.raw_state(m, seed=100 + seed, n_tracers=1)
        base = {**m, **raw, **IS.init_coupled_diagnostics(m, raw)}
        om = O.Mesh(base)
        st = IS.prepare_state(m, raw, diagnostics=lambda u, h: O.solve_diagnostics(om, dt, 0.5, None, u, h),
                              reconstruct=lambda u: O.reconstruct(om, u))
        states.append({**m, **st})
    return m, states


def test_union_split_round_trip():
    m, states = _members()
    us = union(states, [m] * 3)
    t = {k: torch.as_tensor(v) for k, v in us.items() if isinstance(v, np.ndarray) and v.ndim >= 1}
    parts = split(t, [m] * 3)
    for i, st in enumerate(states):
        for k in ("u", "w", "theta_m", "rho_zz", "scalars", "zz", "exner"):
            assert np.array_equal(parts[i][k].numpy(), np.asarray(st[k])), (i, k)
        # connectivity is offset into the union's ranges
        off = i * int(m["nCells"])
        live = np.asarray(st["cellsOnEdge"])[: int(m["nEdges"])]
        assert np.array_equal(parts[i]["cellsOnEdge"].numpy()[: int(m["nEdges"])], np.where(live > 0, live + off, 0))


def test_union_step_equals_members_stepped_alone():
    from oracle import oracle as O, srk3 as OS

    m, states = _members()
    us = union(states, [m] * 3)
    dt = 900.0
    ld = float(m["nominalMinDc"])
    out_u = OS.srk3(us, O.Mesh(us), dt, len_disp=ld)
    nc, ne = int(m["nCells"]), int(m["nEdges"])
    for i, st in enumerate(states):
        out_i = OS.srk3(st, O.Mesh(st), dt, len_disp=ld)
        for k, rows in (("u", ne), ("w", nc), ("theta_m", nc), ("rho_zz", nc), ("scalars", nc), ("exner", nc)):
            a = np.asarray(out_u[k])[i * rows:(i + 1) * rows]
            b = np.asarray(out_i[k])[:rows]
            assert np.array_equal(a, b), (i, k, float(np.max(np.abs(a - b))))
