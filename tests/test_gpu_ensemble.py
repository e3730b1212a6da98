"""Batched uneven ensembles on the GPU: each member equals itself stepped alone, bit for bit."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _prepared(level, L, seed, dt, dm_cache):
    from mpas_dycore_pytorch_amd import initial_state as IS
    from mpas_dycore_pytorch_amd import kernels as K
    from mpas_dycore_pytorch_amd import mesh as M
    from mpas_dycore_pytorch_amd.device import DeviceMesh

    key = (level, L)
    if key not in dm_cache:
        dm_cache[key] = M.build_mesh(level, L, stretch="quadratic")
    m = dm_cache[key]
    raw = IS.raw_state(m, seed=seed, n_tracers=1)
    base = {**m, **raw, **IS.init_coupled_diagnostics(m, raw)}
    dm = DeviceMesh(base, torch.device("cuda"))
    st = IS.prepare_state(
        m, raw,
        diagnostics=lambda u, h: {k: v.cpu().numpy() for k, v in K.solve_diagnostics(dm, dt, 0.5, None, u, h).items()},
        reconstruct=lambda u: {k: v.cpu().numpy() for k, v in K.reconstruct(dm, u).items()})
    return m, {**m, **st}


_MESH_BASE: dict = {}


def _mesh_with_base(m, st):
    """The member's mesh with its base-state fields, one object per mesh (so members group)."""
    key = id(m)
    if key not in _MESH_BASE:
        _MESH_BASE[key] = {**m, **{k: st[k] for k in ("rho_base", "rtheta_base", "exner_base", "theta_base")
                                   if k in st}}
    return _MESH_BASE[key]


@pytest.mark.parametrize("graph", [False, True])
def test_uneven_ensemble_members_equal_single_runs(graph):
    from mpas_dycore_pytorch_amd.ensemble import EnsembleStepper
    from mpas_dycore_pytorch_amd.time_integration import Srk3Stepper

    cache, members, cfgs = {}, [], {}
    for level, dt, seeds in ((2, 900.0, (1, 2, 3)), (1, 1800.0, (4, 5))):
        for s in seeds:
            m, st = _prepared(level, 12, s, dt, cache)
            mesh = _mesh_with_base(m, st)
            cfg = cfgs.setdefault(id(mesh), {"config_time_integration_order": 3,
                                             "config_len_disp": float(m["nominalMinDc"])})
            members.append((st, mesh, dt, cfg))
    ens = EnsembleStepper(members, reorder=True)
    assert len(ens.groups) == 2
    ens.step(1)
    if graph:
        ens.capture_graphs()
    ens.step(2)
    torch.cuda.synchronize()
    got = ens.states()
    for i, (st, mesh, dt, cfg) in enumerate(members):
        single = Srk3Stepper(st, mesh, cfg, dt, reorder=False)
        single.step(3)
        ref = single.state_dict()
        for k in ("u", "w", "theta_m", "rho_zz", "scalars", "exner", "pressure_p", "ru", "pv_edge"):
            a, b = got[i][k].cpu().numpy(), ref[k].cpu().numpy()
            assert np.array_equal(a, b), (i, k, float(np.max(np.abs(a - b))))
    print(f"[ensemble] {len(members)} members in {len(ens.groups)} groups, graph={graph}: bitwise equal", flush=True)
