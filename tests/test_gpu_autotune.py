"""Launch shapes and block orders from the autotuner change timing only: results stay bitwise equal."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import srk3_case

pytestmark = pytest.mark.gpu


def test_launch_shapes_do_not_change_results():
    from mpas_dycore_pytorch_amd import autotune as AT
    from mpas_dycore_pytorch_amd.time_integration import Srk3Stepper

    st, _, meta = srk3_case("l2_L20_tracer")
    cfg = {"config_time_integration_order": 3, "config_len_disp": float(st["nominalMinDc"])}
    outs = []
    try:
        for waves, order in ((4, 0), (1, 0), (2, 1), (16, 0), (4, 1)):
            AT.install({k: {"waves": waves, "order": order} for k in AT.WAVE_KERNELS})
            sp = Srk3Stepper(st, st, cfg, meta["dt"])
            sp.step(2)
            outs.append({k: v.cpu().numpy() for k, v in sp.state_dict().items()})
        sp = Srk3Stepper(st, st, cfg, meta["dt"])
        table = AT.autotune(sp, steps=1)
        # every kernel the step launches is tuned; the option variants are not launched here
        assert {"k_dyn_edges_a", "k_acoustic_edges", "k_scalar_hflux", "k_mono_final"} <= set(table)
        assert set(table) <= set(AT.WAVE_KERNELS) - {"k_dyn_edges_a_opt", "k_dyn_cells_b_opt"}
        assert all(v["waves"] in AT.CANDIDATES for v in table.values())
        assert all(v["order"] in AT.ORDERS for v in table.values())
    finally:
        AT.install({k: 0 for k in AT.WAVE_KERNELS})
    for o in outs[1:]:
        for k in ("u", "w", "theta_m", "rho_zz", "scalars", "pv_edge"):
            assert np.array_equal(o[k], outs[0][k]), k


def test_adjoint_launch_shapes_do_not_change_gradients():
    """The adjoint and tangent kernels take their shape and block order from the same table; the
    gradients and tangents of a large step stay bitwise equal under every choice (the adjoints
    are atomic-free)."""
    from mpas_dycore_pytorch_amd import autotune as AT
    from mpas_dycore_pytorch_amd.adjoint import STEP_INPUTS, step_jvp, step_vjp
    from mpas_dycore_pytorch_amd.time_integration import Srk3Stepper, resolve_config

    st, _, meta = srk3_case("l2_L20_tracer")
    config = {"config_time_integration_order": 3, "config_len_disp": float(st["nominalMinDc"])}
    cfg = resolve_config(config, st, st, meta["dt"], int(np.asarray(st["scalars"]).shape[2]))
    sp = Srk3Stepper(st, st, config, meta["dt"])
    ckpt = {k: sp.t[k].clone() for k in STEP_INPUTS}
    gen = torch.Generator().manual_seed(11)
    ybar = {k: torch.randn(tuple(ckpt[k].shape), generator=gen, dtype=torch.float64).to(ckpt[k].device)
            for k in ("u", "w", "theta_m")}
    for k in ybar:
        ybar[k][-1] = 0.0

    tdot = {k: torch.randn(tuple(ckpt[k].shape), generator=gen, dtype=torch.float64).to(ckpt[k].device)
            for k in ("u", "theta_m")}
    for k in tdot:
        tdot[k][-1] = 0.0

    def grads():
        _, g = step_vjp(sp.mesh, ckpt, cfg, meta["dt"], ybar)
        _, jt = step_jvp(sp.mesh, ckpt, cfg, meta["dt"], tdot)
        out = {k: v.cpu().numpy() for k, v in g.items() if v is not None}
        out.update({"tangent_" + k: v.cpu().numpy() for k, v in jt.items() if v is not None})
        return out

    outs = []
    both = list(AT.ADJOINT_KERNELS) + list(AT.TANGENT_KERNELS)
    try:
        for waves, order in ((4, 0), (1, 1), (2, 0), (4, 1)):
            AT.install({k: {"waves": waves, "order": order} for k in both})
            outs.append(grads())
        table = AT.autotune_reverse(grads, kernels=("k_adj_dyn_a3w", "k_adj_diag_g4"))
        assert set(table) == {"k_adj_dyn_a3w", "k_adj_diag_g4"}
        table.update(AT.autotune_tangent(grads, kernels=("k_tan_dyn_j4",)))
        assert "k_tan_dyn_j4" in table
        assert all(v["waves"] in AT.ADJOINT_CANDIDATES and v["order"] in AT.ORDERS for v in table.values())
    finally:
        AT.install({k: 0 for k in both})
    for o in outs[1:]:
        assert o.keys() == outs[0].keys()
        for k in o:
            assert np.array_equal(o[k], outs[0][k]), k

