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
