"""A large step captured in a HIP graph replays the eager step bit for bit.

The graph is replayed several times with eager work (torch kernels and host
synchronisations) between replays; every replay is compared with an eager
stepper advanced in lockstep.
"""

from __future__ import annotations

import pytest
import torch

from conftest import srk3_case

pytestmark = pytest.mark.gpu


def test_graph_replay_matches_eager(gpu_available):
    from mpas_dycore_pytorch_amd.time_integration import Srk3Stepper

    print("graph replay l2_L20_tracer", flush=True)
    st, _, meta = srk3_case("l2_L20_tracer")
    cfg = {"config_time_integration_order": 3, "config_len_disp": float(st["nominalMinDc"])}
    dev = torch.device("cuda", 0)
    eager = Srk3Stepper(st, st, cfg, meta["dt"], device=dev, reorder=True)
    graph = Srk3Stepper(st, st, cfg, meta["dt"], device=dev, reorder=True)
    eager.step(1)
    graph.capture_graph()  # runs the first (refreshing) step eagerly, then captures
    for rep in range(4):
        junk = torch.randn(1 << 16, device=dev, dtype=torch.float64).sum()  # eager work in between
        torch.cuda.synchronize()
        eager.step(1)
        graph.step(1)
        torch.cuda.synchronize()
        a, b = eager.state_dict(), graph.state_dict()
        for k in a:
            assert torch.equal(a[k], b[k]), f"replay {rep}: {k}"
        assert torch.isfinite(junk)
