"""The allocation-free large step (``mpas_hip_srk3_step_noalloc``) against ``mpas_hip_srk3_step``.

Two calls in a row through the C ABI, the second on a smaller mesh (so it
runs on a static workspace that still holds the first call's contents): each
result must equal the workspace entry point's bit for bit.  A mesh whose
workspace exceeds the static array is rejected before anything runs.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from conftest import srk3_case

pytestmark = pytest.mark.gpu


def _prepared_level3():
    from mpas_dycore_pytorch_amd import initial_state as IS
    from mpas_dycore_pytorch_amd import mesh as M
    from mpas_dycore_pytorch_amd import run_simulation

    m = M.build_mesh(3, 30, stretch="quadratic")
    raw = IS.raw_state(m, seed=31, n_tracers=1)
    cfg = {"config_time_integration_order": 3}
    prepared = run_simulation(raw, m, cfg, 600.0, nsteps=0, device="cuda")
    return prepared, m, cfg, 600.0


def _golden_level2():
    st, _, meta = srk3_case("l2_L20_tracer")
    return st, st, {"config_time_integration_order": 3, "config_len_disp": float(st["nominalMinDc"])}, meta["dt"]


def _noalloc_step(lib, stepper):
    from mpas_dycore_pytorch_amd import _native as N
    from mpas_dycore_pytorch_amd.device import stream_handle

    cfg = N.ConfigStruct()
    C.pointer(cfg)[0] = stepper.cfg
    cfg.diagnostics_current = 0
    with torch.cuda.device(stepper.device):
        rc = lib.mpas_hip_srk3_step_noalloc(stepper.mesh.ref, C.byref(stepper._st), C.byref(cfg),
                                            C.c_void_p(stream_handle(stepper.device)))
    N.check(rc, "mpas_hip_srk3_step_noalloc")


def test_noalloc_step_equals_the_workspace_step_twice_shrinking():
    from mpas_dycore_pytorch_amd import _native as N
    from mpas_dycore_pytorch_amd.time_integration import Srk3Stepper

    lib = N.load()
    dev = torch.device("cuda", 0)
    cases = [("level 3 x 30", _prepared_level3()), ("level 2 x 20", _golden_level2())]
    for name, (state, mesh, cfg, dt) in cases:
        print(f"[noalloc] {name}", flush=True)
        ref = Srk3Stepper(state, mesh, cfg, dt, device=dev)
        got = Srk3Stepper(state, mesh, cfg, dt, device=dev)
        need = lib.mpas_hip_workspace_doubles(ref.mesh.nc, ref.mesh.ne, ref.mesh.nv, ref.mesh.L, ref.n_scalars)
        assert need <= lib.mpas_hip_srk3_step_noalloc_capacity()
        ref.step(1)
        _noalloc_step(lib, got)
        torch.cuda.synchronize()
        a, b = ref.state_dict(), got.state_dict()
        for k in N.STATE_FIELDS:
            assert torch.equal(a[k], b[k]), f"{name}: {k}"
        assert bool(torch.isfinite(b["theta_m"]).all())


def test_noalloc_rejects_a_mesh_beyond_its_workspace():
    from mpas_dycore_pytorch_amd import _native as N

    lib = N.load()
    m = N.MeshStruct()
    m.nCells, m.nEdges, m.nVertices, m.nLevels = 10242, 30720, 20480, 55
    m.maxEdges, m.maxEdges2, m.vertexDegree, m.maxAdv = 6, 12, 3, 15
    cfg = N.ConfigStruct()
    cfg.nScalars, cfg.moist_start, cfg.moist_end, cfg.ns, cfg.split, cfg.order = 1, 1, 1, 2, 3, 3
    st = N.StateStruct()
    rc = lib.mpas_hip_srk3_step_noalloc(C.byref(m), C.byref(st), C.byref(cfg), None)
    assert rc != 0
    msg = lib.mpas_hip_last_error().decode()
    assert "KiB of workspace" in msg and "mpas_hip_srk3_step" in msg
    assert np.isfinite(float(lib.mpas_hip_srk3_step_noalloc_capacity()))
