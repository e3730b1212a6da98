"""C ABI library and host logic, without a GPU.

* the built library loads and exports every function ``include/*.h`` declares;
* the Python struct layouts agree with the compiled ones;
* workspace sizing and the host-side validation (configuration capability
  gate, connectivity bounds, level limits) raise the reference's error types
  before anything reaches a kernel.
"""

from __future__ import annotations

import ctypes as C
import glob
import os
import re

import numpy as np
import pytest
import torch

from conftest import ROOT, srk3_case

from mpas_dycore_pytorch_amd import _native as N
from mpas_dycore_pytorch_amd.device import DeviceMesh
from mpas_dycore_pytorch_amd.errors import (MpasConfigError, MpasMeshMismatchError, MpasMissingFieldError,
                                            MpasUnsupportedCapabilityError)
from mpas_dycore_pytorch_amd.time_integration import resolve_config


def _declared_functions() -> set[str]:
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        if h.endswith("mpas_hip_mixed.h"):
            continue  # the mixed-precision library's header (its own test below)
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(mpas_hip_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    lib = N.load()
    declared = _declared_functions()
    assert len(declared) >= 16
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    assert declared == set(N.SIGNATURES), "binding and header disagree"


def test_struct_layouts_agree():
    lib = N.load()
    sizes = [C.c_int32() for _ in range(4)]
    lib.mpas_hip_struct_sizes(*[C.byref(s) for s in sizes])
    assert [s.value for s in sizes] == [C.sizeof(N.MeshStruct), C.sizeof(N.StateStruct),
                                        C.sizeof(N.WorkspaceStruct), C.sizeof(N.ConfigStruct)]


def test_workspace_sizes_scale_with_mesh():
    lib = N.load()
    a = lib.mpas_hip_workspace_doubles(2562, 7680, 5120, 26, 0)
    b = lib.mpas_hip_workspace_doubles(10242, 30720, 20480, 26, 0)
    c = lib.mpas_hip_workspace_doubles(2562, 7680, 5120, 26, 2)
    assert b > 3.9 * a
    assert c - a >= 2 * 2 * 2563 * 26  # two scalar copies more per species
    # level 7, 55 levels, one tracer: well under one GPU's HBM
    assert lib.mpas_hip_workspace_doubles(163842, 491520, 327680, 55, 1) * 8 < 8e9


def _case():
    st, _, meta = srk3_case("l2_L20_tracer")
    return st, meta


def test_config_gate():
    st, meta = _case()
    ok = {"config_time_integration_order": 3, "config_len_disp": float(st["nominalMinDc"])}
    cfg = resolve_config(ok, st, st, meta["dt"], 2)
    assert cfg.ns == 2 and cfg.split == 3 and cfg.monotonic == 1 and cfg.nScalars == 2
    assert cfg.moist_start == 1 and cfg.moist_end == 1
    assert cfg.order == 3
    # RK2 is the reference's default order (config_defaults.py:34-35)
    assert resolve_config({**ok, "config_time_integration_order": 2}, st, st, meta["dt"], 2).order == 2
    assert resolve_config({"config_len_disp": 1.0}, st, st, meta["dt"], 2).order == 2
    for bad in (1, 4, 2.5, "three", True):
        with pytest.raises(MpasConfigError):
            resolve_config({**ok, "config_time_integration_order": bad}, st, st, meta["dt"], 2)
    # horizontal mixing (dyn_tend.py:3367-3430): "2d_fixed" takes its own coefficients, an
    # unknown mode is a configuration error; the fixed coefficients are ignored under Smagorinsky
    cfg = resolve_config({**ok, "horiz_mixing": "2d_fixed", "config_h_theta_eddy_visc2": 3.0,
                          "config_h_mom_eddy_visc4": 5.0, "config_h_theta_eddy_visc4": 7.0}, st, st, meta["dt"], 2)
    assert (cfg.mix_fixed, cfg.h_theta_eddy_visc2, cfg.h_mom_eddy_visc4, cfg.h_theta_eddy_visc4) == (1, 3.0, 5.0, 7.0)
    assert resolve_config({**ok, "config_h_mom_eddy_visc4": 5.0}, st, st, meta["dt"], 2).mix_fixed == 0
    with pytest.raises(MpasConfigError):
        resolve_config({**ok, "config_horiz_mixing": "3d_smagorinsky"}, st, st, meta["dt"], 2)
    # CAM damping (dyn_tend.py:3434-3458): on only for a positive coefficient and level count;
    # 4 levels by default (config_defaults.py:104-105)
    ld = float(st["nominalMinDc"])
    cfg = resolve_config({**ok, "config_mpas_cam_coef": 0.5}, st, st, meta["dt"], 2)
    assert cfg.cam_levels == 4 and cfg.cam_visc2 == 4.0 * 2.0833 * ld * 0.5
    assert resolve_config({**ok, "mpas_cam_coef": 0.5, "number_cam_damping_levels": 0}, st, st,
                          meta["dt"], 2).cam_levels == 0
    assert resolve_config(ok, st, st, meta["dt"], 2).cam_levels == 0
    with pytest.raises(MpasConfigError):
        resolve_config({**ok, "config_mpas_cam_coef": 0.5, "config_number_cam_damping_levels": 99}, st, st,
                       meta["dt"], 2)
    # a full key name takes precedence over its short alias (config_defaults.py:745-774)
    assert resolve_config({**ok, "epssm": 0.3, "config_epssm": 0.2}, st, st, meta["dt"], 2).epssm == 0.2
    # curvature terms (dyn_tend.py:2471-2505, :4241-4267) need the edge angles and latitudes
    assert resolve_config({**ok, "curvature_u": True}, st, st, meta["dt"], 2).curvature == 1
    no_angles = {k: v for k, v in st.items() if k not in ("angleEdge", "latEdge")}
    with pytest.raises(MpasMissingFieldError):
        resolve_config({**ok, "config_curvature_u": True}, no_angles, no_angles, meta["dt"], 2)
    # vertical eddy viscosities (dyn_tend.py:2731-2781, :4633-4647, :5060-5097)
    cfg = resolve_config({**ok, "config_v_mom_eddy_visc2": 4.0, "config_v_theta_eddy_visc2": 2.0, "mix_full": False},
                         st, st, meta["dt"], 2)
    assert (cfg.v_mom_eddy_visc2, cfg.v_theta_eddy_visc2, cfg.mix_full) == (4.0, 2.0, 0)
    assert resolve_config(ok, st, st, meta["dt"], 2).mix_full == 1
    # positive-definite transport selects the limited final stage (time_integration.py:3980-3987);
    # scalar advection can be switched off; the advection-order keys are inert integers
    # (config_defaults.py:604-614)
    pd = resolve_config({**ok, "monotonic": False, "positive_definite": True}, st, st, meta["dt"], 2)
    assert pd.monotonic == 1 and pd.scalar_advection == 1
    assert resolve_config({**ok, "config_monotonic": False}, st, st, meta["dt"], 2).monotonic == 0
    assert resolve_config({**ok, "advect_scalars": False}, st, st, meta["dt"], 2).scalar_advection == 0
    assert resolve_config({**ok, "config_scalar_adv_order": 2, "config_w_adv_order": "4"}, st, st, meta["dt"],
                          2).monotonic == 1
    with pytest.raises(MpasConfigError):
        resolve_config({**ok, "config_theta_adv_order": "third"}, st, st, meta["dt"], 2)
    # one dynamics substep means the reference's coupled (in-stage) transport
    with pytest.raises(MpasUnsupportedCapabilityError):
        resolve_config({**ok, "config_dynamics_split_steps": 1}, st, st, meta["dt"], 2)
    assert resolve_config({**ok, "config_dynamics_split_steps": 1, "config_scalar_advection": False}, st, st,
                          meta["dt"], 2).split == 1
    for key, value in (("config_split_dynamics_transport", False), ("config_apply_lbcs", True)):
        with pytest.raises(MpasUnsupportedCapabilityError):
            resolve_config({**ok, key: value}, st, st, meta["dt"], 2)
    # Rayleigh damping of u: the reference's coefficient 1 / (levels * days * 86400)
    # (dyn_tend.py:10066-10072) and defaults (6 levels, 5 days; config_defaults.py:100-102)
    assert resolve_config(ok, st, st, meta["dt"], 2).rayleigh_levels == 0
    cfg = resolve_config({**ok, "config_rayleigh_damp_u": True}, st, st, meta["dt"], 2)
    assert cfg.rayleigh_levels == 6 and cfg.rayleigh_coef_inverse == 1.0 / (6.0 * (5.0 * 86400.0))
    cfg = resolve_config({**ok, "rayleigh_damp_u": True, "number_rayleigh_damp_u_levels": 3,
                          "rayleigh_damp_u_timescale_days": 0.5}, st, st, meta["dt"], 2)
    assert cfg.rayleigh_levels == 3 and cfg.rayleigh_coef_inverse == 1.0 / (3.0 * (0.5 * 86400.0))
    for bad in ({"config_number_rayleigh_damp_u_levels": 0}, {"config_number_rayleigh_damp_u_levels": 99},
                {"config_rayleigh_damp_u_timescale_days": 0.0}):
        with pytest.raises(MpasConfigError):
            resolve_config({**ok, "config_rayleigh_damp_u": True, **bad}, st, st, meta["dt"], 2)
    with pytest.raises(MpasConfigError):
        resolve_config(ok, st, st, -1.0, 2)
    # config_len_disp <= 0 (the reference default) resolves from nominalMinDc, then from
    # the smallest dcEdge of rows 1..nEdges (initialization.py:365-380, time_integration.py:1882-1940)
    assert resolve_config({}, st, st, meta["dt"], 2).len_disp == float(st["nominalMinDc"])
    no_nominal = {k: v for k, v in st.items() if k != "nominalMinDc"}
    ne = int(st["nEdges"])
    want = float(np.min(st["dcEdge"][1: ne + 1][st["dcEdge"][1: ne + 1] > 0]))
    with pytest.warns(RuntimeWarning):
        assert resolve_config({}, no_nominal, no_nominal, meta["dt"], 2).len_disp == want
    bare = {k: v for k, v in no_nominal.items() if k != "dcEdge"}
    with pytest.raises(MpasConfigError):
        resolve_config({}, bare, bare, meta["dt"], 2)
    # short aliases are honoured as in the reference
    cfg = resolve_config({**ok, "smdiv": 0.2, "monotonic": False}, st, st, meta["dt"], 2)
    assert cfg.smdiv == 0.2 and cfg.monotonic == 0


def test_mesh_validation_on_host():
    st, _ = _case()
    dm = DeviceMesh(st, torch.device("cpu"))
    assert dm.s.flatTerrain == 1
    assert int(dm.t["cellsOnEdge"][: dm.ne].min()) >= 0
    assert int(dm.t["cellsOnEdge"][dm.ne].max()) == -1  # padding row: "none"
    assert int(dm.t["edgesOnCell"][: dm.nc].max()) < dm.ne
    bad = dict(st)
    coe = np.array(st["cellsOnEdge"])
    coe[3, 1] = dm.nc + 5  # out of range
    bad["cellsOnEdge"] = coe
    with pytest.raises(MpasMeshMismatchError):
        DeviceMesh(bad, torch.device("cpu"))
    bad = dict(st)
    eoc = np.array(st["edgesOnCell"])
    eoc[0, 0] = 0  # a "none" inside the live slots
    bad["edgesOnCell"] = eoc
    with pytest.raises(MpasMeshMismatchError):
        DeviceMesh(bad, torch.device("cpu"))
    bad = dict(st)
    bad["nVertLevels"] = 64
    with pytest.raises(MpasMeshMismatchError):
        DeviceMesh(bad, torch.device("cpu"))


def test_product_path_has_no_oracle_import():
    pkg = os.path.join(ROOT, "mpas_dycore_pytorch_amd")
    for path in glob.glob(os.path.join(pkg, "*.py")):
        text = open(path).read()
        assert "import oracle" not in text and "from oracle" not in text, path


def test_inverse_tables_transpose_the_slot_tables():
    """CSR inverses of advCellsForEdge / edgesOnEdge / cellsOnCell list every (row, slot) naming a
    target, in order."""
    from mpas_dycore_pytorch_amd.device import build_inverse

    st, _ = _case()
    dm = DeviceMesh(st, torch.device("cpu"))
    for table, count, n_rows, n_targets in (("advCellsForEdge", "nAdvCellsForEdge", dm.ne, dm.nc),
                                            ("edgesOnEdge", "nEdgesOnEdge", dm.ne, dm.ne),
                                            ("cellsOnCell", "nEdgesOnCell", dm.nc, dm.nc)):
        tab = dm.t[table][:n_rows].numpy()
        cnt = dm.t[count][:n_rows].numpy().astype(np.int64)
        start, src, slot = build_inverse(tab, cnt, n_targets)
        assert start[0] == 0 and start[-1] == int(cnt.sum()) and np.all(np.diff(start) >= 0)
        seen = set()
        for t in range(n_targets):
            pairs = list(zip(src[start[t]:start[t + 1]].tolist(), slot[start[t]:start[t + 1]].tolist()))
            assert pairs == sorted(pairs)
            for e, j in pairs:
                assert j < cnt[e] and tab[e, j] == t
                seen.add((e, j))
        assert len(seen) == int(cnt.sum())


def test_inverse_table_skips_entries_naming_no_target():
    """A slot naming a padding row (index n_targets) is left out of the transpose."""
    from mpas_dycore_pytorch_amd.device import build_inverse

    tab = np.array([[1, 3, 0], [2, 0, 3], [1, 2, 0]])
    cnt = np.array([3, 2, 2])
    start, src, slot = build_inverse(tab, cnt, 3)
    assert start.tolist() == [0, 2, 4, 6]
    assert list(zip(src.tolist(), slot.tolist())) == [(0, 2), (1, 1), (0, 0), (2, 0), (1, 0), (2, 1)]


def test_launch_shape_table():
    lib = N.load()
    assert lib.mpas_hip_tuning_get(b"k_dyn_edges_a") == 4
    assert lib.mpas_hip_tuning_set(b"k_dyn_edges_a", 2) == 0
    assert lib.mpas_hip_tuning_get(b"k_dyn_edges_a") == 2
    assert lib.mpas_hip_tuning_set(b"k_dyn_edges_a", 3) != 0  # not 1, 2, 4, 8 or 16
    assert lib.mpas_hip_tuning_set(b"k_dyn_edges_a", 0) == 0
    assert lib.mpas_hip_tuning_get(b"k_dyn_edges_a") == 4
    assert lib.mpas_hip_tuning_get_order(b"k_dyn_edges_a") == 0
    assert lib.mpas_hip_tuning_set_order(b"k_dyn_edges_a", 1) == 0
    assert lib.mpas_hip_tuning_get_order(b"k_dyn_edges_a") == 1
    assert lib.mpas_hip_tuning_get(b"k_dyn_edges_a") == 4  # the order leaves the shape alone
    assert lib.mpas_hip_tuning_set_order(b"k_dyn_edges_a", 2) != 0  # 0 or 1
    assert lib.mpas_hip_tuning_set_order(b"k_dyn_edges_a", 0) == 0
    assert lib.mpas_hip_tuning_get_order(b"k_dyn_edges_a") == 0


def test_run_simulation_jvp_validates_the_perturbation():
    """run_simulation.jvp rejects perturbations as the reference does (atmosphere_model.py:2063-2087),
    before anything runs."""
    from mpas_dycore_pytorch_amd.atmosphere_model import run_simulation

    state = {"u": torch.zeros(4, 3, dtype=torch.float64), "w": torch.zeros(2, 4, dtype=torch.float64),
             "nCells": 1, "raw": torch.zeros(2, dtype=torch.float64)}
    with pytest.raises(TypeError):
        run_simulation.jvp(state, {}, {"u": np.zeros((4, 3))})
    with pytest.raises(KeyError):
        run_simulation.jvp(state, {}, {"theta_m": torch.zeros(2, 3, dtype=torch.float64)})
    with pytest.raises(TypeError):
        run_simulation.jvp(state, {}, {"nCells": torch.zeros(1, dtype=torch.float64)})
    with pytest.raises(MpasConfigError):
        run_simulation.jvp(state, {}, {"u": torch.zeros(4, 2, dtype=torch.float64)})
    with pytest.raises(MpasConfigError):
        run_simulation.jvp(state, {}, {"u": torch.zeros(4, 3, dtype=torch.float32)})
    with pytest.raises(MpasUnsupportedCapabilityError):
        run_simulation.jvp(state, {}, {"raw": torch.zeros(2, dtype=torch.float64)})


def test_mixed_library_exports_its_header():
    """The mixed-precision library loads, exports what include/mpas_hip_mixed.h declares, and its
    state struct matches the binding."""
    text = open(os.path.join(ROOT, "include", "mpas_hip_mixed.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    declared = set(re.findall(r"\b(mpas_hip_mixed_\w+)\s*\(", text))
    lib = N.load_mixed()
    assert declared and all(hasattr(lib, n) for n in declared), declared
    assert declared <= set(N.MIXED_SIGNATURES)
    assert lib.mpas_hip_mixed_state_size() == C.sizeof(N.MixedStateStruct)
    # the edge/vertex fields halve: the mixed workspace is smaller than the plain one
    plain = N.load().mpas_hip_workspace_doubles(2562, 7680, 5120, 26, 1)
    mixed = lib.mpas_hip_mixed_workspace_doubles(2562, 7680, 5120, 26, 1)
    assert 0 < mixed < plain


def test_missing_library_fails_loudly():
    # A fresh interpreter whose binding points at a library that does not exist:
    # loading must raise the device error, not fall back to anything else.
    import subprocess
    import sys
    code = (
        "import mpas_dycore_pytorch_amd._native as N\n"
        "from mpas_dycore_pytorch_amd.errors import MpasDeviceError\n"
        "N.LIB_PATH = N.LIB_PATH + '.absent'\n"
        "N.MIXED_LIB_PATH = N.MIXED_LIB_PATH + '.absent'\n"
        "for f in (N.load, N.load_mixed):\n"
        "    try:\n"
        "        f()\n"
        "    except MpasDeviceError as e:\n"
        "        assert 'not built' in str(e)\n"
        "    else:\n"
        "        raise SystemExit(3)\n"
        "print('ok')\n"
    )
    root = __import__("pathlib").Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120, cwd=root)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "ok"
