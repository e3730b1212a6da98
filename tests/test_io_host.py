"""On-disk formats (mesh, state, restart, history) on the host: round trips, identities, tamper checks."""

from __future__ import annotations

import os

import numpy as np
import pytest
import torch

from conftest import srk3_case

from mpas_dycore_pytorch_amd import io as IO
from mpas_dycore_pytorch_amd.errors import MpasConfigError, MpasMissingFieldError


def test_container_round_trip_and_tamper_detection(tmp_path):
    p = tmp_path / "c.mpb"
    arrays = {"a": (np.arange(12, dtype=np.float64).reshape(3, 4), ("x", "y")),
              "i": (np.array([1, -2, 3], dtype=np.int32), ("x",))}
    digest = IO.write_container(p, "test", arrays, {"x": 3, "y": 4}, {"note": "t"})
    got, header = IO.read_container(p)
    assert header["payload_sha256"] == digest and header["attributes"]["note"] == "t"
    assert np.array_equal(got["a"], arrays["a"][0]) and got["i"].dtype == np.int32
    with pytest.raises(FileExistsError):
        IO.write_container(p, "test", arrays, {"x": 3, "y": 4})
    raw = bytearray(p.read_bytes())
    raw[-9] ^= 0x01  # one bit of the last array
    p.write_bytes(bytes(raw))
    with pytest.raises(MpasConfigError):
        IO.read_container(p)
    p.write_bytes(bytes(raw[:-16]))
    with pytest.raises(MpasConfigError):
        IO.read_container(p)
    with pytest.raises(MpasConfigError):
        IO.write_container(tmp_path / "bad.mpb", "t", {"a": (np.zeros((2, 2)), ("x", "y"))}, {"x": 3, "y": 2})
    with pytest.raises(MpasConfigError):
        IO.write_container(tmp_path / "nan.mpb", "t", {"a": (np.full(2, np.nan), ("x",))}, {"x": 2})


def test_mesh_and_state_round_trip(tmp_path):
    st, _, _ = srk3_case("l2_L20_tracer")
    digest = IO.write_mesh(tmp_path / "m.mpb", st)
    m = IO.read_mesh(tmp_path / "m.mpb")
    assert IO.mesh_digest(m) == digest == IO.mesh_digest(st)
    for k, v in st.items():
        if isinstance(v, np.ndarray) and v.dtype.kind in "biuf":
            assert np.array_equal(np.asarray(m[k]), v), k
    IO.write_state(tmp_path / "s.mpb", {"theta_m": st["theta_m"], "u": st["u"]}, mesh=st)
    s = IO.read_state(tmp_path / "s.mpb", mesh=st)
    assert torch.equal(s["theta_m"], torch.as_tensor(st["theta_m"]))
    other = dict(st)
    other["dvEdge"] = np.asarray(st["dvEdge"]) * 2.0
    with pytest.raises(MpasConfigError):
        IO.read_state(tmp_path / "s.mpb", mesh=other)


def test_restart_round_trip_identities_and_errors(tmp_path):
    st, _, meta = srk3_case("l2_L20_tracer")
    cfg = {"config_dt": meta["dt"], "config_len_disp": float(st["nominalMinDc"])}
    state = {k: torch.as_tensor(np.asarray(v)) for k, v in st.items()
             if isinstance(v, np.ndarray) and v.dtype.kind == "f" and k in IO.RESTART_REQUIRED + IO.RESTART_OPTIONAL}
    p = tmp_path / "r.mpb"
    IO.write_restart(state, st, p, "2000-01-01_06:00:00", config=cfg, dt=meta["dt"], model_time_seconds=21600.0)
    back, xtime = IO.read_restart(p, mesh=st, config=cfg, dt=meta["dt"])
    assert xtime == "2000-01-01_06:00:00" and back["_restart_model_time_seconds"] == 21600.0
    for k, v in state.items():
        assert back[k].shape == v.shape, k
        assert torch.equal(back[k][:-1], v[:-1]), k
        assert not bool(back[k][-1].any()), k  # zero padding row
    other = dict(st)
    other["nEdgesOnCell"] = np.asarray(st["nEdgesOnCell"]).copy()
    other["nEdgesOnCell"][0] += 0
    other["fEdge"] = np.asarray(st["fEdge"]) + 1.0
    with pytest.raises(MpasConfigError):
        IO.read_restart(p, mesh=other)
    with pytest.raises(MpasConfigError):
        IO.read_restart(p, dt=meta["dt"] * 2)
    with pytest.raises(MpasConfigError):
        IO.read_restart(p, config={**cfg, "config_len_disp": 1.0})
    with pytest.raises(MpasMissingFieldError):
        IO.write_restart({k: v for k, v in state.items() if k != "rw"}, st, tmp_path / "x.mpb", "2000-01-01_00:00:00")
    with pytest.raises(MpasConfigError):
        IO.write_restart(state, st, tmp_path / "y.mpb", "2000-13-01_00:00:00")
    with pytest.raises(FileExistsError):
        IO.write_restart(state, st, p, "2000-01-01_06:00:00")


def test_history_fields_and_file(tmp_path):
    st, _, _ = srk3_case("l2_L20_tracer")
    st = dict(st)
    nc, L = int(st["nCells"]), int(st["nVertLevels"])
    st.setdefault("pressure_base", np.full((nc + 1, L), 5.0e4))
    f = IO.history_fields(st, st)
    assert set(f) == set(IO.HISTORY_FIELDS)
    qv = np.asarray(st["scalars"])[:nc, :, 0]
    assert np.array_equal(f["theta"].numpy(), np.asarray(st["theta_m"])[:nc] / (1.0 + IO.RVORD * qv))
    assert np.array_equal(f["rho"].numpy(), np.asarray(st["rho_zz"])[:nc] * np.asarray(st["zz"])[:nc])
    assert np.array_equal(f["pressure"].numpy(), np.asarray(st["pressure_p"])[:nc] + st["pressure_base"][:nc])
    assert f["w"].shape == (nc, L + 1) and f["u"].shape == (int(st["nEdges"]), L)
    assert f["vorticity"].shape == (int(st["nVertices"]), L)
    p = tmp_path / "h.mpb"
    IO.write_history(p, st, st, xtime="2000-01-02_00:00:00", fields=["u", "theta", "pressure"])
    h = IO.read_history(p)
    assert set(h["fields"]) == {"u", "theta", "pressure", "time"}
    assert h["fields"]["theta"].shape == (1, nc, L) and h["dimensions"]["Time"] is None
    assert h["fields"]["time"][0] == 946771200.0 and h["attributes"]["units"]["pressure"] == "Pa"
    with pytest.raises(MpasConfigError):
        IO.history_fields(st, st, ["not_a_field"])
    with pytest.raises(MpasMissingFieldError):
        IO.history_fields({k: v for k, v in st.items() if k != "pressure_base"}, {}, ["pressure"])


def test_mesh_generator_command(tmp_path):
    from mpas_dycore_pytorch_amd.mesh import build_mesh

    out = tmp_path / "m1.mpb"
    assert IO._main(["mesh", "--level", "1", "--levels", "10", "--stretch", "uniform", "-o", str(out)]) == 0
    m = IO.read_mesh(out)
    assert IO.mesh_digest(m) == IO.mesh_digest(build_mesh(1, 10, stretch="uniform"))
    assert os.path.getsize(out) > 0
