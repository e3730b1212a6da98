"""Host-side checks of the measurement layer: the one-GPU guard of ``bench.py``,
the algorithmic byte model behind its roofline figures, and that every kernel
name the byte model and the autotuner refer to is a kernel of the HIP sources."""
from __future__ import annotations

import pathlib
import re
import subprocess
import sys

from mpas_dycore_pytorch_amd import autotune, bytes_model as BM

ROOT = pathlib.Path(__file__).resolve().parents[1]
CSRC = ROOT / "mpas_dycore_pytorch_amd" / "csrc"

# icosahedral level-7 mesh: nC = 10*4^7 + 2, nE = 30*4^7, nV = 20*4^7
NC7, NE7, NV7 = 10 * 4**7 + 2, 30 * 4**7, 20 * 4**7


def _hip_kernels() -> set[str]:
    names = set()
    for p in CSRC.glob("*.hip"):
        names.update(re.findall(r"void\s+(k_\w+)\s*\(", p.read_text()))
    for p in CSRC.glob("*.hpp"):
        names.update(re.findall(r"KERNEL\((k_\w+)\s*,", p.read_text()))
    return names


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_array_classes_are_their_closed_forms():
    L, S = 55, 2
    assert BM.class_bytes("C", NC7, NE7, NV7, L, S) == NC7 * L * 8
    assert BM.class_bytes("W", NC7, NE7, NV7, L, S) == NC7 * (L + 1) * 8
    assert BM.class_bytes("E", NC7, NE7, NV7, L, S) == NE7 * L * 8
    assert BM.class_bytes("V", NC7, NE7, NV7, L, S) == NV7 * L * 8
    assert BM.class_bytes("ES", NC7, NE7, NV7, L, S) == NE7 * L * S * 8


def test_level7_step_bytes():
    # the figure the level-7 x 55, one-scalar step roofline is quoted against
    assert BM.step_bytes(NC7, NE7, NV7, 55, 1) == 123135044400
    assert BM.kernel_bytes("k_acoustic_cells", NC7, NE7, NV7, 55, 1) == 2131911032
    # a first acoustic substep reads no rho_pp, rtheta_pp (2 C) and no rw_p, wwAvg (2 W)
    first = BM.kernel_bytes("k_acoustic_cells", NC7, NE7, NV7, 55, 1, "first")
    assert 2131911032 - first == 2 * NC7 * 55 * 8 + 2 * NC7 * 56 * 8
    # RK3 with 2 acoustic substeps: per split step the stages take 1, 1, 2 substeps
    sched = BM.launch_schedule("k_acoustic_cells")
    assert sched == ["first", "first", "first", "later"] * 3
    assert BM.launch_schedule("k_acoustic_cells", order=2) == ["first", "first", "first", "later"] * 3
    assert BM.launch_schedule("k_acoustic_cells", ns=4) == ["first", "first", "later", "first", "later",
                                                             "later", "later"] * 3
    assert BM.launch_schedule("k_dyn_edges_a") == ["rk1", "rk23", "rk23"] * 3
    # launch-weighted bytes of one step's 12 acoustic cell passes: 9 first + 3 later
    assert BM.schedule_bytes("k_acoustic_cells", sched, NC7, NE7, NV7, 55, 1) == 9 * first + 3 * 2131911032
    parts = BM.step_bytes_breakdown(NC7, NE7, NV7, 55, 1)
    assert sum(parts.values()) == BM.step_bytes(NC7, NE7, NV7, 55, 1)
    # every function's bytes grow linearly with the level count (a column of L
    # levels per entity) up to the L+1 interface arrays and the connectivity
    b1, b2, b3 = (BM.step_bytes(NC7, NE7, NV7, L, 1) for L in (10, 20, 30))
    assert b2 - b1 == b3 - b2


def test_modelled_and_tuned_kernels_exist():
    hip = _hip_kernels()
    assert "k_acoustic_cells" in hip
    missing = sorted(set(BM.KERNELS) - hip) + sorted(set(autotune.WAVE_KERNELS) - hip)
    assert not missing, f"kernel names with no HIP kernel: {missing}"


def test_profile_summary_labels_dispatches_by_variant(tmp_path):
    """The rocprofv3 summarizer charges each dispatch its own variant's bytes."""
    from mpas_dycore_pytorch_amd import profile_summary as PS

    d = tmp_path / "run" / "host"
    d.mkdir(parents=True)
    rows = ["Kind,Dispatch_Id,Kernel_Name,Start_Timestamp,End_Timestamp"]
    t = 1000
    for i in range(24):  # two large steps of k_acoustic_cells, plus another kernel between
        dur = 500_000 if BM.launch_schedule("k_acoustic_cells")[i % 12] == "first" else 600_000
        rows.append(f"KERNEL_DISPATCH,{2 * i},\"mpas::k_acoustic_cells(mpas_hip_mesh, ...)\",{t},{t + dur}")
        rows.append(f"KERNEL_DISPATCH,{2 * i + 1},mpas::k_acoustic_cells_other(int),{t},{t + 7}")
        t += 1_000_000
    (d / "kt_kernel_trace.csv").write_text("\n".join(rows) + "\n")
    text = PS.summarize(str(tmp_path / "run"), "k_acoustic_cells", 7, 55, 1)
    first = BM.kernel_bytes("k_acoustic_cells", NC7, NE7, NV7, 55, 1, "first")
    later = BM.kernel_bytes("k_acoustic_cells", NC7, NE7, NV7, 55, 1, "later")
    assert "| first | 18 | 500.0 |" in text and "| later | 6 | 600.0 |" in text
    whole = (18 * first + 6 * later) / (18 * 500e-6 + 6 * 600e-6) / 8e12
    assert f"{whole:.4f} |" in text.splitlines()[-1]
