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
    assert BM.step_bytes(NC7, NE7, NV7, 55, 1) == 134188386048
    assert BM.kernel_bytes("k_acoustic_cells", NC7, NE7, NV7, 55, 1) == 2131911032
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
