"""Build the gfx950 dynamics libraries in tree.

``python -m mpas_dycore_pytorch_amd.build`` compiles every ``csrc/*.hip``
with hipcc for gfx950 and links ``libmpas_dycore_hip.so`` next to this file,
then compiles the forward-step sources a second time with ``-DMPAS_MIXED=1``
(edge- and vertex-located fields stored in binary32) into
``libmpas_dycore_hip_mixed.so``.  Floating-point contraction is off so every
expression rounds as written (the parity tests compare bit patterns with the
CPU float64 reference).
"""

from __future__ import annotations

import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB_NAME = "libmpas_dycore_hip.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
MIXED_LIB_PATH = os.path.join(HERE, "libmpas_dycore_hip_mixed.so")
BUILD_DIR = os.path.join(HERE, "_build")
# sources of the forward step (the mixed-precision library has no derivative kernels)
FORWARD_SOURCES = ("kernels_acoustic.hip", "kernels_column.hip", "kernels_diag.hip", "kernels_dyn.hip",
                   "kernels_init.hip", "kernels_transport.hip", "step.hip")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{INCLUDE}",
]


def sources() -> list[str]:
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def _compile(src: str, mixed: bool = False) -> str:
    out_dir = os.path.join(BUILD_DIR, "mixed") if mixed else BUILD_DIR
    obj = os.path.join(out_dir, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *CFLAGS, *(["-DMPAS_MIXED=1"] if mixed else []), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def _link(objs: list[str], path: str) -> None:
    tmp = path + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, path)


def build(verbose: bool = False) -> str:
    """Compile and link both libraries; always rebuilds."""
    os.makedirs(os.path.join(BUILD_DIR, "mixed"), exist_ok=True)
    srcs = sources()
    mixed_srcs = [os.path.join(CSRC, f) for f in FORWARD_SOURCES]
    jobs = [(f, False) for f in srcs] + [(f, True) for f in mixed_srcs]
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        objs = list(ex.map(lambda j: _compile(*j), jobs))
    _link(objs[: len(srcs)], LIB_PATH)
    _link(objs[len(srcs):], MIXED_LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}")
        print(f"built {MIXED_LIB_PATH}")
    return LIB_PATH


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
