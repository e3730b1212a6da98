"""Launch-shape autotuner for the wavefront-per-entity kernels.

Each wavefront-per-entity kernel (one entity per wavefront, lane = level)
can be launched with 1, 2, 4, 8 or 16 wavefronts per block; the best choice depends
on the kernel's register and LDS footprint, on the mesh size and on how the
launch spreads over the 8 XCDs.  :func:`autotune` times every candidate of
every kernel on the real workload with HIP events around that kernel's
launches (``mpas_hip_timer_arm``), one kernel at a time (coordinate search),
and installs the fastest in the library's launch table
(``mpas_hip_tuning_set``).  Tuning runs whole large steps, so it advances the
stepper's state: run it as part of the warm-up.
"""

from __future__ import annotations

import ctypes as C
import json
from typing import Iterable

import torch

from . import _native as N

# forward kernels launched one wavefront per entity, with their launches per large step
WAVE_KERNELS: dict[str, int] = {
    "k_dyn_edges_a": 9, "k_dyn_edge_adv": 9, "k_dyn_cells_b": 9, "k_dyn_cells_a": 9, "k_dyn_kdiff": 3,
    "k_acoustic_edges": 12, "k_recover_edges": 9, "k_recover_w": 9, "k_diag_edges2": 9, "k_diag_cells": 9,
    "k_diag_vertices": 9,
}
CANDIDATES = (1, 2, 4, 8, 16)


def current() -> dict[str, int]:
    lib = N.load()
    return {k: int(lib.mpas_hip_tuning_get(k.encode())) for k in WAVE_KERNELS}


def install(table: dict[str, int]) -> None:
    lib = N.load()
    for k, w in table.items():
        N.check(lib.mpas_hip_tuning_set(k.encode(), int(w)), "tuning_set")


def autotune(stepper, kernels: Iterable[str] | None = None, candidates=CANDIDATES, steps: int = 2) -> dict:
    """Tune the launch shape of each kernel on ``stepper``'s workload.

    Returns ``{kernel: {"waves": best, "avg_ms": {waves: ms}}}`` and leaves
    the best shapes installed."""
    lib = N.load()
    device = stepper.device
    result = {}
    for name in (kernels or WAVE_KERNELS):
        per_step = WAVE_KERNELS.get(name, 16)
        times = {}
        for w in candidates:
            N.check(lib.mpas_hip_tuning_set(name.encode(), int(w)), "tuning_set")
            torch.cuda.synchronize(device)
            N.check(lib.mpas_hip_timer_arm(name.encode(), per_step * steps + 16), "timer_arm")
            stepper.step(steps)
            tot, n = C.c_double(), C.c_int32()
            N.check(lib.mpas_hip_timer_read(C.byref(tot), C.byref(n)), "timer_read")
            if n.value > 0:
                times[w] = tot.value / n.value
        best = min(times, key=times.get) if times else 4
        N.check(lib.mpas_hip_tuning_set(name.encode(), int(best)), "tuning_set")
        result[name] = {"waves": best, "avg_ms": {str(k): round(v, 5) for k, v in times.items()}}
    return result


def save(path: str, table: dict[str, int]) -> None:
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def load(path: str) -> dict[str, int]:
    with open(path) as f:
        table = {k: int(v) for k, v in json.load(f).items()}
    install(table)
    return table
