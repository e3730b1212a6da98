"""Launch-shape autotuner for the wavefront-per-entity kernels.

Each wavefront-per-entity kernel (one entity per wavefront, lane = level)
has two launch choices: 1, 2, 4, 8 or 16 wavefronts per block, and the block
order -- contiguous chunks of entities dealt to the 8 XCDs (neighbours share an
L2; the default) or the hardware's round-robin order.  The best choice depends
on the kernel's register and LDS footprint, its gather pattern and the mesh
size.  :func:`autotune` times every candidate of every kernel on the real
workload with HIP events around that kernel's launches
(``mpas_hip_timer_arm``), one kernel and one dimension at a time (coordinate
search: the block shape first, then the order under the best shape), and
installs the fastest in the library's launch table (``mpas_hip_tuning_set``,
``mpas_hip_tuning_set_order``).  Tuning runs whole large steps, so it advances
the stepper's state: run it as part of the warm-up.
"""

from __future__ import annotations

import ctypes as C
import json
from typing import Iterable

import torch

from . import _native as N

# forward kernels launched one wavefront per entity, with their launches per large step
# (k_divdamp is absent: the fused step folds divergence damping into k_acoustic_edges)
WAVE_KERNELS: dict[str, int] = {
    "k_dyn_edges_a": 9, "k_dyn_edge_adv": 9, "k_dyn_cells_b": 9, "k_dyn_cells_a": 9, "k_dyn_kdiff": 3,
    "k_acoustic_edges": 12, "k_recover_edges": 9, "k_recover_w": 9, "k_diag_edges2": 9, "k_diag_cells": 9,
    "k_diag_vertices": 9, "k_dyn_edges_b": 3, "k_dyn_del4_vc": 3,
    "k_scalar_hflux": 2, "k_scalar_update": 2, "k_mono_hflux": 1, "k_mono_density": 1, "k_mono_bounds": 1,
    "k_mono_scale": 1, "k_mono_final": 1,
    # the variants with the option terms: launched only when an option is on
    "k_dyn_edges_a_opt": 9, "k_dyn_cells_b_opt": 9,
}
# adjoint kernels launched one wavefront per entity (at most 4 per block), with their
# launches per large step of a VJP
ADJOINT_KERNELS: dict[str, int] = {
    "k_adj_dyn_a3w": 9, "k_adj_dyn_a3_cells": 9, "k_adj_dyn_a3_edges": 9, "k_adj_dyn_a1": 9,
    "k_adj_dyn_a2_cells": 3, "k_adj_dyn_a2_edges": 3, "k_adj_dyn_a6": 3, "k_adj_dyn_a7": 9,
    "k_adj_dyn_fwd_edges": 9, "k_adj_dyn_fwd_cells": 9, "k_adj_diag_g1": 10, "k_adj_diag_g2": 10,
    "k_adj_diag_g3": 10, "k_adj_diag_g4": 10, "k_adj_acoustic_edge_flux": 12, "k_adj_acoustic_theta": 12,
    "k_adj_recover_cell": 9, "k_adj_recover_edge": 9, "k_adj_recover_w": 9, "k_adj_smlstep_edges": 9,
    "k_adj_smlstep_cells": 9,
}
# tangent kernels launched the same way, with their launches per large step of a JVP
TANGENT_KERNELS: dict[str, int] = {
    "k_tan_dyn_j1": 9, "k_tan_dyn_j2": 9, "k_tan_dyn_j3": 9, "k_tan_dyn_j4": 9, "k_tan_dyn_j5": 9,
    "k_tan_dyn_fwd_edges": 9, "k_tan_dyn_fwd_cells": 9, "k_tan_diag_edge": 10, "k_tan_diag_cell": 10,
    "k_tan_diag_vertex": 10, "k_tan_acoustic_edge": 12, "k_tan_recover_cell": 9, "k_tan_recover_edge": 9,
    "k_tan_recover_w": 9, "k_tan_smlstep": 9, "k_tan_scal_edge": 2, "k_tan_scal_cell": 2,
}
ADJOINT_CANDIDATES = (1, 2, 4)
CANDIDATES = (1, 2, 4, 8, 16)
ORDERS = {0: "xcd", 1: "linear"}


def current() -> dict[str, dict[str, int]]:
    lib = N.load()
    return {k: {"waves": int(lib.mpas_hip_tuning_get(k.encode())),
                "order": int(lib.mpas_hip_tuning_get_order(k.encode()))} for k in WAVE_KERNELS}


def install(table: dict) -> None:
    """Install ``{kernel: waves}`` or ``{kernel: {"waves": w, "order": o}}``."""
    lib = N.load()
    for k, v in table.items():
        entry = v if isinstance(v, dict) else {"waves": v}
        N.check(lib.mpas_hip_tuning_set(k.encode(), int(entry.get("waves", 0))), "tuning_set")
        N.check(lib.mpas_hip_tuning_set_order(k.encode(), int(entry.get("order", 0))), "tuning_set_order")


def _time_kernel(lib, run, name: str, capacity: int) -> float | None:
    """Average duration (ms) of ``name``'s launches during ``run()``."""
    torch.cuda.synchronize()
    N.check(lib.mpas_hip_timer_arm(name.encode(), capacity), "timer_arm")
    run()
    tot, n = C.c_double(), C.c_int32()
    N.check(lib.mpas_hip_timer_read(C.byref(tot), C.byref(n)), "timer_read")
    return tot.value / n.value if n.value > 0 else None


def autotune(stepper, kernels: Iterable[str] | None = None, candidates=CANDIDATES, steps: int = 2,
             orders=tuple(ORDERS)) -> dict:
    """Tune the launch shape, then the block order, of each kernel on ``stepper``'s workload.

    Returns ``{kernel: {"waves": best, "order": best, "avg_ms": {waves: ms},
    "order_ms": {order: ms}}}`` and leaves the best choices installed."""
    return tune(lambda: stepper.step(steps), {k: WAVE_KERNELS.get(k, 16) * steps for k in (kernels or WAVE_KERNELS)},
                candidates, orders, lib=getattr(stepper, "lib", None))


def autotune_reverse(run_vjp, kernels: Iterable[str] | None = None, candidates=ADJOINT_CANDIDATES,
                     orders=tuple(ORDERS), steps: int = 1) -> dict:
    """The same search over the adjoint kernels, timed on ``run_vjp()`` (a VJP of
    ``steps`` large steps); as :func:`autotune`, it leaves the best choices installed."""
    return tune(run_vjp, {k: ADJOINT_KERNELS.get(k, 16) * steps for k in (kernels or ADJOINT_KERNELS)},
                candidates, orders)


def autotune_tangent(run_jvp, kernels: Iterable[str] | None = None, candidates=ADJOINT_CANDIDATES,
                     orders=tuple(ORDERS), steps: int = 1) -> dict:
    """The same search over the tangent kernels, timed on ``run_jvp()`` (a JVP of
    ``steps`` large steps)."""
    return tune(run_jvp, {k: TANGENT_KERNELS.get(k, 16) * steps for k in (kernels or TANGENT_KERNELS)},
                candidates, orders)


def tune(run, kernels: dict[str, int], candidates, orders=tuple(ORDERS), lib=None) -> dict:
    """Coordinate search over ``kernels`` (name: launches per ``run()``): each kernel's
    shape under the default order, then its order under the best shape.  ``lib`` is the
    library whose kernels ``run()`` launches (the FP64 one by default; the mixed-precision
    library keeps a table of its own)."""
    lib = lib or N.load()
    result = {}
    for name, launches in kernels.items():
        capacity = launches + 16
        N.check(lib.mpas_hip_tuning_set_order(name.encode(), 0), "tuning_set_order")
        times = {}
        for w in candidates:
            N.check(lib.mpas_hip_tuning_set(name.encode(), int(w)), "tuning_set")
            t = _time_kernel(lib, run, name, capacity)
            if t is not None:
                times[w] = t
        if not times:  # not launched on this workload: keep the defaults
            N.check(lib.mpas_hip_tuning_set(name.encode(), 0), "tuning_set")
            continue
        best = min(times, key=times.get)
        N.check(lib.mpas_hip_tuning_set(name.encode(), int(best)), "tuning_set")
        order_times = {0: times[best]} if best in times else {}
        for o in orders:
            if o in order_times:
                continue
            N.check(lib.mpas_hip_tuning_set_order(name.encode(), int(o)), "tuning_set_order")
            t = _time_kernel(lib, run, name, capacity)
            if t is not None:
                order_times[o] = t
        best_order = min(order_times, key=order_times.get) if order_times else 0
        N.check(lib.mpas_hip_tuning_set_order(name.encode(), int(best_order)), "tuning_set_order")
        result[name] = {"waves": best, "order": best_order,
                        "avg_ms": {str(k): round(v, 5) for k, v in times.items()},
                        "order_ms": {ORDERS[k]: round(v, 5) for k, v in order_times.items()}}
    return result


def save(path: str, table: dict) -> None:
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def load(path: str) -> dict:
    with open(path) as f:
        table = json.load(f)
    install(table)
    return table
