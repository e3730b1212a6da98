#!/usr/bin/env python3
"""Throughput of the GPU split-explicit large step (``atm_srk3``) on one MI355X.

Workload (default): one large step = the reference's RK3 split-explicit
step (3 dynamics substeps x 3 RK stages, 2 acoustic substeps, monotonic FCT
transport of the water-vapour slot) on a synthetic icosahedral-Voronoi mesh
of level 7 (163842 cells, 491520 edges, 327680 vertices) with 55 levels
stretched quadratically to 30 km, flat terrain, dt = 300 s, FP64.  The
state is the seeded balanced jet with a smoothed theta perturbation,
prepared on the GPU.  Inputs are resident in HBM before timing starts.

Prints one JSON line: throughput in cell-level-steps/s, the roofline of the
dominant kernel (its average duration measured with HIP events on the launch
stream during the timed steps, against its algorithmic bytes per launch),
the whole-step algorithmic bandwidth, and a CPU baseline (the oracle, a
single-threaded C restatement of the reference, timed on a bounded sample).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-level-steps/s at 163842x55 FP64; VJP/fwd cost ratio; % HBM roofline"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def prepare(level: int, n_levels: int, stretch: str, n_tracers: int, dt: float, seed: int, device, m=None):
    """Mesh (built unless given), raw synthetic fields and the cold-start chain (diagnostics on the GPU)."""
    import numpy as np

    from mpas_dycore_pytorch_amd import initial_state as IS
    from mpas_dycore_pytorch_amd import kernels as K
    from mpas_dycore_pytorch_amd import mesh as M
    from mpas_dycore_pytorch_amd.device import DeviceMesh

    t0 = time.time()
    if m is None:
        m = M.build_mesh(level, n_levels, stretch=stretch)
    raw = IS.raw_state(m, seed=seed, n_tracers=n_tracers)
    base = {**m, **raw, **IS.init_coupled_diagnostics(m, raw)}
    dm = DeviceMesh(base, device)

    def diag(u, h):
        d = K.solve_diagnostics(dm, dt, 0.5, None, u, h)
        return {k: v.cpu().numpy() for k, v in d.items()}

    def recon(u):
        return {k: v.cpu().numpy() for k, v in K.reconstruct(dm, u).items()}

    st = IS.prepare_state(m, raw, diagnostics=diag, reconstruct=recon)
    del dm
    log(f"prepared level {level} x {n_levels} ({m['nCells']} cells) in {time.time() - t0:.1f} s")
    return m, st, np


def traffic_from_profile(kernel: str, workload: str):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    mm = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(mm.group(1)) if mm else None


def cpu_worker(level: int, n_levels: int, dt: float, steps: int, seed: int) -> None:
    """One single-threaded oracle run of the CPU sample; prints its work and time as JSON."""
    from mpas_dycore_pytorch_amd import initial_state as IS
    from mpas_dycore_pytorch_amd import mesh as M
    from oracle import oracle as O
    from oracle import srk3 as OS

    m = M.build_mesh(level, n_levels, stretch="quadratic")
    raw = IS.raw_state(m, seed=seed, n_tracers=0)
    om = O.Mesh({**m, **IS.init_coupled_diagnostics(m, raw), **raw})
    st = IS.prepare_state(m, raw, diagnostics=lambda u, h: O.solve_diagnostics(om, dt, 0.5, None, u, h),
                          reconstruct=lambda u: O.reconstruct(om, u))
    om = O.Mesh(st)
    cur = st
    t0 = time.perf_counter()
    for step in range(steps):
        cur = OS.srk3(cur, om, dt, len_disp=float(m["nominalMinDc"]), diagnostics_current=(step > 0))
    el = time.perf_counter() - t0
    print(json.dumps({"units": int(m["nCells"]) * n_levels * steps, "seconds": el, "cells": int(m["nCells"])}),
          flush=True)


def cpu_baseline(level: int, n_levels: int, dt: float, steps: int, seed: int) -> dict:
    """The oracle (C restatement of the reference kernels) on every host core this process may use:
    one single-threaded run of the sample per core, started together as child processes (the GPU
    is not touched by them); throughput = all the runs' work over the slowest run's time."""
    import subprocess

    cores = len(os.sched_getaffinity(0))
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-worker", str(level), str(n_levels), str(dt),
           str(steps), str(seed)]
    env = dict(os.environ, OMP_NUM_THREADS="1")
    procs = [subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env)
             for _ in range(cores)]
    runs = []
    for p in procs:
        out, err = p.communicate(timeout=600)
        if p.returncode != 0:
            raise RuntimeError(f"cpu baseline worker failed: {err[-2000:]}")
        runs.append(json.loads(out.strip().splitlines()[-1]))
    units = sum(r["units"] for r in runs)
    slowest = max(r["seconds"] for r in runs)
    one = runs[0]
    ref = reference_cpu_timing()
    return {"value": units / slowest, "unit": "cell-level-steps/s", "cores": cores, "kind": "port",
            "sample": (f"oracle srk3 (single-threaded C kernels, Python orchestration), {cores} concurrent runs, "
                       f"one per host core, each level {level} ({one['cells']} cells) x {n_levels} levels, "
                       f"dt {dt:g} s, {steps} large steps; slowest run {slowest:.1f} s, single-run throughput "
                       f"{one['units'] / one['seconds']:.3e}; the unit is per cell-level, so it carries to the "
                       f"level-7 workload by the cell count (not measured at level 7)"),
            "reference_measured": ref}


def reference_cpu_timing() -> dict | None:
    """The reference's own CPU float64 step time, measured in the build container when its golden
    summary was written (tests/golden/summary_cfg1_l4_L26.json)."""
    path = os.path.join(ROOT, "tests", "golden", "summary_cfg1_l4_L26.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        summ = json.load(f)
    secs = summ["reference_seconds_per_step"]
    n = 2562 if summ["level"] == 4 else None
    best = min(secs)
    return {"seconds_per_step": secs, "threads": summ.get("torch_threads"), "level": summ["level"],
            "levels": summ["nVertLevels"], "cell_level_steps_per_s": (n * summ["nVertLevels"] / best) if n else None,
            "where": "reference atm_srk3 on the build container's CPU (Intel Xeon 2.7 GHz), torch float64"}


def vjp_leg(m, stepper, config, dt: float, steps: int, fwd_ms_per_step: float, device, jvp: bool = True,
            tune: bool = False) -> dict:
    # m: the prepared mapping (mesh, base state and initial fields, host arrays)
    """Forward + VJP of ``steps`` large steps from the stepper's current state.

    Loss = sum over cells of area * sum_k (theta_m - theta_m0)^2 after the
    steps; gradients with respect to u, w, theta_m, rho_zz through the
    hand-written adjoint kernels, one checkpoint per large step, acoustic
    substeps recomputed in the backward (the last step keeps the tape of its
    forward).  The ratio divides the wall time of
    forward + VJP by ``steps`` plain forward steps."""
    import torch

    from mpas_dycore_pytorch_amd.atmosphere_model import run_simulation
    from mpas_dycore_pytorch_amd.boundary import mark_complete_step_boundary

    nc = int(m["nCells"])
    # the stepper's state (device tensors) as a bound step boundary of this mesh, config and dt
    state = {k: v.clone() for k, v in stepper.state_dict().items()}
    mark_complete_step_boundary(state, mesh=m, config=config, dt=dt)
    area = torch.as_tensor(1.0 / m["invAreaCell"][:nc], device=device)
    th0 = state["theta_m"].detach().clone()

    def objective(out):
        d = out["theta_m"][:nc] - th0[:nc]
        return torch.sum(area[:, None] * d * d)

    keys = ("u", "w", "theta_m", "rho_zz")
    # untimed warm-up: device mesh, inverse tables and allocator pools
    run_simulation.vjp(state, m, objective, track_keys=keys, config=config, dt=dt, nsteps=1, device=device)
    tuning = None
    if tune:
        # the adjoint kernels' launch shapes and block orders, timed on 1-step VJPs (untimed)
        from mpas_dycore_pytorch_amd import autotune as AT

        t_tune = time.time()
        tuning = AT.autotune_reverse(lambda: run_simulation.vjp(state, m, objective, track_keys=keys, config=config,
                                                                dt=dt, nsteps=1, device=device))
        log(f"autotuned {len(tuning)} adjoint kernels in {time.time() - t_tune:.1f} s: "
            + ", ".join(f"{k}={v['waves']}{'L' if v['order'] else ''}" for k, v in tuning.items()))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    grads = run_simulation.vjp(state, m, objective, track_keys=keys, config=config, dt=dt, nsteps=steps,
                               device=device)
    torch.cuda.synchronize()
    total_s = time.perf_counter() - t0
    for k, g in grads.items():
        if not bool(torch.isfinite(g).all()):
            raise RuntimeError(f"non-finite gradient of {k}")
    peak = torch.cuda.max_memory_allocated(device) / 2**30
    out = {"steps": steps, "fwd_plus_vjp_ms_per_step": round(1e3 * total_s / steps, 2),
           "ratio": round(1e3 * total_s / steps / fwd_ms_per_step, 3), "peak_hbm_gib": round(peak, 1),
           "checkpoint": "per large step (acoustic substeps recomputed); the last step keeps its forward tape",
           "adjoint_launch": ({k: f"{v['waves']}{' linear' if v['order'] else ''}" for k, v in tuning.items()}
                              if tuning else "default 4 waves")}
    if not jvp:
        return out
    # forward mode on the same state: seeded tangents N(0, 0.1) on u and N(0, 0.01) on w (interior
    # interfaces), padding rows zero, through run_simulation.jvp
    gen = torch.Generator(device="cpu").manual_seed(20261001)
    tu = (0.1 * torch.randn(tuple(state["u"].shape), generator=gen, dtype=torch.float64)).to(device)
    tw = (0.01 * torch.randn(tuple(state["w"].shape), generator=gen, dtype=torch.float64)).to(device)
    tu[-1] = 0.0
    tw[-1] = 0.0
    tw[:, 0] = 0.0
    tw[:, -1] = 0.0
    run_simulation.jvp(state, m, {"u": tu, "w": tw}, config=config, dt=dt, nsteps=1, device=device)
    tan_tuning = None
    if tune:
        from mpas_dycore_pytorch_amd import autotune as AT

        t_tune = time.time()
        tan_tuning = AT.autotune_tangent(lambda: run_simulation.jvp(state, m, {"u": tu, "w": tw}, config=config,
                                                                    dt=dt, nsteps=1, device=device))
        log(f"autotuned {len(tan_tuning)} tangent kernels in {time.time() - t_tune:.1f} s: "
            + ", ".join(f"{k}={v['waves']}{'L' if v['order'] else ''}" for k, v in tan_tuning.items()))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tan = run_simulation.jvp(state, m, {"u": tu, "w": tw}, config=config, dt=dt, nsteps=steps, device=device)
    torch.cuda.synchronize()
    jvp_s = time.perf_counter() - t0
    if not bool(torch.isfinite(tan["theta_m"]).all()):
        raise RuntimeError("non-finite theta_m tangent")
    out.update({"fwd_plus_jvp_ms_per_step": round(1e3 * jvp_s / steps, 2),
                "jvp_ratio": round(1e3 * jvp_s / steps / fwd_ms_per_step, 3),
                "tangent_launch": ({k: f"{v['waves']}{' linear' if v['order'] else ''}" for k, v in tan_tuning.items()}
                                   if tan_tuning else "default 4 waves")})
    return out


ENSEMBLE_DT = {4: 1800.0, 5: 900.0, 6: 450.0, 7: 300.0, 8: 150.0}


def mixed_leg(st, config, dt: float, steps: int, err_steps: int, fp64_ms_per_step: float, reorder: bool,
              device, tune: bool = False) -> dict:
    """The mixed-precision step (binary32 edge/vertex diagnostics, tendencies and fluxes; binary32
    arithmetic with compensated sums in the PV flux, diagnostics and tracer transport) on the same
    workload: time per step,
    speed-up over the FP64 step, and per-field relative L2/Linf error against the FP64 GPU step after
    ``err_steps`` large steps from the same state."""
    import torch

    from mpas_dycore_pytorch_amd.time_integration import Srk3Stepper

    mix = Srk3Stepper(st, st, config, dt, device=device, reorder=reorder, precision="mixed")
    ref = Srk3Stepper(st, st, config, dt, device=device, reorder=reorder)
    mix.step(err_steps)
    ref.step(err_steps)
    torch.cuda.synchronize()
    a, b = ref.state_dict(), mix.state_dict()
    errors = {}
    for k in ("u", "w", "theta_m", "rho_zz", "exner", "pressure_p", "rtheta_p", "pv_edge", "ke", "scalars"):
        x, y = a[k][:-1], b[k][:-1]
        if x.numel() == 0:
            continue
        d = x - y
        errors[k] = {"rel_l2": float(torch.linalg.vector_norm(d) / max(float(torch.linalg.vector_norm(x)), 1e-300)),
                     "rel_linf": float(d.abs().max() / max(float(x.abs().max()), 1e-300))}
    del ref, a, b
    if tune:
        # the mixed library's own launch table, tuned as the FP64 one was (untimed)
        from mpas_dycore_pytorch_amd import autotune as AT

        AT.autotune(mix)
    stream = torch.cuda.current_stream(device)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    mix.step(steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    units = mix.mesh.nc * mix.mesh.L
    return {"ms_per_step": ms, "value": units / (ms / 1e3), "speedup_vs_fp64": fp64_ms_per_step / ms,
            "error_steps": err_steps, "errors_vs_fp64_gpu": errors,
            "launch": "autotuned (own table)" if tune else "default 4 waves",
            "storage": "binary32: v, pv_edge, rho_edge, vorticity, edge tendencies, cqu, tracer fluxes, edge/"
                       "vertex scratch; binary64: u and the mass fluxes, every cell/interface field, the column "
                       "solve, exner recovery and base state",
            "arithmetic": "binary32 with Neumaier-compensated slot sums: PV flux and KE gradient of the u "
                          "tendency, KE / tangential velocity / APVM of the diagnostic solve, tracer transport; "
                          "binary64: mass, theta and w flux divergences, divergence, vorticity, column work"}


def ensemble_bench(args, device) -> int:
    """Uneven ensemble: members on several mesh levels, each with its own seeded
    theta perturbation and the level's dt.  Batched (one grid per kernel per
    mesh, HIP graphs) against the members stepped one by one."""
    import torch

    from mpas_dycore_pytorch_amd.ensemble import EnsembleStepper
    from mpas_dycore_pytorch_amd.time_integration import Srk3Stepper

    spec = [tuple(int(x) for x in part.split("x")) for part in args.ensemble.split(",")]
    members = []
    for level, count in spec:
        dt = ENSEMBLE_DT[level]
        m, st0, _ = prepare(level, args.levels, "quadratic", 1, dt, args.seed, device)
        mesh = {**m, **{k: st0[k] for k in ("rho_base", "rtheta_base", "exner_base") if k in st0}}
        cfg = {"config_time_integration_order": 3, "config_len_disp": float(m["nominalMinDc"])}
        members.append((st0, mesh, dt, cfg))
        for i in range(1, count):
            _, st, _ = prepare(level, args.levels, "quadratic", 1, dt, args.seed + 1000 * level + i, device, m=m)
            members.append((st, mesh, dt, cfg))
        log(f"level {level}: {count} members prepared")
    cell_levels = sum(int(mem[1]["nCells"]) * args.levels for mem in members)
    steps = args.steps

    ens = EnsembleStepper(members, reorder=bool(args.reorder))
    ens.step(1)
    ens.capture_graphs()
    ens.step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step(steps)
    torch.cuda.synchronize()
    batched_s = time.perf_counter() - t0
    batched = cell_levels * steps / batched_s
    log(f"batched: {1e3 * batched_s / steps:.2f} ms per ensemble step, {batched:.4e} cell-level-steps/s")
    del ens
    torch.cuda.empty_cache()

    singles = []
    for st, mesh, dt, cfg in members:
        sp = Srk3Stepper(st, mesh, cfg, dt, device=device, reorder=bool(args.reorder))
        sp.step(1)
        sp.capture_graph()
        singles.append(sp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for sp in singles:
        sp.step(steps)
    torch.cuda.synchronize()
    single_s = time.perf_counter() - t0
    one_by_one = cell_levels * steps / single_s
    log(f"one by one: {1e3 * single_s / steps:.2f} ms per ensemble step, {one_by_one:.4e} cell-level-steps/s")
    out = {
        "metric": "cell-level-steps/s of an uneven ensemble (batched vs one by one)", "value": batched,
        "unit": "cell-level-steps/s", "n_gpus": 1, "steps": steps, "warmup": 2,
        "ms_per_step": 1e3 * batched_s / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "uneven ensemble " + ", ".join(f"{c} x level {lv} (dt {ENSEMBLE_DT[lv]:g} s)"
                                                                for lv, c in spec) + f", {args.levels} levels",
                   "members": len(members), "cell_levels_per_step": cell_levels},
        "one_by_one": one_by_one, "speedup_vs_one_by_one": round(batched / one_by_one, 3),
    }
    print(json.dumps(out), flush=True)
    return 0


def layout_probe(args, lib, N) -> int:
    """Column streams over as many columns as the level has edges, one wavefront
    per column, at row pitch L (the device layout), L + 1 (whole 64-byte
    sectors, the spare lane writing the padding) and 64 (whole 128-byte lines):
    a three-read two-write stream and a gather of the own and six neighbour
    columns with one write.  Algorithmic bytes count the L live levels only."""
    n = 30 * 4 ** args.level
    L = args.levels
    rows = []
    for variant, shape, arrays in ((0, "stream: 3 column reads + 2 column writes", 5),
                                   (1, "gather: own + 6 neighbour columns read, 1 written", 2)):
        for pitch in (L, L + 1, 64):
            ms = C.c_double()
            N.check(lib.mpas_hip_layout_probe(n, L, pitch, variant, 20, C.byref(ms)), "mpas_hip_layout_probe")
            alg = n * L * 8 * arrays
            rows.append({"shape": shape, "pitch": pitch, "ms": round(ms.value, 4),
                         "algorithmic_bytes": alg, "achieved_gbs": round(alg / ms.value / 1e6, 1)})
            log(f"{shape}, pitch {pitch}: {ms.value:.4f} ms, {alg / ms.value / 1e6:.1f} GB/s")
    print(json.dumps({"metric": "column-stream time by level pitch", "columns": n, "levels": L, "rows": rows}))
    return 0


def boundary_bench(args, device) -> int:
    """``atm_srk3`` called as the reference's driver calls it, with host-resident
    (NumPy) state and mesh dicts: every call uploads the state, runs one large
    step and returns its fields to host memory.  The mesh's device copy is kept
    between calls (``DeviceMesh.get``); the state crosses PCIe both ways."""
    import torch

    from mpas_dycore_pytorch_amd.time_integration import atm_srk3

    m, st, _ = prepare(args.level, args.levels, "quadratic", 0, args.dt, args.seed, device)
    cfg = {"config_time_integration_order": 3, "config_len_disp": float(m["nominalMinDc"])}
    state = dict(st)
    for _ in range(max(args.warmup, 1)):
        state = atm_srk3(state, st, cfg, args.dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        state = atm_srk3(state, st, cfg, args.dt)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    work = m["nCells"] * args.levels
    log(f"atm_srk3 with host dicts: {1e3 * sec:.2f} ms per call")
    print(json.dumps({"metric": "cell-level-steps/s of atm_srk3 with host-resident state (PCIe both ways)",
                      "value": work / sec, "unit": "cell-level-steps/s", "ms_per_call": 1e3 * sec,
                      "steps": args.steps, "warmup": args.warmup,
                      "config": {"workload": f"level {args.level} ({m['nCells']} cells) x {args.levels} levels, "
                                             f"dt {args.dt:g} s, host NumPy state in and out"}}), flush=True)
    return 0


def main() -> int:
    if len(sys.argv) > 1 and sys.argv[1] == "--cpu-worker":
        level, n_levels, dt, steps, seed = sys.argv[2:7]
        cpu_worker(int(level), int(n_levels), float(dt), int(steps), int(seed))
        return 0
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--level", type=int, default=7)
    ap.add_argument("--levels", type=int, default=55)
    ap.add_argument("--dt", type=float, default=300.0)
    ap.add_argument("--tracers", type=int, default=0,
                    help="passive tracers beside the water-vapour slot (the level-6 drop-in workload has 1)")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--roofline-kernel", default="k_acoustic_cells")
    ap.add_argument("--cpu-level", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--reorder", type=int, default=1, help="1: space-filling-curve mesh order on the device")
    ap.add_argument("--graph", type=int, default=0, help="1: replay the large step from a captured HIP graph")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: tune the launch shape of each wavefront-per-entity kernel during warm-up")
    ap.add_argument("--ensemble", type=str, default="",
                    help="measure an uneven ensemble instead, e.g. '5x32,6x16,7x4' (level x members): the batched "
                         "form (one grid per kernel per mesh, HIP graphs, groups on their own streams) against the "
                         "members stepped one by one")
    ap.add_argument("--mixed-steps", type=int, default=6,
                    help="large steps of the mixed-precision leg's error check against the FP64 step (its "
                         "timing uses --steps); 0 skips the leg")
    ap.add_argument("--no-jvp", action="store_true", help="skip the forward-mode part of the derivative leg")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="step of the main leg: fp64 (the workload's precision) or the mixed-precision step "
                         "(the roofline bytes stay the FP64 model's)")
    ap.add_argument("--layout-probe", action="store_true",
                    help="measure column streams at level pitch 55, 56 and 64 over the level's edge count instead")
    ap.add_argument("--boundary", action="store_true",
                    help="time atm_srk3 called with host-resident state and mesh dicts (PCIe both ways) instead")
    ap.add_argument("--vjp-steps", type=int, default=24,
                    help="large steps of the reverse-mode leg (forward + VJP of an area-weighted squared "
                         "theta perturbation, one checkpoint per large step; 24 is the differentiation "
                         "workload's run length); 0 skips it")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one process, one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    from mpas_dycore_pytorch_amd import _native as N
    from mpas_dycore_pytorch_amd import bytes_model as B
    from mpas_dycore_pytorch_amd.time_integration import Srk3Stepper

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        return 1
    device = torch.device("cuda", 0)
    torch.cuda.set_device(device)
    lib = N.load()

    if args.ensemble:
        return ensemble_bench(args, device)
    if args.layout_probe:
        return layout_probe(args, lib, N)
    if args.boundary:
        return boundary_bench(args, device)

    m, st, np = prepare(args.level, args.levels, "quadratic", args.tracers, args.dt, args.seed, device)
    nC, nE, nV, L = m["nCells"], m["nEdges"], m["nVertices"], args.levels
    cfg = {"config_time_integration_order": 3, "config_len_disp": float(m["nominalMinDc"])}
    stepper = Srk3Stepper(st, st, cfg, args.dt, device=device, reorder=bool(args.reorder), precision=args.precision)
    S = stepper.n_scalars
    tlib = stepper.lib  # the library whose kernels the stepper runs (and whose timer sees them)

    stepper.step(args.warmup)
    tuning = None
    if args.autotune:
        from mpas_dycore_pytorch_amd import autotune as AT

        t_tune = time.time()
        tuning = AT.autotune(stepper)
        log(f"autotuned {len(tuning)} kernels in {time.time() - t_tune:.1f} s: "
            + ", ".join(f"{k}={v['waves']}{'L' if v['order'] else ''}" for k, v in tuning.items()))
        log("block order, ms per launch (xcd chunks / linear): "
            + ", ".join(f"{k} {v['order_ms'].get('xcd', 0):.4f}/{v['order_ms'].get('linear', 0):.4f}"
                        for k, v in tuning.items()))
    if args.graph:
        stepper.capture_graph()
    torch.cuda.synchronize()
    log(f"warmup {args.warmup} steps done")

    schedule = B.launch_schedule(args.roofline_kernel)  # variant of each launch in one large step
    per_step = len(schedule)
    if not args.graph:
        N.check(tlib.mpas_hip_timer_arm(args.roofline_kernel.encode(), per_step * args.steps), "timer_arm")
    stream = torch.cuda.current_stream(device)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    stepper.step(args.steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    total_ms = ev0.elapsed_time(ev1)
    timed_steps = args.steps
    if args.graph:
        # a replayed graph carries no timer events: time the kernel in one eager step
        N.check(tlib.mpas_hip_timer_arm(args.roofline_kernel.encode(), per_step), "timer_arm")
        with torch.cuda.device(device):
            stepper._launch()
        torch.cuda.synchronize()
        timed_steps = 1
    each = (C.c_double * (per_step * timed_steps))()
    kn = C.c_int32()
    N.check(tlib.mpas_hip_timer_read_each(each, per_step * timed_steps, C.byref(kn)), "timer_read_each")
    if kn.value != per_step * timed_steps:
        raise RuntimeError(f"{args.roofline_kernel}: {kn.value} launches timed, {per_step * timed_steps} expected")
    for k, v in stepper.state_dict().items():
        if not bool(torch.isfinite(v).all()):
            raise RuntimeError(f"non-finite {k} after the timed steps")
    ms_per_step = total_ms / args.steps
    units = nC * L
    value = units * args.steps / (total_ms / 1e3)
    log(f"{args.steps} steps: {ms_per_step:.3f} ms/step, {value:.4e} cell-level-steps/s")

    # each launch is charged the algorithmic bytes of its own variant (first or later acoustic
    # substep, first or later RK stage), in launch order
    variants = schedule * timed_steps
    ms = [each[i] for i in range(kn.value)]
    k_total_ms = sum(ms)
    k_avg_ms = k_total_ms / kn.value
    k_bytes = B.schedule_bytes(args.roofline_kernel, variants, nC, nE, nV, L, S) / kn.value
    k_gbs = k_bytes / (k_avg_ms / 1e3) / 1e9
    by_variant = {}
    for v in sorted(set(variants)):
        vm = [t for t, vv in zip(ms, variants) if vv == v]
        vb = B.kernel_bytes(args.roofline_kernel, nC, nE, nV, L, S, v)
        vavg = sum(vm) / len(vm)
        by_variant[v] = {"launches": len(vm), "avg_ms": round(vavg, 5), "algorithmic_bytes": vb,
                         "achieved_gbs": round(vb / (vavg / 1e3) / 1e9, 1)}
    workload_tag = f"l{args.level}_L{args.levels}"
    roof = {"bound": "hbm", "kernel": args.roofline_kernel, "achieved": round(k_gbs, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(k_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profile(args.roofline_kernel, workload_tag),
            "algorithmic_bytes_per_launch": round(k_bytes), "avg_launch_ms": round(k_avg_ms, 5),
            "launches_timed": kn.value, "variants": by_variant}
    sb = B.step_bytes(nC, nE, nV, L, S)
    step_gbs = sb / (ms_per_step / 1e3) / 1e9
    step_roof = {"algorithmic_bytes": sb, "achieved": round(step_gbs, 1), "unit": "GB/s",
                 "frac": round(step_gbs / HBM_PEAK_GBS, 4)}

    mixed = None
    if args.mixed_steps > 0 and args.precision == "fp64":
        mixed = mixed_leg(st, cfg, args.dt, args.steps, args.mixed_steps, ms_per_step, bool(args.reorder), device,
                          tune=bool(args.autotune))
        log(f"mixed precision: {mixed['ms_per_step']:.3f} ms/step, x{mixed['speedup_vs_fp64']:.2f}; errors "
            + ", ".join(f"{k} {v['rel_l2']:.1e}" for k, v in mixed["errors_vs_fp64_gpu"].items()))

    vjp = None
    if args.vjp_steps > 0:
        vjp = vjp_leg(st, stepper, cfg, args.dt, args.vjp_steps, ms_per_step, device, jvp=not args.no_jvp,
                      tune=bool(args.autotune))
        log(f"reverse mode: {vjp}")

    cpu = None
    if not args.no_cpu_baseline:
        del stepper
        torch.cuda.empty_cache()
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(args.cpu_level, args.levels, {5: 900.0, 4: 1800.0, 6: 450.0}.get(args.cpu_level, 300.0),
                           args.cpu_steps, args.seed)
        log(f"cpu baseline {cpu['value']:.4e} cell-level-steps/s")

    out = {
        "metric": METRIC, "value": value, "unit": "cell-level-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32/f64 mixed", "data": "synthetic",
        "config": {"workload": (f"atm_srk3 large step: icosahedral level {args.level} ({nC} cells, {nE} edges) x "
                                f"{L} quadratic levels to 30 km, dt {args.dt:g} s, RK3 split 3 x 2 acoustic "
                                f"substeps, monotonic FCT on {S} scalar(s), flat terrain"),
                   "cells": nC, "levels": L, "dt": args.dt, "parallelism": "single GPU",
                   "mesh_order": "hilbert" if args.reorder else "generator",
                   "launch": "hip graph" if args.graph else "stream",
                   "waves_per_block": ({k: v["waves"] for k, v in tuning.items()} if tuning else "default 4"),
                   "linear_block_order": ([k for k, v in tuning.items() if v["order"]] if tuning else [])},
        "roofline": roof, "step_roofline": step_roof, "vjp_fwd_ratio": (vjp or {}).get("ratio"), "vjp": vjp, "mixed_precision": mixed, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
