"""Reverse-mode differentiation of the large step through the hand-written adjoint kernels.

The reference differentiates ``atm_srk3`` with torch autograd
(``differentiable=True``, ``mpas_pytorch/dynamics/time_integration.py:4897``)
and drives it with ``run_simulation.vjp`` (``atmosphere_model.py:2101-2155``),
checkpointing each large step (``autodiff/checkpoint.py:106``).  Here:

* :func:`taped_step` runs one large step as a chain of per-function GPU
  kernels (the orchestration of ``_atm_srk3_impl``, ``:3420-4894``), never
  mutating a tensor, and records each call (inputs and outputs) on a
  :class:`Tape`;
* :meth:`Tape.backward` walks the tape in reverse and calls the matching
  ``mpas_hip_*_vjp`` entry point for every kernel call, accumulating the
  adjoint of each tensor in a buffer keyed by the tensor's identity;
* :func:`step_vjp` is the VJP of one large step; the multi-step driver in
  ``atmosphere_model`` stores the state at every large step and recomputes
  one step's tape at a time in the backward sweep (acoustic substeps are
  recomputed, never stored across steps);
* :meth:`Tape.forward_tangents` walks the same tape forward through the
  ``mpas_hip_*_jvp`` entry points (the tangent-linear model), and
  :func:`step_jvp` is the JVP of one large step (the reference's forward-mode
  ``run_simulation.jvp``, ``atmosphere_model.py:2041-2098``).

Derivatives are those of the hard primal (branch selections frozen) with
torch's conventions at switching points, as the reference's primitives define
them (``autodiff/soft_limiter.py:72-186``).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Any, Callable, Mapping

import torch

from . import _native as N
from . import kernels as K
from .device import DeviceMesh, stream_handle

# outputs of the cold start that carry a derivative (the base-state ones do not)
ICD_DIFF_OUT = ("theta_m", "rho_zz", "ru", "rw", "rho_p", "rtheta_p", "exner", "pressure_p")

# slot order of each adjoint entry point: inputs, then outputs (include/mpas_hip.h)
SLOTS: dict[str, tuple[tuple[str, ...], tuple[str, ...]]] = {
    "moist_coefficients": (("scalars",), ("cqw", "cqu", "qtot")),
    "vert_imp_coefs": (("cqw", "p", "t", "rt", "qtot"), K.VIMP_OUT),
    "solve_diagnostics": (("u", "h", "v_in"), K.DIAG_OUT),
    "dyn_tend": (K.DYN_INPUTS + tuple(k + "_in" for k in K.DYN_CARRIED), K.DYN_CARRIED + K.DYN_OUTPUTS),
    "smlstep_pert": (("u_tend", "w_tend_in"), ("w_tend_out",)),
    "acoustic_step": (K.ACOUSTIC_IN + tuple(k + "_in" for k in K.ACOUSTIC_IO), K.ACOUSTIC_IO),
    "divergence_damping": (("theta_m", "rtheta_pp", "rtheta_pp_old", "ru_p_in"), ("ru_p_out",)),
    "recover_large_step": (K.RECOVER_IN + ("wwAvg_in", "ruAvg_in", "exner_in", "pressure_p_in"),
                           ("wwAvg", "ruAvg", "rho_p", "rho_zz", "rtheta_p", "theta_m", "exner", "pressure_p",
                            "rw", "w", "ru", "u")),
    "advance_scalars": (("scalar_old", "scalar_new_in", "rho_zz_old", "rho_zz_new", "uhAvg", "wwAvg"),
                        ("scalar_new_out", "horiz_flux")),
    "advance_scalars_mono": (("scalars_old", "scalars_new_in", "rho_zz_old", "uhAvg", "wwAvg"),
                             ("scalars_new_out", "rho_zz_int")),
    "reconstruct": (("u",), K.RECON_OUT),
    # the cold start (raw slots 0-2 or derived slots 3-4 are None in the other mode)
    "init_coupled": (("theta", "rho", "qv", "theta_m_in", "rho_zz_in", "u", "w"), ICD_DIFF_OUT),
}
# output adjoints the entry points use as accumulators (must be non-null)
REQUIRED_OUT = {
    "vert_imp_coefs": ("cofwz", "coftz", "cofwt"),
    "dyn_tend": ("h_divergence", "kdiff"),
    "acoustic_step": ("ru_p",),
    "recover_large_step": ("rho_zz", "ru"),
    "advance_scalars_mono": ("rho_zz_int",),
}
# inputs that depend on the mesh only (no adjoint wanted)
CONSTANT_IN = {"vert_imp_coefs": (), "acoustic_step": ("cofrz", "cofwr")}


@dataclass
class Op:
    kind: str                       # kernel name, or "add", "scale", "copy", "mask_pad"
    args: tuple = ()
    ins: list = field(default_factory=list)
    outs: list = field(default_factory=list)


class Tape:
    """Kernel calls of one forward pass, with every input and output tensor kept alive."""

    def __init__(self, dm: DeviceMesh):
        self.dm = dm
        self.ops: list[Op] = []

    # ---- recording helpers (pure tensor ops of the orchestration)
    def kernel(self, name: str, args: tuple, ins: Mapping[str, torch.Tensor | None],
               outs: Mapping[str, torch.Tensor]):
        """Record one kernel call; an input slot the call does not read is None."""
        si, so = SLOTS[name]
        self.ops.append(Op(name, args, [ins.get(k) for k in si], [outs[k] for k in so]))

    def add(self, a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
        out = a + b
        self.ops.append(Op("add", (), [a, b], [out]))
        return out

    def scale(self, a: torch.Tensor, c: float) -> torch.Tensor:
        out = a * c
        self.ops.append(Op("scale", (c,), [a], [out]))
        return out

    def copy(self, a: torch.Tensor) -> torch.Tensor:
        out = a.clone()
        self.ops.append(Op("copy", (), [a], [out]))
        return out

    def mask_pad(self, a: torch.Tensor, row: int) -> torch.Tensor:
        out = a.clone()
        out[row] = 0.0
        self.ops.append(Op("mask_pad", (row,), [a], [out]))
        return out

    def depending_on(self, sources) -> set[int]:
        """Ids of the recorded tensors that depend on any of ``sources`` (a forward scan)."""
        dep = {id(t) for t in sources}
        for op in self.ops:
            if any(t is not None and id(t) in dep for t in op.ins):
                dep.update(id(t) for t in op.outs)
        return dep

    # ---- reverse sweep
    def backward(self, seeds: Mapping[int, torch.Tensor], scratch: torch.Tensor,
                 wrt: set[int] | None = None) -> dict[int, torch.Tensor]:
        """Adjoints of every tensor reached, keyed by ``id(tensor)``.

        ``seeds`` maps ``id(output tensor)`` to its adjoint (cotangent).  With
        ``wrt`` (ids of the tensors that depend on the inputs of interest, see
        :meth:`depending_on`) no adjoint is formed for any other tensor: the
        kernels get a null slot and skip that work."""
        bars: dict[int, torch.Tensor] = {k: v.clone() for k, v in seeds.items()}
        lib = N.load()
        dm = self.dm
        stream = C.c_void_p(stream_handle(dm.device))

        def get(t: torch.Tensor, create: bool) -> torch.Tensor | None:
            b = bars.get(id(t))
            if b is None and create:
                b = torch.zeros_like(t)
                bars[id(t)] = b
            return b

        for op in reversed(self.ops):
            if op.kind == "add":
                g = bars.get(id(op.outs[0]))
                if g is not None:
                    for x in op.ins:
                        get(x, True).add_(g)
                continue
            if op.kind == "scale":
                g = bars.get(id(op.outs[0]))
                if g is not None:
                    get(op.ins[0], True).add_(g, alpha=op.args[0])
                continue
            if op.kind == "copy":
                g = bars.get(id(op.outs[0]))
                if g is not None:
                    get(op.ins[0], True).add_(g)
                continue
            if op.kind == "mask_pad":
                g = bars.get(id(op.outs[0]))
                if g is not None:
                    g = g.clone()
                    g[op.args[0]] = 0.0
                    get(op.ins[0], True).add_(g)
                continue
            si, so = SLOTS[op.kind]
            out_bars = [bars.get(id(t)) for t in op.outs]
            if all(b is None for b in out_bars):
                continue  # nothing downstream depends on this call
            if wrt is not None and not any(t is not None and id(t) in wrt for t in op.ins):
                for t in op.outs:
                    bars.pop(id(t), None)
                continue  # no input of interest upstream of this call
            req = REQUIRED_OUT.get(op.kind, ())
            for j, name in enumerate(so):
                if out_bars[j] is None and name in req:
                    out_bars[j] = get(op.outs[j], True)
            const = CONSTANT_IN.get(op.kind, ())
            in_bars: list[torch.Tensor | None] = []
            pending: list[tuple[torch.Tensor, torch.Tensor]] = []
            seen: set[int] = set()
            for name, t in zip(si, op.ins):
                if t is None or name in const or (wrt is not None and id(t) not in wrt):
                    in_bars.append(None)
                    continue
                if id(t) in seen:
                    # the same tensor in two slots: accumulate separately, then add
                    tmp = torch.zeros_like(t)
                    pending.append((t, tmp))
                    in_bars.append(tmp)
                    continue
                seen.add(id(t))
                in_bars.append(get(t, True))
            fwd = (C.c_void_p * (len(op.ins) + len(op.outs)))(
                *[(x.data_ptr() if x is not None else None) for x in op.ins + op.outs])
            bar = (C.c_void_p * (len(op.ins) + len(op.outs)))(
                *[(b.data_ptr() if b is not None else None) for b in in_bars + out_bars])
            _call_vjp(lib, op, dm, fwd, bar, scratch, stream)
            for t, tmp in pending:
                get(t, True).add_(tmp)
            # each output adjoint is consumed by its producer exactly once
            for t in op.outs:
                bars.pop(id(t), None)
        return bars


    # ---- forward (tangent) sweep
    def forward_tangents(self, seeds: Mapping[int, torch.Tensor], scratch: torch.Tensor,
                         keep: set[int] | None = None) -> dict[int, torch.Tensor]:
        """Tangents of every tensor reached from ``seeds``, keyed by ``id(tensor)``.

        ``seeds`` maps ``id(input tensor)`` to its tangent.  A tangent is
        dropped after the last call that reads it unless its id is in
        ``keep``; tensors no seed reaches get no tangent."""
        dots: dict[int, torch.Tensor] = dict(seeds)
        lib = N.load()
        dm = self.dm
        stream = C.c_void_p(stream_handle(dm.device))
        keep = set(keep or ())
        last: dict[int, int] = {}
        for i, op in enumerate(self.ops):
            for t in op.ins:
                if t is not None:
                    last[id(t)] = i
        for i, op in enumerate(self.ops):
            tin = [(dots.get(id(t)) if t is not None else None) for t in op.ins]
            if op.kind in ("add", "scale", "copy", "mask_pad"):
                live = [d for d in tin if d is not None]
                if live:
                    if op.kind == "add":
                        out = live[0] if len(live) == 1 else live[0] + live[1]
                    elif op.kind == "scale":
                        out = live[0] * op.args[0]
                    elif op.kind == "copy":
                        out = live[0]
                    else:
                        out = live[0].clone()
                        out[op.args[0]] = 0.0
                    dots[id(op.outs[0])] = out
            else:
                si, so = SLOTS[op.kind]
                const = CONSTANT_IN.get(op.kind, ())
                tin = [None if name in const else d for name, d in zip(si, tin)]
                if any(d is not None for d in tin):
                    tout = [torch.empty_like(t) for t in op.outs]
                    fwd = (C.c_void_p * (len(op.ins) + len(op.outs)))(
                        *[(x.data_ptr() if x is not None else None) for x in op.ins + op.outs])
                    dot = (C.c_void_p * (len(op.ins) + len(op.outs)))(
                        *[(d.data_ptr() if d is not None else None) for d in tin + tout])
                    _call_jvp(lib, op, dm, fwd, dot, scratch, stream)
                    for t, d in zip(op.outs, tout):
                        dots[id(t)] = d
            for t in op.ins:
                if t is not None and last.get(id(t)) == i and id(t) not in keep:
                    dots.pop(id(t), None)
        return dots


def _call_jvp(lib, op: Op, dm: DeviceMesh, fwd, dot, scratch: torch.Tensor, stream) -> None:
    a = op.args
    sp = C.c_void_p(scratch.data_ptr())
    k = op.kind
    if k == "moist_coefficients":
        rc = lib.mpas_hip_moist_coefficients_jvp(dm.ref, *a, fwd, dot, stream)
    elif k == "vert_imp_coefs":
        rc = lib.mpas_hip_vert_imp_coefs_jvp(dm.ref, *a, fwd, dot, stream)
    elif k == "solve_diagnostics":
        rc = lib.mpas_hip_solve_diagnostics_jvp(dm.ref, *a, fwd, dot, sp, stream)
    elif k == "dyn_tend":
        rc = lib.mpas_hip_dyn_tend_jvp(dm.ref, C.byref(a[0]), fwd, dot, sp, stream)
    elif k == "smlstep_pert":
        rc = lib.mpas_hip_smlstep_pert_jvp(dm.ref, fwd, dot, stream)
    elif k == "acoustic_step":
        rc = lib.mpas_hip_acoustic_step_jvp(dm.ref, C.byref(a[0]), fwd, dot, sp, stream)
    elif k == "divergence_damping":
        rc = lib.mpas_hip_divergence_damping_jvp(dm.ref, *a, fwd, dot, stream)
    elif k == "recover_large_step":
        rc = lib.mpas_hip_recover_large_step_jvp(dm.ref, *a, fwd, dot, stream)
    elif k == "advance_scalars":
        rc = lib.mpas_hip_advance_scalars_jvp(dm.ref, *a, fwd, dot, stream)
    elif k == "advance_scalars_mono":
        rc = lib.mpas_hip_advance_scalars_mono_jvp(dm.ref, *a, fwd, dot, sp, stream)
    elif k == "reconstruct":
        rc = lib.mpas_hip_reconstruct_jvp(dm.ref, fwd, dot, stream)
    elif k == "init_coupled":
        rc = lib.mpas_hip_init_coupled_diagnostics_jvp(dm.ref, int(a[0]), a[1].data_ptr(), fwd, dot, stream)
    else:  # pragma: no cover - every recorded kind is handled above
        raise KeyError(k)
    N.check(rc, k + "_jvp")


def _call_vjp(lib, op: Op, dm: DeviceMesh, fwd, bar, scratch: torch.Tensor, stream) -> None:
    a = op.args
    sp = C.c_void_p(scratch.data_ptr())
    k = op.kind
    if k == "moist_coefficients":
        rc = lib.mpas_hip_moist_coefficients_vjp(dm.ref, *a, fwd, bar, stream)
    elif k == "vert_imp_coefs":
        rc = lib.mpas_hip_vert_imp_coefs_vjp(dm.ref, *a, fwd, bar, stream)
    elif k == "solve_diagnostics":
        rc = lib.mpas_hip_solve_diagnostics_vjp(dm.ref, *a, fwd, bar, sp, C.byref(dm.inverse_tables()), stream)
    elif k == "dyn_tend":
        rc = lib.mpas_hip_dyn_tend_vjp(dm.ref, C.byref(a[0]), fwd, bar, sp, C.byref(dm.inverse_tables()), stream)
    elif k == "smlstep_pert":
        rc = lib.mpas_hip_smlstep_pert_vjp(dm.ref, fwd, bar, stream)
    elif k == "acoustic_step":
        rc = lib.mpas_hip_acoustic_step_vjp(dm.ref, C.byref(a[0]), fwd, bar, sp, stream)
    elif k == "divergence_damping":
        rc = lib.mpas_hip_divergence_damping_vjp(dm.ref, *a, fwd, bar, stream)
    elif k == "recover_large_step":
        rc = lib.mpas_hip_recover_large_step_vjp(dm.ref, *a, fwd, bar, stream)
    elif k == "advance_scalars":
        rc = lib.mpas_hip_advance_scalars_vjp(dm.ref, *a, fwd, bar, sp, C.byref(dm.inverse_tables()), stream)
    elif k == "advance_scalars_mono":
        rc = lib.mpas_hip_advance_scalars_mono_vjp(dm.ref, *a, fwd, bar, sp, C.byref(dm.inverse_tables()),
                                                   stream)
    elif k == "reconstruct":
        rc = lib.mpas_hip_reconstruct_vjp(dm.ref, fwd, bar, stream)
    elif k == "init_coupled":
        need = lib.mpas_hip_init_coupled_vjp_scratch_doubles(dm.ne, dm.L)
        if scratch.numel() < need:
            raise ValueError(f"init_coupled_vjp needs {need} doubles of scratch, got {scratch.numel()}")
        rc = lib.mpas_hip_init_coupled_diagnostics_vjp(dm.ref, int(a[0]), a[1].data_ptr(), fwd, bar, sp, stream)
    else:  # pragma: no cover - every recorded kind is handled above
        raise KeyError(k)
    N.check(rc, k + "_vjp")


def vjp_scratch(dm: DeviceMesh, n_scalars: int) -> torch.Tensor:
    n = N.load().mpas_hip_vjp_scratch_doubles(dm.nc, dm.ne, dm.nv, dm.L, n_scalars)
    return torch.empty(int(n), dtype=torch.float64, device=dm.device)


# ---------------------------------------------------------------------------
# the taped large step
# ---------------------------------------------------------------------------
DIAG_KEYS = ("h_edge", "v", "divergence", "vorticity", "ke", "pv_edge")
STEP_INPUTS = ("u", "v", "w", "ru", "rw", "theta_m", "rho_zz", "rho_edge", "divergence", "vorticity", "ke",
               "pv_edge", "pressure_p", "exner", "rtheta_p", "rho_p", "scalars", "rt_diabatic_tend", "ur_cell",
               "vr_cell")


def _explicit_weights(rk_step: int, order: int = 3) -> tuple[float, float]:
    """Old/new density weights of an explicit transport stage (advance_scalars.py:271-286)."""
    wn = {1: 1.0 / 2.0 if order == 2 else 1.0 / 3.0, 2: 1.0 / 2.0}.get(rk_step, 1.0)
    return 1.0 - wn, wn


def rk_weights(dt_dyn: float, ns: int, order: int) -> tuple[list, list, list]:
    """(stage lengths, acoustic substep lengths, substep counts) of the three
    stages, as ``compute_rk_weights`` forms them (time_integration.py:2787-2856)."""
    half = max(1, ns // 2)
    if order == 2:
        return [dt_dyn / 2.0, dt_dyn / 2.0, dt_dyn], [dt_dyn / float(ns)] * 3, [half, half, ns]
    return [dt_dyn / 3.0, dt_dyn / 2.0, dt_dyn], [dt_dyn / 3.0, dt_dyn / float(ns), dt_dyn / float(ns)], [1, half, ns]


def taped_step(dm: DeviceMesh, state: Mapping[str, torch.Tensor], cfg: Any, dt: float,
               tape: Tape) -> dict[str, torch.Tensor]:
    """One large step through per-function kernels, recorded on ``tape``.

    ``state`` holds device tensors for :data:`STEP_INPUTS`; ``cfg`` is the
    resolved configuration (``time_integration.resolve_config``).  Returns the
    new state (prognostics and the diagnostics ``atm_srk3`` returns)."""
    nc = dm.nc
    st = {k: state[k] for k in STEP_INPUTS}
    S = int(st["scalars"].shape[2])
    ns, split = int(cfg.ns), int(cfg.split)
    len_disp, epssm, smdiv, apvm, coef3 = cfg.len_disp, cfg.epssm, cfg.smdiv, cfg.apvm, cfg.coef3
    rt_diab = st["rt_diabatic_tend"]
    st["theta_m"] = tape.mask_pad(st["theta_m"], nc)
    theta_m_init = st["theta_m"]
    rho_zz_old_split = st["rho_zz"]
    scalars_old = st["scalars"]
    scalars_new = st["scalars"]
    ru_save, rw_save = st["ru"], st["rw"]
    rtheta_p_save, rho_p_save = st["rtheta_p"], st["rho_p"]
    theta_m_save, rr_save = st["theta_m"], st["rho_p"]
    # Zero inputs (the acoustic state a step starts from, the carried
    # tendencies the first RK stage records, the cold tangential velocity) are
    # read only: one zero tensor per shape serves them all.  They depend on no
    # step input, so the sweeps give them no adjoint or tangent.
    zeros: dict[tuple[int, ...], torch.Tensor] = {}

    def z(*shape: int) -> torch.Tensor:
        if shape not in zeros:
            zeros[shape] = torch.zeros(shape, dtype=torch.float64, device=dm.device)
        return zeros[shape]

    L = dm.L
    acoustic = {"ru_p": z(dm.ne + 1, L), "ruAvg": z(dm.ne + 1, L), "rho_pp": z(nc + 1, L),
                "rtheta_pp": z(nc + 1, L), "rtheta_pp_old": z(nc + 1, L), "rw_p": z(nc + 1, L + 1),
                "wwAvg": z(nc + 1, L + 1)}
    order = 2 if int(cfg.order) == 2 else 3
    rk_timestep, rk_sub, nsub = rk_weights(dt / split, ns, order)
    moist = vcoef = carried = None
    ruAvg_split = wwAvg_split = None
    ms, me = int(cfg.moist_start), int(cfg.moist_end)
    for sub in range(1, split + 1):
        for rk in (1, 2, 3):
            if rk == 1 and sub == 1:
                moist = K.moist_coefficients(dm, st["scalars"], ms, me)
                tape.kernel("moist_coefficients", (S, ms, me), {"scalars": st["scalars"]}, moist)
                if not cfg.diagnostics_current:
                    v0 = z(dm.ne + 1, L)
                    d = K.solve_diagnostics(dm, dt, apvm, None, st["u"], st["rho_zz"], v0)
                    tape.kernel("solve_diagnostics", (float(dt), float(apvm), -1),
                                {"u": st["u"], "h": st["rho_zz"], "v_in": v0}, d)
                    st.update({k: d[k] for k in DIAG_KEYS})
                    st["rho_edge"] = d["h_edge"]
            if rk == 1 or (rk == 2 and order == 3):
                vin = {"cqw": moist["cqw"], "p": st["exner"], "t": st["theta_m"], "rt": st["rtheta_p"],
                       "qtot": moist["qtot"]}
                vcoef = K.vert_imp_coefs(dm, rk_sub[rk - 1], epssm, *vin.values())
                tape.kernel("vert_imp_coefs", (float(rk_sub[rk - 1]), float(epssm)), vin, vcoef)
            fields = {"u": st["u"], "v": st["v"], "w": st["w"], "ru": st["ru"], "rw": st["rw"],
                      "ru_save": ru_save, "rw_save": rw_save, "theta_m": st["theta_m"],
                      "theta_m_save": theta_m_save, "rho_zz": st["rho_zz"], "rho_edge": st["rho_edge"],
                      "divergence": st["divergence"], "vorticity": st["vorticity"], "ke": st["ke"],
                      "pv_edge": st["pv_edge"], "pp": st["pressure_p"], "rr_save": rr_save,
                      "cqw": moist["cqw"], "cqu": moist["cqu"], "rt_diabatic_tend": rt_diab, "qtot": moist["qtot"],
                      "ur_cell": st["ur_cell"], "vr_cell": st["vr_cell"]}
            params = K.dyn_params(rk, dt, len_disp, c_s=cfg.c_s, visc4_2dsmag=cfg.visc4_2dsmag,
                                  coef_3rd_order=coef3, del4u_div_factor=cfg.del4u_div_factor,
                                  rayleigh_levels=cfg.rayleigh_levels,
                                  rayleigh_coef_inverse=cfg.rayleigh_coef_inverse, mix_fixed=bool(cfg.mix_fixed),
                                  h_theta_eddy_visc2=cfg.h_theta_eddy_visc2, h_mom_eddy_visc4=cfg.h_mom_eddy_visc4,
                                  h_theta_eddy_visc4=cfg.h_theta_eddy_visc4, cam_levels=cfg.cam_levels,
                                  cam_visc2=cfg.cam_visc2, curvature=bool(cfg.curvature),
                                  v_mom_eddy_visc2=cfg.v_mom_eddy_visc2, v_theta_eddy_visc2=cfg.v_theta_eddy_visc2,
                                  mix_full=bool(cfg.mix_full))
            if rk == 1:
                carried_in = {k: z(*t.shape) for k, t in
                              zip(K.DYN_CARRIED, (st["rho_zz"], st["u"], st["w"], st["rho_zz"], st["rho_zz"]))}
            else:
                carried_in = carried
            tend = K.dyn_tend(dm, params, fields, carried_in if rk > 1 else None)
            ins = dict(fields)
            ins.update({k + "_in": carried_in[k] for k in K.DYN_CARRIED})
            tape.kernel("dyn_tend", (params,), ins, tend)
            carried = {k: tend[k] for k in K.DYN_CARRIED}
            tend_rw = K.smlstep_pert(dm, tend["tend_u"], tend["tend_w"])
            tape.kernel("smlstep_pert", (), {"u_tend": tend["tend_u"], "w_tend_in": tend["tend_w"]},
                        {"w_tend_out": tend_rw})
            afields = dict(vcoef)
            afields.update({"rho_zz": st["rho_zz"], "theta_m": theta_m_init, "exner": st["exner"],
                            "cqu": moist["cqu"], "tend_ru": tend["tend_u"], "tend_rho": tend["tend_rho"],
                            "tend_rt": tend["tend_theta"], "tend_rw": tend_rw, "w": st["w"], "rw": st["rw"],
                            "rw_save": rw_save})
            for small in range(1, nsub[rk - 1] + 1):
                afields.update(acoustic)
                ap = K.acoustic_params(small, rk_sub[rk - 1], epssm)
                new = K.acoustic_step(dm, ap, afields)
                ins = {k: afields[k] for k in K.ACOUSTIC_IN}
                ins.update({k + "_in": afields[k] for k in K.ACOUSTIC_IO})
                tape.kernel("acoustic_step", (ap,), ins, new)
                ru_p = K.divergence_damping(dm, rk_sub[rk - 1], smdiv, len_disp, theta_m_init, new["rtheta_pp"],
                                            new["rtheta_pp_old"], new["ru_p"])
                tape.kernel("divergence_damping", (float(rk_sub[rk - 1]), float(smdiv), float(len_disp)),
                            {"theta_m": theta_m_init, "rtheta_pp": new["rtheta_pp"],
                             "rtheta_pp_old": new["rtheta_pp_old"], "ru_p_in": new["ru_p"]}, {"ru_p_out": ru_p})
                new["ru_p"] = ru_p
                acoustic = new
            rfields = {"rho_p_save": rho_p_save, "rho_pp": acoustic["rho_pp"], "rtheta_p_save": rtheta_p_save,
                       "rtheta_pp": acoustic["rtheta_pp"], "rt_diabatic_tend": rt_diab, "rw_save": rw_save,
                       "rw_p": acoustic["rw_p"], "ru_save": ru_save, "ru_p": acoustic["ru_p"],
                       "wwAvg": acoustic["wwAvg"], "ruAvg": acoustic["ruAvg"]}
            rec = K.recover_large_step(dm, rk, rk_timestep[rk - 1], nsub[rk - 1], rfields,
                                       exner_in=st["exner"], pressure_p_in=st["pressure_p"])
            ins = {k: rfields[k] for k in K.RECOVER_IN}
            ins.update({"wwAvg_in": rfields["wwAvg"], "ruAvg_in": rfields["ruAvg"], "exner_in": st["exner"],
                        "pressure_p_in": st["pressure_p"]})
            tape.kernel("recover_large_step", (int(rk), float(rk_timestep[rk - 1]), int(nsub[rk - 1])), ins, rec)
            acoustic = dict(acoustic)
            acoustic["wwAvg"], acoustic["ruAvg"] = rec["wwAvg"], rec["ruAvg"]
            for k in ("rho_p", "rho_zz", "rtheta_p", "theta_m", "exner", "pressure_p", "rw", "w", "ru", "u"):
                st[k] = rec[k]
            d = K.solve_diagnostics(dm, dt, apvm, rk, st["u"], st["rho_zz"], st["v"])
            tape.kernel("solve_diagnostics", (float(dt), float(apvm), int(rk)),
                        {"u": st["u"], "h": st["rho_zz"], "v_in": st["v"]}, d)
            st.update({k: d[k] for k in DIAG_KEYS})
            st["rho_edge"] = d["h_edge"]
        if sub == 1:
            ruAvg_split, wwAvg_split = acoustic["ruAvg"], acoustic["wwAvg"]
        else:
            ruAvg_split = tape.add(ruAvg_split, acoustic["ruAvg"])
            wwAvg_split = tape.add(wwAvg_split, acoustic["wwAvg"])
        if sub < split:
            st["theta_m"] = tape.mask_pad(st["theta_m"], nc)
            ru_save, rw_save = st["ru"], st["rw"]
            rtheta_p_save, rho_p_save = st["rtheta_p"], st["rho_p"]
            theta_m_init = st["theta_m"]
            rr_save = rho_p_save
            theta_m_save = st["theta_m"]
    inv_split = 1.0 / float(split)
    ruAvg = tape.scale(ruAvg_split, inv_split)
    wwAvg = tape.scale(wwAvg_split, inv_split)
    rho_zz_new = st["rho_zz"]
    if S > 0 and cfg.scalar_advection:
        for rk_t in (1, 2, 3):
            dt_t = [dt / 2.0 if order == 2 else dt / 3.0, dt / 2.0, dt][rk_t - 1]
            if rk_t == 3 and cfg.monotonic:
                out, rzi = K.advance_scalars_mono(dm, dt_t, coef3, scalars_old, scalars_new, rho_zz_old_split,
                                                  ruAvg, wwAvg)
                tape.kernel("advance_scalars_mono", (S, float(dt_t), float(coef3)),
                            {"scalars_old": scalars_old, "scalars_new_in": scalars_new,
                             "rho_zz_old": rho_zz_old_split, "uhAvg": ruAvg, "wwAvg": wwAvg},
                            {"scalars_new_out": out, "rho_zz_int": rzi})
            else:
                wo, wn = _explicit_weights(rk_t, order) if not (rk_t == 3) else (0.0, 1.0)
                out, hf = K.advance_scalars(dm, dt_t, coef3, wo, wn, scalars_old, scalars_new, rho_zz_old_split,
                                            rho_zz_new, ruAvg, wwAvg)
                tape.kernel("advance_scalars", (S, float(dt_t), float(coef3), float(wo), float(wn)),
                            {"scalar_old": scalars_old, "scalar_new_in": scalars_new, "rho_zz_old": rho_zz_old_split,
                             "rho_zz_new": rho_zz_new, "uhAvg": ruAvg, "wwAvg": wwAvg},
                            {"scalar_new_out": out, "horiz_flux": hf})
            scalars_new = out
    st["scalars"] = scalars_new
    st["ruAvg"], st["wwAvg"] = ruAvg, wwAvg
    rc = K.reconstruct(dm, st["u"])
    tape.kernel("reconstruct", (), {"u": st["u"]}, rc)
    for k in K.RECON_OUT:
        st[k] = rc[k]
    st["ur_cell"] = tape.copy(rc["uReconstructZonal"])
    st["vr_cell"] = tape.copy(rc["uReconstructMeridional"])
    # h_edge is rho_edge under its other name (atmosphere_model.py:176-200): the
    # same tensor, so a cotangent on either reaches the diagnostic solve
    st["h_edge"] = st["rho_edge"]
    return st


def step_vjp(dm: DeviceMesh, state: Mapping[str, torch.Tensor], cfg: Any, dt: float,
             cotangents: Mapping[str, torch.Tensor], scratch: torch.Tensor | None = None
             ) -> tuple[dict[str, torch.Tensor], dict[str, torch.Tensor]]:
    """(new state, adjoints of the step inputs) for output adjoints ``cotangents``.

    The forward is recomputed from ``state`` (the large step's checkpoint);
    adjoints are returned for every key of :data:`STEP_INPUTS` (zero where
    the step does not depend on the input)."""
    tape = Tape(dm)
    out = taped_step(dm, state, cfg, dt, tape)
    return out, tape_vjp(tape, state, out, cotangents, scratch)


def tape_vjp(tape: Tape, state: Mapping[str, torch.Tensor], out: Mapping[str, torch.Tensor],
             cotangents: Mapping[str, torch.Tensor], scratch: torch.Tensor | None = None
             ) -> dict[str, torch.Tensor]:
    """Adjoints of the step inputs ``state`` from a large step already
    recorded on ``tape`` (``out`` is what :func:`taped_step` returned)."""
    dm = tape.dm
    seeds = {}
    for k, g in cotangents.items():
        if g is None or k not in out:
            continue
        t = out[k]
        if id(t) in seeds:
            seeds[id(t)] = seeds[id(t)] + g
        else:
            seeds[id(t)] = g.to(device=t.device, dtype=torch.float64).contiguous()
    if scratch is None:
        scratch = vjp_scratch(dm, int(state["scalars"].shape[2]))
    bars = tape.backward(seeds, scratch, tape.depending_on(state[k] for k in STEP_INPUTS))
    grads = {}
    for k in STEP_INPUTS:
        b = bars.get(id(state[k]))
        grads[k] = b if b is not None else torch.zeros_like(state[k])
    return grads


def step_jvp(dm: DeviceMesh, state: Mapping[str, torch.Tensor], cfg: Any, dt: float,
             tangents: Mapping[str, torch.Tensor], scratch: torch.Tensor | None = None
             ) -> tuple[dict[str, torch.Tensor], dict[str, torch.Tensor]]:
    """(new state, tangents of the new state) for input tangents ``tangents``.

    ``tangents`` maps keys of :data:`STEP_INPUTS` to their tangents; the step
    is run on a tape from ``state`` and the tape is swept forward.  Output
    keys that no input tangent reaches get no entry."""
    tape = Tape(dm)
    out = taped_step(dm, state, cfg, dt, tape)
    return out, tape_jvp(tape, state, out, tangents, scratch)


def tape_jvp(tape: Tape, state: Mapping[str, torch.Tensor], out: Mapping[str, torch.Tensor],
             tangents: Mapping[str, torch.Tensor], scratch: torch.Tensor | None = None
             ) -> dict[str, torch.Tensor]:
    """Tangents of the new state ``out`` of a large step already recorded on
    ``tape`` from ``state``, for input tangents ``tangents``."""
    dm = tape.dm
    seeds = {}
    for k, t in tangents.items():
        if t is None or k not in STEP_INPUTS:
            continue
        x = state[k]
        seeds[id(x)] = t.to(device=x.device, dtype=torch.float64).contiguous()
    if scratch is None:
        scratch = vjp_scratch(dm, int(state["scalars"].shape[2]))
    keep = {id(v) for v in out.values() if isinstance(v, torch.Tensor)}
    dots = tape.forward_tangents(seeds, scratch, keep)
    res = {}
    for k, v in out.items():
        if isinstance(v, torch.Tensor) and id(v) in dots:
            res[k] = dots[id(v)]
    return res


# ---------------------------------------------------------------------------
# the taped cold start
# ---------------------------------------------------------------------------
COLD_START_INPUTS = ("theta", "rho", "qv", "theta_m", "rho_zz", "u", "w")


def taped_cold_start(dm: DeviceMesh, fields: Mapping[str, torch.Tensor], theta_base: torch.Tensor, dt: float,
                     apvm: float, tape: Tape) -> dict[str, torch.Tensor]:
    """The cold start of a differentiable run, recorded on ``tape``.

    The reference prepares the state inside its differentiable
    ``run_simulation`` (``atmosphere_model.py:1589-1597`` ->
    ``_prepare_timestep_state`` ``:782``): ``atm_init_coupled_diagnostics``,
    the padding sentinels, the first diagnostic solve and the vector
    reconstruction.  ``fields`` holds ``u``, ``w`` and either ``theta_m`` and
    ``rho_zz`` (used as given, the derived path) or raw ``theta``, ``rho``,
    ``qv``.  Returns the prepared fields; the base-state outputs
    (``rtheta_base``, ``exner_base``, ``pressure_base``) are constants."""
    derived = "theta_m" in fields and "rho_zz" in fields
    u, w = fields["u"], fields["w"]
    if derived:
        icd = K.init_coupled_diagnostics(dm, None, None, None, u, w, theta_base, fields["theta_m"], fields["rho_zz"])
        ins = {"theta_m_in": fields["theta_m"], "rho_zz_in": fields["rho_zz"], "u": u, "w": w}
    else:
        icd = K.init_coupled_diagnostics(dm, fields["theta"], fields["rho"], fields["qv"], u, w, theta_base)
        ins = {"theta": fields["theta"], "rho": fields["rho"], "qv": fields["qv"], "u": u, "w": w}
    tape.kernel("init_coupled", (1 if derived else 0, theta_base), ins, icd)
    out = dict(icd)
    v0 = torch.zeros_like(u)
    d = K.solve_diagnostics(dm, dt, apvm, None, u, icd["rho_zz"], v0)
    tape.kernel("solve_diagnostics", (float(dt), float(apvm), -1), {"u": u, "h": icd["rho_zz"], "v_in": v0}, d)
    out.update({k: d[k] for k in DIAG_KEYS})
    out["rho_edge"] = d["h_edge"]
    rc = K.reconstruct(dm, u)
    tape.kernel("reconstruct", (), {"u": u}, rc)
    for k in K.RECON_OUT:
        out[k] = rc[k]
    out["ur_cell"] = tape.copy(rc["uReconstructZonal"])
    out["vr_cell"] = tape.copy(rc["uReconstructMeridional"])
    return out
