"""``atm_srk3`` on the GPU: the drop-in for the reference's large step.

``atm_srk3(state, mesh, config, dt, strict=True, differentiable=False)``
keeps the signature, the dict contract and the error types of the reference
(``mpas_pytorch/dynamics/time_integration.py:4897-4956``): it does not
mutate the caller's mappings, it returns a new dict holding the advanced
prognostics (``u, w, theta_m, rho_zz, scalars, rt_diabatic_tend``) and the
diagnostics carried between steps (``atmosphere_model.py:157-201``).

The whole large step (setup, 3 dynamics substeps x 3 RK stages with their
acoustic substeps, split transport, reconstruction) is one call into
``libmpas_dycore_hip.so`` on the current HIP stream; nothing runs on the CPU
between kernels.  ``Srk3Stepper`` keeps the state resident across steps for
callers that integrate many steps (``run_simulation`` style).
"""

from __future__ import annotations

import ctypes as C
import warnings
from typing import Any, Mapping

import numpy as np
import torch

from . import _native as N
from .device import DeviceMesh, require_gpu, stream_handle, to_numpy
from .errors import MpasConfigError, MpasMissingFieldError, MpasUnsupportedCapabilityError

# Reference defaults (mpas_pytorch/config_defaults.py:16-136)
DEFAULTS = {
    "config_time_integration_order": 2,
    "config_split_dynamics_transport": True,
    "config_dynamics_split_steps": 3,
    "config_number_of_sub_steps": 2,
    "config_epssm": 0.1,
    "config_smdiv": 0.1,
    "config_apvm_upwinding": 0.5,
    "config_coef_3rd_order": 0.25,
    "config_smagorinsky_coef": 0.125,
    "config_visc4_2dsmag": 0.05,
    "config_del4u_div_factor": 10.0,
    "config_horiz_mixing": "2d_smagorinsky",
    "config_monotonic": True,
    "config_len_disp": 0.0,
    "config_rayleigh_damp_u": False,
    "config_rayleigh_damp_u_timescale_days": 5.0,
    "config_number_rayleigh_damp_u_levels": 6,
    "config_h_theta_eddy_visc2": 0.0,
    "config_h_mom_eddy_visc4": 0.0,
    "config_h_theta_eddy_visc4": 0.0,
    "config_mpas_cam_coef": 0.0,
    "config_number_cam_damping_levels": 4,
    "config_curvature_u": False,
    "config_v_mom_eddy_visc2": 0.0,
    "config_v_theta_eddy_visc2": 0.0,
    "config_mix_full": True,
    "config_scalar_advection": True,
    "config_positive_definite": False,
}
# Short aliases the reference also accepts (config_defaults.py:149-160)
ALIASES = {"smdiv": "config_smdiv", "len_disp": "config_len_disp", "monotonic": "config_monotonic",
           "time_integration_order": "config_time_integration_order",
           "del4u_div_factor": "config_del4u_div_factor", "visc4_2dsmag": "config_visc4_2dsmag",
           "positive_definite": "config_positive_definite", "apply_lbcs": "config_apply_lbcs",
           "rayleigh_damp_u": "config_rayleigh_damp_u",
           "rayleigh_damp_u_timescale_days": "config_rayleigh_damp_u_timescale_days",
           "number_rayleigh_damp_u_levels": "config_number_rayleigh_damp_u_levels",
           "horiz_mixing": "config_horiz_mixing", "mpas_cam_coef": "config_mpas_cam_coef",
           "number_cam_damping_levels": "config_number_cam_damping_levels",
           "apvm_upwinding": "config_apvm_upwinding", "epssm": "config_epssm",
           "coef_3rd_order": "config_coef_3rd_order", "smagorinsky_coef": "config_smagorinsky_coef",
           "c_s": "config_smagorinsky_coef", "nAcousticSteps": "config_number_of_sub_steps",
           "number_of_sub_steps": "config_number_of_sub_steps",
           "split_dynamics_transport": "config_split_dynamics_transport",
           "dynamics_split_steps": "config_dynamics_split_steps", "mix_full": "config_mix_full",
           "curvature_u": "config_curvature_u", "scalar_advection": "config_scalar_advection",
           "advect_scalars": "config_scalar_advection"}

# Options whose non-default values select code the GPU step does not have.
REQUIRED_VALUES = {
    "config_split_dynamics_transport": True,
    "config_apply_lbcs": False,
}
# Advection-order keys the reference validates as integers and otherwise ignores
# (config_defaults.py:138-146, :604-614): its numerics are third order throughout.
INERT_ORDER_KEYS = ("config_w_adv_order", "config_theta_adv_order", "config_scalar_adv_order",
                    "config_u_vadv_order", "config_w_vadv_order", "config_theta_vadv_order",
                    "config_scalar_vadv_order")

PROGNOSTIC_KEYS = ("u", "w", "theta_m", "rho_zz", "scalars", "rt_diabatic_tend")
CARRIED_KEYS = ("rho_p", "rtheta_p", "exner", "pressure_p", "ru", "rw")
DIAGNOSTIC_KEYS = ("v", "divergence", "vorticity", "ke", "pv_edge", "rho_edge")
OUTPUT_ONLY_KEYS = ("ur_cell", "vr_cell", "uReconstructX", "uReconstructY", "uReconstructZ",
                    "uReconstructZonal", "uReconstructMeridional", "ruAvg", "wwAvg")


def _cfg(config: Mapping[str, Any] | None) -> dict[str, Any]:
    """The defaults overlaid with ``config``; a full key name takes precedence
    over its short alias when both are given (config_defaults.py:745-774)."""
    out = dict(DEFAULTS)
    given = dict(config or {})
    for k, v in given.items():
        if k in ALIASES and ALIASES[k] not in given:
            out[ALIASES[k]] = v
    for k, v in given.items():
        if k not in ALIASES:
            out[k] = v
    return out


def _same(a: Any, b: Any) -> bool:
    if isinstance(b, bool) or isinstance(b, str):
        return a == b
    try:
        return float(a) == float(b)
    except (TypeError, ValueError):
        return False


def resolve_len_disp(config_len_disp: float, mesh: Mapping[str, Any], state: Mapping[str, Any] | None = None,
                     strict: bool = True) -> float:
    """The horizontal length scale of the damping and mixing coefficients.

    A positive ``config_len_disp`` is used as given; otherwise the mesh's
    ``nominalMinDc``; failing that the smallest positive ``dcEdge`` over rows
    1..nEdges, with a ``RuntimeWarning`` (the reference's fallbacks,
    ``initialization.py:365-380`` and ``time_integration.py:1882-1940``)."""
    if config_len_disp > 0.0:
        return float(config_len_disp)
    src = dict(state or {})
    src.update(mesh)
    nominal = src.get("nominalMinDc")
    if nominal is not None:
        v = float(np.asarray(to_numpy(nominal)).reshape(-1)[0])
        if v > 0.0:
            return v
    dc = src.get("dcEdge")
    if dc is not None:
        a = np.asarray(to_numpy(dc), dtype=np.float64).reshape(-1)
        ne = src.get("nEdges")
        if ne is not None and a.shape[0] >= int(to_numpy(ne)) + 1:
            a = a[1: int(to_numpy(ne)) + 1]
        pos = a[a > 0.0]
        if pos.size:
            fallback = float(pos.min())
            warnings.warn(f"config_len_disp <= 0 and no nominalMinDc: using the smallest dcEdge {fallback:.6f}",
                          RuntimeWarning, stacklevel=3)
            return fallback
    msg = ("config_len_disp <= 0 and the mesh has neither nominalMinDc nor a positive dcEdge to fall back on; "
           "set config_len_disp > 0")
    if strict:
        raise MpasConfigError(msg)
    return 0.0


def resolve_config(config: Mapping[str, Any] | None, state: Mapping[str, Any], mesh: Mapping[str, Any],
                   dt: float, n_scalars: int, strict: bool = True) -> N.ConfigStruct:
    """Check the configuration against what the GPU step implements and pack it."""
    c = _cfg(config)
    for key, want in REQUIRED_VALUES.items():
        if key in c and not _same(c[key], want):
            raise MpasUnsupportedCapabilityError(capability=key, requested=c[key])
    for key in INERT_ORDER_KEYS:
        if key in c:
            try:
                int(c[key])
            except (TypeError, ValueError):
                raise MpasConfigError(f"Config key '{key}' must be integer-like, got {c[key]!r}.") from None
    order = c["config_time_integration_order"]
    try:
        order_ok = not isinstance(order, bool) and float(order) == int(order) and int(order) in (2, 3)
    except (TypeError, ValueError):
        order_ok = False
    if not order_ok:
        raise MpasConfigError(f"config_time_integration_order must be 2 or 3, got {order!r}")
    dt = float(dt)
    if not dt > 0.0:
        raise MpasConfigError(f"dt must be > 0, got {dt}")
    ns, split = int(c["config_number_of_sub_steps"]), int(c["config_dynamics_split_steps"])
    if ns < 1 or split < 1:
        raise MpasConfigError("config_number_of_sub_steps and config_dynamics_split_steps must be >= 1")
    if split == 1 and n_scalars > 0 and bool(c["config_scalar_advection"]):
        # with one dynamics substep the reference skips its split transport loop
        # (time_integration.py:3953) and transports in every RK stage instead
        raise MpasUnsupportedCapabilityError(
            "config_dynamics_split_steps=1 transports the scalars inside every RK stage (the reference's "
            "coupled transport), which the GPU step does not implement", capability="config_dynamics_split_steps",
            requested=split)
    len_disp = resolve_len_disp(float(c["config_len_disp"]), mesh, state, strict=strict)
    coef3 = float(c["config_coef_3rd_order"])
    if "coef_3rd_order" in state and float(to_numpy(state["coef_3rd_order"])) != coef3:
        raise MpasUnsupportedCapabilityError(
            "state coef_3rd_order differs from config_coef_3rd_order", capability="coef_3rd_order",
            requested=float(to_numpy(state["coef_3rd_order"])))
    c_s = float(c["config_smagorinsky_coef"])
    if "c_s" in state:
        c_s = float(to_numpy(state["c_s"]))
    s = N.ConfigStruct()
    s.dt, s.len_disp, s.epssm, s.smdiv = dt, len_disp, float(c["config_epssm"]), float(c["config_smdiv"])
    s.apvm, s.coef3, s.c_s = float(c["config_apvm_upwinding"]), coef3, c_s
    s.visc4_2dsmag, s.del4u_div_factor = float(c["config_visc4_2dsmag"]), float(c["config_del4u_div_factor"])
    # positive-definite transport selects the same limited final stage as the monotonic one
    # (time_integration.py:3980-3987, :7942-7945)
    limited = bool(c["config_monotonic"]) or bool(c["config_positive_definite"])
    s.ns, s.split, s.monotonic, s.nScalars = ns, split, 1 if limited else 0, n_scalars
    s.scalar_advection = 1 if bool(c["config_scalar_advection"]) else 0
    s.moist_start = int(to_numpy(state.get("moist_start", 1)))
    s.moist_end = int(to_numpy(state.get("moist_end", n_scalars)))
    if s.moist_start < 1 or s.moist_end > n_scalars:
        raise MpasConfigError(f"moist species range {s.moist_start}..{s.moist_end} outside 1..{n_scalars}")
    s.diagnostics_current = 0
    s.order = int(order)
    n_levels = int(to_numpy(mesh["nVertLevels"])) if "nVertLevels" in mesh else None
    s.rayleigh_levels, s.rayleigh_coef_inverse = rayleigh_damping(c, n_levels)
    mixing = str(c["config_horiz_mixing"]).strip()
    if mixing not in ("2d_smagorinsky", "2d_fixed"):
        raise MpasConfigError(f"Unsupported config_horiz_mixing={c['config_horiz_mixing']!r}. Supported modes are "
                              "'2d_smagorinsky' and '2d_fixed'.")
    s.mix_fixed = 1 if mixing == "2d_fixed" else 0
    s.h_theta_eddy_visc2 = float(c["config_h_theta_eddy_visc2"])
    s.h_mom_eddy_visc4 = float(c["config_h_mom_eddy_visc4"])
    s.h_theta_eddy_visc4 = float(c["config_h_theta_eddy_visc4"])
    s.cam_levels, s.cam_visc2 = cam_damping(c, len_disp, n_levels)
    s.curvature = 1 if bool(c["config_curvature_u"]) else 0
    s.v_mom_eddy_visc2 = float(c["config_v_mom_eddy_visc2"])
    s.v_theta_eddy_visc2 = float(c["config_v_theta_eddy_visc2"])
    s.mix_full = 1 if bool(c["config_mix_full"]) else 0
    if (s.v_mom_eddy_visc2 > 0.0 or s.v_theta_eddy_visc2 > 0.0) and "zgrid" not in mesh and "zgrid" not in state:
        raise MpasMissingFieldError(fields=["zgrid"], source="mesh")
    if s.curvature and not ("angleEdge" in mesh and "latEdge" in mesh):
        raise MpasMissingFieldError(fields=[k for k in ("angleEdge", "latEdge") if k not in mesh], source="mesh")
    return s


def cam_damping(c: Mapping[str, Any], len_disp: float, n_levels: int | None) -> tuple[int, float]:
    """(damped top levels, 4 * 2.0833 * len_disp * coef) of the CAM damping of kdiff
    (dyn_tend.py:3434-3458); (0, 0.0) unless config_mpas_cam_coef > 0 and
    config_number_cam_damping_levels > 0, as the reference gates it.  More damped
    levels than the model has are rejected."""
    coef = float(c["config_mpas_cam_coef"])
    levels = c["config_number_cam_damping_levels"]
    if isinstance(levels, bool) or int(levels) != levels:
        raise MpasConfigError(f"config_number_cam_damping_levels must be an integer, got {levels!r}")
    levels = int(levels)
    if not (coef > 0.0 and levels > 0):
        return 0, 0.0
    if n_levels is not None and levels > n_levels:
        raise MpasConfigError(f"config_number_cam_damping_levels {levels} exceeds the {n_levels} model levels")
    return levels, 4.0 * 2.0833 * len_disp * coef


def rayleigh_damping(c: Mapping[str, Any], n_levels: int | None) -> tuple[int, float]:
    """(damped top levels, 1 / (levels * timescale in s)) of the Rayleigh damping of u
    (dyn_tend.py:10066-10072); (0, 0.0) when config_rayleigh_damp_u is off.  The reference
    divides by the level count and the timescale, so zero or negative values are rejected."""
    if not bool(c["config_rayleigh_damp_u"]):
        return 0, 0.0
    levels = c["config_number_rayleigh_damp_u_levels"]
    days = float(c["config_rayleigh_damp_u_timescale_days"])
    if isinstance(levels, bool) or int(levels) != levels or int(levels) < 1:
        raise MpasConfigError(f"config_number_rayleigh_damp_u_levels must be a positive integer, got {levels!r}")
    levels = int(levels)
    if n_levels is not None and levels > n_levels:
        raise MpasConfigError(f"config_number_rayleigh_damp_u_levels {levels} exceeds the {n_levels} model levels")
    if not days > 0.0:
        raise MpasConfigError(f"config_rayleigh_damp_u_timescale_days must be > 0, got {days}")
    return levels, 1.0 / (float(levels) * (days * 86400.0))


def _device_tensor(v: Any, device: torch.device, shape: tuple[int, ...], name: str) -> torch.Tensor:
    if isinstance(v, torch.Tensor):
        t = v.detach().to(device=device, dtype=torch.float64).clone()
    else:
        t = torch.as_tensor(np.array(v, dtype=np.float64), device=device)
    if tuple(t.shape) != shape:
        raise MpasMissingFieldError(f"{name} has shape {tuple(t.shape)}, expected {shape}", fields=[name],
                                    source="state")
    return t.contiguous()


_REORDER_CACHE: dict[int, tuple[Any, Any, dict]] = {}


def _reordered_mesh(mesh: Mapping[str, Any]):
    """(permutation, permuted mesh) of a mesh mapping, cached per mapping."""
    from .reorder import compute_sfc_permutation, permute_mapping

    hit = _REORDER_CACHE.get(id(mesh))
    if hit is not None and hit[0] is mesh:
        return hit[1], hit[2]
    perm = compute_sfc_permutation(mesh)
    dims = tuple(int(to_numpy(mesh[k])) for k in ("nCells", "nEdges", "nVertices"))
    pm = permute_mapping(mesh, perm, dims)
    if len(_REORDER_CACHE) > 4:
        _REORDER_CACHE.clear()
    _REORDER_CACHE[id(mesh)] = (mesh, perm, pm)
    return perm, pm


PRECISIONS = ("fp64", "mixed")


class Srk3Stepper:
    """Device-resident state advanced in place by successive large steps.

    With ``reorder=True`` the mesh and state are held on the device in
    space-filling-curve order (``reorder.py``); ``state_dict`` maps the
    results back to the caller's order, bit for bit the same values.

    ``precision="mixed"`` runs the mixed-precision step
    (``libmpas_dycore_hip_mixed.so``): the edge- and vertex-located state
    fields (``_native.MIXED_F32_FIELDS``) are held in float32 on the device,
    every cell and interface field in float64; ``state_dict`` widens the
    float32 fields back to float64.
    """

    def __init__(self, state: Mapping[str, Any], mesh: Mapping[str, Any], config: Mapping[str, Any] | None,
                 dt: float, device: torch.device | str | None = None, strict: bool = True,
                 reorder: bool = False, precision: str | None = None, rebind_base: bool = True):
        if precision is None:
            # the build's own configuration key: "fp64" (default) or "mixed"
            precision = str(_cfg(config).get("config_precision", "fp64"))
        if precision not in PRECISIONS:
            raise MpasConfigError(f"precision must be one of {PRECISIONS}, got {precision!r}")
        self.precision = precision
        self.mixed = precision == "mixed"
        self.device = require_gpu(device)
        if "scalars" not in state:
            raise MpasMissingFieldError(fields=["scalars"], source="state")
        sc = state["scalars"]
        n_scalars = int(sc.shape[2]) if len(sc.shape) == 3 else 0
        self.n_scalars = n_scalars
        self._inv_rows: dict[str, torch.Tensor] | None = None
        if reorder:
            from .reorder import permute_mapping

            perm, mesh_p = _reordered_mesh(mesh)
            dims = tuple(int(to_numpy(mesh[k])) for k in ("nCells", "nEdges", "nVertices"))
            state = permute_mapping(state, perm, dims)
            mesh = mesh_p
            nc, ne, nv = dims
            self._inv_rows = {
                "cell": torch.as_tensor(np.concatenate([perm.cell_old_to_new, [nc]]), device=self.device),
                "edge": torch.as_tensor(np.concatenate([perm.edge_old_to_new, [ne]]), device=self.device),
                "vertex": torch.as_tensor(np.concatenate([perm.vertex_old_to_new, [nv]]), device=self.device),
            }
        self.mesh = DeviceMesh.get(mesh, self.device, state, rebind_base=rebind_base)
        self.cfg = resolve_config(config, state, mesh, dt, n_scalars, strict=strict)
        dm = self.mesh
        missing = [k for k in PROGNOSTIC_KEYS[:5] + CARRIED_KEYS if k not in state]
        if missing:
            raise MpasMissingFieldError(fields=missing, source="state")
        t: dict[str, torch.Tensor] = {}
        with torch.cuda.device(self.device):
            for k in PROGNOSTIC_KEYS + CARRIED_KEYS + DIAGNOSTIC_KEYS + OUTPUT_ONLY_KEYS:
                shape = dm.shape_of(k, n_scalars)
                src = state.get(k)
                if src is None and k == "rho_edge":
                    src = state.get("h_edge")
                if src is None:
                    t[k] = torch.zeros(shape, dtype=torch.float64, device=self.device)
                else:
                    t[k] = _device_tensor(src, self.device, shape, k)
                if self.mixed and k in N.MIXED_F32_FIELDS:
                    t[k] = t[k].to(torch.float32).contiguous()
            self.work = dm.workspace(n_scalars, self.mixed)
        self.t = t
        self._diagnostics_current = False
        self._st = N.StateStruct(*[t[k].data_ptr() for k in N.STATE_FIELDS])
        self._ws = N.WorkspaceStruct(self.work.data_ptr(), self.work.numel())

    @property
    def lib(self):
        """The library whose step this stepper runs (its tuning table and timer)."""
        return N.load_mixed() if self.mixed else N.load()

    def _launch(self) -> None:
        self.cfg.diagnostics_current = 1 if self._diagnostics_current else 0
        stream = C.c_void_p(stream_handle(self.device))
        if self.mixed:
            lib = N.load_mixed()
            N.check(lib.mpas_hip_mixed_srk3_step(self.mesh.ref, C.byref(self._st), C.byref(self._ws),
                                                 C.byref(self.cfg), stream), "mpas_hip_mixed_srk3_step", lib)
        else:
            lib = N.load()
            N.check(lib.mpas_hip_srk3_step(self.mesh.ref, C.byref(self._st), C.byref(self._ws), C.byref(self.cfg),
                                           stream), "mpas_hip_srk3_step")
        self._diagnostics_current = True

    def step(self, n: int = 1) -> None:
        """Advance ``n`` large steps in place on the current stream (no host sync)."""
        with torch.cuda.device(self.device):
            for _ in range(n):
                if self._graph is not None and self._diagnostics_current:
                    self._graph.replay()
                else:
                    self._launch()

    _graph = None

    def capture_graph(self) -> None:
        """Capture one steady-state large step in a HIP graph; later ``step``
        calls replay it.  The step's kernels allocate nothing, never
        synchronise with the host and clear memory with a fill kernel, so
        the captured sequence is the eager one.  The first step of a stepper
        (diagnostic refresh) always runs eagerly."""
        with torch.cuda.device(self.device):
            if not self._diagnostics_current:
                self._launch()
            torch.cuda.synchronize(self.device)
            g = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream(self.device)
            side.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.stream(side):
                # capture only: a captured launch does not execute
                with torch.cuda.graph(g, stream=side):
                    self._launch()
            torch.cuda.current_stream(self.device).wait_stream(side)
            torch.cuda.synchronize(self.device)
            self._graph = g

    def state_dict(self) -> dict[str, torch.Tensor]:
        """Device tensors in the caller's row order (views when not reordered)."""
        if self._inv_rows is None:
            out = dict(self.t)
        else:
            dm = self.mesh
            kind = {dm.nc + 1: "cell", dm.ne + 1: "edge", dm.nv + 1: "vertex"}
            out = {k: v.index_select(0, self._inv_rows[kind[v.shape[0]]]) for k, v in self.t.items()}
        if self.mixed:
            out = {k: (v.to(torch.float64) if v.dtype == torch.float32 else v) for k, v in out.items()}
        out["h_edge"] = out["rho_edge"]
        return out


class _Srk3Step(torch.autograd.Function):
    """One large step as a single autograd node, in both modes.

    By default the forward runs the fused step and keeps only the step's
    input state (the per-large-step checkpoint); the backward recomputes the
    step from it on a tape of per-function kernels, acoustic substeps
    included, and applies the hand-written adjoint kernels
    (``adjoint.step_vjp``); under ``torch.autograd.forward_ad`` the jvp runs
    the same tape forward through the tangent-linear kernels
    (``adjoint.step_jvp``).

    With ``box["keep_tape"]`` the forward runs the step on the tape instead
    (bit for bit the fused step's results) and keeps the tape until the
    first backward or jvp uses it, which then skips the recompute; a later
    use (``retain_graph``) recomputes."""

    @staticmethod
    def forward(box: dict, *inputs: torch.Tensor):
        from .adjoint import STEP_INPUTS, Tape, taped_step

        state = dict(box["state"])
        state.update({k: v.detach() for k, v in zip(box["in_keys"], inputs)})
        u = state["u"]
        out_device = u.device if isinstance(u, torch.Tensor) else torch.device("cpu")
        gpu = out_device if out_device.type == "cuda" else None
        stepper = Srk3Stepper(state, box["mesh"], box["config"], box["dt"], device=gpu, strict=box["strict"],
                              rebind_base=box.get("rebind_base", True))
        # the checkpoint is binary64 (widened from the mixed step's binary32 fields): the
        # backward recomputes the step in binary64 around the trajectory the forward took
        ckpt = {k: stepper.t[k].to(torch.float64).clone() for k in STEP_INPUTS}
        cfg = N.ConfigStruct()
        C.pointer(cfg)[0] = stepper.cfg
        cfg.diagnostics_current = 0
        kept = None
        if box.get("keep_tape", False) and not stepper.mixed:
            tape = Tape(stepper.mesh)
            with torch.cuda.device(stepper.device):
                taped = taped_step(stepper.mesh, ckpt, cfg, float(box["dt"]), tape)
            # the caller gets copies: the tape's tensors stay private to the tape
            new = {k: taped[k].clone() for k in N.STATE_FIELDS}
            new["h_edge"] = taped["rho_edge"].clone()
            kept = (tape, taped)
        else:
            stepper.step(1)
            new = stepper.state_dict()
            new["h_edge"] = new["h_edge"].clone()
        out_keys = sorted(new)
        box["out_keys"] = out_keys
        box["_ctx"] = (stepper.mesh, cfg, float(box["dt"]), ckpt, list(box["in_keys"]), out_keys,
                       [x.device for x in inputs], kept)
        if out_device.type != "cuda":
            torch.cuda.synchronize(stepper.device)
        return tuple(new[k].to(out_device) for k in out_keys)

    @staticmethod
    def setup_context(ctx, inputs, output):
        box = inputs[0]
        ctx.dm, ctx.cfg, ctx.dt, ctx.ckpt, ctx.in_keys, ctx.out_keys, ctx.in_devices, ctx.kept = box.pop("_ctx")
        ctx.out_like = [(o.shape, o.dtype, o.device) for o in output]

    @staticmethod
    def _take_tape(ctx):
        """(tape, taped outputs) recorded by the forward, handed out once."""
        kept, ctx.kept = ctx.kept, None
        return kept

    @staticmethod
    def backward(ctx, *grads: torch.Tensor):
        from .adjoint import step_vjp, tape_vjp

        dev = ctx.dm.device
        cot = {k: g.to(device=dev, dtype=torch.float64) for k, g in zip(ctx.out_keys, grads) if g is not None}
        kept = _Srk3Step._take_tape(ctx)
        with torch.cuda.device(dev):
            if kept is not None:
                gin = tape_vjp(kept[0], ctx.ckpt, kept[1], cot)
            else:
                _, gin = step_vjp(ctx.dm, ctx.ckpt, ctx.cfg, ctx.dt, cot)
        return (None,) + tuple(gin[k].to(d) for k, d in zip(ctx.in_keys, ctx.in_devices))

    @staticmethod
    def jvp(ctx, _box_tangent, *tangents: torch.Tensor):
        from .adjoint import step_jvp, tape_jvp

        dev = ctx.dm.device
        tan = {k: t.to(device=dev, dtype=torch.float64) for k, t in zip(ctx.in_keys, tangents) if t is not None}
        kept = _Srk3Step._take_tape(ctx)
        with torch.cuda.device(dev):
            if kept is not None:
                tout = tape_jvp(kept[0], ctx.ckpt, kept[1], tan)
            else:
                _, tout = step_jvp(ctx.dm, ctx.ckpt, ctx.cfg, ctx.dt, tan)
        res = []
        for k, (shape, dtype, device) in zip(ctx.out_keys, ctx.out_like):
            t = tout.get(k)
            res.append(t.to(device=device, dtype=dtype) if t is not None
                       else torch.zeros(shape, dtype=dtype, device=device))
        return tuple(res)


def _atm_srk3_differentiable(state: Mapping[str, Any], mesh: Mapping[str, Any], config: Mapping[str, Any],
                             dt: float, strict: bool, keep_tape: bool = False,
                             rebind_base: bool = True) -> dict[str, Any]:
    """The differentiable step.  The forward records the step's tape when
    ``keep_tape`` is set (the caller knows a backward follows) or when an
    input carries a forward-mode tangent (the jvp runs right after the
    forward); otherwise the backward recomputes it from the checkpoint.
    ``rebind_base=False`` keeps the base state the device mesh holds (a run
    passes it for every step after its first)."""
    from torch.autograd import forward_ad

    from .adjoint import STEP_INPUTS

    in_keys = [k for k in STEP_INPUTS if isinstance(state.get(k), torch.Tensor)]
    dual = any(forward_ad.unpack_dual(state[k]).tangent is not None for k in in_keys)
    box = {"state": {k: v for k, v in state.items() if k not in in_keys}, "mesh": mesh, "config": config,
           "dt": dt, "strict": strict, "in_keys": in_keys, "keep_tape": bool(keep_tape or dual),
           "rebind_base": bool(rebind_base)}
    outs = _Srk3Step.apply(box, *[state[k] for k in in_keys])
    result = dict(state)
    result.update(zip(box["out_keys"], outs))
    return result


def atm_srk3(state: Mapping[str, Any], mesh: Mapping[str, Any], config: Mapping[str, Any], dt: float,
             strict: bool = True, differentiable: bool = False) -> dict[str, Any]:
    """One split-explicit RK3 large step on the GPU (reference ``atm_srk3``).

    Inputs may be CPU or GPU tensors (or NumPy arrays) in the reference
    layout; the step runs on the current ROCm device and the results come
    back on the device of ``state["u"]``.  With ``differentiable=True`` the
    returned tensors carry an autograd graph whose backward applies the
    hand-written adjoint kernels, and forward-mode dual numbers
    (``torch.autograd.forward_ad``) get their tangents from the
    tangent-linear kernels.
    """
    if differentiable:
        return _atm_srk3_differentiable(state, mesh, config, dt, strict)
    u = state.get("u")
    out_device = u.device if isinstance(u, torch.Tensor) else torch.device("cpu")
    gpu = out_device if out_device.type == "cuda" else None
    stepper = Srk3Stepper(state, mesh, config, dt, device=gpu, strict=strict)
    stepper.step(1)
    new = stepper.state_dict()
    result = dict(state)
    if out_device.type != "cuda":
        torch.cuda.synchronize(stepper.device)
        new = {k: v.to(out_device) for k, v in new.items()}
    result.update(new)
    return result
