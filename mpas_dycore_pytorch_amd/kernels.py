"""Per-kernel GPU entry points (one per reference ``atm_*_work`` function).

Each function takes a :class:`DeviceMesh` and float64 device tensors in the
reference layout, launches the matching ``mpas_hip_*`` entry point on the
current stream, and returns new device tensors in a dict keyed by the
reference's output names.  The large step itself does not go through these
wrappers (``atm_srk3`` makes one library call); they exist for per-kernel
parity checks and timings.
"""

from __future__ import annotations

import ctypes as C
from typing import Mapping

import torch

from . import _native as N
from .device import DeviceMesh, stream_handle

DYN_INPUTS = ("u", "v", "w", "ru", "rw", "ru_save", "rw_save", "theta_m", "theta_m_save", "rho_zz",
              "rho_edge", "divergence", "vorticity", "ke", "pv_edge", "pp", "rr_save", "cqw", "cqu",
              "rt_diabatic_tend", "qtot", "ur_cell", "vr_cell")
DYN_CARRIED = ("tend_rho", "tend_u_euler", "tend_w_euler", "tend_theta_euler", "kdiff")
DYN_OUTPUTS = ("tend_u", "tend_w", "tend_theta", "h_divergence", "rthdynten")
ACOUSTIC_IN = ("cofrz", "cofwr", "cofwz", "coftz", "cofwt", "a_tri", "alpha_tri", "gamma_tri",
               "rho_zz", "theta_m", "exner", "cqu", "tend_ru", "tend_rho", "tend_rt", "tend_rw",
               "w", "rw", "rw_save")
ACOUSTIC_IO = ("ru_p", "ruAvg", "rho_pp", "rtheta_pp", "rtheta_pp_old", "rw_p", "wwAvg")
RECOVER_IN = ("rho_p_save", "rho_pp", "rtheta_p_save", "rtheta_pp", "rt_diabatic_tend", "rw_save",
              "rw_p", "ru_save", "ru_p")
VIMP_OUT = ("cofrz", "cofwr", "cofwz", "coftz", "cofwt", "a_tri", "alpha_tri", "gamma_tri")
DIAG_OUT = ("h_edge", "v", "divergence", "vorticity", "ke", "pv_edge")
RECON_OUT = ("uReconstructX", "uReconstructY", "uReconstructZ", "uReconstructZonal", "uReconstructMeridional")


def _d(t: torch.Tensor, dm: DeviceMesh) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(t)
    return t.to(device=dm.device, dtype=torch.float64).contiguous()


def _p(t: torch.Tensor) -> int:
    return t.data_ptr()


def _z(dm: DeviceMesh, *shape: int) -> torch.Tensor:
    return torch.zeros(shape, dtype=torch.float64, device=dm.device)


def _e(dm: DeviceMesh, *shape: int) -> torch.Tensor:
    """An output buffer its kernel writes in full (every row and level)."""
    return torch.empty(shape, dtype=torch.float64, device=dm.device)


def _ep(dm: DeviceMesh, *shape: int) -> torch.Tensor:
    """An output buffer its kernel writes in full except the padding row,
    which is zero."""
    t = torch.empty(shape, dtype=torch.float64, device=dm.device)
    t[shape[0] - 1].zero_()
    return t


def _s(dm: DeviceMesh) -> C.c_void_p:
    return C.c_void_p(stream_handle(dm.device))


def dyn_params(rk_step: int, dt: float, len_disp: float, c_s: float = 0.125, visc4_2dsmag: float = 0.05,
               coef_3rd_order: float = 0.25, del4u_div_factor: float = 10.0,
               prandtl: float = 1.0, rayleigh_levels: int = 0,
               rayleigh_coef_inverse: float = 0.0, mix_fixed: bool = False, h_theta_eddy_visc2: float = 0.0,
               h_mom_eddy_visc4: float = 0.0, h_theta_eddy_visc4: float = 0.0, cam_levels: int = 0,
               cam_visc2: float = 0.0, curvature: bool = False, v_mom_eddy_visc2: float = 0.0,
               v_theta_eddy_visc2: float = 0.0, mix_full: bool = True) -> N.DynParamsStruct:
    """Scalar set-up of the dynamics tendency (dyn_tend.py:3367-3458, :9616-9626).

    ``mix_fixed`` selects the "2d_fixed" horizontal mixing: kdiff is
    ``h_theta_eddy_visc2`` everywhere and the del4 coefficients are the given
    ones; otherwise Smagorinsky kdiff and ``visc4_2dsmag * len_disp**3``.  The
    CAM damping floors kdiff on the top ``cam_levels`` levels (0: off) at
    ``cam_visc2 * (1 - (nLevels - (k + 1)) / cam_levels)``.  The Rayleigh
    damping of u acts on ``rayleigh_levels`` top levels (0: off) with its
    coefficient 1 / (levels * timescale) (dyn_tend.py:10066-10072).  ``curvature``
    adds the curvature terms of u and w (config_curvature_u, dyn_tend.py:2471-2505,
    :4241-4267), which read the ``ur_cell`` / ``vr_cell`` inputs.  The vertical
    eddy viscosities mix u, w and theta on the first RK stage, without
    ``mix_full`` as departures from the reference profiles (dyn_tend.py:2731-2781,
    :4633-4647, :5060-5097)."""
    p = N.DynParamsStruct()
    p.curvature = 1 if curvature else 0
    p.v_mom_eddy_visc2 = float(v_mom_eddy_visc2)
    p.v_theta_eddy_visc2 = float(v_theta_eddy_visc2)
    p.mix_full = 1 if mix_full else 0
    p.rk_step = int(rk_step)
    p.dt = float(dt)
    p.coef_3rd_order = float(coef_3rd_order)
    inv_dt = 1.0 / float(dt)
    p.kdiff_coef = (c_s * len_disp) ** 2
    p.kdiff_cap = (0.01 * len_disp ** 2) * inv_dt
    if mix_fixed:
        p.h_mom_eddy_visc4 = float(h_mom_eddy_visc4)
        p.h_theta_eddy_visc4 = float(h_theta_eddy_visc4)
    else:
        p.h_mom_eddy_visc4 = visc4_2dsmag * len_disp ** 3
        p.h_theta_eddy_visc4 = p.h_mom_eddy_visc4
    p.del4u_div_factor = float(del4u_div_factor)
    p.prandtl_inv = 1.0 / prandtl
    p.rayleigh_levels = int(rayleigh_levels)
    p.rayleigh_coef_inverse = float(rayleigh_coef_inverse)
    p.mix_fixed = 1 if mix_fixed else 0
    p.kdiff_fixed = float(h_theta_eddy_visc2)
    p.cam_levels = int(cam_levels)
    p.cam_visc2 = float(cam_visc2)
    return p


def acoustic_params(small_step: int, dts: float, epssm: float = 0.1) -> N.AcousticParamsStruct:
    """Scalars of the acoustic substep (acoustic_step.py:1297-1340)."""
    p = N.AcousticParamsStruct()
    p.small_step = int(small_step)
    p.dts = float(dts)
    p.epssm = float(epssm)
    p.resm = (1.0 - epssm) / (1.0 + epssm)
    rgas, cp = 287.0, 1004.5
    p.c2 = cp * (rgas / (cp - rgas))
    return p


def moist_coefficients(dm: DeviceMesh, scalars, moist_start: int, moist_end: int) -> dict:
    sc = _d(scalars, dm)
    S = sc.shape[2]
    out = {"cqw": _e(dm, dm.nc + 1, dm.L + 1), "cqu": _e(dm, dm.ne + 1, dm.L), "qtot": _e(dm, dm.nc + 1, dm.L)}
    N.check(N.load().mpas_hip_moist_coefficients(dm.ref, _p(sc), S, int(moist_start), int(moist_end),
                                                 _p(out["cqw"]), _p(out["cqu"]), _p(out["qtot"]), _s(dm)),
            "moist_coefficients")
    return out


def vert_imp_coefs(dm: DeviceMesh, dts: float, epssm: float, cqw, p, t, rt, qtot) -> dict:
    ins = [_d(x, dm) for x in (cqw, p, t, rt, qtot)]
    nc, L = dm.nc, dm.L
    out = {"cofrz": _e(dm, L), "cofwr": _e(dm, nc + 1, L), "cofwz": _e(dm, nc + 1, L),
           "coftz": _e(dm, nc + 1, L + 1), "cofwt": _e(dm, nc + 1, L), "a_tri": _e(dm, nc + 1, L),
           "alpha_tri": _e(dm, nc + 1, L), "gamma_tri": _e(dm, nc + 1, L)}
    N.check(N.load().mpas_hip_vert_imp_coefs(dm.ref, float(dts), float(epssm), *[_p(x) for x in ins],
                                             *[_p(out[k]) for k in VIMP_OUT], _s(dm)), "vert_imp_coefs")
    return out


def solve_diagnostics(dm: DeviceMesh, dt: float, apvm: float, rk_step, u, h, v_in=None) -> dict:
    u, h = _d(u, dm), _d(h, dm)
    v_in = _d(v_in, dm) if v_in is not None else _z(dm, dm.ne + 1, dm.L)
    out = {"h_edge": _e(dm, dm.ne + 1, dm.L), "v": _e(dm, dm.ne + 1, dm.L), "divergence": _e(dm, dm.nc + 1, dm.L),
           "vorticity": _e(dm, dm.nv + 1, dm.L), "ke": _e(dm, dm.nc + 1, dm.L), "pv_edge": _e(dm, dm.ne + 1, dm.L)}
    rk = -1 if rk_step is None else int(rk_step)
    N.check(N.load().mpas_hip_solve_diagnostics(dm.ref, float(dt), float(apvm), rk, _p(u), _p(h), _p(v_in),
                                                *[_p(out[k]) for k in DIAG_OUT], _p(dm.kernel_scratch()), _s(dm)),
            "solve_diagnostics")
    return out


def dyn_tend(dm: DeviceMesh, params: N.DynParamsStruct, fields: Mapping, carried: Mapping | None = None) -> dict:
    nc, ne, L = dm.nc, dm.ne, dm.L
    # the cell winds are read only by the curvature terms
    ins = [_d(fields[k], dm) if (k in fields or k not in ("ur_cell", "vr_cell")) else _z(dm, nc + 1, L)
           for k in DYN_INPUTS]
    carried = carried or {}
    shapes = {"tend_rho": (nc + 1, L), "tend_u_euler": (ne + 1, L), "tend_w_euler": (nc + 1, L + 1),
              "tend_theta_euler": (nc + 1, L), "kdiff": (nc + 1, L)}
    # without carried inputs (the first RK stage) the kernels write the carried
    # tendencies in full but for their padding rows, and kdiff in full
    io = {k: (_d(carried[k], dm).clone() if k in carried else (_e if k == "kdiff" else _ep)(dm, *shapes[k]))
          for k in DYN_CARRIED}
    out = {"tend_u": _e(dm, ne + 1, L), "tend_w": _e(dm, nc + 1, L + 1), "tend_theta": _e(dm, nc + 1, L),
           "h_divergence": _e(dm, nc + 1, L), "rthdynten": _e(dm, nc + 1, L)}
    N.check(N.load().mpas_hip_dyn_tend(dm.ref, C.byref(params), *[_p(x) for x in ins],
                                       *[_p(io[k]) for k in DYN_CARRIED], *[_p(out[k]) for k in DYN_OUTPUTS],
                                       _p(dm.kernel_scratch()), _s(dm)), "dyn_tend")
    out.update(io)
    return out


def smlstep_pert(dm: DeviceMesh, u_tend, w_tend) -> torch.Tensor:
    wt = _d(w_tend, dm).clone()
    ut = _d(u_tend, dm)
    N.check(N.load().mpas_hip_smlstep_pert(dm.ref, _p(ut), _p(wt), _s(dm)), "smlstep_pert")
    return wt


def acoustic_step(dm: DeviceMesh, params: N.AcousticParamsStruct, fields: Mapping) -> dict:
    ins = [_d(fields[k], dm) for k in ACOUSTIC_IN]
    io = {k: _d(fields[k], dm).clone() for k in ACOUSTIC_IO}
    N.check(N.load().mpas_hip_acoustic_step(dm.ref, C.byref(params), *[_p(x) for x in ins],
                                            *[_p(io[k]) for k in ACOUSTIC_IO], _s(dm)), "acoustic_step")
    return io


def divergence_damping(dm: DeviceMesh, dts, smdiv, len_disp, theta_m, rtheta_pp, rtheta_pp_old, ru_p):
    out = _d(ru_p, dm).clone()
    ins = [_d(x, dm) for x in (theta_m, rtheta_pp, rtheta_pp_old)]
    N.check(N.load().mpas_hip_divergence_damping(dm.ref, float(dts), float(smdiv), float(len_disp),
                                                 *[_p(x) for x in ins], _p(out), _s(dm)), "divergence_damping")
    return out


def recover_large_step(dm: DeviceMesh, rk_step: int, dt: float, ns: int, fields: Mapping,
                       exner_in=None, pressure_p_in=None) -> dict:
    ins = [_d(fields[k], dm) for k in RECOVER_IN]
    nc, ne, L = dm.nc, dm.ne, dm.L
    ww = _d(fields["wwAvg"], dm).clone()
    ra = _d(fields["ruAvg"], dm).clone()
    # written in full but for the padding rows (rho_zz in full); rw leaves its
    # boundary interfaces to the zero fill
    out = {k: _ep(dm, nc + 1, L) for k in ("rho_p", "rtheta_p", "theta_m")}
    out["rho_zz"] = _e(dm, nc + 1, L)
    out["exner"] = _d(exner_in, dm).clone() if exner_in is not None else _z(dm, nc + 1, L)
    out["pressure_p"] = _d(pressure_p_in, dm).clone() if pressure_p_in is not None else _z(dm, nc + 1, L)
    out["rw"] = _z(dm, nc + 1, L + 1)
    out["w"] = _ep(dm, nc + 1, L + 1)
    out["ru"] = _ep(dm, ne + 1, L)
    out["u"] = _ep(dm, ne + 1, L)
    names = ("rho_p", "rho_zz", "rtheta_p", "theta_m", "exner", "pressure_p", "rw", "w", "ru", "u")
    N.check(N.load().mpas_hip_recover_large_step(dm.ref, int(rk_step), float(dt), int(ns), *[_p(x) for x in ins],
                                                 _p(ww), _p(ra), *[_p(out[k]) for k in names], _s(dm)),
            "recover_large_step")
    out["wwAvg"] = ww
    out["ruAvg"] = ra
    return out


def advance_scalars(dm: DeviceMesh, dt, coef3, weight_old, weight_new, scalar_old, scalar_new,
                    rho_zz_old, rho_zz_new, uhAvg, wwAvg):
    so, sn = _d(scalar_old, dm), _d(scalar_new, dm)
    S = so.shape[2]
    out = torch.empty_like(sn)
    hf = _ep(dm, dm.ne + 1, dm.L, S)
    ins = [_d(x, dm) for x in (rho_zz_old, rho_zz_new, uhAvg, wwAvg)]
    N.check(N.load().mpas_hip_advance_scalars(dm.ref, S, float(dt), float(coef3), float(weight_old),
                                              float(weight_new), _p(so), _p(sn), *[_p(x) for x in ins],
                                              _p(out), _p(hf), _s(dm)), "advance_scalars")
    # the kernel writes the live rows; the padding row carries the input
    out[dm.nc] = sn[dm.nc]
    return out, hf


def advance_scalars_mono(dm: DeviceMesh, dt, coef3, scalars_old, scalars_new, rho_zz_old, uhAvg, wwAvg):
    so, sn = _d(scalars_old, dm), _d(scalars_new, dm)
    S = so.shape[2]
    out = sn.clone()
    rzi = _ep(dm, dm.nc + 1, dm.L)
    ins = [_d(x, dm) for x in (rho_zz_old, uhAvg, wwAvg)]
    N.check(N.load().mpas_hip_advance_scalars_mono(dm.ref, S, float(dt), float(coef3), _p(so), _p(sn),
                                                   *[_p(x) for x in ins],
                                                   _p(out), _p(rzi), _p(dm.kernel_scratch()), _s(dm)),
            "advance_scalars_mono")
    return out, rzi


ICD_OUT = ("theta_m", "rho_zz", "ru", "rw", "rho_p", "rtheta_base", "rtheta_p", "exner", "exner_base",
           "pressure_p", "pressure_base")


def init_coupled_diagnostics(dm: DeviceMesh, theta, rho, qv, u, w, theta_base, theta_m=None,
                             rho_zz=None) -> dict:
    """Cold-start coupling (``atm_init_coupled_diagnostics``) of padded fields.

    From raw ``theta``, ``rho``, ``qv``; or, when ``theta_m`` and ``rho_zz``
    are given, from those (the reference's derived path,
    ``init_coupled_diagnostics.py:256-262``), ``theta``/``rho``/``qv`` unused."""
    derived = theta_m is not None and rho_zz is not None
    nc, ne, L = dm.nc, dm.ne, dm.L
    out = {k: _z(dm, nc + 1, L) for k in ICD_OUT}
    out["ru"] = _z(dm, ne + 1, L)
    out["rw"] = _z(dm, nc + 1, L + 1)
    uw = [_d(x, dm) for x in (u, w, theta_base)]
    if derived:
        out["theta_m"].copy_(_d(theta_m, dm))
        out["rho_zz"].copy_(_d(rho_zz, dm))
        raw = [None, None, None]
    else:
        keep = [_d(x, dm) for x in (theta, rho, qv)]  # alive across the launch
        raw = [_p(x) for x in keep]
    N.check(N.load().mpas_hip_init_coupled_diagnostics(dm.ref, *raw, *[_p(x) for x in uw],
                                                       *[_p(out[k]) for k in ICD_OUT], _s(dm)),
            "init_coupled_diagnostics")
    return out


def reconstruct(dm: DeviceMesh, u) -> dict:
    out = {k: _e(dm, dm.nc + 1, dm.L) for k in RECON_OUT}
    ud = _d(u, dm)
    N.check(N.load().mpas_hip_reconstruct(dm.ref, _p(ud), *[_p(out[k]) for k in RECON_OUT], _s(dm)),
            "reconstruct")
    out["ur_cell"] = out["uReconstructZonal"].clone()
    out["vr_cell"] = out["uReconstructMeridional"].clone()
    return out
