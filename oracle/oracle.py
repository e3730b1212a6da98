"""Python binding of the CPU oracle (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use this module, and only as the checker / timed baseline.  It
wraps ``mpas_oracle.c`` (the loop-level restatement of the reference
kernels) and restates the orchestration of ``atm_srk3``
(``mpas_pytorch/dynamics/time_integration.py:3420-4894``, stage body
``execute_rk_step`` :5000-9154) in Python.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpas_oracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

MESH_FIELDS = [
    ("nCells", C.c_int32), ("nEdges", C.c_int32), ("nVertices", C.c_int32), ("nLevels", C.c_int32),
    ("maxEdges", C.c_int32), ("maxEdges2", C.c_int32), ("vertexDegree", C.c_int32), ("maxAdv", C.c_int32),
    ("cellsOnEdge", _ip), ("verticesOnEdge", _ip), ("nEdgesOnCell", _ip), ("edgesOnCell", _ip),
    ("cellsOnCell", _ip), ("verticesOnCell", _ip), ("kiteForCell", _ip), ("edgesOnVertex", _ip),
    ("nEdgesOnEdge", _ip), ("edgesOnEdge", _ip), ("nAdvCellsForEdge", _ip), ("advCellsForEdge", _ip),
    ("edgesOnCell_sign", _dp), ("edgesOnVertex_sign", _dp), ("weightsOnEdge", _dp),
    ("kiteAreasOnVertex", _dp), ("dcEdge", _dp), ("dvEdge", _dp), ("invDcEdge", _dp), ("invDvEdge", _dp),
    ("invAreaCell", _dp), ("invAreaTriangle", _dp), ("fEdge", _dp), ("fVertex", _dp),
    ("meshScalingDel2", _dp), ("meshScalingDel4", _dp), ("adv_coefs", _dp), ("adv_coefs_3rd", _dp),
    ("defc_a", _dp), ("defc_b", _dp), ("coeffs_reconstruct", _dp), ("latCell", _dp), ("lonCell", _dp),
    ("zz", _dp), ("zxu", _dp), ("dss", _dp), ("zb_cell", _dp), ("zb3_cell", _dp),
    ("rdzw", _dp), ("rdzu", _dp), ("fzm", _dp), ("fzp", _dp), ("rdnw", _dp), ("fnm", _dp), ("fnp", _dp),
    ("cf1", C.c_double), ("cf2", C.c_double), ("cf3", C.c_double),
    ("rho_base", _dp), ("rtheta_base", _dp), ("exner_base", _dp),
]

# connectivity fields converted from 1-based (0 = none) to 0-based (-1 = none)
_ONE_BASED = {"cellsOnEdge", "verticesOnEdge", "edgesOnCell", "cellsOnCell", "verticesOnCell",
              "kiteForCell", "edgesOnVertex", "edgesOnEdge", "advCellsForEdge"}
_COUNTS = {"nEdgesOnCell", "nEdgesOnEdge", "nAdvCellsForEdge"}


class OrcMesh(C.Structure):
    _fields_ = MESH_FIELDS


class OrcDynParams(C.Structure):
    _fields_ = [("rk_step", C.c_int), ("dt", C.c_double), ("coef_3rd_order", C.c_double),
                ("kdiff_coef", C.c_double), ("kdiff_cap", C.c_double),
                ("h_mom_eddy_visc4", C.c_double), ("h_theta_eddy_visc4", C.c_double),
                ("del4u_div_factor", C.c_double), ("prandtl_inv", C.c_double),
                ("rayleigh_levels", C.c_int), ("rayleigh_coef_inverse", C.c_double),
                ("mix_fixed", C.c_int), ("kdiff_fixed", C.c_double), ("cam_levels", C.c_int),
                ("cam_visc2", C.c_double), ("curvature", C.c_int), ("v_mom_eddy_visc2", C.c_double),
                ("v_theta_eddy_visc2", C.c_double), ("mix_full", C.c_int)]


class OrcAcousticParams(C.Structure):
    _fields_ = [("small_step", C.c_int), ("dts", C.c_double), ("epssm", C.c_double),
                ("resm", C.c_double), ("c2", C.c_double)]


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (no fused multiply-add)."""
    src = os.path.join(HERE, "mpas_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        cmd = ["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
               "-o", LIB_PATH, src, "-lm"]
        subprocess.run(cmd, check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        assert _lib.orc_mesh_sizeof() == C.sizeof(OrcMesh), "oracle mesh struct mismatch"
        assert _lib.orc_dyn_params_sizeof() == C.sizeof(OrcDynParams)
        assert _lib.orc_acoustic_params_sizeof() == C.sizeof(OrcAcousticParams)
    return _lib


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _ptr(a: np.ndarray):
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


class Mesh:
    """Oracle view of a reference-layout mesh/state dict (NumPy arrays)."""

    def __init__(self, d: dict):
        self.keep: dict[str, np.ndarray] = {}
        s = OrcMesh()
        nc, ne, nv = int(d["nCells"]), int(d["nEdges"]), int(d["nVertices"])
        L = int(d["nVertLevels"])
        s.nCells, s.nEdges, s.nVertices, s.nLevels = nc, ne, nv, L
        s.maxEdges = int(np.asarray(d["edgesOnCell"]).shape[1])
        s.maxEdges2 = int(np.asarray(d["edgesOnEdge"]).shape[1])
        s.vertexDegree = 3
        s.maxAdv = int(np.asarray(d["advCellsForEdge"]).shape[1])
        for name, typ in MESH_FIELDS[8:]:
            if typ is C.c_double:
                setattr(s, name, float(d[name]))
                continue
            if name == "zxu" and "zxu" not in d:
                a = np.zeros((ne + 1, L))
            else:
                a = np.asarray(d[name])
            if typ is _ip:
                a = a.astype(np.int64)
                if name in _ONE_BASED:
                    a = a - 1
                a = np.ascontiguousarray(a.astype(np.int32))
            else:
                a = _f64(a)
            self.keep[name] = a
            setattr(s, name, _ptr(a))
        self.s = s
        self.nc, self.ne, self.nv, self.L = nc, ne, nv, L
        # Coriolis factors of the curvature terms as the reference forms them
        # (dyn_tend.py:2494-2502, :4258-4265): (2 omega cos angleEdge) cos latEdge per
        # edge and 2 omega cos latCell per cell, omega = 7.29212e-5 (constants.py:88)
        two_omega = 2.0 * 7.29212e-5
        if "angleEdge" in d and "latEdge" in d:
            self.curv_edge = _f64(two_omega * np.cos(_f64(d["angleEdge"])) * np.cos(_f64(d["latEdge"])))
        else:
            self.curv_edge = np.zeros(ne + 1)
        self.curv_cell = _f64(two_omega * np.cos(_f64(d["latCell"])))
        # vertical mixing: interface heights and the reference profiles of the
        # departures (zeros where the mapping has none)
        self.zgrid = _f64(d["zgrid"]) if "zgrid" in d else np.zeros((nc + 1, L + 1))
        self.t_init = _f64(d["t_init"]) if "t_init" in d else np.zeros((nc + 1, L))
        self.u_init = _f64(d["u_init"]) if "u_init" in d else np.zeros(L)
        self.v_init = _f64(d["v_init"]) if "v_init" in d else np.zeros(L)
        ang = _f64(d["angleEdge"]) if "angleEdge" in d else np.zeros(ne + 1)
        self.cos_edge, self.sin_edge = _f64(np.cos(ang)), _f64(np.sin(ang))

    @property
    def ref(self):
        return C.byref(self.s)


def _out(shape):
    return np.zeros(shape, dtype=np.float64)


# ---------------------------------------------------------------------------
# kernel wrappers (argument names follow the reference kernels)
# ---------------------------------------------------------------------------
def moist_coefficients(m: Mesh, scalars, moist_start, moist_end):
    sc = _f64(scalars)
    S = sc.shape[2]
    cqw, cqu, qtot = _out((m.nc + 1, m.L + 1)), _out((m.ne + 1, m.L)), _out((m.nc + 1, m.L))
    lib().orc_moist_coefficients(m.ref, _ptr(sc), C.c_int(S), C.c_int(moist_start), C.c_int(moist_end),
                                 _ptr(cqw), _ptr(cqu), _ptr(qtot))
    return {"cqw": cqw, "cqu": cqu, "qtot": qtot}


def vert_imp_coefs(m: Mesh, dts, epssm, cqw, p, t, rt, qtot):
    ins = [_f64(x) for x in (cqw, p, t, rt, qtot)]
    nc, L = m.nc, m.L
    out = {"cofrz": _out(L), "cofwr": _out((nc + 1, L)), "cofwz": _out((nc + 1, L)),
           "coftz": _out((nc + 1, L + 1)), "cofwt": _out((nc + 1, L)), "a_tri": _out((nc + 1, L)),
           "alpha_tri": _out((nc + 1, L)), "gamma_tri": _out((nc + 1, L))}
    lib().orc_vert_imp_coefs(m.ref, C.c_double(dts), C.c_double(epssm), *[_ptr(x) for x in ins],
                             *[_ptr(out[k]) for k in ("cofrz", "cofwr", "cofwz", "coftz", "cofwt",
                                                       "a_tri", "alpha_tri", "gamma_tri")])
    return out


def solve_diagnostics(m: Mesh, dt, apvm, rk_step, u, h, v_in=None):
    u, h = _f64(u), _f64(h)
    v_in = _f64(v_in) if v_in is not None else np.zeros((m.ne + 1, m.L))
    out = {"h_edge": _out((m.ne + 1, m.L)), "v": _out((m.ne + 1, m.L)),
           "divergence": _out((m.nc + 1, m.L)), "vorticity": _out((m.nv + 1, m.L)),
           "ke": _out((m.nc + 1, m.L)), "pv_edge": _out((m.ne + 1, m.L))}
    lib().orc_solve_diagnostics(m.ref, C.c_double(dt), C.c_double(apvm),
                                C.c_int(-1 if rk_step is None else int(rk_step)),
                                _ptr(u), _ptr(h), _ptr(v_in),
                                *[_ptr(out[k]) for k in ("h_edge", "v", "divergence", "vorticity", "ke", "pv_edge")])
    return out


DYN_INPUTS = ("u", "v", "w", "ru", "rw", "ru_save", "rw_save", "theta_m", "theta_m_save", "rho_zz",
              "rho_edge", "divergence", "vorticity", "ke", "pv_edge", "pp", "rr_save", "cqw", "cqu",
              "rt_diabatic_tend", "qtot")


def dyn_params(rk_step, dt, len_disp, c_s=0.125, visc4_2dsmag=0.05, coef_3rd_order=0.25,
               del4u_div_factor=10.0, prandtl=1.0, rayleigh_levels=0, rayleigh_timescale_days=5.0,
               horiz_mixing="2d_smagorinsky", h_theta_eddy_visc2=0.0, h_mom_eddy_visc4=0.0,
               h_theta_eddy_visc4=0.0, cam_coef=0.0, cam_levels=4, curvature=False, v_mom_eddy_visc2=0.0,
               v_theta_eddy_visc2=0.0, mix_full=True) -> OrcDynParams:
    """Scalar set-up of dyn_tend.py:3367-3458 and :9616-9626, evaluated in Python:
    Smagorinsky or "2d_fixed" mixing coefficients, the CAM damping of kdiff when
    ``cam_coef`` > 0 and ``cam_levels`` > 0, Rayleigh damping of u over
    ``rayleigh_levels`` top levels (0: off) with the coefficient of
    dyn_tend.py:10066-10072; ``curvature`` adds the curvature terms of u and w
    (config_curvature_u, dyn_tend.py:2471-2505, :4241-4267); the vertical eddy
    viscosities mix u, w and theta on the first RK stage, with ``mix_full`` False
    on the departures from u_init/v_init/t_init (dyn_tend.py:2731-2781, :4633-4647,
    :5060-5097, :10178-10197)."""
    p = OrcDynParams()
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
    if horiz_mixing == "2d_smagorinsky":
        p.h_mom_eddy_visc4 = visc4_2dsmag * len_disp ** 3
        p.h_theta_eddy_visc4 = p.h_mom_eddy_visc4
    elif horiz_mixing == "2d_fixed":
        p.h_mom_eddy_visc4 = float(h_mom_eddy_visc4)
        p.h_theta_eddy_visc4 = float(h_theta_eddy_visc4)
    else:
        raise ValueError(f"config_horiz_mixing {horiz_mixing!r}")
    p.mix_fixed = 1 if horiz_mixing == "2d_fixed" else 0
    p.kdiff_fixed = float(h_theta_eddy_visc2)
    if cam_coef > 0.0 and cam_levels > 0:
        p.cam_levels = int(cam_levels)
        p.cam_visc2 = 4.0 * 2.0833 * len_disp * cam_coef
    else:
        p.cam_levels, p.cam_visc2 = 0, 0.0
    p.del4u_div_factor = float(del4u_div_factor)
    p.prandtl_inv = 1.0 / prandtl
    p.rayleigh_levels = int(rayleigh_levels)
    p.rayleigh_coef_inverse = (1.0 / (float(rayleigh_levels) * (rayleigh_timescale_days * 86400.0))
                               if rayleigh_levels > 0 else 0.0)
    return p


def dyn_tend(m: Mesh, params: OrcDynParams, fields: dict, carried: dict | None = None):
    """``fields`` holds DYN_INPUTS; ``carried`` the rk>1 inputs (tend_rho, *_euler, kdiff)."""
    ins = [_f64(fields[k]) for k in DYN_INPUTS]
    nc, ne, L = m.nc, m.ne, m.L
    if carried is None:
        carried = {}
    io = {
        "tend_rho": _f64(carried.get("tend_rho", np.zeros((nc + 1, L)))).copy(),
        "tend_u_euler": _f64(carried.get("tend_u_euler", np.zeros((ne + 1, L)))).copy(),
        "tend_w_euler": _f64(carried.get("tend_w_euler", np.zeros((nc + 1, L + 1)))).copy(),
        "tend_theta_euler": _f64(carried.get("tend_theta_euler", np.zeros((nc + 1, L)))).copy(),
        "kdiff": _f64(carried.get("kdiff", np.zeros((nc + 1, L)))).copy(),
    }
    out = {"tend_u": _out((ne + 1, L)), "tend_w": _out((nc + 1, L + 1)), "tend_theta": _out((nc + 1, L)),
           "h_divergence": _out((nc + 1, L)), "rthdynten": _out((nc + 1, L))}
    # the step input's reconstructed winds (read only with curvature on)
    ur = _f64(fields.get("ur_cell", np.zeros((nc + 1, L))))
    vr = _f64(fields.get("vr_cell", np.zeros((nc + 1, L))))
    lib().orc_dyn_tend(m.ref, C.byref(params), *[_ptr(x) for x in ins], _ptr(ur), _ptr(vr),
                       _ptr(m.curv_edge), _ptr(m.curv_cell), _ptr(m.zgrid), _ptr(m.t_init), _ptr(m.u_init),
                       _ptr(m.v_init), _ptr(m.cos_edge), _ptr(m.sin_edge),
                       *[_ptr(io[k]) for k in ("tend_rho", "tend_u_euler", "tend_w_euler", "tend_theta_euler", "kdiff")],
                       *[_ptr(out[k]) for k in ("tend_u", "tend_w", "tend_theta", "h_divergence", "rthdynten")])
    out.update(io)
    return out


def smlstep_pert(m: Mesh, u_tend, w_tend):
    wt = _f64(w_tend).copy()
    lib().orc_smlstep_pert(m.ref, _ptr(_f64(u_tend)), _ptr(wt))
    return wt


ACOUSTIC_IN = ("cofrz", "cofwr", "cofwz", "coftz", "cofwt", "a_tri", "alpha_tri", "gamma_tri",
               "rho_zz", "theta_m", "exner", "cqu", "tend_ru", "tend_rho", "tend_rt", "tend_rw",
               "w", "rw", "rw_save")
ACOUSTIC_IO = ("ru_p", "ruAvg", "rho_pp", "rtheta_pp", "rtheta_pp_old", "rw_p", "wwAvg")


def acoustic_params(small_step, dts, epssm=0.1) -> OrcAcousticParams:
    p = OrcAcousticParams()
    p.small_step = int(small_step)
    p.dts = float(dts)
    p.epssm = float(epssm)
    p.resm = (1.0 - epssm) / (1.0 + epssm)
    rgas, cp = 287.0, 1004.5
    rcv = rgas / (cp - rgas)
    p.c2 = cp * rcv
    return p


def acoustic_step(m: Mesh, params: OrcAcousticParams, fields: dict):
    ins = [_f64(fields[k]) for k in ACOUSTIC_IN]
    io = {k: _f64(fields[k]).copy() for k in ACOUSTIC_IO}
    lib().orc_acoustic_step(m.ref, C.byref(params), *[_ptr(x) for x in ins], *[_ptr(io[k]) for k in ACOUSTIC_IO])
    return io


def divergence_damping(m: Mesh, dts, smdiv, len_disp, theta_m, rtheta_pp, rtheta_pp_old, ru_p):
    out = _f64(ru_p).copy()
    lib().orc_divergence_damping(m.ref, C.c_double(dts), C.c_double(smdiv), C.c_double(len_disp),
                                 _ptr(_f64(theta_m)), _ptr(_f64(rtheta_pp)), _ptr(_f64(rtheta_pp_old)), _ptr(out))
    return out


RECOVER_IN = ("rho_p_save", "rho_pp", "rtheta_p_save", "rtheta_pp", "rt_diabatic_tend", "rw_save",
              "rw_p", "ru_save", "ru_p")


def recover_large_step(m: Mesh, rk_step, dt, ns, fields: dict, exner_in=None, pressure_p_in=None):
    ins = [_f64(fields[k]) for k in RECOVER_IN]
    nc, ne, L = m.nc, m.ne, m.L
    ww = _f64(fields["wwAvg"]).copy()
    ra = _f64(fields["ruAvg"]).copy()
    out = {k: _out((nc + 1, L)) for k in ("rho_p", "rho_zz", "rtheta_p", "theta_m")}
    out["exner"] = _f64(exner_in).copy() if exner_in is not None else _out((nc + 1, L))
    out["pressure_p"] = _f64(pressure_p_in).copy() if pressure_p_in is not None else _out((nc + 1, L))
    out["rw"] = _out((nc + 1, L + 1))
    out["w"] = _out((nc + 1, L + 1))
    out["ru"] = _out((ne + 1, L))
    out["u"] = _out((ne + 1, L))
    lib().orc_recover_large_step(m.ref, C.c_int(rk_step), C.c_double(dt), C.c_int(ns), *[_ptr(x) for x in ins],
                                 _ptr(ww), _ptr(ra),
                                 *[_ptr(out[k]) for k in ("rho_p", "rho_zz", "rtheta_p", "theta_m", "exner",
                                                          "pressure_p", "rw", "w", "ru", "u")])
    out["wwAvg"] = ww
    out["ruAvg"] = ra
    return out


def advance_scalars(m: Mesh, dt, coef3, weight_old, weight_new, scalar_old, scalar_new,
                    rho_zz_old, rho_zz_new, uhAvg, wwAvg):
    so, sn = _f64(scalar_old), _f64(scalar_new)
    S = so.shape[2]
    out = sn.copy()
    hf = _out((m.ne + 1, m.L, S))
    lib().orc_advance_scalars(m.ref, C.c_int(S), C.c_double(dt), C.c_double(coef3), C.c_double(weight_old),
                              C.c_double(weight_new), _ptr(so), _ptr(sn), _ptr(_f64(rho_zz_old)),
                              _ptr(_f64(rho_zz_new)), _ptr(_f64(uhAvg)), _ptr(_f64(wwAvg)), _ptr(out), _ptr(hf))
    return out, hf


def advance_scalars_mono(m: Mesh, dt, coef3, scalars_old, scalars_new, rho_zz_old, uhAvg, wwAvg):
    so, sn = _f64(scalars_old), _f64(scalars_new)
    S = so.shape[2]
    out = sn.copy()
    rzi = _out((m.nc + 1, m.L))
    lib().orc_advance_scalars_mono(m.ref, C.c_int(S), C.c_double(dt), C.c_double(coef3), _ptr(so), _ptr(sn),
                                   _ptr(_f64(rho_zz_old)), _ptr(_f64(uhAvg)), _ptr(_f64(wwAvg)),
                                   _ptr(out), _ptr(rzi))
    return out, rzi


def reconstruct(m: Mesh, u):
    names = ("uReconstructX", "uReconstructY", "uReconstructZ", "uReconstructZonal", "uReconstructMeridional")
    out = {k: _out((m.nc + 1, m.L)) for k in names}
    lib().orc_reconstruct(m.ref, _ptr(_f64(u)), *[_ptr(out[k]) for k in names])
    out["ur_cell"] = out["uReconstructZonal"].copy()
    out["vr_cell"] = out["uReconstructMeridional"].copy()
    return out
