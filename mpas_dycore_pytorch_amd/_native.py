"""ctypes binding of ``libmpas_dycore_hip.so`` (the C ABI in ``include/mpas_hip.h``).

The library is built in tree by ``mpas_dycore_pytorch_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback for the product
path.
"""

from __future__ import annotations

import ctypes as C
import os

from .errors import MpasDeviceError, MpasError

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpas_dycore_hip.so")

_dp = C.c_void_p  # device pointers travel as plain addresses
_vp = C.c_void_p

MESH_INT_FIELDS = ("nCells", "nEdges", "nVertices", "nLevels", "maxEdges", "maxEdges2", "vertexDegree",
                   "maxAdv", "flatTerrain", "reserved0")
MESH_INDEX_FIELDS = ("cellsOnEdge", "verticesOnEdge", "nEdgesOnCell", "edgesOnCell", "cellsOnCell",
                     "verticesOnCell", "kiteForCell", "edgesOnVertex", "nEdgesOnEdge", "edgesOnEdge",
                     "nAdvCellsForEdge", "advCellsForEdge")
MESH_REAL_FIELDS_A = ("edgesOnCell_sign", "edgesOnVertex_sign", "weightsOnEdge", "kiteAreasOnVertex",
                      "dcEdge", "dvEdge", "invDcEdge", "invDvEdge", "invAreaCell", "invAreaTriangle",
                      "fEdge", "fVertex", "meshScalingDel2", "meshScalingDel4", "adv_coefs", "adv_coefs_3rd",
                      "defc_a", "defc_b", "coeffs_reconstruct", "latCell", "lonCell", "zz", "zxu", "dss",
                      "zb_cell", "zb3_cell", "rdzw", "rdzu", "fzm", "fzp", "rdnw", "fnm", "fnp")
MESH_SCALAR_FIELDS = ("cf1", "cf2", "cf3", "reserved1")
MESH_REAL_FIELDS_B = ("rho_base", "rtheta_base", "exner_base")
MESH_REAL_FIELDS_C = ("curvEdge", "curvCell", "zgrid", "t_init", "u_init", "v_init", "cosAngleEdge", "sinAngleEdge")

STATE_FIELDS = ("u", "w", "theta_m", "rho_zz", "scalars", "rt_diabatic_tend",
                "rho_p", "rtheta_p", "exner", "pressure_p", "ru", "rw",
                "v", "divergence", "vorticity", "ke", "pv_edge", "rho_edge",
                "ur_cell", "vr_cell",
                "uReconstructX", "uReconstructY", "uReconstructZ",
                "uReconstructZonal", "uReconstructMeridional",
                "ruAvg", "wwAvg")


class MeshStruct(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in MESH_INT_FIELDS] + [(n, _dp) for n in MESH_INDEX_FIELDS]
                + [(n, _dp) for n in MESH_REAL_FIELDS_A] + [(n, C.c_double) for n in MESH_SCALAR_FIELDS]
                + [(n, _dp) for n in MESH_REAL_FIELDS_B] + [(n, _dp) for n in MESH_REAL_FIELDS_C])


class StateStruct(C.Structure):
    _fields_ = [(n, _dp) for n in STATE_FIELDS]


class WorkspaceStruct(C.Structure):
    _fields_ = [("base", _dp), ("n_doubles", C.c_int64)]


class ConfigStruct(C.Structure):
    _fields_ = [("dt", C.c_double), ("len_disp", C.c_double), ("epssm", C.c_double), ("smdiv", C.c_double),
                ("apvm", C.c_double), ("coef3", C.c_double), ("c_s", C.c_double), ("visc4_2dsmag", C.c_double),
                ("del4u_div_factor", C.c_double), ("ns", C.c_int32), ("split", C.c_int32),
                ("monotonic", C.c_int32), ("nScalars", C.c_int32), ("moist_start", C.c_int32),
                ("moist_end", C.c_int32), ("diagnostics_current", C.c_int32), ("order", C.c_int32),
                ("rayleigh_levels", C.c_int32), ("rayleigh_coef_inverse", C.c_double),
                ("mix_fixed", C.c_int32), ("h_theta_eddy_visc2", C.c_double), ("h_mom_eddy_visc4", C.c_double),
                ("h_theta_eddy_visc4", C.c_double), ("cam_levels", C.c_int32), ("cam_visc2", C.c_double),
                ("curvature", C.c_int32), ("v_mom_eddy_visc2", C.c_double), ("v_theta_eddy_visc2", C.c_double),
                ("mix_full", C.c_int32), ("scalar_advection", C.c_int32)]


class DynParamsStruct(C.Structure):
    _fields_ = [("rk_step", C.c_int32), ("dt", C.c_double), ("coef_3rd_order", C.c_double),
                ("kdiff_coef", C.c_double), ("kdiff_cap", C.c_double), ("h_mom_eddy_visc4", C.c_double),
                ("h_theta_eddy_visc4", C.c_double), ("del4u_div_factor", C.c_double),
                ("prandtl_inv", C.c_double), ("rayleigh_levels", C.c_int32), ("rayleigh_coef_inverse", C.c_double),
                ("mix_fixed", C.c_int32), ("kdiff_fixed", C.c_double), ("cam_levels", C.c_int32),
                ("cam_visc2", C.c_double), ("curvature", C.c_int32), ("v_mom_eddy_visc2", C.c_double),
                ("v_theta_eddy_visc2", C.c_double), ("mix_full", C.c_int32)]


class AcousticParamsStruct(C.Structure):
    _fields_ = [("small_step", C.c_int32), ("dts", C.c_double), ("epssm", C.c_double), ("resm", C.c_double),
                ("c2", C.c_double)]


class InverseTablesStruct(C.Structure):
    _fields_ = [(n, _dp) for n in ("advStart", "advEdge", "advSlot", "eoeStart", "eoeEdge", "eoeSlot",
                                   "cocStart", "cocCell", "cocSlot")]


# Entry points: name -> (restype, argtypes).  "M" marks the mesh pointer.
_I32, _F64, _I64 = C.c_int32, C.c_double, C.c_int64
_PM = C.POINTER(MeshStruct)
_PP = C.POINTER(C.c_void_p)  # host array of device pointers
SIGNATURES = {
    "mpas_hip_workspace_doubles": (_I64, [_I32, _I32, _I32, _I32, _I32]),
    "mpas_hip_srk3_step": (C.c_int, [_PM, C.POINTER(StateStruct), C.POINTER(WorkspaceStruct),
                                     C.POINTER(ConfigStruct), _vp]),
    "mpas_hip_struct_sizes": (C.c_int, [C.POINTER(_I32)] * 4),
    "mpas_hip_srk3_step_noalloc": (C.c_int, [_PM, C.POINTER(StateStruct), C.POINTER(ConfigStruct), _vp]),
    "mpas_hip_srk3_step_noalloc_capacity": (_I64, []),
    "mpas_hip_kernel_scratch_doubles": (_I64, [_I32, _I32, _I32, _I32]),
    "mpas_hip_moist_coefficients": (C.c_int, [_PM, _dp, _I32, _I32, _I32, _dp, _dp, _dp, _vp]),
    "mpas_hip_vert_imp_coefs": (C.c_int, [_PM, _F64, _F64] + [_dp] * 13 + [_vp]),
    "mpas_hip_solve_diagnostics": (C.c_int, [_PM, _F64, _F64, _I32] + [_dp] * 10 + [_vp]),
    "mpas_hip_dyn_tend": (C.c_int, [_PM, C.POINTER(DynParamsStruct)] + [_dp] * 34 + [_vp]),
    "mpas_hip_smlstep_pert": (C.c_int, [_PM, _dp, _dp, _vp]),
    "mpas_hip_acoustic_step": (C.c_int, [_PM, C.POINTER(AcousticParamsStruct)] + [_dp] * 26 + [_vp]),
    "mpas_hip_divergence_damping": (C.c_int, [_PM, _F64, _F64, _F64, _dp, _dp, _dp, _dp, _vp]),
    "mpas_hip_recover_large_step": (C.c_int, [_PM, _I32, _F64, _I32] + [_dp] * 21 + [_vp]),
    "mpas_hip_advance_scalars": (C.c_int, [_PM, _I32, _F64, _F64, _F64, _F64] + [_dp] * 8 + [_vp]),
    "mpas_hip_advance_scalars_mono": (C.c_int, [_PM, _I32, _F64, _F64] + [_dp] * 8 + [_vp]),
    "mpas_hip_reconstruct": (C.c_int, [_PM] + [_dp] * 6 + [_vp]),
    "mpas_hip_init_coupled_diagnostics": (C.c_int, [_PM] + [_dp] * 17 + [_vp]),
    "mpas_hip_vjp_scratch_doubles": (_I64, [_I32, _I32, _I32, _I32, _I32]),
    "mpas_hip_moist_coefficients_vjp": (C.c_int, [_PM, _I32, _I32, _I32, _PP, _PP, _vp]),
    "mpas_hip_vert_imp_coefs_vjp": (C.c_int, [_PM, _F64, _F64, _PP, _PP, _vp]),
    "mpas_hip_solve_diagnostics_vjp": (C.c_int, [_PM, _F64, _F64, _I32, _PP, _PP, _dp,
                                                 C.POINTER(InverseTablesStruct), _vp]),
    "mpas_hip_dyn_tend_vjp": (C.c_int, [_PM, C.POINTER(DynParamsStruct), _PP, _PP, _dp,
                                        C.POINTER(InverseTablesStruct), _vp]),
    "mpas_hip_smlstep_pert_vjp": (C.c_int, [_PM, _PP, _PP, _vp]),
    "mpas_hip_acoustic_step_vjp": (C.c_int, [_PM, C.POINTER(AcousticParamsStruct), _PP, _PP, _dp, _vp]),
    "mpas_hip_divergence_damping_vjp": (C.c_int, [_PM, _F64, _F64, _F64, _PP, _PP, _vp]),
    "mpas_hip_recover_large_step_vjp": (C.c_int, [_PM, _I32, _F64, _I32, _PP, _PP, _vp]),
    "mpas_hip_advance_scalars_vjp": (C.c_int, [_PM, _I32, _F64, _F64, _F64, _F64, _PP, _PP, _dp,
                                               C.POINTER(InverseTablesStruct), _vp]),
    "mpas_hip_advance_scalars_mono_vjp": (C.c_int, [_PM, _I32, _F64, _F64, _PP, _PP, _dp,
                                                    C.POINTER(InverseTablesStruct), _vp]),
    "mpas_hip_reconstruct_vjp": (C.c_int, [_PM, _PP, _PP, _vp]),
    "mpas_hip_layout_probe": (C.c_int, [_I64, _I32, _I32, _I32, _I32, C.POINTER(C.c_double)]),
    "mpas_hip_tridiagonal_factorize": (C.c_int, [_I64, _I32] + [_vp] * 5 + [_vp]),
    "mpas_hip_tridiagonal_factorize_vjp": (C.c_int, [_I64, _I32] + [_vp] * 9 + [_vp]),
    "mpas_hip_tridiagonal_factorize_jvp": (C.c_int, [_I64, _I32] + [_vp] * 9 + [_vp]),
    "mpas_hip_tridiagonal_solve": (C.c_int, [_I64, _I32] + [_vp] * 5 + [_vp]),
    "mpas_hip_tridiagonal_solve_vjp": (C.c_int, [_I64, _I32] + [_vp] * 10 + [_vp]),
    "mpas_hip_tridiagonal_solve_jvp": (C.c_int, [_I64, _I32] + [_vp] * 10 + [_vp]),
    "mpas_hip_moist_coefficients_jvp": (C.c_int, [_PM, _I32, _I32, _I32, _PP, _PP, _vp]),
    "mpas_hip_vert_imp_coefs_jvp": (C.c_int, [_PM, _F64, _F64, _PP, _PP, _vp]),
    "mpas_hip_solve_diagnostics_jvp": (C.c_int, [_PM, _F64, _F64, _I32, _PP, _PP, _dp, _vp]),
    "mpas_hip_dyn_tend_jvp": (C.c_int, [_PM, C.POINTER(DynParamsStruct), _PP, _PP, _dp, _vp]),
    "mpas_hip_smlstep_pert_jvp": (C.c_int, [_PM, _PP, _PP, _vp]),
    "mpas_hip_acoustic_step_jvp": (C.c_int, [_PM, C.POINTER(AcousticParamsStruct), _PP, _PP, _dp, _vp]),
    "mpas_hip_divergence_damping_jvp": (C.c_int, [_PM, _F64, _F64, _F64, _PP, _PP, _vp]),
    "mpas_hip_recover_large_step_jvp": (C.c_int, [_PM, _I32, _F64, _I32, _PP, _PP, _vp]),
    "mpas_hip_advance_scalars_jvp": (C.c_int, [_PM, _I32, _F64, _F64, _F64, _F64, _PP, _PP, _vp]),
    "mpas_hip_advance_scalars_mono_jvp": (C.c_int, [_PM, _I32, _F64, _F64, _PP, _PP, _dp, _vp]),
    "mpas_hip_reconstruct_jvp": (C.c_int, [_PM, _PP, _PP, _vp]),
    "mpas_hip_init_coupled_vjp_scratch_doubles": (_I64, [_I32, _I32]),
    "mpas_hip_init_coupled_diagnostics_vjp": (C.c_int, [_PM, _I32, _dp, _PP, _PP, _dp, _vp]),
    "mpas_hip_init_coupled_diagnostics_jvp": (C.c_int, [_PM, _I32, _dp, _PP, _PP, _vp]),
    "mpas_hip_history_fields": (C.c_int, [_I32, _I32, _I32, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _vp]),
    "mpas_hip_tuning_set": (C.c_int, [C.c_char_p, _I32]),
    "mpas_hip_tuning_get": (_I32, [C.c_char_p]),
    "mpas_hip_tuning_set_order": (C.c_int, [C.c_char_p, _I32]),
    "mpas_hip_tuning_get_order": (_I32, [C.c_char_p]),
    "mpas_hip_last_error": (C.c_char_p, []),
    "mpas_hip_timer_arm": (C.c_int, [C.c_char_p, _I32]),
    "mpas_hip_timer_read": (C.c_int, [C.POINTER(C.c_double), C.POINTER(_I32)]),
    "mpas_hip_timer_read_each": (C.c_int, [C.POINTER(C.c_double), _I32, C.POINTER(_I32)]),
}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library once and check that the ABI structs agree."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MpasDeviceError(
            f"HIP library not built: {LIB_PATH} is missing "
            "(run `python -m mpas_dycore_pytorch_amd.build`)"
        )
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    sizes = [C.c_int32() for _ in range(4)]
    lib.mpas_hip_struct_sizes(*[C.byref(s) for s in sizes])
    want = [C.sizeof(MeshStruct), C.sizeof(StateStruct), C.sizeof(WorkspaceStruct), C.sizeof(ConfigStruct)]
    got = [s.value for s in sizes]
    if got != want:
        raise MpasError(f"ABI struct size mismatch: library {got}, binding {want}")
    _lib = lib
    return lib


def check(rc: int, what: str, lib: C.CDLL | None = None) -> None:
    if rc != 0:
        msg = (lib or load()).mpas_hip_last_error()
        raise MpasError(f"{what} failed: {msg.decode() if msg else 'unknown error'}")


# ---------------------------------------------------------------------------
# mixed-precision library (include/mpas_hip_mixed.h)
# ---------------------------------------------------------------------------
MIXED_LIB_PATH = os.path.join(HERE, "libmpas_dycore_hip_mixed.so")
# state fields stored in binary32 by the mixed-precision step
MIXED_F32_FIELDS = ("v", "vorticity", "pv_edge", "rho_edge")
MixedStateStruct = StateStruct  # same field order; every field is a device address

MIXED_SIGNATURES = {
    "mpas_hip_mixed_state_size": (_I32, []),
    "mpas_hip_mixed_workspace_doubles": (_I64, [_I32, _I32, _I32, _I32, _I32]),
    "mpas_hip_mixed_srk3_step": (C.c_int, [_PM, C.POINTER(StateStruct), C.POINTER(WorkspaceStruct),
                                           C.POINTER(ConfigStruct), _vp]),
    "mpas_hip_tuning_set": (C.c_int, [C.c_char_p, _I32]),
    "mpas_hip_tuning_get": (_I32, [C.c_char_p]),
    "mpas_hip_tuning_set_order": (C.c_int, [C.c_char_p, _I32]),
    "mpas_hip_tuning_get_order": (_I32, [C.c_char_p]),
    "mpas_hip_last_error": (C.c_char_p, []),
    "mpas_hip_timer_arm": (C.c_int, [C.c_char_p, _I32]),
    "mpas_hip_timer_read": (C.c_int, [C.POINTER(C.c_double), C.POINTER(_I32)]),
}

_mixed = None


def load_mixed() -> C.CDLL:
    """Load the mixed-precision step library once (fails loudly when missing)."""
    global _mixed
    if _mixed is not None:
        return _mixed
    if not os.path.exists(MIXED_LIB_PATH):
        raise MpasDeviceError(
            f"HIP library not built: {MIXED_LIB_PATH} is missing "
            "(run `python -m mpas_dycore_pytorch_amd.build`)"
        )
    lib = C.CDLL(MIXED_LIB_PATH)
    for name, (res, args) in MIXED_SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.mpas_hip_mixed_state_size() != C.sizeof(MixedStateStruct):
        raise MpasError("ABI struct size mismatch for mpas_hip_mixed_state")
    _mixed = lib
    return lib
