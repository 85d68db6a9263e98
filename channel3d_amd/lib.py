"""ctypes binding of libchannel3d.so (the C ABI declared in include/channel3d.h).

Loading fails loudly when the HIP library has not been built: there is no
CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libchannel3d.so")

C3D_OK = 0
C3D_ERR_ARG, C3D_ERR_HIP, C3D_ERR_NOMEM, C3D_ERR_UNSUPPORTED, C3D_ERR_IO = -1, -2, -3, -4, -5
NCHASTAT = 35
NGRADSTAT = 98
NSPEC1D = 18
BC_PERIOD, BC_NOSLIP, BC_FREESLIP = 0, -1, -2
FP64, FP32 = 0, 1
HALO, INTERIOR, WALL = 0, 1, 2

# state field ids
(UX, UY, UZ, PR, HISTX, HISTY, HISTZ, ARR1, ARR2, HALOARR, UY_YM, DUY_YM, NUT) = range(13)
STATE_LAYOUT = {UX: HALO, UY: HALO, UZ: HALO, PR: HALO, HISTX: INTERIOR, HISTY: INTERIOR,
                HISTZ: INTERIOR, ARR1: INTERIOR, ARR2: INTERIOR, HALOARR: HALO,
                UY_YM: WALL, DUY_YM: WALL, NUT: HALO}

# metric ids
(M_YP, M_YC, M_DYP, M_DYC, M_RDYP, M_RDYC, M_YINTERP, M_AM2PR, M_AC2PR, M_AP2PR,
 M_AM2C, M_AC2C, M_AP2C, M_AM2P, M_AC2P, M_AP2P, M_AM2CCN, M_AP2CCN,
 M_WAVEX, M_WAVEZ, M_AC2CCN, M_DELTA) = range(22)

EXPORTS = (
    "c3d_create", "c3d_destroy", "c3d_last_error", "c3d_sync", "c3d_mesh_coords",
    "c3d_get_metric", "c3d_set_dt", "c3d_field_alloc", "c3d_field_free", "c3d_upload",
    "c3d_download", "c3d_field_copy", "c3d_state_field", "c3d_PMcoeUpdate",
    "c3d_Update_uy_ym", "c3d_clcRhsX", "c3d_clcRhsY", "c3d_clcRhsZ", "c3d_clcRhsXYZ", "c3d_clcU1Hat", "c3d_clcU2Hat", "c3d_clcU3Hat",
    "c3d_SetBC_and_UpdateHaloForPrSrc", "c3d_SetBC_and_UpdateHalo_pr",
    "c3d_SetBC_and_UpdateHalo", "c3d_clcPrSrc", "c3d_clcPPE", "c3d_FluidVelUpdate",
    "c3d_PressureUpdate", "c3d_CalcMaxCFL", "c3d_CheckDivergence", "c3d_CalcVmax",
    "c3d_CalcUxAver", "c3d_substep", "c3d_step", "c3d_step_diagnostics", "c3d_project",
    "c3d_ppe", "c3d_profile_enable", "c3d_profile_reset", "c3d_profile_count",
    "c3d_profile_get", "c3d_device_bytes",
    "c3d_write_restart", "c3d_read_restart", "c3d_restart_name",
    "c3d_clcStat_CH", "c3d_stat_get", "c3d_stat_reset", "c3d_write_stat", "c3d_append_prgrad",
    "c3d_set_ppe_refine", "c3d_clcStat", "c3d_stat_set_grad", "c3d_stat_get_grad",
    "c3d_write_vgrad", "c3d_set_les", "c3d_Clc_SGS_Vis",
    "c3d_clcSpectra1D", "c3d_spectra_get", "c3d_spectra_reset", "c3d_write_spectra",
    "c3d_clcSpectra2D", "c3d_spectra2d_get", "c3d_write_spectra2d", "c3d_set_graph",
    "c3d_closure_nparam", "c3d_closure_create", "c3d_closure_destroy", "c3d_closure_set_params",
    "c3d_closure_get_params", "c3d_closure_size", "c3d_closure_add_samples",
    "c3d_closure_get_samples", "c3d_closure_predict", "c3d_closure_grad", "c3d_closure_train",
    "c3d_closure_adam_step", "c3d_closure_dsm", "c3d_set_les_closure", "c3d_closure_fallbacks",
    "c3d_field_bytes", "c3d_field_alloc_batch", "c3d_ppe_batch", "c3d_helmholtz_batch",
    "c3d_ppe_batch_noalloc", "c3d_closure_set_mode",
)


class Params(ctypes.Structure):
    _fields_ = [
        ("nxc", ctypes.c_int), ("nyc", ctypes.c_int), ("nzc", ctypes.c_int),
        ("xlx", ctypes.c_double), ("yly", ctypes.c_double), ("zlz", ctypes.c_double),
        ("xnu", ctypes.c_double), ("dt", ctypes.c_double),
        ("bc", ctypes.c_int * 6),
        ("uxbc", ctypes.c_double * 6), ("uybc", ctypes.c_double * 6),
        ("uzbc", ctypes.c_double * 6),
        ("gravity", ctypes.c_double * 3),
        ("is_ux_const", ctypes.c_int),
        ("ubulk", ctypes.c_double), ("ybulk1", ctypes.c_double), ("ybulk2", ctypes.c_double),
        ("ischeme", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("order", ctypes.c_int),
        ("u_crf", ctypes.c_double),
    ]


class C3DError(RuntimeError):
    pass


_lib = None
_VP = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_FP = ctypes.POINTER(ctypes.c_float)


def load(path: str | None = None):
    """Load the library; raises when it is missing (build it with channel3d_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise C3DError(f"{path} not found: build the HIP library first "
                       "(python -m channel3d_amd.build)")
    L = ctypes.CDLL(path)
    I = ctypes.c_int
    sig = {
        "c3d_create": ([ctypes.POINTER(Params), _DP, ctypes.POINTER(_VP)], I),
        "c3d_destroy": ([_VP], I),
        "c3d_last_error": ([], ctypes.c_char_p),
        "c3d_sync": ([_VP], I),
        "c3d_mesh_coords": ([I, ctypes.c_double, I, _DP, _IP, _IP, _IP, _DP, _DP], I),
        "c3d_get_metric": ([_VP, I, I, I, _DP], I),
        "c3d_set_dt": ([_VP, ctypes.c_double], I),
        "c3d_field_alloc": ([_VP, I, ctypes.POINTER(_VP)], I),
        "c3d_field_free": ([_VP, _VP], I),
        "c3d_upload": ([_VP, I, _VP, _DP], I),
        "c3d_download": ([_VP, I, _VP, _DP], I),
        "c3d_field_copy": ([_VP, I, _VP, _VP], I),
        "c3d_state_field": ([_VP, I, ctypes.POINTER(_VP)], I),
        "c3d_PMcoeUpdate": ([_VP, I], I),
        "c3d_Update_uy_ym": ([_VP, _VP, _VP], I),
        "c3d_clcRhsX": ([_VP] * 7, I),
        "c3d_clcRhsY": ([_VP] * 7, I),
        "c3d_clcRhsZ": ([_VP] * 7, I),
        "c3d_clcRhsXYZ": ([_VP] * 11, I),
        "c3d_clcU1Hat": ([_VP] * 3, I),
        "c3d_clcU2Hat": ([_VP] * 4, I),
        "c3d_clcU3Hat": ([_VP] * 3, I),
        "c3d_SetBC_and_UpdateHaloForPrSrc": ([_VP] * 5, I),
        "c3d_SetBC_and_UpdateHalo_pr": ([_VP] * 2, I),
        "c3d_SetBC_and_UpdateHalo": ([_VP] * 5, I),
        "c3d_clcPrSrc": ([_VP] * 6 + [_DP], I),
        "c3d_clcPPE": ([_VP] * 3, I),
        "c3d_FluidVelUpdate": ([_VP] * 5, I),
        "c3d_PressureUpdate": ([_VP] * 3, I),
        "c3d_CalcMaxCFL": ([_VP] * 4 + [_DP], I),
        "c3d_CheckDivergence": ([_VP] * 4 + [_DP], I),
        "c3d_CalcVmax": ([_VP] * 4 + [_DP], I),
        "c3d_CalcUxAver": ([_VP, _VP, _DP], I),
        "c3d_substep": ([_VP, I], I),
        "c3d_step": ([_VP], I),
        "c3d_step_diagnostics": ([_VP, _DP], I),
        "c3d_project": ([_VP], I),
        "c3d_ppe": ([_VP] * 3, I),
        "c3d_profile_enable": ([_VP, I], I),
        "c3d_profile_reset": ([_VP], I),
        "c3d_profile_count": ([_VP], I),
        "c3d_profile_get": ([_VP, I, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_long),
                             _DP, _DP], I),
        "c3d_device_bytes": ([_VP], ctypes.c_size_t),
        "c3d_write_restart": ([_VP, ctypes.c_char_p] + [_VP] * 7, I),
        "c3d_read_restart": ([_VP, ctypes.c_char_p] + [_VP] * 7, I),
        "c3d_restart_name": ([ctypes.c_char_p, ctypes.c_char_p, I, ctypes.c_char_p,
                              ctypes.c_size_t], I),
        "c3d_clcStat_CH": ([_VP] * 5, I),
        "c3d_stat_get": ([_VP, _DP, _IP, _DP], I),
        "c3d_stat_reset": ([_VP], I),
        "c3d_write_stat": ([_VP, ctypes.c_char_p, I, I], I),
        "c3d_append_prgrad": ([_VP, ctypes.c_char_p, I], I),
        "c3d_set_ppe_refine": ([_VP, I], I),
        "c3d_set_graph": ([_VP, I], I),
        "c3d_closure_nparam": ([], I),
        "c3d_closure_create": ([ctypes.c_longlong, ctypes.POINTER(_VP)], I),
        "c3d_closure_destroy": ([_VP], I),
        "c3d_closure_set_params": ([_VP, _FP], I),
        "c3d_closure_get_params": ([_VP, _FP], I),
        "c3d_closure_size": ([_VP, ctypes.POINTER(ctypes.c_longlong)], I),
        "c3d_closure_add_samples": ([_VP, _VP, I, I, I, ctypes.POINTER(ctypes.c_longlong)], I),
        "c3d_closure_get_samples": ([_VP, ctypes.c_longlong, ctypes.c_longlong, _FP, _FP, _FP], I),
        "c3d_closure_predict": ([_VP, ctypes.c_longlong, ctypes.c_longlong, _FP], I),
        "c3d_closure_grad": ([_VP, ctypes.c_longlong, ctypes.c_longlong, _FP, _FP], I),
        "c3d_closure_train": ([_VP, ctypes.c_longlong, ctypes.c_longlong, I, I, ctypes.c_float,
                               ctypes.c_uint, _FP], I),
        "c3d_closure_adam_step": ([_VP, _FP, ctypes.c_float], I),
        "c3d_closure_dsm": ([_VP, _VP, I, I, I, _DP, _DP, _DP], I),
        "c3d_set_les_closure": ([_VP, _VP], I),
        "c3d_closure_fallbacks": ([_VP, ctypes.POINTER(ctypes.c_longlong)], I),
        "c3d_clcStat": ([_VP] * 5, I),
        "c3d_stat_set_grad": ([_VP, I], I),
        "c3d_stat_get_grad": ([_VP, _DP], I),
        "c3d_write_vgrad": ([_VP, ctypes.c_char_p], I),
        "c3d_set_les": ([_VP, I, I], I),
        "c3d_field_bytes": ([_VP, I], ctypes.c_size_t),
        "c3d_field_alloc_batch": ([_VP, I, I, ctypes.POINTER(_VP)], I),
        "c3d_ppe_batch": ([_VP, I, _VP, _VP, I], I),
        "c3d_ppe_batch_noalloc": ([_VP, I, _VP, _VP, I], I),
        "c3d_closure_set_mode": ([_VP, I], I),
        "c3d_helmholtz_batch": ([_VP, I, I, I, _VP, _VP, I], I),
        "c3d_Clc_SGS_Vis": ([_VP] * 5, I),
        "c3d_clcSpectra1D": ([_VP] * 5 + [I, I], I),
        "c3d_spectra_get": ([_VP, _DP, _DP, _IP], I),
        "c3d_spectra_reset": ([_VP], I),
        "c3d_write_spectra": ([_VP, ctypes.c_char_p, ctypes.c_char_p], I),
        "c3d_clcSpectra2D": ([_VP] * 5 + [I], I),
        "c3d_spectra2d_get": ([_VP, _DP, _IP], I),
        "c3d_write_spectra2d": ([_VP, ctypes.c_char_p], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != C3D_OK:
        msg = load().c3d_last_error()
        raise C3DError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def dptr(a):
    import numpy as np
    assert isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.f_contiguous
    return a.ctypes.data_as(_DP)
