"""ctypes binding of ``simp_amd/lib/libsimp_amd.so`` (the C ABI of
``include/simp.h``).  Loading fails loudly when the library is missing:
there is no CPU fallback."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsimp_amd.so")

SIMP_OK = 0
SIMP_E_ARG = 1
SIMP_E_HIP = 2
SIMP_E_CAPACITY = 3
SIMP_E_NAN = 4
SIMP_E_UNSUPPORTED = 5
SIMP_E_NOMEM = 6
SIMP_E_STATE = 7
SIMP_FP64 = 0
SIMP_MIXED = 1
SIMP_WIND_UNIFORM = 0
SIMP_WIND_GRID = 1
SIMP_CONFIG_VERSION = 3
SIMP_SLE_KEEP = -1
SIMP_OFF_BOUNDARY_PRESSURE = 1
SIMP_OFF_PRESSURE_FORCE = 2
SIMP_OFF_WIND = 4
SIMP_OFF_WATER = 8
SIMP_OFF_VISCOSITY = 16

#: every entry point declared in include/simp.h
EXPORTS = (
    "simp_config_default", "simp_create", "simp_destroy", "simp_set_wind_grid",
    "simp_single_step", "simp_single_step_noalloc", "simp_kernel_values", "simp_nnps", "simp_upload", "simp_step_pc",
    "simp_download", "simp_set_diagnostics", "simp_download_diagnostics",
    "simp_sync", "simp_stream", "simp_profile_steps", "simp_strerror",
    "simp_batch_create", "simp_batch_destroy", "simp_batch_add", "simp_batch_remove",
    "simp_batch_update", "simp_interaction_stats", "simp_batch_interaction_stats",
    "simp_set_search", "simp_search_stats", "simp_set_tuning", "simp_get_tuning",
    "simp_tune", "simp_batch_set_tuning", "simp_batch_tune",
    "simp_batch_step", "simp_batch_sync", "simp_batch_status", "simp_batch_download",
    "simp_batch_count", "simp_batch_list", "simp_instance_footprint",
    "simp_output_create", "simp_output_append", "simp_output_read_last",
    "simp_snapshot_size", "simp_snapshot_save", "simp_snapshot_load",
    "simp_snapshot_info", "simp_kernel_gradients", "simp_set_module_state",
)


class SimpConfig(C.Structure):
    _fields_ = [("version", C.c_int32), ("precision", C.c_int32),
                ("int_algorithm", C.c_int32), ("nnps_algorithm", C.c_int32),
                ("sle", C.c_int32), ("tse", C.c_int32), ("rheology", C.c_int32),
                ("bfa", C.c_int32), ("eos_alg", C.c_int32), ("pa_sph", C.c_int32),
                ("skf", C.c_int32), ("den_alg", C.c_int32), ("wind_mode", C.c_int32),
                ("max_particles", C.c_int32), ("max_boundary", C.c_int32),
                ("reserved", C.c_int32), ("wind", C.c_double * 2),
                ("current", C.c_double * 2), ("max_hsml", C.c_double),
                ("artificial_visc", C.c_int32), ("art_visc_alg", C.c_int32),
                ("xsph_regularization", C.c_int32), ("boundary_friction", C.c_int32),
                ("boundary_pressure", C.c_int32), ("pressure_force", C.c_int32),
                ("wind_force", C.c_int32), ("water_force", C.c_int32),
                ("viscosity", C.c_int32)]


class SimpInstanceDesc(C.Structure):
    _fields_ = [("n", C.c_int32), ("nb", C.c_int32),
                ("x", C.c_void_p), ("vx", C.c_void_p), ("mass", C.c_void_p),
                ("rho", C.c_void_p), ("hsml", C.c_void_p), ("hth", C.c_void_p),
                ("A", C.c_void_p), ("xb", C.c_void_p), ("massb", C.c_void_p),
                ("dt", C.c_double), ("current_time", C.c_double),
                ("itimestep", C.c_int32), ("nsteps", C.c_int32), ("tse", C.c_int32),
                ("wind_mode", C.c_int32), ("wind", C.c_double * 2),
                ("current", C.c_double * 2), ("max_hsml", C.c_double),
                ("grid_nx", C.c_int32), ("grid_ny", C.c_int32),
                ("grid_x0", C.c_double), ("grid_y0", C.c_double),
                ("grid_dx", C.c_double), ("grid_dy", C.c_double),
                ("grid_u", C.c_void_p), ("grid_v", C.c_void_p),
                ("den_alg", C.c_int32), ("pa_sph", C.c_int32), ("eos_alg", C.c_int32),
                ("rheology", C.c_int32), ("skf", C.c_int32), ("reserved2", C.c_int32),
                ("sle", C.c_int32), ("xsph", C.c_int32), ("art_visc", C.c_int32),
                ("art_visc_alg", C.c_int32), ("bfa", C.c_int32),
                ("boundary_friction", C.c_int32), ("b_normal", C.c_void_p),
                ("forces_off", C.c_int32), ("reserved3", C.c_int32)]


class SimpTuning(C.Structure):
    _fields_ = [("passA_batch", C.c_int32), ("passB_batch", C.c_int32),
                ("reserved", C.c_int32), ("skin", C.c_double)]


class SimpError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        msg = lib().simp_strerror(status).decode() if _lib is not None else str(status)
        super().__init__(f"{where}: {msg} (status {status})")


_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with "
                          "`make -C simp_amd/csrc` (or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    L.simp_config_default.argtypes = [C.POINTER(SimpConfig)]
    L.simp_config_default.restype = None
    L.simp_strerror.argtypes = [C.c_int]
    L.simp_strerror.restype = C.c_char_p
    L.simp_create.argtypes = [C.POINTER(SimpConfig), C.POINTER(_vp)]
    L.simp_destroy.argtypes = [_vp]
    L.simp_set_wind_grid.argtypes = [_vp, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.c_double, _vp, _vp]
    L.simp_single_step.argtypes = ([_vp, C.c_int] + [_vp] * 7 + [C.c_int, _vp, _vp,
                                   C.c_double, C.c_int] + [_vp] * 9 + [_vp, _vp])
    L.simp_single_step_noalloc.argtypes = L.simp_single_step.argtypes
    L.simp_set_search.argtypes = [_vp, C.c_int]
    L.simp_set_tuning.argtypes = [_vp, _vp]
    L.simp_get_tuning.argtypes = [_vp, _vp]
    L.simp_tune.argtypes = [_vp, C.c_int, _vp, _vp, _vp]
    L.simp_batch_set_tuning.argtypes = [_vp, _vp]
    L.simp_batch_tune.argtypes = [_vp, C.c_int, _vp, _vp, _vp]
    L.simp_search_stats.argtypes = [_vp, _vp, _vp, _vp]
    L.simp_kernel_values.argtypes = [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp]
    L.simp_kernel_gradients.argtypes = [_vp, C.c_int, _vp, _vp, _vp, _vp]
    L.simp_set_module_state.argtypes = [_vp, C.c_double, C.c_int, _vp]
    L.simp_nnps.argtypes = [_vp, C.c_int, _vp, _vp, C.c_int, _vp, C.c_int, _vp, _vp, _vp,
                            _vp, _vp, C.c_int, _vp, _vp, _vp, _vp]
    L.simp_upload.argtypes = ([_vp, C.c_int] + [_vp] * 7 + [C.c_int, _vp, _vp, C.c_double,
                              C.c_double, C.c_int])
    L.simp_step_pc.argtypes = [_vp, C.c_int, C.c_int]
    L.simp_download.argtypes = [_vp] + [_vp] * 9
    L.simp_set_diagnostics.argtypes = [_vp, C.c_int]
    L.simp_download_diagnostics.argtypes = [_vp] + [_vp] * 5
    L.simp_sync.argtypes = [_vp]
    L.simp_stream.argtypes = [_vp]
    L.simp_stream.restype = _vp
    L.simp_profile_steps.argtypes = [_vp, C.c_int, C.c_int, C.c_char_p, C.c_int, _vp, _vp]
    L.simp_batch_create.argtypes = [C.c_int, C.POINTER(_vp)]
    L.simp_batch_destroy.argtypes = [_vp]
    L.simp_batch_add.argtypes = [_vp, C.c_int, C.POINTER(SimpInstanceDesc), _vp]
    L.simp_batch_remove.argtypes = [_vp, C.c_int, _vp]
    L.simp_interaction_stats.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.simp_batch_interaction_stats.argtypes = [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp]
    L.simp_batch_update.argtypes = [_vp, C.c_int, _vp, C.c_int, C.POINTER(SimpInstanceDesc), _vp]
    L.simp_batch_step.argtypes = [_vp, C.c_int, C.c_int]
    L.simp_batch_sync.argtypes = [_vp]
    L.simp_batch_status.argtypes = [_vp, C.c_int, _vp, _vp, _vp, _vp]
    L.simp_batch_download.argtypes = [_vp, C.c_int] + [_vp] * 9
    L.simp_batch_count.argtypes = [_vp, _vp, _vp]
    L.simp_batch_list.argtypes = [_vp, C.c_int, _vp, _vp, _vp, _vp, _vp]
    L.simp_instance_footprint.argtypes = [C.POINTER(SimpInstanceDesc)]
    L.simp_instance_footprint.restype = C.c_longlong
    L.simp_snapshot_size.argtypes = [_vp]
    L.simp_snapshot_size.restype = C.c_longlong
    L.simp_snapshot_save.argtypes = [_vp, _vp, C.c_longlong]
    L.simp_snapshot_load.argtypes = [_vp, _vp, C.c_longlong]
    L.simp_snapshot_info.argtypes = [_vp, C.c_longlong, _ip, _ip, _ip, _dp]
    L.simp_output_create.argtypes = [C.c_char_p]
    L.simp_output_append.argtypes = [C.c_char_p, C.c_int] + [_vp] * 9 + [C.c_int, C.c_double]
    L.simp_output_read_last.argtypes = [C.c_char_p, C.c_int, _vp, _vp, _vp] + [_vp] * 9
    _lib = L
    return L


def ptr(a):
    """Data pointer of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the C ABI must be contiguous"
    return C.c_void_p(a.ctypes.data)


def check(status: int, where: str) -> None:
    if status != SIMP_OK:
        raise SimpError(status, where)
