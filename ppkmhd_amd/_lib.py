"""ctypes binding of libppkmhd_amd.so (the C ABI declared in include/ppkmhd_amd.h).

The library is loaded from the package directory (built in-tree by ``python -m
ppkmhd_amd.build`` or ``__graft_entry__.build()``).  There is no fallback: if the library is
missing or a call fails, an exception is raised.
"""
import ctypes as C
import os

from . import _abi as A

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libppkmhd_amd.so")

# every entry point of include/ppkmhd_amd.h: name -> (restype, argtypes)
_P = C.POINTER(A.Params)
_PR = C.POINTER(A.Problem)
_H = C.c_void_p
_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)
# ppk_ini_lookup(ctx, section, name, value, value_size) -> 1 found / 0 absent
INI_LOOKUP = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_char),
                         C.c_int)

SIGNATURES = {
    "ppk_mhd_params_size": (C.c_int, []),
    "ppk_mhd_problem_size": (C.c_int, []),
    "ppk_mhd_version": (C.c_char_p, []),
    "ppk_mhd_device_count": (C.c_int, []),
    "ppk_mhd_params_from_ini": (C.c_int, [C.c_char_p, _P, _PR, C.c_char_p, C.c_int]),
    "ppk_mhd_params_from_ini_text": (C.c_int, [C.c_char_p, _P, _PR, C.c_char_p, C.c_int]),
    "ppk_mhd_params_from_lookup": (C.c_int, [INI_LOOKUP, C.c_void_p, _P, _PR, C.c_char_p,
                                             C.c_int]),
    "ppk_mhd_params_init": (C.c_int, [_P]),
    "ppk_mhd_create": (C.c_int, [_P, C.POINTER(_H)]),
    "ppk_mhd_destroy": (None, [_H]),
    "ppk_mhd_last_error": (C.c_char_p, [_H]),
    "ppk_mhd_init_problem": (C.c_int, [_H, _PR]),
    "ppk_mhd_init_problem_device": (C.c_int, [_H, _PR]),
    "ppk_mhd_initial_state": (C.c_int, [_P, _PR, _DP, C.c_size_t]),
    "ppk_mhd_upload": (C.c_int, [_H, _DP, C.c_size_t]),
    "ppk_mhd_download": (C.c_int, [_H, _DP, C.c_size_t]),
    "ppk_mhd_device_state": (C.c_int, [_H, C.POINTER(_DP)]),
    "ppk_mhd_stream": (C.c_int, [_H, C.POINTER(C.c_void_p)]),
    "ppk_mhd_make_boundaries": (C.c_int, [_H]),
    "ppk_mhd_compute_dt": (C.c_int, [_H, _DP]),
    "ppk_mhd_compute_dt_noalloc": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ppk_mhd_step": (C.c_int, [_H, C.c_double, C.c_int]),
    "ppk_mhd_run": (C.c_int, [_H, C.c_int, C.c_double, C.c_int]),
    "ppk_mhd_get_time": (C.c_int, [_H, _DP, _DP, C.POINTER(C.c_int64)]),
    "ppk_mhd_set_time": (C.c_int, [_H, C.c_double, C.c_int64]),
    "ppk_mhd_finished": (C.c_int, [_H, C.POINTER(C.c_int)]),
    "ppk_mhd_synchronize": (C.c_int, [_H]),
    "ppk_mhd_timers_enable": (C.c_int, [_H, C.c_uint32]),
    "ppk_mhd_timers_read": (C.c_int, [_H, C.c_int, _DP, C.POINTER(C.c_int64)]),
    "ppk_mhd_timer_count": (C.c_int, []),
    "ppk_mhd_timer_name": (C.c_char_p, [C.c_int]),
    "ppk_mhd_divb_max": (C.c_int, [_H, _DP]),
    "ppk_mhd_conserved_sums": (C.c_int, [_H, _DP]),
    "ppk_mhd_math_selfcheck": (C.c_int, [_DP, C.c_int64, _DP]),
    "ppk_mhd_write_vtk": (C.c_int, [C.POINTER(A.Params), _DP, C.c_char_p, C.c_char_p, C.c_int,
                                    C.c_int]),
    "ppk_mhd_save_vtk": (C.c_int, [_H, C.c_char_p, C.c_char_p, C.c_int, C.c_int]),
    "ppk_mhd_write_restart": (C.c_int, [C.POINTER(A.Params), _DP, C.c_double, C.c_int64,
                                        C.c_int64, C.c_char_p]),
    "ppk_mhd_read_restart": (C.c_int, [C.POINTER(A.Params), C.c_char_p, _DP, C.c_size_t, _DP,
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "ppk_mhd_save_state": (C.c_int, [_H, C.c_char_p, C.c_int64]),
    "ppk_mhd_load_state": (C.c_int, [_H, C.c_char_p, C.POINTER(C.c_int64)]),
    "ppk_mhd_batch_create": (C.c_int, [_P, C.c_int, C.POINTER(_H)]),
    "ppk_mhd_batch_destroy": (None, [_H]),
    "ppk_mhd_batch_last_error": (C.c_char_p, [_H]),
    "ppk_mhd_batch_size": (C.c_int, [_H]),
    "ppk_mhd_batch_init_problem": (C.c_int, [_H, _PR]),
    "ppk_mhd_batch_upload": (C.c_int, [_H, C.c_int, _DP, C.c_size_t]),
    "ppk_mhd_batch_download": (C.c_int, [_H, C.c_int, _DP, C.c_size_t]),
    "ppk_mhd_batch_run": (C.c_int, [_H, C.c_int]),
    "ppk_mhd_batch_get_time": (C.c_int, [_H, C.c_int, _DP, _DP, C.POINTER(C.c_int64)]),
    "ppk_mhd_batch_synchronize": (C.c_int, [_H]),
    "ppk_mhd_batch_set_tiles": (C.c_int, [_H, C.c_int, C.c_int]),
    "ppk_mhd_batch_get_tiles": (C.c_int, [_H, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ppk_mhd_batch_autotune": (C.c_int, [_H, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         _DP]),
    "ppk_mlp_param_count": (C.c_int, []),
    "ppk_mlp_create": (C.c_int, [C.c_uint64, C.POINTER(_H)]),
    "ppk_mlp_destroy": (None, [_H]),
    "ppk_mlp_last_error": (C.c_char_p, [_H]),
    "ppk_mlp_get_params": (C.c_int, [_H, _FP, C.c_size_t]),
    "ppk_mlp_set_params": (C.c_int, [_H, _FP, C.c_size_t]),
    "ppk_faceset_create": (C.c_int, [C.c_size_t, C.POINTER(_H)]),
    "ppk_faceset_destroy": (None, [_H]),
    "ppk_faceset_last_error": (C.c_char_p, [_H]),
    "ppk_faceset_size": (C.c_int, [_H, C.POINTER(C.c_size_t)]),
    "ppk_faceset_clear": (C.c_int, [_H]),
    "ppk_faceset_append": (C.c_int, [_H, _FP, _FP, C.c_size_t]),
    "ppk_faceset_read": (C.c_int, [_H, C.c_size_t, C.c_size_t, _FP, _FP]),
    "ppk_mhd_dump_faces": (C.c_int, [_H, _H, C.c_int, C.c_int, C.POINTER(C.c_size_t)]),
    "ppk_mlp_standardize": (C.c_int, [_H, _H]),
    "ppk_mlp_train": (C.c_int, [_H, _H, C.c_int, C.c_int, C.c_float, C.c_uint64]),
    "ppk_mlp_gradient": (C.c_int, [_H, _FP, _FP, C.c_size_t, _FP, C.c_size_t]),
    "ppk_mlp_infer": (C.c_int, [_H, _FP, C.c_size_t, _FP]),
    "ppk_mlp_evaluate": (C.c_int, [_H, _H, _DP]),
    "ppk_mhd_set_surrogate": (C.c_int, [_H, _H, C.c_double, C.c_double]),
    "ppk_mhd_surrogate_counts": (C.c_int, [_H, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                           C.c_int]),
    "ppk_hydro_create": (C.c_int, [_P, C.POINTER(_H)]),
    "ppk_hydro_destroy": (None, [_H]),
    "ppk_hydro_last_error": (C.c_char_p, [_H]),
    "ppk_hydro_nbvar": (C.c_int, [_H]),
    "ppk_hydro_init_problem": (C.c_int, [_H, _PR]),
    "ppk_hydro_initial_state": (C.c_int, [_P, _PR, _DP, C.c_size_t]),
    "ppk_hydro_upload": (C.c_int, [_H, _DP, C.c_size_t]),
    "ppk_hydro_download": (C.c_int, [_H, _DP, C.c_size_t]),
    "ppk_hydro_device_state": (C.c_int, [_H, C.POINTER(_DP)]),
    "ppk_hydro_stream": (C.c_int, [_H, C.POINTER(C.c_void_p)]),
    "ppk_hydro_make_boundaries": (C.c_int, [_H]),
    "ppk_hydro_compute_dt": (C.c_int, [_H, _DP]),
    "ppk_hydro_step": (C.c_int, [_H, C.c_double, C.c_int]),
    "ppk_hydro_run": (C.c_int, [_H, C.c_int, C.c_double, C.c_int]),
    "ppk_hydro_get_time": (C.c_int, [_H, _DP, _DP, C.POINTER(C.c_int64)]),
    "ppk_hydro_set_time": (C.c_int, [_H, C.c_double, C.c_int64]),
    "ppk_hydro_finished": (C.c_int, [_H, C.POINTER(C.c_int)]),
    "ppk_hydro_synchronize": (C.c_int, [_H]),
    "ppk_hydro_timers_enable": (C.c_int, [_H, C.c_uint32]),
    "ppk_hydro_timers_read": (C.c_int, [_H, C.c_int, _DP, C.POINTER(C.c_int64)]),
    "ppk_hydro_timer_count": (C.c_int, []),
    "ppk_hydro_timer_name": (C.c_char_p, [C.c_int]),
    "ppk_hydro_save_vtk": (C.c_int, [_H, C.c_char_p, C.c_char_p, C.c_int, C.c_int]),
}

_lib = None


class PPKError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"ppkmhd_amd error {code}: {msg}")
        self.code = code


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} not found: build it with `python -m ppkmhd_amd.build` "
                "(the HIP kernels are required; there is no CPU fallback)")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def check(rc, handle=None):
    if rc != A.PPK_OK:
        msg = lib().ppk_mhd_last_error(handle)
        raise PPKError(rc, msg.decode() if msg else "")
    return rc


def check_with(rc, handle, last_error):
    if rc != A.PPK_OK:
        msg = last_error(handle)
        raise PPKError(rc, msg.decode() if msg else "")
    return rc


def check_batch(rc, handle=None):
    if rc != A.PPK_OK:
        msg = lib().ppk_mhd_batch_last_error(handle)
        raise PPKError(rc, msg.decode() if msg else "")
    return rc
