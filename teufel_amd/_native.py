"""ctypes binding of the C ABI in include/teufel_gpu.h.

The library is teufel_amd/lib/libteufel_gpu.so, built in-tree by
teufel_amd/build.py.  There is no fallback: if it is missing, loading raises.
"""
import ctypes
import os

from .build import LIB_PATH

TG_OK, TG_ERR_ARG, TG_ERR_OOM, TG_ERR_HIP, TG_ERR_RANGE, TG_ERR_STATE = range(6)
TG_FP64, TG_MIXED = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)


class TgTuning(ctypes.Structure):
    _fields_ = [("bins_per_window", ctypes.c_int), ("particles_per_chunk", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("window_groups", ctypes.c_int), ("walk", ctypes.c_int),
                ("spill", ctypes.c_int)]


# (name, restype, argtypes) for every entry point of include/teufel_gpu.h
_i, _d, _vp, _cp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_char_p
SIGNATURES = [
    ("tg_create", _i, [_i, ctypes.POINTER(_vp)]),
    ("tg_destroy", _i, [_vp]),
    ("tg_last_error", _cp, [_vp]),
    ("tg_device_info", _i, [_vp, _cp, _i, _ip]),
    ("tg_set_trajectories", _i, [_vp, _i, _i] + [_dp] * 11),
    ("tg_set_trajectories_device", _i, [_vp, _i, _i] + [_vp] * 11),
    ("tg_add_screen", _i, [_vp, _i, _dp, _dp, _d, _i, _ip]),
    ("tg_zero_screen", _i, [_vp, _i]),
    ("tg_remove_screens", _i, [_vp]),
    ("tg_screen_count", _i, [_vp, _ip]),
    ("tg_integrate", _i, [_vp, _i, ctypes.POINTER(TgTuning)]),
    ("tg_synchronize", _i, [_vp]),
    ("tg_integrate_range", _i, [_vp, _i, ctypes.POINTER(TgTuning), _i, _i, _i]),
    ("tg_integrate_range_static", _i, [_vp, _i, ctypes.POINTER(TgTuning), _i, _i, _i]),
    ("tg_read_traces", _i, [_vp, _i, _dp, _i]),
    ("tg_screen_traces_device", _i, [_vp, _i, ctypes.POINTER(_vp)]),
    ("tg_screen_poynting", _i, [_vp, _i, _dp]),
    ("tg_count_in_window", _i, [_vp, _i, ctypes.POINTER(ctypes.c_longlong)]),
    ("tg_screen_spectra", _i, [_vp, _i, _dp, _dp, _dp, _dp, _dp]),
    ("tg_deposit_only", _i, [_vp, _i, _i, _dp, _dp, _d, _d, _i, _dp]),
    ("tg_fields_at", _i, [_vp, _i, _dp, _dp, _i, _i, _dp, _i]),
    ("tg_autotune", _i, [_vp, _i, ctypes.POINTER(TgTuning)]),
    ("tg_last_timing", _i, [_vp, _fp, _fp]),
    ("tg_last_stats", _i, [_vp] + [ctypes.POINTER(ctypes.c_longlong)] * 3),
    ("tg_track_undulator", _i, [_vp, _i, _dp, _dp, _d, _d, _d, _d, _i, _d, _d, _i, _dp]),
    ("tg_track_lattice", _i, [_vp, _i, _dp, _dp, _d, _d, _d, _d, _i, _d, _d, _i, _dp, _dp, _dp]),
    ("tg_get_trajectories", _i, [_vp, _i, _i] + [_dp] * 11),
    ("tg_trajectory_shape", _i, [_vp, _ip, _ip]),
]

_lib = None


def load(path=None):
    """Load the HIP library; raises if it has not been built.  TEUFEL_AMD_LIB
    in the environment names another build of the same library."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("TEUFEL_AMD_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(
            "teufel_amd: HIP library %s is not built (run __graft_entry__.build() "
            "or python -m teufel_amd.build); there is no CPU fallback" % path)
    lib = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib
