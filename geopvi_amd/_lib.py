"""ctypes binding of libgeopvi_fwi2d.so (the C ABI in include/geopvi_fwi2d.h).

The HIP library is the only compute path: if it is missing or fails to
load, every solver entry point raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

from .forward.fwi2d.params import CParams

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgeopvi_fwi2d.so")

GPVI_PHASE_ALL = 0
GPVI_PHASE_FORWARD = 1
GPVI_PHASE_ADJOINT = 2
GPVI_FLAG_NO_GRAPH = 1
GPVI_FLAG_NO_STORE = 2
GPVI_FLAG_ONE_STREAM = 4
GPVI_ERR_HIP = -5  # HIP runtime failure (include/geopvi_fwi2d.h)
GPVI_ERR_CAPACITY = -7  # engine unit/tile tables full (include/geopvi_fwi2d.h)

# Every exported symbol and its ctypes signature: (restype, argtypes).
_P = ctypes.POINTER(CParams)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_sz = ctypes.c_size_t
SIGNATURES = {
    "gpvi_last_error": (ctypes.c_char_p, []),
    "gpvi_abi_version": (_i, []),
    "gpvi_fwi2d_read_params": (_i, [ctypes.c_char_p, _P]),
    "gpvi_fwi2d_validate": (_i, [_P]),
    "gpvi_fwi2d_plain_workspace_bytes": (_sz, [_P]),
    "gpvi_fwi2d_fwi_device": (_i, [_P, _vp, _vp, _vp, _vp, _i, _vp, _sz, _vp]),
    "gpvi_fwi2d_forward_device": (_i, [_P, _vp, _vp, _vp, _sz, _vp]),
    "gpvi_fwi2d_batch_workspace_bytes": (_sz, [_P, _i, _i]),
    "gpvi_fwi2d_model_workspace_bytes": (_sz, [_P, _i]),
    "gpvi_fwi2d_batch": (_i, [_P, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _vp, _sz, _vp]),
    "gpvi_fwi2d_batch_static": (_i, [_P, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp]),
    "gpvi_fwi2d_set_tuning": (_i, [_i, _i, _i]),
    "gpvi_fwi2d_get_tuning": (_i, [ctypes.POINTER(_i)] * 3),
    "gpvi_fwi2d_tile_shape": (_i, [_i, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "gpvi_fwi2d_set_time_blocking": (_i, [_i]),
    "gpvi_fwi2d_get_time_blocking": (_i, []),
    "gpvi_fwi2d_set_lanes": (_i, [_i]),
    "gpvi_fwi2d_get_lanes": (_i, []),
    "gpvi_engine_create": (_i, [_vp, _i, _i, ctypes.POINTER(_vp)]),
    "gpvi_engine_destroy": (_i, [_vp]),
    "gpvi_engine_submit": (_i, [_vp, _P, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _sz,
                                ctypes.POINTER(ctypes.c_longlong)]),
    "gpvi_engine_advance": (_i, [_vp, _i]),
    "gpvi_engine_job_status": (_i, [_vp, ctypes.c_longlong]),
    "gpvi_engine_release": (_i, [_vp, ctypes.c_longlong]),
    "gpvi_engine_cancel": (_i, [_vp, ctypes.c_longlong]),
    "gpvi_engine_active_jobs": (_i, [_vp]),
    "gpvi_engine_active_units": (_i, [_vp]),
    "gpvi_engine_pending_steps": (_i, [_vp]),
    "gpvi_engine_sync": (_i, [_vp]),
    "fwi_2D": (None, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                      ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _i, _i]),
    "forward_2D": (None, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_float),
                          ctypes.POINTER(ctypes.c_float), _i, _i]),
    "gpvi_compat_last_status": (_i, []),
}

_lib = None


class GpviError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code
        self.msg = msg


def load(path: str = LIB_PATH):
    """Load the library (once) and bind every signature.  Raises if missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"geopvi_amd HIP library not built: {path} is missing "
            "(run `python -m geopvi_amd.csrc.build` or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().gpvi_last_error()
    return msg.decode() if msg else ""


def check(rc: int) -> int:
    if rc < 0:
        raise GpviError(rc, last_error())
    return rc


def set_tuning(variant: int, grid: int, graph_steps: int) -> None:
    check(load().gpvi_fwi2d_set_tuning(int(variant), int(grid), int(graph_steps)))


def get_tuning():
    v, g, s = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    load().gpvi_fwi2d_get_tuning(ctypes.byref(v), ctypes.byref(g), ctypes.byref(s))
    return v.value, g.value, s.value


def set_time_blocking(steps_per_launch: int) -> None:
    check(load().gpvi_fwi2d_set_time_blocking(int(steps_per_launch)))


def get_time_blocking() -> int:
    return int(load().gpvi_fwi2d_get_time_blocking())


def set_lanes(lanes: int) -> None:
    """1 or 2: batched calls advance two halves of their units on two streams."""
    check(load().gpvi_fwi2d_set_lanes(int(lanes)))


def get_lanes() -> int:
    return int(load().gpvi_fwi2d_get_lanes())


def tile_shape(variant: int):
    tx, tz = ctypes.c_int(), ctypes.c_int()
    check(load().gpvi_fwi2d_tile_shape(int(variant), ctypes.byref(tx), ctypes.byref(tz)))
    return tx.value, tz.value
