"""ctypes binding of libflr_hip.so (C ABI declared in include/flr.h).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

from ._build import LIB_PATH

FLR_OK = 0
FLR_ERR_INVALID_ARGUMENT = 1
FLR_ERR_HIP = 2
FLR_ERR_WINDOW = 3
FLR_ERR_NO_ORBITS = 4
FLR_ERR_UNSUPPORTED = 5

FLR_FLAG_EXACT = 0x1
FLR_FLAG_T_UNIFORM = 0x2
FLR_FLAG_ZERO_EDGES = 0x4
FLR_FLAG_OVERWRITE = 0x8
FLR_FLAG_AET = 0x10
FLR_FLAG_MIXED = 0x20
FLR_FLAG_Q31 = 0x40
FLR_FLAG_C128 = 0x80

FLR_GB_COMPLEX128 = 0
FLR_GB_COMPLEX64 = 1
FLR_GB_Q31 = 2

# name -> (restype, argtypes); kept in the order of include/flr.h
_vp, _ci, _cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
SIGNATURES = {
    "flr_last_error": (ctypes.c_char_p, []),
    "flr_version": (ctypes.c_char_p, []),
    "flr_device_count": (_ci, [ctypes.POINTER(_ci)]),
    "flr_orbits_create": (_ci, [_cd, _ci, _vp, _vp, _vp, _vp, _vp, _vp, _cd, ctypes.POINTER(_vp)]),
    "flr_orbits_destroy": (None, [_vp]),
    "flr_orbits_extent": (_ci, [_vp, ctypes.POINTER(_cd), ctypes.POINTER(_cd)]),
    "flr_orbits_eval": (_ci, [_vp, _vp, _ci, _ci, _ci, _vp, _vp, _vp]),
    "flr_response_create": (_ci, [_vp, ctypes.POINTER(_vp)]),
    "flr_response_destroy": (None, [_vp]),
    "flr_get_response": (_ci, [_vp, _vp, _vp, _vp, _vp, _vp, _cd, _ci, _vp, _ci, _ci, _cd, _ci,
                               _vp, _cd, _ci, _vp, _ci, _vp]),
    "flr_get_tdi_delays": (_ci, [_vp, _vp, _vp, _ci, _ci, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _ci,
                                 _ci, _ci, _cd, _ci, _vp, _cd, _ci, _vp, _ci, _vp]),
    "flr_set_lagrange": (_ci, [_vp, _ci, _vp, _cd, _ci, _vp]),
    "flr_set_tdi": (_ci, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _ci, _ci]),
    "flr_projection": (_ci, [_vp, _vp, _vp, _ci, _vp, _cd, _cd, _cd, _vp, _vp, _vp, _ci, _ci, _vp]),
    "flr_tdi": (_ci, [_vp, _vp, _vp, _ci, _vp, _cd, _cd, _cd, _ci, _ci, _vp]),
    "flr_check_errors": (_ci, [_vp, _vp]),
    "flr_batch_create": (_ci, [_vp, _vp, _ci, _vp, _vp, _vp, _cd, _cd, _ci, ctypes.POINTER(_vp)]),
    "flr_fused_create": (_ci, [_vp, _vp, _ci, _vp, _vp, _cd, _cd, _ci, ctypes.POINTER(_vp)]),
    "flr_fused_reserve": (_ci, [_vp, _ci, _cd, _cd, _ci, ctypes.POINTER(_vp)]),
    "flr_fused_bind": (_ci, [_vp, _vp, _ci, _vp, _vp, _vp]),
    "flr_batch_run": (_ci, [_vp, _vp]),
    "flr_fused_inner": (_ci, [_vp, _vp, _vp, _vp]),
    "flr_batch_set_launch": (_ci, [_vp, _ci, _ci]),
    "flr_batch_samples": (ctypes.c_int64, [_vp]),
    "flr_batch_destroy": (None, [_vp]),
    "flr_fused_run_static": (_ci, [_vp, _vp, _ci, _vp, _vp, _cd, _cd, _ci, _ci, _ci, _vp]),
    "flr_gb_generate": (_ci, [_vp, _ci, ctypes.c_int64, _cd, _ci, _vp, _vp]),
    "flr_response_set_exact": (_ci, [_vp, _ci]),
}


class FlrSource(ctypes.Structure):
    """flr_source of include/flr.h: one source of a packed batch."""
    _fields_ = [
        ("offset", ctypes.c_int64),
        ("num_pts", ctypes.c_int32),
        ("projections_start_ind", ctypes.c_int32),
        ("tdi_start_ind", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("t0", ctypes.c_double),
        ("k", ctypes.c_double * 3),
        ("u", ctypes.c_double * 3),
        ("v", ctypes.c_double * 3),
        ("scale", ctypes.c_double),
    ]

_lib = None


class FLRError(RuntimeError):
    """A failing libflr_hip call (HIP runtime error or device-side check)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[flr status {code}] {message}")
        self.code = code


def library_path() -> str:
    return os.environ.get("FLR_HIP_LIBRARY", LIB_PATH)


def lib():
    """The loaded libflr_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is missing: build the HIP library first "
                "(python __graft_entry__.py build, or fastlisaresponse_amd._build.build())"
            )
        L = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    """Raise the Python exception matching a nonzero status (ValueError for bad arguments, like
    the reference's std::invalid_argument, FLRError otherwise)."""
    if status == FLR_OK:
        return
    msg = lib().flr_last_error().decode(errors="replace")
    if status in (FLR_ERR_INVALID_ARGUMENT, FLR_ERR_NO_ORBITS):
        raise ValueError(msg)
    raise FLRError(status, msg)
