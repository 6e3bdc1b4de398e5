"""Loader and signatures of liblensdiff_hip.so (the C ABI of include/lensdiff_hip.h).

There is no fallback: if the HIP library is missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblensdiff_hip.so")

# every symbol the header declares, with (restype, argtypes)
_F = C.POINTER(C.c_float)
_P = C.POINTER(_abi.Params)
_CTX = C.c_void_p
_BANK = C.c_void_p
SIGNATURES = {
    "ldh_params_default": (None, [_P]),
    "ldh_context_create": (C.c_int32, [C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "ldh_context_destroy": (C.c_int32, [_CTX]),
    "ldh_status_string": (C.c_char_p, [C.c_int32]),
    "ldh_last_error": (C.c_char_p, [_CTX]),
    "ldh_synchronize": (C.c_int32, [_CTX]),
    "ldh_resolve_sizes": (C.c_int32, [_P, C.POINTER(_abi.Sizes)]),
    "ldh_pupil_amplitude": (C.c_int32, [_CTX, _P, _F]),
    "ldh_pupil_phase": (C.c_int32, [_CTX, _P, _F]),
    "ldh_raw_psf": (C.c_int32, [_CTX, _F, _F, C.c_int32, C.c_int32, _F]),
    "ldh_raw_psf_window": (C.c_int32, [_CTX, _F, _F, C.c_int32, C.c_int32, C.c_int32, _F]),
    "ldh_first_minimum_radius": (C.c_int32, [_CTX, _F, C.c_int32, _F]),
    "ldh_resample_kernel": (C.c_int32, [_CTX, _F, C.c_int32, C.c_float, C.c_int32, _F]),
    "ldh_shape_kernels": (C.c_int32, [_CTX, _P, _F, C.c_int32, _F, _F, _F, C.POINTER(C.c_int32)]),
    "ldh_sequential_sum": (C.c_int32, [_CTX, _F, C.c_int64, _F]),
    "ldh_sequential_sums": (C.c_int32, [_CTX, _F, C.c_int64, C.c_int32, C.c_int32, _F]),
    "ldh_fft2d": (C.c_int32, [_CTX, _F, C.c_int32, C.c_int32, C.c_int32]),
    "ldh_sequential_sums_noalloc": (C.c_int32, [_CTX, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]),
    "ldh_bank_acquire": (C.c_int32, [_CTX, _P, C.POINTER(C.c_void_p)]),
    "ldh_bank_release": (C.c_int32, [_CTX, _BANK]),
    "ldh_bank_build_batch": (C.c_int32, [_CTX, _P, C.c_int32, C.c_int32, C.POINTER(C.c_void_p),
                                         C.POINTER(_abi.BatchStats)]),
    "ldh_bank_info_get": (C.c_int32, [_CTX, _BANK, C.POINTER(_abi.BankInfo)]),
    "ldh_bank_download": (C.c_int32, [_CTX, _BANK, C.c_int32, C.c_int32, _F, C.c_int64]),
    "ldh_render": (C.c_int32, [_CTX, C.POINTER(_abi.RenderRequest), _P]),
    "ldh_render_device": (C.c_int32, [_CTX, _P, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32]),
    "ldh_server_start": (C.c_int32, [_CTX, C.c_int32, C.c_uint64]),
    "ldh_server_stop": (C.c_int32, [_CTX]),
    "ldh_submit": (C.c_int32, [_CTX, _P, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int64)]),
    "ldh_poll": (C.c_int32, [_CTX, C.c_int64, C.POINTER(_abi.RequestStats)]),
    "ldh_wait": (C.c_int32, [_CTX, C.c_int64, C.POINTER(_abi.RequestStats)]),
    "ldh_server_stats_get": (C.c_int32, [_CTX, C.POINTER(_abi.ServerStats)]),
    "ldh_device_alloc": (C.c_int32, [_CTX, C.c_uint64, C.POINTER(C.c_void_p)]),
    "ldh_device_free": (C.c_int32, [_CTX, C.c_void_p]),
    "ldh_memcpy_h2d": (C.c_int32, [_CTX, C.c_void_p, C.c_void_p, C.c_uint64]),
    "ldh_memcpy_d2h": (C.c_int32, [_CTX, C.c_void_p, C.c_void_p, C.c_uint64]),
    "ldh_profiling_enable": (C.c_int32, [_CTX, C.c_int32]),
    "ldh_profiling_read": (C.c_int32, [_CTX, C.POINTER(_abi.KernelStat), C.c_int32]),
    "ldh_profiling_reset": (C.c_int32, [_CTX]),
    "ldh_tuning_report": (C.c_int32, [C.c_char_p, C.c_int32]),
    "ldh_tuning_force": (None, [C.c_int32]),
    "ldh_debug_inject_failure": (None, [C.c_char_p, C.c_int32]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library (once). Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            "liblensdiff_hip.so is not built (%s); run `python -c \"import __graft_entry__ as g; "
            "g.build()\"` or `make -C lensdiff_amd/csrc`" % path)
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class LensDiffError(RuntimeError):
    def __init__(self, status: int, tag: str):
        super().__init__("%s (%s)" % (tag, status))
        self.status = status
        self.tag = tag
