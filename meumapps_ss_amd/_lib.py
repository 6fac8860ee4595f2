"""ctypes binding of the C ABI in include/meumapps_ss.h.

The library is the product: there is no CPU fallback.  Loading fails loudly
when ``libmeumapps_ss.so`` has not been built.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int32, c_size_t, c_void_p

from .params import MMAdjGrads, MMDiag, MMParams

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmeumapps_ss.so")

_lib = None


class MMError(RuntimeError):
    """A nonzero return code of the C ABI."""

    def __init__(self, code: int, where: str):
        msg = load().mm_errstr(code).decode() if _lib is not None else str(code)
        super().__init__(f"{where}: error {code}: {msg}")
        self.code = code


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    lib = ctypes.CDLL(LIB_PATH)
    D = POINTER(c_double)
    sigs = {
        "mm_create": (c_int32, [POINTER(MMParams), c_int32, POINTER(c_void_p)]),
        "mm_destroy": (None, [c_void_p]),
        "mm_upload": (c_int32, [c_void_p, D, D, D, D, D]),
        "mm_download": (c_int32, [c_void_p, D, D, D, D, D]),
        "mm_step": (c_int32, [c_void_p, c_int32, POINTER(MMDiag)]),
        "mm_get_step": (c_int32, [c_void_p]),
        "mm_set_step": (c_int32, [c_void_p, c_int32]),
        "mm_field": (c_int32, [c_void_p, c_int32, POINTER(c_void_p), POINTER(c_int32), POINTER(c_int32)]),
        "mm_field_to_host": (c_int32, [c_void_p, c_int32, D]),
        "mm_fft_create": (c_int32, [c_int32, c_int32, c_int32, c_int32, POINTER(c_void_p)]),
        "mm_fft_destroy": (None, [c_void_p]),
        "mm_fft_forward": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
        "mm_fft_inverse": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
        "mm_fft_spectrum_elems": (c_size_t, [c_void_p]),
        "mm_fft_forward_noalloc": (c_int32, [c_int32, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p]),
        "mm_ran2": (c_double, [POINTER(c_int32), POINTER(c_int32)]),
        "mm_ran2_fill": (c_int32, [POINTER(c_int32), POINTER(c_int32), ctypes.c_int64, D]),
        "mm_langevin": (c_double, [POINTER(c_int32), POINTER(c_int32), c_double]),
        "mm_noise_accumulate": (c_int32, [POINTER(c_int32), POINTER(c_int32), c_double, c_int32,
                                          ctypes.c_int64, D]),
        "mm_set_noise": (c_int32, [c_void_p, c_double, c_int32, POINTER(c_int32)]),
        "mm_get_noise_state": (c_int32, [c_void_p, POINTER(c_int32), POINTER(c_int32)]),
        "mm_ic_uniform": (c_int32, [c_int32, c_int32, c_int32, c_int32, c_int32, c_int32, c_double,
                                    D, D, D]),
        "mm_errstr": (ctypes.c_char_p, [c_int32]),
        "mm_probe": (c_int32, [c_void_p, c_int32]),
        "mm_probe_read": (c_int32, [c_void_p, POINTER(c_double), POINTER(ctypes.c_int64), POINTER(c_double)]),
        "mm_get_tuning": (c_int32, [c_void_p, POINTER(c_int32), POINTER(ctypes.c_float)]),
        "mm_set_graph": (c_int32, [c_void_p, c_int32, c_int32]),
        "mm_batch_create": (c_int32, [c_int32, c_void_p, c_int32, POINTER(c_void_p)]),
        "mm_batch_member": (c_int32, [c_void_p, c_int32, POINTER(c_void_p)]),
        "mm_batch_step": (c_int32, [c_void_p, c_int32, c_void_p]),
        "mm_batch_destroy": (None, [c_void_p]),
        "mm_adj_enable": (c_int32, [c_void_p]),
        "mm_adj_zero": (c_int32, [c_void_p]),
        "mm_adj_set": (c_int32, [c_void_p, D, D, D]),
        "mm_adj_get": (c_int32, [c_void_p, D, D, D]),
        "mm_state_bytes": (c_int32, [c_void_p, POINTER(ctypes.c_int64)]),
        "mm_state_save": (c_int32, [c_void_p, c_void_p]),
        "mm_state_load": (c_int32, [c_void_p, c_void_p]),
        "mm_adj_tape_step": (c_int32, [c_void_p, POINTER(MMDiag)]),
        "mm_adj_step": (c_int32, [c_void_p, POINTER(MMAdjGrads)]),
        "mm_branch_bytes": (c_int32, [c_void_p, POINTER(ctypes.c_int64)]),
        "mm_branch_tape": (c_int32, [c_void_p, c_int32, c_void_p, c_int32]),
        "mm_branch_flips": (c_int32, [c_void_p, POINTER(ctypes.c_int64)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, where: str) -> None:
    if rc != 0:
        raise MMError(rc, where)


EXPORTED = [
    "mm_create", "mm_destroy", "mm_upload", "mm_download", "mm_step", "mm_get_step",
    "mm_set_step", "mm_field", "mm_field_to_host", "mm_fft_create", "mm_fft_destroy",
    "mm_fft_forward", "mm_fft_inverse", "mm_fft_spectrum_elems", "mm_fft_forward_noalloc", "mm_ran2",
    "mm_ic_uniform",
    "mm_errstr", "mm_probe", "mm_probe_read", "mm_get_tuning", "mm_set_graph",
    "mm_batch_create", "mm_batch_member", "mm_batch_step", "mm_batch_destroy",
    "mm_adj_enable", "mm_adj_zero", "mm_adj_set", "mm_adj_get", "mm_state_bytes", "mm_state_save",
    "mm_state_load", "mm_adj_tape_step", "mm_adj_step", "mm_langevin", "mm_set_noise",
    "mm_get_noise_state", "mm_noise_accumulate", "mm_ran2_fill", "mm_branch_bytes", "mm_branch_tape",
    "mm_branch_flips",
]
