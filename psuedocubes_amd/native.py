"""ctypes binding of libpcube.so (declared in include/pcube.h).

The HIP library is the only compute path: importing this module on a machine
without the built library raises immediately; there is no CPU substitute.
"""
from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Any

import numpy as np

LIBRARY = Path(__file__).resolve().parent / 'libpcube.so'

OK, OVERFLOW, INVALID, UNDERFLOW, NOMEM, DEVICE, RANGE, MISMATCH = range(8)

u64 = ctypes.c_uint64
u32 = ctypes.c_uint32
c_int = ctypes.c_int
P_u64 = ctypes.POINTER(ctypes.c_uint64)
P_u32 = ctypes.POINTER(ctypes.c_uint32)
P_dbl = ctypes.POINTER(ctypes.c_double)
P_int = ctypes.POINTER(ctypes.c_int)

# name -> argument types; every function returns int (status)
SIGNATURES: dict[str, list[Any]] = {
    'pcube_device_count': [P_int],
    'pcube_init': [c_int, u64, ctypes.c_double, P_u64, P_u64, P_u64],
    'pcube_memory': [P_u64, P_u64, P_u64],
    'pcube_memory_peak': [],
    'pcube_device_info': [ctypes.c_char_p, ctypes.c_size_t, P_int],
    'pcube_kernel_info': [c_int, ctypes.c_char_p, ctypes.c_size_t, P_int, P_int, P_int],
    'pcube_alloc': [u64, P_u64],
    'pcube_free': [u64],
    'pcube_upload': [u64, ctypes.c_void_p, u64],
    'pcube_download': [ctypes.c_void_p, u64, u64],
    'pcube_pinned_alloc': [u64, ctypes.POINTER(ctypes.c_void_p)],
    'pcube_pinned_free': [ctypes.c_void_p],
    'pcube_host_focus_window': [P_u64, u64, P_u64, u64, u64, u64, u64, u64, P_u64, u64, P_u64],
    'pcube_focus_create': [u64, P_u64, P_u32, u32, u32, P_int, P_u64, P_u64],
    'pcube_focus_arrays': [c_int, P_u64, P_u64],
    'pcube_focus_destroy': [c_int],
    'pcube_focus_window': [c_int, u64, u64, u64, u64, P_u64],
    'pcube_focus_wheel': [c_int, P_u32, P_u32, P_u32, u32, P_u64, P_u32],
    'pcube_focus_windows': [c_int, u32, P_u64, P_u64, u64, u64, P_u64],
    'pcube_tile_ranges': [u64, u64, u64, u64, u64, u64, u64, u64, u64, P_u64, P_u32],
    'pcube_tile_ranges_many': [u64, u64, u64, u64, u64, u64, u32, P_u64, P_u64, P_u64, P_u64, u64, P_u32],
    'pcube_prepare_range': [c_int, c_int, u64, u64, u64, u64, u64, u64, u64, u64, u64, P_u64, P_u32, u64, u64, u64,
                            u64, P_u64, P_u64, P_u64, P_u64],
    'pcube_prepare_match': [c_int, c_int, u64, u64, u64, u64, u64, u64, u64, u64, u64, P_u64, P_u32, u64, u64, u64,
                            u64, P_u64, P_u64, P_u64, P_u64, c_int, u64, c_int, u64, P_u64],
    'pcube_match_adoptable': [c_int, u64, P_int],
    'pcube_prepare': [c_int, c_int, u32, P_u64, u64, u64, u64, u64, P_u64, P_u64, P_u64],
    'pcube_searchsorted': [u64, u64, P_u64, u64, P_u64],
    'pcube_matcher_create': [u64, u64, u64, u64, u64, u64, P_u32, u32, P_int],
    'pcube_matcher_destroy': [c_int],
    'pcube_matcher_table_fill': [c_int, c_int],
    'pcube_match': [c_int, u64, u64, u64, u64, u64, u64, u32, u64, c_int, P_u64, P_u64, u64, P_dbl],
    'pcube_match_start': [c_int, u64, u64, u64, u64, u64, u64, u32, u64, c_int],
    'pcube_match_finish': [c_int, P_u64, P_u64, u64, P_dbl],
    'pcube_lease': [c_int, c_int, c_int, u32, P_u32, P_u64, u64, u64, u64, P_u64, P_u64, P_u64, P_u64, u64, P_dbl],
    'pcube_batch_create': [P_int, u32, c_int, c_int, u64, u32, c_int, P_int],
    'pcube_batch_destroy': [c_int],
    'pcube_batch_run': [c_int, u32, P_u32, P_u32, P_u64, P_u64, P_u32, u32, P_u64, P_u64, P_u64, P_u64, u64, P_dbl,
                        P_u64],
    'pcube_host_verify': [P_u64, u64, u32, ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_uint8), P_u64],
    'pcube_batch_values': [u64, P_u64, u32, P_u64, P_u64, P_u64, P_u32, u64, u64, u64, u64, u64, u64,
                           ctypes.POINTER(ctypes.c_uint8), P_u64, ctypes.c_void_p, u64, P_u64, P_u32],
    'pcube_floor_div': [P_u64, P_u64, u64],
    'pcube_bit_transpose': [P_u64, P_u64, u64],
    'pcube_scope_records': [u32, P_u64, P_u64, u32, u32, P_u64, ctypes.c_char_p, ctypes.c_void_p, u64, P_u64],
    'pcube_scope_records_many': [u32, P_u32, P_u64, P_u64, u32, u32, P_u64, ctypes.c_char_p, ctypes.c_void_p, u64,
                                 P_u64],
    'pcube_arithmetic': [P_u64, P_u64, u64],
    'pcube_copy_probe': [u64, P_dbl],
    'pcube_lds_probe': [u32, u32, P_dbl, P_dbl],
    'pcube_valu_probe': [u32, P_dbl, P_dbl],
    'pcube_cube_count': [u64, u64, u32, P_u64, P_u32, u64, P_dbl],
    'pcube_cube_count_noalloc': [u64, u64, u32, P_u64, P_u32, u64],
}

_lib: ctypes.CDLL | None = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            raise ImportError(f'{LIBRARY} is missing: build it with python -m psuedocubes_amd.build '
                              '(there is no CPU fallback)')
        handle = ctypes.CDLL(str(LIBRARY))
        for name, argtypes in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = argtypes
            fn.restype = ctypes.c_int
        handle.pcube_last_error.argtypes = []
        handle.pcube_last_error.restype = ctypes.c_char_p
        handle.pcube_memory_peak.restype = ctypes.c_void_p
        _lib = handle
    return _lib


def error_text() -> str:
    return lib().pcube_last_error().decode(errors='replace')


def check(code: int, what: str = '') -> int:
    """Map a status code to the reference's exception types."""
    if code == OK:
        return code
    message = f'{what}: {error_text()}' if what else error_text()
    if code == OVERFLOW:
        raise OverflowError(message)
    if code == INVALID:
        raise ValueError(message)
    if code in (UNDERFLOW, RANGE, MISMATCH):
        raise ArithmeticError(message)
    if code == NOMEM:
        raise MemoryError(message)
    raise RuntimeError(message)


def call(name: str, *args: Any) -> int:
    return check(getattr(lib(), name)(*args), name)


def ptr64(array: np.ndarray) -> Any:
    return array.ctypes.data_as(P_u64)


def ptr64_at(address: int) -> Any:
    """A uint64 pointer at a host address (inside an array the caller keeps alive)."""
    return ctypes.cast(address, P_u64)


def ptr32(array: np.ndarray) -> Any:
    return array.ctypes.data_as(P_u32)
