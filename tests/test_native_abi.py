"""The C ABI: libpcube.so builds, loads and exports every symbol include/pcube.h declares.

No device calls are made here; these run on machines without a GPU.
"""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

from psuedocubes_amd import native

HEADER = Path(__file__).resolve().parents[1] / 'include' / 'pcube.h'


def declared() -> list[str]:
    text = re.sub(r'/\*.*?\*/', '', HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r'\b(pcube_[a-z_0-9]+)\s*\(', text)))


def test_header_declares_the_engine_surface() -> None:
    names = declared()
    for required in ('pcube_init', 'pcube_focus_create', 'pcube_focus_window', 'pcube_matcher_create',
                     'pcube_match', 'pcube_lease', 'pcube_arithmetic', 'pcube_last_error'):
        assert required in names


def test_library_exports_every_declared_symbol() -> None:
    lib = ctypes.CDLL(str(native.LIBRARY))
    for name in declared():
        assert hasattr(lib, name), name


def test_python_binding_covers_the_header() -> None:
    bound = set(native.SIGNATURES) | {'pcube_last_error'}
    assert set(declared()) == bound
    native.lib()  # loads and applies every signature


def test_status_codes_map_to_reference_exceptions() -> None:
    import pytest
    for code, exc in ((native.OVERFLOW, OverflowError), (native.INVALID, ValueError),
                      (native.UNDERFLOW, ArithmeticError), (native.RANGE, ArithmeticError),
                      (native.MISMATCH, ArithmeticError),
                      (native.NOMEM, MemoryError), (native.DEVICE, RuntimeError)):
        with pytest.raises(exc):
            native.check(code)
    assert native.check(native.OK) == native.OK
