"""The producer/consumer staging path on the CPU (tools/batch_window_experiment.py:52-82, 256-336).

The host window producer pcube_host_focus_window (the CPU producer of the staging path) is checked
against the oracle's window, against brute force on tiny foci, and against the reference's own CPU
producer review/focus_window_cpu.cpp compiled into oracle/_ref when that build is present.  It runs
on the host only; no GPU call is made here.
"""
from __future__ import annotations

import ctypes
import random
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from psuedocubes_amd import native

ROOT = Path(__file__).resolve().parents[1]
REF_WINDOW = ROOT / 'oracle' / '_ref' / 'focus_window.so'


def host_window(C: np.ndarray, D: np.ndarray, M: int, lo: int, hi: int, base: int = 0,
                cap: int = 1_000_000) -> tuple[int, np.ndarray, int]:
    out = np.empty(max(1, cap), dtype=np.uint64)
    count = ctypes.c_uint64()
    code = native.lib().pcube_host_focus_window(native.ptr64(C), C.size, native.ptr64(D), D.size, M, lo, hi, base,
                                                 native.ptr64(out), cap, ctypes.byref(count))
    return code, out[:min(count.value, cap)].copy(), count.value


def oracle_window(C: np.ndarray, D: np.ndarray, M: int, lo: int, hi: int, base: int = 0) -> np.ndarray:
    out = np.empty(1_000_000, dtype=np.uint64)
    count = ctypes.c_uint64()
    assert O.lib().oracle_window(C, C.size, D, D.size, M, base + lo, base + hi, out, out.size, ctypes.byref(count)) == 0
    return out[:count.value].copy()


def random_halves(rng: random.Random, M: int, nc: int, nd: int) -> tuple[np.ndarray, np.ndarray]:
    C = np.array(sorted(rng.randrange(M) for _ in range(nc)), dtype=np.uint64)
    D = np.array(sorted(rng.randrange(M) for _ in range(nd)), dtype=np.uint64)
    return C, D


def test_host_window_brute_force_small() -> None:
    """Every y = period*M + (c + d) mod M in [lo, hi), by brute force, on tiny random halves."""
    rng = random.Random(619003)
    for trial in range(200):
        M = rng.randrange(2, 500)
        C, D = random_halves(rng, M, rng.randrange(0, 12), rng.randrange(0, 12))
        lo = rng.randrange(0, M + 1)
        hi = rng.randrange(lo, M + 1)
        base = rng.randrange(0, 5) * M
        code, got, n = host_window(C, D, M, lo, hi, base)
        assert code == 0
        assert got.tolist() == O.brute_window(C.tolist(), D.tolist(), M, base + lo, base + hi), (trial, M, lo, hi)


def test_host_window_equals_oracle_on_p619_focus(config619) -> None:
    """The A and B foci of p = 619 (4.2e12 and 6.7e12 classes): windows of a full batch, the
    period's ends, and windows that reach the top of the period."""
    rng = random.Random(5)
    for focus in config619.focuses():
        f = O.Focus(focus)
        M = f.modulus
        windows = [(0, 6_400_000_000), (M - 6_400_000_000, M), (M // 2, M // 2 + 20_000_000_000), (0, 0), (M, M)]
        windows += [(a, a + rng.randrange(1, 10**10)) for a in (rng.randrange(M - 10**10) for _ in range(4))]
        for lo, hi in windows:
            code, got, n = host_window(f.C, f.D, M, lo, hi)
            assert code == 0
            assert np.array_equal(got, oracle_window(f.C, f.D, M, lo, hi)), (M, lo, hi)
        base = 7 * M
        code, got, _ = host_window(f.C, f.D, M, 123, 6_400_000_123, base)
        assert code == 0 and np.array_equal(got, oracle_window(f.C, f.D, M, 123, 6_400_000_123, base))


@pytest.mark.skipif(not REF_WINDOW.exists(), reason='the reference CPU producer is built only where the reference is')
def test_host_window_equals_reference_cpu_producer(config619) -> None:
    """Same multiset as review/focus_window_cpu.cpp (which returns its output unsorted; the
    reference sorts it after the call, tools/batch_window_experiment.py:82)."""
    ref = ctypes.CDLL(str(REF_WINDOW))
    p64 = np.ctypeslib.ndpointer(dtype=np.uint64, ndim=1, flags='C_CONTIGUOUS')
    ref.focus_window.argtypes = [p64, ctypes.c_uint64, p64, ctypes.c_uint64] + [ctypes.c_uint64] * 4 + \
        [p64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    ref.focus_window.restype = ctypes.c_int
    rng = random.Random(11)
    for focus in config619.focuses():
        f = O.Focus(focus)
        M = f.modulus
        for lo, hi in [(0, 6_400_000_000), (M - 6_400_000_000, M)] + [(a, a + 10**10) for a in
                                                                       (rng.randrange(M - 10**10) for _ in range(3))]:
            out = np.empty(1_000_000, dtype=np.uint64)
            count = ctypes.c_uint64()
            assert ref.focus_window(f.C, f.C.size, f.D, f.D.size, M, lo, hi, 3 * M, out, out.size,
                                    ctypes.byref(count)) == 0
            code, got, _ = host_window(f.C, f.D, M, lo, hi, 3 * M)
            assert code == 0 and np.array_equal(got, np.sort(out[:count.value]))
    # overflow and invalid arguments report like the reference
    f = O.Focus(config619.focuses()[1])
    out = np.empty(10, dtype=np.uint64)
    count = ctypes.c_uint64()
    assert ref.focus_window(f.C, f.C.size, f.D, f.D.size, f.modulus, 0, 10**9, 0, out, 10, ctypes.byref(count)) == 1
    code, _, n = host_window(f.C, f.D, f.modulus, 0, 10**9, 0, cap=10)
    assert code == 1 and n == count.value == 11


def test_host_window_limits() -> None:
    C = np.array([1, 5, 9], dtype=np.uint64)
    D = np.array([0, 2, 7], dtype=np.uint64)
    assert host_window(C, D, 10, 5, 4)[0] == native.INVALID            # lo > hi
    assert host_window(C, D, 10, 0, 11)[0] == native.INVALID           # hi > M
    assert host_window(C, D, 1 << 63, 0, 1)[0] == native.INVALID       # M >= 2^63
    code, got, _ = host_window(C, D, 10, 0, 10)
    assert code == 0 and got.tolist() == O.brute_window(C.tolist(), D.tolist(), 10, 0, 10)
    code, got, n = host_window(C, D, 10, 0, 10, cap=3)
    assert code == native.OVERFLOW and n == 4
    empty = np.empty(0, dtype=np.uint64)
    assert host_window(empty, D, 10, 0, 10)[1].size == 0
