"""The allocation-free cube count (pcube_cube_count_noalloc) against the
allocating pcube_cube_count, both called through the C ABI.

Two listed calls in a row, the second on a smaller range, so a stale counter
or root word left in the shared static array by the first would show; the
full-domain count needs no bitset and has no size limit; a listed range past
the array's maximum is refused; the library's memory pool does not move.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from psuedocubes_amd import native

pytestmark = pytest.mark.gpu

GOOD_WORDS_MAX = 1 << 16


def _odd_words(first: int, last: int) -> int:
    r0 = first | 1
    odds = 0 if r0 > last else (last - r0) // 2 + 1
    return (odds + 31) // 32


def _allocating(first: int, last: int, p: int, listed: bool) -> tuple[int, np.ndarray | None]:
    count, seconds = ctypes.c_uint64(), ctypes.c_double()
    good = np.zeros(max(1, _odd_words(first, last)), dtype=np.uint32) if listed else None
    native.call('pcube_cube_count', first, last, p, ctypes.byref(count),
                native.ptr32(good) if listed else None, good.size if listed else 0, ctypes.byref(seconds))
    return count.value, good


def _noalloc(first: int, last: int, p: int, listed: bool) -> tuple[int, np.ndarray | None]:
    count = ctypes.c_uint64()
    # sentinel words: every word the range covers must be overwritten
    good = np.full(max(1, _odd_words(first, last)), 0xA5A5A5A5, dtype=np.uint32) if listed else None
    native.call('pcube_cube_count_noalloc', first, last, p, ctypes.byref(count),
                native.ptr32(good) if listed else None, good.size if listed else 0)
    return count.value, good


def _reserved() -> int:
    reserved, peak, cap = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    native.call('pcube_memory', ctypes.byref(reserved), ctypes.byref(peak), ctypes.byref(cap))
    return reserved.value


def test_noalloc_equals_allocating_on_two_calls(runtime) -> None:
    # the first range fills most of the static bitset (61,036 of 65,536 words);
    # the second is smaller and starts elsewhere
    cases = [(1000000001, 1003906000, 619), (3000000001, 3000200000, 433)]
    assert _odd_words(*cases[0][:2]) <= GOOD_WORDS_MAX and _odd_words(*cases[1][:2]) < _odd_words(*cases[0][:2])
    for first, last, p in cases:
        expected, expected_good = _allocating(first, last, p, True)
        before = _reserved()
        got, got_good = _noalloc(first, last, p, True)
        assert _reserved() == before, 'the allocation-free entry point moved the memory pool'
        assert got == expected, (first, last, p)
        assert np.array_equal(got_good, expected_good), (first, last, p)
        assert int(np.unpackbits(got_good.view(np.uint8)).sum()) == got


def test_noalloc_full_domain_and_limits(runtime) -> None:
    got, _ = _noalloc(1000000001, 1442249570, 619, False)
    assert got == 38322775
    assert got == _allocating(1000000001, 1442249570, 619, False)[0]
    # a single root, even and odd, after the large calls
    for first, last, p in ((2, 2, 2), (3, 3, 3), (1000000007, 1000000007, 619)):
        assert _noalloc(first, last, p, True)[0] == _allocating(first, last, p, True)[0]
    # a listed range one word past the static bitset is refused before any work
    first = 1000000001
    last = first + 2 * (GOOD_WORDS_MAX * 32)
    assert _odd_words(first, last) == GOOD_WORDS_MAX + 1
    with pytest.raises(ValueError):
        _noalloc(first, last, 619, True)
    # an invalid prime bound is refused as by the allocating entry point
    with pytest.raises(ValueError):
        _noalloc(1, 10, 2001, False)
