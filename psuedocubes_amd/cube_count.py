"""Full-domain control on the device: roots coprime to every prime <= p.

Mirrors the reference's independent cube counts (`review/cube_count.c:12-38`,
`review/cube_count_independent.c:9-35`, run by `tools/check_cpu.py:24-29`):
the number of roots r in [first, last] with gcd(r, p#) = 1, i.e. the genuine
cubes r^3 the sieve domain holds (38,322,775 for [1000000001, 1442249570] at
p = 619).  Same argument limits, the same JSON fields, and the root list for
ranges of at most 100,000.
"""
from __future__ import annotations

import ctypes
from typing import Any

import numpy as np

from . import native

LIST_LIMIT = 100_000
SEGMENT_BYTES = (1 << 18) // 8   # one block's LDS bitset of odd roots


def cube_count(first: int, last: int, p: int, list_roots: bool = False) -> dict[str, Any]:
    """Count (and with list_roots, list) the roots in [first, last] coprime to p#."""
    if first < 1 or last < first or last > 2**32 - 1 or p < 2 or p > 2000 or (list_roots and last - first > LIST_LIMIT):
        raise ValueError('cube count needs 1 <= first <= last <= 2^32 - 1, 2 <= p <= 2000 and short listed ranges')
    count, seconds = ctypes.c_uint64(), ctypes.c_double()
    r0 = first | 1
    odds = 0 if r0 > last else (last - r0) // 2 + 1
    good = np.zeros(max(1, (odds + 31) // 32), dtype=np.uint32) if list_roots else None
    native.call('pcube_cube_count', first, last, p, ctypes.byref(count),
                native.ptr32(good) if good is not None else None, 0 if good is None else good.size,
                ctypes.byref(seconds))
    row: dict[str, Any] = dict(algorithm='gpu-odd-bit-segment', root_first=str(first), root_last=str(last), p=p)
    if good is not None:
        bits = np.unpackbits(good.view(np.uint8), bitorder='little')[:odds]
        row['roots'] = [str(r0 + 2 * int(i)) for i in np.flatnonzero(bits)]
    row['count'] = str(count.value)
    row['working_bytes'] = SEGMENT_BYTES
    row['kernel_seconds'] = seconds.value
    return row
