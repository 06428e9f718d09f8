"""Native host verification of emitted values (C++ in libpcube, CPU only).

Same semantics as the Python predicate (`pseudocube/oracle.py:36-41`,
`review/verify_cube.py:8-20`): every value must satisfy local3(n, p), and
each noncube becomes the witness record `verify(n, p)` would produce.  Used
where emitted values are numerous (small prime bounds in the batched and
served forms), where the pure-Python exponentiation would dominate.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import native

CLAIM = 'Candidate validity only; no minimality claim.'


def native_check(values: list[int], p: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(local3 flags, perfect-cube flags, floor cube roots) for each value."""
    n = len(values)
    mask = (1 << 64) - 1
    limbs = np.array([x for v in values for x in (v & mask, v >> 64)] or [0, 0], dtype=np.uint64)
    ok = np.zeros(max(1, n), dtype=np.uint8)
    cube = np.zeros(max(1, n), dtype=np.uint8)
    root = np.zeros(max(1, n), dtype=np.uint64)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    native.check(native.lib().pcube_host_verify(native.ptr64(limbs), n, p, ok.ctypes.data_as(u8p),
                                                cube.ctypes.data_as(u8p), native.ptr64(root)), 'pcube_host_verify')
    return ok[:n], cube[:n], root[:n]


def native_witnesses(values: list[int], p: int) -> list[dict[str, object]]:
    """Drop-in for batch_experiment.python_witnesses using the native predicate."""
    if not values:
        return []
    ok, cube, root = native_check(values, p)
    if not ok.all():
        raise ArithmeticError('GPU/CPU disagreement')
    # every value passed local3(n, p), so n > 0, n mod 9 is 1 or 8, and no prime <= p
    # divides n or (for q = 1 mod 3) fails the cubic-residue test: verify() would list
    # no failing prime, and a noncube among them is a valid pseudocube
    out = []
    for n, local, c, r in zip(values, ok.tolist(), cube.tolist(), root.tolist()):
        if c:
            continue
        out.append({'n': str(n), 'p': p, 'cube_root_floor': str(r), 'perfect_cube': False, 'failing_primes': [],
                    'valid_pseudocube': bool(local), 'claim': CLAIM})
    return out
