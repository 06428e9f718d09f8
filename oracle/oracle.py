"""CPU oracle for the batch engine: ctypes wrapper over liboracle.so plus
pure-Python brute force for tiny configurations.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker and never as the thing
measured or shipped.  The product package never imports it.

What it restates (reference file:line):
  * focus arrays C/D and windows: pseudocube/gpu.py:70-114,
    review/focus_window_cpu.cpp:4-25
  * batch geometry: pseudocube/batch_experiment.py:27-48
  * matcher counters/outputs: pseudocube/batch_experiment.py:104-157, 184-235
  * lease accumulation and transcript: pseudocube/optimized.py:40-78
  * brute force: pseudocube/focus.py:170-179 (enumerate_cpu) and the
    complete-set recipe of tests/test_pseudocube_math.py:74-88
  * full-domain cube count: review/cube_count.c:12-38

Pinning (each function has at least one independent check in tests/):
  * oracle_window    - brute-force multisets (tests/test_pseudocube_batch_experiment.py:18-34
                       recipe) and the reference's own compiled
                       focus_window_cpu.cpp under oracle/_ref when present
  * oracle_match     - complete sets vs local3 brute force at p in {7,13,19,31};
                       the 256-batch lease fixture transcript and output hash
                       (results/continuation-2026-09-15/positive-lease-replay.json)
  * oracle_partial   - Focus.at on sampled indices; focus_u/focus_v fixture sums
  * oracle_products  - 5 extremes + 2000 seeded cases vs Python ints and oracle/_ref/wide.so
  * oracle_cube_count - gcd brute force on seeded short ranges; the reference's
                       cube_count.c and cube_count_independent.c compiled under
                       oracle/_ref (counts and root lists); the reference's
                       known answer 38,322,775 (results/current.json:25)
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
from math import prod
from pathlib import Path
from typing import Any

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / 'liboracle.so'

_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, ndim=1, flags='C_CONTIGUOUS')
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, ndim=1, flags='C_CONTIGUOUS')
_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f'{LIB_PATH} missing: run make -C oracle')
        handle = ctypes.CDLL(str(LIB_PATH))
        handle.oracle_partial.argtypes = [_u64p, _u32p, ctypes.c_uint32, ctypes.c_uint64, _u64p, ctypes.c_uint64]
        handle.oracle_window.argtypes = [_u64p, ctypes.c_uint64, _u64p, ctypes.c_uint64, ctypes.c_uint64,
                                         ctypes.c_uint64, ctypes.c_uint64, _u64p, ctypes.c_uint64,
                                         ctypes.POINTER(ctypes.c_uint64)]
        handle.oracle_match.argtypes = ([_u64p, ctypes.c_uint64, _u64p, ctypes.c_uint64] + [ctypes.c_uint64] * 6
                                        + [_u32p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64,
                                           ctypes.c_uint32, ctypes.c_uint64, _u64p, _u64p, ctypes.c_int])
        handle.oracle_products.argtypes = [ctypes.c_uint64] * 4 + [ctypes.POINTER(ctypes.c_uint64)] * 2
        handle.oracle_cube_count.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                             ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64,
                                             ctypes.POINTER(ctypes.c_uint64)]
        handle.oracle_cube_count.restype = ctypes.c_uint64
        for name in ('oracle_partial', 'oracle_window', 'oracle_match', 'oracle_products'):
            getattr(handle, name).restype = ctypes.c_int
        _lib = handle
    return _lib


# Worker threads per process: the CPUs this process may run on, at most 16 (one GPU
# box grants a process 16 CPUs whatever os.cpu_count() reports there).
THREAD_LIMIT = 16


def available_cpus() -> int:
    """CPUs this process may run on (its affinity set), else os.cpu_count()."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def threads() -> int:
    return max(1, min(THREAD_LIMIT, available_cpus()))


# ---------------------------------------------------------------------------
# Focus arrays and windows

class Focus:
    """Sorted half arrays C and D of one focus, built by the C oracle."""

    def __init__(self, focus: Any) -> None:
        first, second = focus.halves()
        self.modulus = focus.modulus
        self.C = np.sort(partial(first, focus.modulus))
        self.D = np.sort(partial(second, focus.modulus))

    def window(self, low: int, high: int, cap: int = 1_000_000) -> np.ndarray:
        if not 0 <= low <= high < 1 << 64:
            raise ValueError('invalid window')
        out = np.empty(max(1, cap), dtype=np.uint64)
        count = ctypes.c_uint64()
        code = lib().oracle_window(self.C, len(self.C), self.D, len(self.D), self.modulus,
                                   low, high, out, cap, ctypes.byref(count))
        if code == 1:
            raise OverflowError('oracle window overflow')
        if code:
            raise ValueError('invalid oracle window arguments')
        return out[:count.value].copy()


def partial(rings: tuple[tuple[int, ...], ...], modulus: int) -> np.ndarray:
    terms = np.array([t for ring in rings for t in ring] or [0], dtype=np.uint64)
    sizes = np.array([len(ring) for ring in rings] or [1], dtype=np.uint32)
    n = prod(len(ring) for ring in rings)
    out = np.empty(n, dtype=np.uint64)
    if not rings:
        out[0] = 0
        return out
    if lib().oracle_partial(terms, sizes, len(rings), modulus, out, n):
        raise ValueError('invalid partial-sum arguments')
    return out


def brute_window(C: list[int], D: list[int], modulus: int, low: int, high: int) -> list[int]:
    """All period*M + (c + d) mod M in [low, high), by brute force."""
    out = []
    if high <= low:
        return out
    for period in range(low // modulus, (high - 1) // modulus + 1):
        for c in C:
            for d in D:
                y = period * modulus + (c + d) % modulus
                if low <= y < high:
                    out.append(y)
    return sorted(out)


# ---------------------------------------------------------------------------
# Batch matcher

def tile_ranges(config: Any, first: int, end: int, width: int) -> list[tuple[int, int, int, int]]:
    """(first, end, u_min, u_end) per subtile (pseudocube/batch_experiment.py:27-32)."""
    if not 0 <= first < end <= config.B:
        raise ValueError('bad tile')
    if width <= 0 or (end - first + width - 1) // width > 64:
        raise ValueError('at most 64 subtiles')
    out = []
    for v in range(first, end, width):
        stop = min(v + width, end)
        out.append((v, stop, (config.lo + v * config.A) // config.B + 1,
                    (config.hi + (stop - 1) * config.A) // config.B + 1))
    return out


def match(config: Any, us: np.ndarray, vs: np.ndarray, v_first: int, width: int, ntiles: int,
          cap: int = 8192, nthreads: int | None = None) -> tuple[np.ndarray, np.ndarray]:
    """(counters[ntiles, ng+3], values[ntiles, cap, 2]) exactly as the masked kernel defines them."""
    order = config.filter_order()
    moduli = np.array(order, dtype=np.uint32)
    ng = (len(order) + 3) // 4
    counters = np.zeros(ntiles * (ng + 3), dtype=np.uint64)
    values = np.zeros(ntiles * max(1, cap) * 2, dtype=np.uint64)
    mask = (1 << 64) - 1
    code = lib().oracle_match(np.ascontiguousarray(us, dtype=np.uint64), len(us),
                              np.ascontiguousarray(vs, dtype=np.uint64), len(vs),
                              config.A, config.B, config.lo & mask, config.lo >> 64,
                              config.hi & mask, config.hi >> 64, moduli, len(order),
                              v_first, width, ntiles, cap, counters, values, nthreads or threads())
    if code:
        raise ValueError('invalid oracle match arguments')
    return counters.reshape(ntiles, ng + 3), values.reshape(ntiles, max(1, cap), 2)


def _encoded(value: Any) -> bytes:
    def fix(v: Any) -> Any:
        if type(v) is int and abs(v) >= 1 << 53:
            return str(v)
        if isinstance(v, dict):
            return {str(k): fix(x) for k, x in v.items()}
        if isinstance(v, (list, tuple)):
            return [fix(x) for x in v]
        return v
    return (json.dumps(fix(value), sort_keys=True, separators=(',', ':'), allow_nan=False) + '\n').encode()


def _digest(value: Any) -> str:
    return hashlib.sha256(_encoded(value)).hexdigest()


def local3(n: int, p: int) -> bool:
    """Trial-division restatement of the local predicate (pseudocube/oracle.py:36-41)."""
    if n <= 0 or n % 9 not in (1, 8):
        return False
    for q in range(2, p + 1):
        if any(q % d == 0 for d in range(2, int(q ** 0.5) + 1)):
            continue
        if n % q == 0 or (q % 3 == 1 and pow(n, (q - 1) // 3, q) != 1):
            return False
    return True


def batch(config: Any, fa: Focus, fb: Focus, first: int, end: int, width: int = 100_000_000,
          cap: int = 8192, nthreads: int | None = None) -> dict[str, Any]:
    """One compute batch: windows, bounds, match, per-subtile scopes.

    Returns the fields of the reference's batch row that the lease transcript
    and compare() cover (pseudocube/batch_experiment.py:145-164).
    """
    ranges = tile_ranges(config, first, end, width)
    us = fa.window(ranges[0][2], ranges[-1][3])
    vs = fb.window(first, end)
    bounds = []
    for v0, v1, u0, u1 in ranges:
        a, b = np.searchsorted(us, np.array([u0, u1], dtype=np.uint64))
        c, d = np.searchsorted(vs, np.array([v0, v1], dtype=np.uint64))
        if (int(b) - int(a)) * (int(d) - int(c)) > 50_000_000:
            raise OverflowError('per-subtile opportunity limit')
        bounds.append((int(a), int(b), int(c), int(d)))
    counters, values = match(config, us, vs, first, width, len(ranges), cap, nthreads)
    ng = counters.shape[1] - 3
    if counters[:, ng + 2].any():
        raise ArithmeticError('underflow')
    if (counters[:, ng + 1] > cap).any():
        raise OverflowError('output overflow')
    scopes, all_values = [], []
    for row, (v0, v1, _, _) in enumerate(ranges):
        number = int(counters[row, ng + 1])
        got = sorted(int(values[row, s, 0]) + (int(values[row, s, 1]) << 64) for s in range(number))
        all_values.extend(got)
        scopes.append(dict(first=str(v0), end=str(v1), valid_focused_candidates=int(counters[row, 0]),
                           group_survivors=[int(n) for n in counters[row, 1:ng + 1]],
                           local_survivors=number, output_sha256=_digest([str(n) for n in got])))
    groups = [int(n) for n in counters[:, 1:ng + 1].sum(axis=0)]
    return dict(focus_u=sum(b - a for a, b, _, _ in bounds), focus_v=len(vs),
                cartesian_opportunities=sum((b - a) * (d - c) for a, b, c, d in bounds),
                valid_focused_candidates=int(counters[:, 0].sum()), local_survivors=len(all_values),
                subtile_scopes=scopes, group_survivors=groups, filter_groups=config.filter_groups(),
                all_values=[str(n) for n in all_values])


def lease(config: Any, fa: Focus, fb: Focus, first: int, end: int, batch_width: int = 6_400_000_000,
          width: int = 100_000_000, on_batch: Any = None) -> dict[str, Any]:
    """Accumulate batches like pseudocube/optimized.py:40-78; returns the compact-row fields."""
    transcript = hashlib.sha256()
    totals = dict(focus_u=0, focus_v=0, cartesian_opportunities=0, valid_focused_candidates=0, local_survivors=0)
    groups: list[int] | None = None
    values: list[str] = []
    batches = 0
    for start in range(first, end, batch_width):
        stop = min(start + batch_width, end)
        row = batch(config, fa, fb, start, stop, width)
        for key in totals:
            totals[key] += row[key]
        groups = row['group_survivors'] if groups is None else [a + b for a, b in zip(groups, row['group_survivors'])]
        values.extend(row['all_values'])
        encoded = _encoded(row['subtile_scopes'])
        transcript.update(encoded)
        if on_batch is not None:
            on_batch(start, stop, row, encoded)
        batches += 1
    return dict(totals, group_survivors=groups, all_values=values, compute_batches=batches,
                transcript_sha256=transcript.hexdigest(), output_sha256=_digest(values))


def cube_count(first: int, last: int, p: int, list_roots: bool = False) -> tuple[int, list[int] | None]:
    """Roots in [first, last] coprime to every prime <= p (review/cube_count.c:12-38):
    the count, and with list_roots the roots themselves."""
    count = ctypes.c_uint64()
    if not list_roots:
        return int(lib().oracle_cube_count(first, last, p, None, 0, ctypes.byref(count))), None
    cap = (last - first) // 2 + 2
    roots = (ctypes.c_uint64 * cap)()
    n = int(lib().oracle_cube_count(first, last, p, roots, cap, ctypes.byref(count)))
    return n, [int(roots[i]) for i in range(min(n, cap))]


def enumerate_small(config: Any, first: int, end: int) -> list[int]:
    """Brute force over the two focus wheels (pseudocube/focus.py:170-179), tiny configs only."""
    af, bf = config.focuses()
    residual = {q: {a for a in range(q) if _allowed(q, a)} for q in config.residuals()}
    us_all = [af.at(i) for i in range(af.count)]
    out = []
    for v in sorted(bf.at(i) for i in range(bf.count)):
        if not first <= v < end:
            continue
        low, high = (config.lo + v * config.A) // config.B + 1, (config.hi + v * config.A) // config.B + 1
        for period in range(low // af.modulus, max(low, high - 1) // af.modulus + 1):
            for u0 in us_all:
                u = period * af.modulus + u0
                if low <= u < high:
                    n = u * config.B - v * config.A
                    if all(n % q in allowed for q, allowed in residual.items()):
                        out.append(n)
    return sorted(out)


def _allowed(q: int, a: int) -> bool:
    if q == 9:
        return a in (1, 8)
    if a == 0:
        return False
    if q % 3 != 1:
        return True
    return pow(a, (q - 1) // 3, q) == 1
