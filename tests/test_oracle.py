"""Pins the CPU oracle before anything is checked against it.

* windows: brute-force multisets with wraps, multi-period windows and
  overflow (recipe of the reference's tests/test_pseudocube_batch_experiment.py:18-34),
  plus the reference's own compiled review/focus_window_cpu.cpp (oracle/_ref)
* partial sums: Focus.at on sampled mixed-radix indices
* matcher: complete sets against the local3 brute force for p in {7, 13, 19, 31}
  (tests/test_pseudocube_math.py:74-88 of the reference), partition additivity
* the whole 256-batch positive lease against the reference's fixture
  (results/continuation-2026-09-15/positive-lease-replay.json): output hash,
  transcript hash (which covers every per-subtile population and output hash),
  every count and group population
"""
from __future__ import annotations

import ctypes
import json
from math import prod
from pathlib import Path
import random

import numpy as np
import pytest

from oracle import oracle as O
from psuedocubes_amd.focus import Config
from psuedocubes_amd.predicate import local3

GOLDEN = Path(__file__).parent / 'golden'


def small_config(p: int, lo: int = 0, hi: int = 100000) -> Config:
    return Config(p, lo, hi, 14 * (19 if p >= 19 else 1), 9 * (13 if p >= 13 else 1) * (31 if p >= 31 else 1))


def reference_producer():
    path = O.HERE / '_ref' / 'focus_window.so'
    if not path.exists():
        return None
    lib = ctypes.CDLL(str(path))
    ptr = np.ctypeslib.ndpointer(dtype=np.uint64, ndim=1, flags='C_CONTIGUOUS')
    fn = lib.focus_window
    fn.argtypes = [ptr, ctypes.c_uint64, ptr, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                   ctypes.c_uint64, ctypes.c_uint64, ptr, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    fn.restype = ctypes.c_int
    return fn


def test_windows_match_brute_force_and_reference_producer() -> None:
    rng = random.Random(61964)
    ref = reference_producer()
    for modulus in (7, 113, 701856356111039402):
        c = sorted({rng.randrange(modulus) for _ in range(30)} | {0, modulus - 1})
        d = sorted({rng.randrange(modulus) for _ in range(30)} | {0, modulus - 1})
        C, D = np.array(c, dtype=np.uint64), np.array(d, dtype=np.uint64)
        for lo, hi in [(0, modulus), (1, modulus - 1), (0, 0), (modulus - 1, modulus + 1),
                       (2 * modulus + 1, 4 * modulus - 1)]:
            out = np.empty(4096, dtype=np.uint64)
            count = ctypes.c_uint64()
            assert O.lib().oracle_window(C, len(C), D, len(D), modulus, lo, hi, out, 4096, ctypes.byref(count)) == 0
            got = out[:count.value].tolist()
            assert got == O.brute_window(c, d, modulus, lo, hi)
            if ref is not None and hi > lo:
                pieces = []
                for period in range(lo // modulus, (hi - 1) // modulus + 1):
                    base = period * modulus
                    buf = np.empty(4096, dtype=np.uint64)
                    n = ctypes.c_uint64()
                    assert ref(C, len(C), D, len(D), modulus, max(0, lo - base), min(modulus, hi - base), base,
                               buf, 4096, ctypes.byref(n)) == 0
                    pieces.extend(buf[:n.value].tolist())
                assert sorted(pieces) == got
        out = np.empty(1, dtype=np.uint64)
        count = ctypes.c_uint64()
        assert O.lib().oracle_window(C, len(C), D, len(D), modulus, 0, modulus, out, 1, ctypes.byref(count)) == 1


def test_partial_sums_follow_focus_at(config619: Config) -> None:
    rng = random.Random(7)
    for focus in config619.focuses():
        first, second = focus.halves()
        for rings in (first, second):
            arr = O.partial(rings, focus.modulus)
            assert arr.size == prod(len(r) for r in rings)
            for i in [0, arr.size - 1] + [rng.randrange(arr.size) for _ in range(200)]:
                value, rest = 0, i
                for ring in rings:
                    rest, digit = divmod(rest, len(ring))
                    value = (value + ring[digit]) % focus.modulus
                assert int(arr[i]) == value


def test_complete_small_sets_and_partitions() -> None:
    for p in (7, 13, 19, 31):
        config = small_config(p)
        fa, fb = (O.Focus(f) for f in config.focuses())
        width = max(1, (config.B + 6) // 7)
        values = []
        for start in range(0, config.B, width * 64):
            values += [int(x) for x in O.batch(config, fa, fb, start, min(config.B, start + width * 64), width)['all_values']]
        expected = [n for n in range(1, 100001) if local3(n, p)]
        assert sorted(values) == expected and len(values) == len(set(values))
        # partition additivity: a different subtile width gives the same set and populations
        other = O.batch(config, fa, fb, 0, config.B, max(1, (config.B + 63) // 64))
        assert sorted(int(x) for x in other['all_values']) == expected


def test_narrow_interval_endpoints() -> None:
    # lo is excluded and hi included: (n-1, n] holds n, (n, n+100] does not
    for n in (1, 125, 343, 1001):
        if not local3(n, 7):
            continue
        inside = small_config(7, 0, n)
        outside = small_config(7, n, n + 100)
        for config, expect in ((inside, True), (outside, False)):
            fa, fb = (O.Focus(f) for f in config.focuses())
            got = O.batch(config, fa, fb, 0, config.B, 1)['all_values']
            assert (str(n) in got) == expect


@pytest.mark.slow
def test_oracle_reproduces_reference_lease(config619: Config, oracle_focuses619) -> None:
    fixture = json.loads((GOLDEN / 'lease_575222931200000000.json').read_text())
    batches = json.loads((GOLDEN / 'lease_575222931200000000_batches.json').read_text())['batches']
    seen = []
    import hashlib

    def keep(start: int, stop: int, row: dict, enc: bytes) -> None:
        seen.append(hashlib.sha256(enc).hexdigest())

    fa, fb = oracle_focuses619
    result = O.lease(config619, fa, fb, int(fixture['first']), int(fixture['end']), on_batch=keep)
    ref = fixture['result']
    for key in ('focus_u', 'focus_v', 'cartesian_opportunities', 'valid_focused_candidates', 'local_survivors',
                'group_survivors', 'all_values'):
        assert result[key] == ref[key], key
    assert result['output_sha256'] == ref['subtile_scopes'][0]['output_sha256']
    assert result['transcript_sha256'] == ref['subtile_scopes'][0]['transcript_sha256']
    assert seen == [b['scopes_sha256'] for b in batches]


def test_prime_bound_nesting_invariants(config619: Config, oracle_focuses619) -> None:
    """The oracle's p-generic path at p = 313 and 433, pinned against p = 619 (which
    the reference lease pins) by what the mathematics fixes for one window:

    * A, B and the focus wheels are the same for every p, so the focused
      positions (focus_u, focus_v, opportunities, valid candidates) are equal;
    * the first eight residual filters are 193 ... 277 for every p, so group
      populations 1 and 2, per subtile, are equal;
    * the survivors at 433 are exactly the survivors at 313 that pass the
      residual filters q in (313, 433] (x mod q an allowed residue), and the
      survivors at 619 those of 433 that pass q in (433, 619];
    * every survivor satisfies local3 at its own p.
    Windows: the reference lease's batches 0 and 34 (one and three survivors at
    p = 619) and the first windows of the seeded mixed-prime stream."""
    from psuedocubes_amd.batched import window_stream
    from psuedocubes_amd.focus import A619, B619
    from psuedocubes_amd.tables import allowed
    configs = {p: Config(p, 10**27, 3 * 10**27, A619, B619) for p in (313, 433, 619)}
    fa, fb = oracle_focuses619
    batches = json.loads((GOLDEN / 'lease_575222931200000000_batches.json').read_text())['batches']
    windows = [(int(batches[i]['first']), int(batches[i]['end'])) for i in (0, 34)]
    windows += [(w.first, w.end) for w in window_stream(config619, 4096, 619003, (1e6, 1e9), (313, 433, 619))[:6]]
    found = {313: 0, 433: 0, 619: 0}
    for first, end in windows:
        rows = {p: O.batch(configs[p], fa, fb, first, end) for p in configs}
        for key in ('focus_u', 'focus_v', 'cartesian_opportunities', 'valid_focused_candidates'):
            assert rows[313][key] == rows[433][key] == rows[619][key], (first, key)
        for p in (313, 433):
            for a, b in zip(rows[p]['subtile_scopes'], rows[619]['subtile_scopes']):
                assert a['valid_focused_candidates'] == b['valid_focused_candidates']
                assert a['group_survivors'][:2] == b['group_survivors'][:2]
        survivors = {p: [int(x) for x in rows[p]['all_values']] for p in configs}
        for small, large in ((313, 433), (433, 619)):
            extra = [q for q in configs[large].residuals() if small < q <= large]
            passing = [x for x in survivors[small] if all(x % q in allowed(q) for q in extra)]
            assert passing == survivors[large], (first, small, large)
        for p in configs:
            assert all(local3(x, p) for x in survivors[p][:200])
            found[p] += len(survivors[p])
    assert found[619] >= 4 and found[313] > found[433] >= found[619]
