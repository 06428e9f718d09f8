"""The batched form's host assembly against the reference's row path, on the CPU.

BatchedEngine.assemble verifies a whole device pass's values in one native call
(pcube_batch_values) and builds each window's row from the verified strings.
Here a pass is built by the oracle instead of the device (counters, bounds and
emitted values of 48 windows of the seeded mixed-prime stream, p in
{313, 433, 619}, including short windows), and every row must equal, field for
field, the row batch_row builds from the same counters and values with the
pure-Python predicate (`pseudocube/batch_experiment.py:121-167`,
`pseudocube/oracle.py:36-45`).  Planted faults in the values (a changed value,
a duplicate, a value under the wrong subtile, a value outside (lo, hi]) must
raise ArithmeticError.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from psuedocubes_amd.batch_experiment import batch_row, python_witnesses, tile_ranges
from psuedocubes_amd.batched import BatchedEngine, window_stream
from psuedocubes_amd.focus import A619, B619, Config

PRIMES = (313, 433, 619)
CONFIGS = {p: Config(p, 10**27, 3 * 10**27, A619, B619) for p in PRIMES}
CAP = 8192


class _Tables:
    def __init__(self, config: Config) -> None:
        self.filter_groups = config.filter_groups()
        self.ng = len(self.filter_groups)


def engine_without_device() -> BatchedEngine:
    """A BatchedEngine with only what assemble() reads (no device tables)."""
    engine = object.__new__(BatchedEngine)
    engine.configs = CONFIGS
    engine.base = CONFIGS[619]
    engine.cap = CAP
    engine.engines = {p: _Tables(c) for p, c in CONFIGS.items()}
    return engine


@pytest.fixture(scope='module')
def oracle_pass(oracle_focuses619):
    """(raw pass as BatchedEngine.launch returns it, per-window per-subtile values)."""
    fa, fb = oracle_focuses619
    windows = window_stream(CONFIGS[619], 4096, 619003, (1e6, 1e9), PRIMES)[:48]
    plans, bounds, sizes, counters, limbs, per_window = [], [], [], [], [], []
    for w in windows:
        config = CONFIGS[w.p]
        ranges = tile_ranges(config, w.first, w.end, w.width)
        us, vs = fa.window(ranges[0][2], ranges[-1][3]), fb.window(w.first, w.end)
        for v0, v1, u0, u1 in ranges:
            a, b = np.searchsorted(us, np.array([u0, u1], dtype=np.uint64))
            c, d = np.searchsorted(vs, np.array([v0, v1], dtype=np.uint64))
            bounds.append((int(a), int(b), int(c), int(d)))
        cnt, val = O.match(config, us, vs, w.first, w.width, len(ranges), CAP)
        ng = cnt.shape[1] - 3
        counters.append(cnt.reshape(-1))
        tiles = []
        for t in range(len(ranges)):
            n = int(cnt[t, ng + 1])
            # the device lists a subtile's values in no particular order
            got = val[t, :n][::-1]
            limbs.append(got.reshape(-1))
            tiles.append([int(lo) + (int(hi) << 64) for lo, hi in got.tolist()])
        per_window.append(tiles)
        sizes += [len(us), len(vs)]
        plans.append((ranges, 0))
    raw = dict(windows=windows, plans=plans, bounds=np.array(bounds, dtype=np.uint64),
               sizes=np.array(sizes, dtype=np.uint64), counters=np.concatenate(counters),
               values=np.concatenate(limbs + [np.zeros(2, dtype=np.uint64)]),
               timings=dict(kernel_seconds=0.0, focus_enumeration_seconds=0.0, transfer_seconds=0.0))
    return raw, per_window


def strip(row: dict) -> dict:
    return {k: v for k, v in row.items() if not k.endswith('_seconds') and k not in ('matching_layout',)}


def test_native_pass_rows_equal_python_rows(oracle_pass) -> None:
    raw, per_window = oracle_pass
    rows = engine_without_device().assemble(raw)
    emitted = 0
    t0 = 0
    offset = 0
    for k, (window, (ranges, _), tiles) in enumerate(zip(raw['windows'], raw['plans'], per_window)):
        config = CONFIGS[window.p]
        ng = len(config.filter_groups())
        nt = len(ranges)
        block = raw['counters'][offset:offset + nt * (ng + 3)].reshape(nt, ng + 3)
        offset += nt * (ng + 3)
        bounds = [tuple(r) for r in raw['bounds'][t0:t0 + nt].tolist()]
        t0 += nt
        expected = batch_row(config, config.filter_groups(), ranges, bounds, int(raw['sizes'][2 * k + 1]), block,
                             tiles, CAP, raw['timings'], 0.0, witnesses=python_witnesses)
        assert strip(rows[k]) == strip(expected), window
        emitted += expected['local_survivors']
    assert emitted > 500   # the p = 313 windows emit hundreds of values: the accepting path ran


def planted(raw: dict, change) -> dict:
    values = raw['values'].copy()
    change(values)
    return dict(raw, values=values)


def test_native_pass_fails_closed_on_planted_faults(oracle_pass) -> None:
    raw, per_window = oracle_pass
    engine = engine_without_device()
    # positions of the first subtile with two or more values in the pass's value array
    counts = [len(t) for tiles in per_window for t in tiles]
    start = 0
    for n in counts:
        if n >= 2:
            break
        start += n

    def change_value(v):
        v[2 * start] ^= np.uint64(1 << 20)

    def duplicate(v):
        v[2 * start + 2:2 * start + 4] = v[2 * start:2 * start + 2]

    def outside(v):
        v[2 * start + 1] = np.uint64(1 << 40)   # far above hi = 3 * 10^27 < 2^92

    for change in (change_value, duplicate, outside):
        with pytest.raises(ArithmeticError):
            engine.assemble(planted(raw, change))
    # a value counted under the next subtile lands in the wrong subtile
    with pytest.raises(ArithmeticError):
        engine.verify_values(raw['windows'], raw['plans'], shifted_counts(engine, raw), raw['values'])


def shifted_counts(engine: BatchedEngine, raw: dict) -> list[np.ndarray]:
    """Per-window kept counts with one value of the first subtile that has values
    counted under the subtile after it."""
    kept = []
    offset = 0
    for window, (ranges, _) in zip(raw['windows'], raw['plans']):
        ng = engine.engines[window.p].ng
        nt = len(ranges)
        block = raw['counters'][offset:offset + nt * (ng + 3)].reshape(nt, ng + 3)
        kept.append(np.minimum(block[:, ng + 1], CAP).astype(np.uint64))
        offset += nt * (ng + 3)
    for k in kept:
        hit = np.flatnonzero(k[:-1])
        if hit.size:
            i = hit[-1]
            k[i] -= 1
            k[i + 1] += 1
            return kept
    raise AssertionError('no subtile with values before another subtile')
