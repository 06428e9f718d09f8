"""The pipelined lease's fold from totals.

`compute_lease` folds each batch of a device pass from its totals
(`BatchTotals`: `pass_totals` over the pass's counters and bounds, the pass's
values checked in one native call, its records formatted in one call) instead
of building its row with `batch_row` and folding that
(`pseudocube/optimized.py:40-61`).  Both must give the same receipt: checked
here on synthetic passes (random bounds and counters, ragged last batch, and a
pass with planted local3 values and genuine cubes), with the row path and the
Python predicate as the reference.  The GPU lease tests then compare whole
leases with the per-batch loop and the reference lease's golden transcript.
"""
from __future__ import annotations

import random

import numpy as np
import pytest

from psuedocubes_amd.accounting import encoded
from psuedocubes_amd.batch_experiment import batch_row, python_witnesses, tile_ranges
from psuedocubes_amd.focus import A619, B619, Config
from psuedocubes_amd.optimized import (BATCH_WIDTH, PROFILE, Accumulator, BatchTotals, pass_rows, pass_sizes,
                                       pass_totals)
from psuedocubes_amd.predicate import local3

FIRST = 575222931200000000
NG = 24


def synthetic_pass(rng: random.Random, config: Config, nbatches: int, tail: int):
    width = PROFILE['tile_width']
    batches = [(FIRST + b * BATCH_WIDTH, FIRST + (b + 1) * BATCH_WIDTH) for b in range(nbatches)]
    batches[-1] = (batches[-1][0], batches[-1][0] + tail)
    ranges = [tile_ranges(config, s, e, width) for s, e in batches]
    nt = sum(len(r) for r in ranges)
    bounds = np.zeros((nt, 4), dtype=np.uint64)
    for k in range(nt):
        a, c = rng.getrandbits(30), rng.getrandbits(30)
        bounds[k] = (a, a + rng.randrange(0, 20000), c, c + rng.randrange(0, 1100))
    counters = np.zeros((nt, NG + 3), dtype=np.uint64)
    counters[:, 0] = [rng.getrandbits(25) for _ in range(nt)]
    counters[:, 1:NG + 1] = np.array([[rng.getrandbits(20) for _ in range(NG)] for _ in range(nt)], dtype=np.uint64)
    sizes = [rng.getrandbits(17) for _ in batches]
    return batches, ranges, bounds, counters, sizes


def test_quiet_fold_equals_row_fold() -> None:
    rng = random.Random(4301)
    config = Config(619, 10**27, 3 * 10**27, A619, B619)
    filters = [[q] for q in range(NG)]
    batches, ranges, bounds, counters, sizes = synthetic_pass(rng, config, 5, 250_000_001)
    per_batch = [len(r) for r in ranges]
    quiet, focus, pairs, sums = pass_totals(per_batch, bounds, counters, NG)
    assert quiet == [True] * 5
    by_rows, by_totals = Accumulator(FIRST, batches[-1][1]), Accumulator(FIRST, batches[-1][1])
    t0 = 0
    for b, ((s, e), nt) in enumerate(zip(batches, per_batch)):
        block = counters[t0:t0 + nt]
        bnd = [tuple(r) for r in bounds[t0:t0 + nt].tolist()]
        row = batch_row(config, filters, ranges[b], bnd, sizes[b], block, [[] for _ in range(nt)], 8192, {}, 0.0)
        assert (row['focus_u'], row['cartesian_opportunities']) == (focus[b], pairs[b])
        assert [row['valid_focused_candidates']] + row['group_survivors'] == sums[b]
        records = encoded(row['subtile_scopes'])
        by_rows.add(s, e, row, records)
        totals = BatchTotals(nt, (ranges[b][0][0], ranges[b][-1][1]), focus[b], sizes[b], pairs[b], sums[b][0],
                            sums[b][1:], filters, {})
        by_totals.add_totals(s, e, totals, records)
        t0 += nt
    strip = lambda r: {k: v for k, v in r.items() if not k.endswith('_seconds')}  # noqa: E731
    assert strip(by_rows.finish(0.0)) == strip(by_totals.finish(0.0))
    # a geometry the row path would refuse is refused here too
    with pytest.raises(ValueError):
        Accumulator(FIRST, FIRST + BATCH_WIDTH).add_totals(
            FIRST, FIRST + BATCH_WIDTH, totals._replace(span=(FIRST, FIRST + BATCH_WIDTH - 1)), records)


def test_pass_totals_marks_emitting_batches() -> None:
    rng = random.Random(4302)
    config = Config(619, 10**27, 3 * 10**27, A619, B619)
    batches, ranges, bounds, counters, sizes = synthetic_pass(rng, config, 4, BATCH_WIDTH)
    per_batch = [len(r) for r in ranges]
    counters[64 + 17, NG + 1] = 1      # batch 1 emitted a value
    counters[3 * 64 + 63, NG + 2] = 1  # batch 3 saw an underflow
    assert pass_totals(per_batch, bounds, counters, NG)[0] == [True, False, True, False]
    bounds[5, 1] = bounds[5, 0] - 1    # a reversed pair: nothing is folded from totals
    assert pass_totals(per_batch, bounds, counters, NG)[0] == [False] * 4


# p = 37 with the lease's batches inside [0, B): B = 9 5 7 ... 31 (about 3.0e11), A = 2,
# 37 the one residual prime
SMALL = Config(37, 10**18, 2 * 10**18, 2, 9 * 5 * 7 * 11 * 13 * 17 * 19 * 23 * 29 * 31)
P = SMALL.p


class _Matcher:
    config = SMALL
    filter_groups = SMALL.filter_groups()
    kernel = dict(registers=96, local_bytes=0)
    layout = 'synthetic'


def emitting_pass(rng: random.Random):
    """A pass of two full batches and a ragged one over v in [0, 12.8e9 + 250000001),
    with local3 values at p = 37 (noncubes and genuine cubes) in some subtiles,
    each listed in the subtile holding its canonical v, in device (unsorted) order."""
    config = SMALL
    ng = len(config.filter_groups())
    width = PROFILE['tile_width']
    batches = [(0, BATCH_WIDTH), (BATCH_WIDTH, 2 * BATCH_WIDTH), (2 * BATCH_WIDTH, 2 * BATCH_WIDTH + 250_000_001)]
    ranges = [tile_ranges(config, s, e, width) for s, e in batches]
    flat = [r for rs in ranges for r in rs]
    end = batches[-1][1]
    picked = set()
    while len(picked) < 90:
        # a position (u, v) of the lease, kept when x = uB - vA passes local3
        v = rng.randrange(0, end)
        u = rng.randrange(*config.u_interval(v))
        n = u * config.B - v * config.A
        if local3(n, P):
            picked.add(n)
    cubes = 0
    for r in range(1_000_001, 1_259_921):
        if r ** 3 <= config.hi and config.canonical(r ** 3)[1] < end and local3(r ** 3, P):
            picked.add(r ** 3)
            cubes += 1
            if cubes == 18:
                break
    # every subtile of batch 1 empty: a quiet batch between emitting ones
    picked = {n for n in picked if not BATCH_WIDTH <= config.canonical(n)[1] < 2 * BATCH_WIDTH}
    per_tile = [[] for _ in flat]
    for n in sorted(picked):
        v = config.canonical(n)[1]
        per_tile[v // width].append(n)
    for values in per_tile:
        rng.shuffle(values)
    nt = len(flat)
    bounds = np.zeros((nt, 4), dtype=np.uint64)
    for k in range(nt):
        a, c = rng.getrandbits(30), rng.getrandbits(30)
        bounds[k] = (a, a + rng.randrange(0, 20000), c, c + rng.randrange(0, 1100))
    counters = np.zeros((nt, ng + 3), dtype=np.uint64)
    counters[:, 0] = [rng.getrandbits(25) + 100 for _ in range(nt)]
    counters[:, 1:ng + 1] = np.array([[rng.getrandbits(20) for _ in range(ng)] for _ in range(nt)],
                                     dtype=np.uint64).reshape(nt, ng)
    counters[:, ng + 1] = [len(v) for v in per_tile]
    limbs = [x for values in per_tile for n in values for x in (n & (2**64 - 1), n >> 64)]
    values = np.array(limbs + [0, 0], dtype=np.uint64)
    sizes = np.array([x for _ in batches for x in (rng.getrandbits(12), rng.getrandbits(17))], dtype=np.uint64)
    tiles = np.array(flat, dtype=np.uint64).reshape(-1)
    counts = [len(r) for r in ranges]
    arrays = (tiles, np.array(counts, dtype=np.uint32), bounds, sizes, counters, values, np.zeros(3))
    return batches, ranges, counts, arrays, per_tile


def row_fold(batches, ranges, counts, arrays, per_tile) -> dict:
    _, _, bounds, sizes, counters, _, _ = arrays
    acc = Accumulator(batches[0][0], batches[-1][1])
    t0 = 0
    for b, (s, e) in enumerate(batches):
        nt = counts[b]
        bnd = [tuple(r) for r in bounds[t0:t0 + nt].tolist()]
        row = batch_row(SMALL, _Matcher.filter_groups, ranges[b], bnd, int(sizes[2 * b + 1]), counters[t0:t0 + nt],
                        per_tile[t0:t0 + nt], 8192, {}, 0.0, witnesses=python_witnesses)
        acc.add(s, e, row, encoded(row['subtile_scopes']))
        t0 += nt
    return acc.finish(0.0)


def pass_fold(batches, counts, arrays, quiet: bool) -> tuple[dict, list[str]]:
    acc = Accumulator(batches[0][0], batches[-1][1])
    kinds = []
    for s, e, row, records in pass_rows(_Matcher, batches, counts, arrays, 0.0, quiet):
        if isinstance(row, BatchTotals):
            acc.add_totals(s, e, row, records)
            kinds.append('totals')
        else:
            assert records == encoded(row['subtile_scopes'])
            acc.add(s, e, row, records)
            kinds.append('row')
    return acc.finish(0.0), kinds


def test_pass_fold_with_values_equals_row_fold() -> None:
    batches, ranges, counts, arrays, per_tile = emitting_pass(random.Random(4303))
    expected = row_fold(batches, ranges, counts, arrays, per_tile)
    assert expected['local_survivors'] >= 45 and expected['genuine_cubes'] >= 8
    assert len(expected['cpu_confirmed_hits']) == expected['local_survivors'] - expected['genuine_cubes']
    strip = lambda r: {k: v for k, v in r.items() if not k.endswith('_seconds')}  # noqa: E731
    by_totals, kinds = pass_fold(batches, counts, arrays, quiet=True)
    assert kinds == ['totals'] * 3
    assert strip(by_totals) == strip(expected)
    by_rows, kinds = pass_fold(batches, counts, arrays, quiet=False)
    assert kinds == ['row'] * 3
    assert strip(by_rows) == strip(expected)


def test_pass_fold_raises_at_the_failing_batch() -> None:
    """A value that fails local3 in the last batch: the pass-level check refuses
    the pass, the per-batch path yields the rows before that batch (equal to the
    row path's) and raises ArithmeticError at it; an underflow raises too."""
    batches, ranges, counts, arrays, per_tile = emitting_pass(random.Random(4304))
    tiles, per_batch, bounds, sizes, counters, values, seconds = arrays
    last = max(k for k, v in enumerate(per_tile) if v)
    assert last >= counts[0] + counts[1]
    n = per_tile[last][0]
    bad = n + SMALL.B   # the same canonical v; n + B is even, so not local3
    assert not local3(bad, P)
    position = sum(len(v) for v in per_tile[:last])
    planted = values.copy()
    planted[2 * position], planted[2 * position + 1] = bad & (2**64 - 1), bad >> 64
    seen = []
    with pytest.raises(ArithmeticError):
        for s, e, row, records in pass_rows(_Matcher, batches, counts, (tiles, per_batch, bounds, sizes, counters,
                                                                        planted, seconds), 0.0, True):
            seen.append((s, e))
    assert seen == batches[:2]
    under = counters.copy()
    under[3, -1] = 1
    with pytest.raises(ArithmeticError):
        list(pass_rows(_Matcher, batches, counts, (tiles, per_batch, bounds, sizes, under, values, seconds), 0.0, True))


def test_pass_sizes_cover_the_lease() -> None:
    """Every lease length splits into passes of 1..64 batches covering it in order:
    a first pass of 8, passes of 64, and a tail halving down to 8."""
    assert pass_sizes(256) == [8, 64, 64, 64, 32, 16, 8]
    assert pass_sizes(4) == [4] and pass_sizes(16) == [8, 8]
    for n in range(1, 257):
        sizes = pass_sizes(n)
        assert sum(sizes) == n and all(1 <= s <= 64 for s in sizes)
