"""The pipelined lease's fold of batches that emitted nothing.

`compute_lease` folds such a batch from its totals (`QuietBatch`, from
`pass_totals` over a device pass's counters and bounds) instead of building
its row with `batch_row` and folding that (`pseudocube/optimized.py:40-61`).
Both must give the same receipt: checked here on synthetic passes (random
bounds and counters, ragged last batch), with the row path as the reference.
The GPU lease tests then compare whole leases with the per-batch loop and the
reference lease's golden transcript.
"""
from __future__ import annotations

import random

import numpy as np
import pytest

from psuedocubes_amd.accounting import encoded
from psuedocubes_amd.batch_experiment import batch_row, tile_ranges
from psuedocubes_amd.focus import A619, B619, Config
from psuedocubes_amd.optimized import (BATCH_WIDTH, PROFILE, Accumulator, QuietBatch, pass_totals)

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
        totals = QuietBatch(nt, (ranges[b][0][0], ranges[b][-1][1]), focus[b], sizes[b], pairs[b], sums[b][0],
                            sums[b][1:], filters, {})
        by_totals.add_quiet(s, e, totals, records)
        t0 += nt
    strip = lambda r: {k: v for k, v in r.items() if not k.endswith('_seconds')}  # noqa: E731
    assert strip(by_rows.finish(0.0)) == strip(by_totals.finish(0.0))
    # a geometry the row path would refuse is refused here too
    with pytest.raises(ValueError):
        Accumulator(FIRST, FIRST + BATCH_WIDTH).add_quiet(
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
