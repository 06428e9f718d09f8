"""prepare() starting the batch's match (pcube_prepare_match), run() collecting it.

With a MaskedMatcher of the configuration alive, prepare() enqueues the
batch's match as soon as its bounds are known and run() collects that match
instead of starting one.  Every receipt must equal the reference lease's
golden batch rows (the same checks as the per-batch golden test), whether
run() collected the started match or started the batch itself: in the plain
prepare-then-run order, when a second batch was prepared while the first was
pending, when another batch ran first, and when a lease used the matcher in
between.
"""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import pytest

from psuedocubes_amd.accounting import encoded
from psuedocubes_amd.batch_experiment import MaskedMatcher, prepare
from psuedocubes_amd.optimized import BATCH_WIDTH, compute_lease

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).parent / 'golden'
LEASE_BATCHES = json.loads((GOLDEN / 'lease_575222931200000000_batches.json').read_text())['batches']


def assert_golden(row: dict, index: int) -> None:
    golden = LEASE_BATCHES[index]
    assert hashlib.sha256(encoded(row['subtile_scopes'])).hexdigest() == golden['scopes_sha256']
    assert row['valid_focused_candidates'] == golden['valid_focused_candidates']
    assert row['group_survivors'] == golden['group_survivors']
    assert row['all_values'] == golden['all_values']
    assert (row['focus_u'], row['focus_v']) == (golden['focus_u'], golden['focus_v'])


def golden_batch(runtime, config, focuses, index: int):
    golden = LEASE_BATCHES[index]
    return prepare(runtime, config, *focuses, int(golden['first']), int(golden['end']))


@pytest.fixture
def counted(runtime, config619):
    """A fresh MaskedMatcher (now the one prepare() starts batches on) whose
    engine counts the batches run() starts itself."""
    matcher = MaskedMatcher(runtime, config619)
    starts = []
    start = matcher.engine.start

    def counting(*args, **kwargs):
        starts.append(args)
        return start(*args, **kwargs)

    matcher.engine.start = counting
    return matcher, starts


@pytest.mark.parametrize('index', [0, 255])
def test_run_collects_the_match_prepare_started(runtime, config619, focuses619, counted, index) -> None:
    matcher, starts = counted
    batch = golden_batch(runtime, config619, focuses619, index)
    assert batch.started is not None and batch.started[0] is matcher.engine
    row = matcher.run(batch)
    assert starts == []
    assert_golden(row, index)
    # collected already: a second run starts the batch itself, same receipt
    again = matcher.run(batch)
    assert len(starts) == 1
    assert_golden(again, index)


def test_second_prepare_while_first_pending(runtime, config619, focuses619, counted) -> None:
    matcher, starts = counted
    first = golden_batch(runtime, config619, focuses619, 1)
    second = golden_batch(runtime, config619, focuses619, 2)
    assert first.started is not None and second.started is None   # one pending batch per matcher
    assert_golden(matcher.run(second), 2)
    assert_golden(matcher.run(first), 1)   # its start was superseded by second's: run again
    assert len(starts) == 2


def test_other_matcher_and_output_cap(runtime, config619, focuses619, counted) -> None:
    matcher, starts = counted
    batch = golden_batch(runtime, config619, focuses619, 3)
    other = MaskedMatcher(runtime, config619)
    assert_golden(other.run(batch), 3)   # a different matcher starts the batch itself
    assert_golden(matcher.run(batch, output_cap=4096), 3)   # another capacity: started again
    assert len(starts) == 1


def test_lease_between_prepare_and_run(runtime, config619, focuses619, counted) -> None:
    matcher, starts = counted
    batch = golden_batch(runtime, config619, focuses619, 4)
    assert batch.started is not None
    first = int(LEASE_BATCHES[0]['first'])
    lease = compute_lease(matcher, *focuses619, first, first + 2 * BATCH_WIDTH)
    assert lease['compute_batches'] == 2
    assert_golden(matcher.run(batch), 4)   # the lease reused the batch buffers: run starts it again
    assert len(starts) == 1


def test_finish_after_lease_is_rejected(runtime, config619, focuses619, counted) -> None:
    matcher, _ = counted
    batch = golden_batch(runtime, config619, focuses619, 5)
    engine = matcher.engine
    engine.start(batch.us, batch.vs, batch.first, batch.width, len(batch.ranges), 8192, matcher.engine_id)
    first = int(LEASE_BATCHES[0]['first'])
    compute_lease(matcher, *focuses619, first, first + BATCH_WIDTH)
    with pytest.raises(ValueError, match='lease'):
        engine.finish()
    assert_golden(matcher.run(batch), 5)
