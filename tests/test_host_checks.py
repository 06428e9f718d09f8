"""Host-side integrity checks that need no device.

* batch_row (`pseudocube/batch_experiment.py:121-167`) fails closed when the
  values handed back disagree with the subtile counters: a counted value that
  is missing, or a value no counter accounts for, is an ArithmeticError, never
  a PASS row with fewer survivors.
* The served loop answers a malformed request (empty or out-of-range v range,
  a prime bound without resident tables) on its own and keeps serving the
  others; here with a stand-in for the device engine, so only the loop's
  admission and bookkeeping run.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from psuedocubes_amd.batch_experiment import batch_row, tile_ranges
from psuedocubes_amd.focus import A619, B619, Config
from psuedocubes_amd.served import MERGED_COUNTS, MERGED_TIMES, Request, Server

GOLDEN = Path(__file__).parent / 'golden'
CONFIG = Config(619, 10**27, 3 * 10**27, A619, B619)
WITNESS = int(json.loads((GOLDEN / 'witness.json').read_text())['n'])


def witness_batch() -> tuple[list, list, np.ndarray, int]:
    """A 3-subtile batch whose middle subtile holds the reference's witness value."""
    v = CONFIG.canonical(WITNESS)[1]
    first = v - 150_000_000
    ranges = tile_ranges(CONFIG, first, first + 300_000_000, 100_000_000)
    bounds = [(0, 10, 0, 10)] * len(ranges)
    ng = len(CONFIG.filter_groups())
    counters = np.zeros((len(ranges), ng + 3), dtype=np.uint64)
    counters[:, 0] = 1000
    counters[:, 1:ng + 1] = 5
    return ranges, bounds, counters, ng


def run_row(ranges, bounds, counters, per_tile):
    return batch_row(CONFIG, CONFIG.filter_groups(), ranges, bounds, 30, counters, per_tile, 8192, {}, 0.0)


def test_batch_row_accepts_consistent_values() -> None:
    ranges, bounds, counters, ng = witness_batch()
    counters[1, ng + 1] = 1
    row = run_row(ranges, bounds, counters, [[], [WITNESS], []])
    assert row['status'] == 'PASS' and row['all_values'] == [str(WITNESS)] and row['local_survivors'] == 1


def test_batch_row_rejects_values_the_counters_disagree_with() -> None:
    ranges, bounds, counters, ng = witness_batch()
    counters[1, ng + 1] = 1
    # the counter reports a value the list lost
    with pytest.raises(ArithmeticError):
        run_row(ranges, bounds, counters, [[], [], []])
    # a value appears that no counter reports
    quiet = counters.copy()
    quiet[1, ng + 1] = 0
    with pytest.raises(ArithmeticError):
        run_row(ranges, bounds, quiet, [[], [WITNESS], []])
    # the value sits under the wrong subtile's count
    with pytest.raises(ArithmeticError):
        run_row(ranges, bounds, counters, [[WITNESS], [], []])
    # fewer subtile lists than subtiles
    with pytest.raises(ArithmeticError):
        run_row(ranges, bounds, counters, [[], [WITNESS]])


class _StandInEngine:
    """The parts of BatchedEngine the served loop calls; rows carry only what merge_rows sums."""

    def __init__(self) -> None:
        self.configs = {619: CONFIG}
        self.runtime = type('R', (), {'cap': 1 << 30})()
        self.launched: list[int] = []

    def footprint(self, window) -> int:
        return 1

    def launch(self, windows):
        self.launched.append(len(windows))
        return windows

    def assemble(self, windows):
        rows = []
        for w in windows:
            row = {k: 0 for k in MERGED_COUNTS + MERGED_TIMES}
            row.update(status='PASS', host_verified=True, focus_v=w.end - w.first,
                       subtile_scopes=[{'first': str(w.first), 'end': str(w.end)}],
                       all_values=[], cpu_confirmed_hits=[], overflow_splits=[], group_survivors=[0] * 24)
            rows.append(row)
        return rows


def test_served_loop_answers_malformed_requests_and_goes_on() -> None:
    good = [Request(i, 10**15 + i * 10**10, 10**15 + i * 10**10 + 1_700_000_000, 0.0) for i in range(5)]
    bad = [Request(100, 50, 50, 0.0), Request(101, B619 - 5, B619 + 5, 0.0), Request(102, 7, 9, 0.0, p=313),
           Request(103, -3, 10, 0.0)]
    server = Server(_StandInEngine(), piece_tiles=8)
    seen = []
    done = server.serve(good[:2] + bad + good[2:], realtime=False, on_complete=seen.append)
    by_id = {c.request.id: c for c in done}
    assert sorted(by_id) == sorted(r.id for r in good + bad) and len(seen) == len(done)
    for request in bad:
        assert by_id[request.id].error and by_id[request.id].row['status'] == 'REJECTED'
    for request in good:
        completed = by_id[request.id]
        assert completed.error is None
        assert completed.row['focus_v'] == request.end - request.first
        assert [s['first'] for s in completed.row['subtile_scopes']][0] == str(request.first)


def test_served_loop_commits_completed_rows_to_the_ledger(tmp_path) -> None:
    """With a ledger, every completed request is coverage of its scope before it is
    reported (one transaction per device step); a rejected request is not; a
    repeated range is recorded as a replay, not as a second unit of coverage."""
    from psuedocubes_amd.accounting import Ledger
    good = [Request(i, 10**15 + i * 10**10, 10**15 + i * 10**10 + 1_700_000_000, 0.0) for i in range(5)]
    repeat = Request(50, good[1].first, good[1].end, 0.0)
    bad = [Request(100, 50, 50, 0.0)]
    ledger = Ledger(tmp_path, {'served': 'p=619 stand-in'}, 'build')
    server = Server(_StandInEngine(), piece_tiles=8, ledger=ledger)
    committed_when_seen = []
    done = server.serve(good + bad + [repeat], realtime=False,
                        on_complete=lambda c: committed_when_seen.append(
                            c.receipt is None or ledger.db.execute('SELECT count(*) FROM receipt_blobs WHERE token=?',
                                                                   (c.receipt,)).fetchone()[0] == 1))
    assert all(committed_when_seen)
    by_id = {c.request.id: c for c in done}
    assert by_id[100].receipt is None and all(by_id[r.id].receipt for r in good + [repeat])
    audit = ledger.audit()
    assert audit['states'] == {'committed': 5}
    assert sorted(r['focus_v'] for r in audit['committed_results']) == [1_700_000_000] * 5
    assert [v['purpose'] for v in audit['validations']] == ['served repeat']
    scopes = sorted(s for (s,) in ledger.db.execute('SELECT scope FROM work'))
    assert [json.loads(s) for s in scopes] == sorted((Server.scope(r) for r in good), key=lambda s: s['first'])
    ledger.db.close()


def test_served_loop_with_a_writer_process_reports_after_commit(tmp_path) -> None:
    """The same through a LedgerWriter: the loop hands each step's rows to the child
    process and reports a request only once the child acknowledged its commit;
    reopening the ledger shows every row as coverage."""
    from psuedocubes_amd.accounting import Ledger, LedgerWriter
    good = [Request(i, 10**15 + i * 10**10, 10**15 + i * 10**10 + 1_700_000_000, 0.0) for i in range(40)]
    config = {'served': 'p=619 stand-in'}
    with LedgerWriter(tmp_path, config, 'build') as writer:
        done = Server(_StandInEngine(), piece_tiles=8, ledger=writer).serve(good, realtime=False)
        assert not writer.pending
    assert sorted(c.request.id for c in done) == list(range(40)) and all(c.receipt for c in done)
    ledger = Ledger(tmp_path, config, 'build')
    audit = ledger.audit()
    assert audit['states'] == {'committed': 40}
    tokens = {row[0] for row in ledger.db.execute("SELECT token FROM work WHERE state='committed'")}
    assert tokens == {c.receipt for c in done}
    ledger.db.close()
    # a writer that cannot open its ledger fails at once
    with pytest.raises(ValueError):
        LedgerWriter(tmp_path, {'another': 'configuration'}, 'build')
