"""Served form: continuous batching of range requests.

Every request's merged row must equal the plain drop-in's row for its whole
range and the row it gets when served alone, while steps mix requests that
arrived at different times.
"""
from __future__ import annotations

import pytest

from psuedocubes_amd.batch_experiment import MaskedMatcher, compare, prepare
from psuedocubes_amd.batched import BatchedEngine
from psuedocubes_amd.focus import A619, B619, Config
from psuedocubes_amd.served import Request, Server, latency_summary, merge_rows, request_stream

pytestmark = pytest.mark.gpu


def strip(row: dict) -> dict:
    return {k: v for k, v in row.items() if not k.endswith('_seconds') and k not in
            ('matching_layout', 'kernel_registers', 'kernel_local_bytes', 'claim')}


@pytest.fixture(scope='module')
def engine(runtime, focuses619):
    return BatchedEngine(runtime, [Config(619, 10**27, 3 * 10**27, A619, B619)], *focuses619)


def test_served_results_same_alone_in_mix_and_plain(runtime, config619, focuses619, engine) -> None:
    requests = request_stream(config619, 24, 619021, rate=400.0)
    requests.append(Request(24, B619 - 6_400_000_000, B619, 0.01))
    requests.append(Request(25, 575222931200000000, 575222931200000000 + 3 * 100_000_000 + 7, 0.02))
    server = Server(engine, piece_tiles=4)
    mixed = {c.request.id: c for c in server.serve(requests)}
    assert sorted(mixed) == [r.id for r in requests]
    assert max(server.step_sizes) > 1          # steps did mix requests
    plain = MaskedMatcher(runtime, config619)
    for request in requests:
        row = mixed[request.id].row
        expected = plain.run(prepare(runtime, config619, *focuses619, request.first, request.end))
        compare(expected, row)
        assert strip(expected) == strip(row), request.id
    for request in requests[:4] + requests[-2:]:
        alone = Server(engine, piece_tiles=4).serve([Request(request.id, request.first, request.end, 0.0)])
        assert strip(alone[0].row) == strip(mixed[request.id].row)
    summary = latency_summary(list(mixed.values()))
    assert 0 < summary['p50'] <= summary['p99']


def test_admission_respects_slots_and_budget(runtime, config619, engine) -> None:
    requests = request_stream(config619, 12, 7, rate=0.0)
    tight = Server(engine, max_slots=2, piece_tiles=2)
    done = tight.serve(requests)
    assert len(done) == 12 and max(tight.step_sizes) <= 2
    small = Server(engine, budget=1, piece_tiles=2)     # the budget admits one request at a time
    assert len(small.serve(requests[:3])) == 3
    assert max(small.step_sizes) == 1


def test_merge_rows_of_pieces_equals_whole(runtime, config619, focuses619) -> None:
    plain = MaskedMatcher(runtime, config619)
    first = 575222931200000000
    whole = plain.run(prepare(runtime, config619, *focuses619, first, first + 1_200_000_000))
    parts = [plain.run(prepare(runtime, config619, *focuses619, s, s + 400_000_000))
             for s in range(first, first + 1_200_000_000, 400_000_000)]
    assert strip(merge_rows(parts)) == strip(whole)


@pytest.mark.parametrize('writer', [False, True])
def test_served_rows_are_committed_to_the_ledger(runtime, config619, engine, tmp_path, writer) -> None:
    """Each completed request's row reaches the ledger as coverage of its scope before
    it is reported, in the serving process or through a writer process, and the
    receipt decodes to the same canonical row."""
    from psuedocubes_amd.accounting import Ledger, LedgerWriter, encoded
    requests = request_stream(config619, 16, 619023, rate=0.0)
    identity = {'served': config619.p, 'lo': str(config619.lo), 'hi': str(config619.hi)}
    if writer:
        with LedgerWriter(tmp_path, identity, 'test-build') as child:
            done = Server(engine, piece_tiles=4, ledger=child).serve(requests, realtime=False)
        ledger = Ledger(tmp_path, identity, 'test-build')
    else:
        ledger = Ledger(tmp_path, identity, 'test-build')
        done = Server(engine, piece_tiles=4, ledger=ledger).serve(requests, realtime=False)
    audit = ledger.audit()
    assert audit['states'] == {'committed': len(requests)}
    by_token = {}
    for token, sha in ledger.db.execute("SELECT token,sha FROM work WHERE state='committed'"):
        by_token[token] = ledger.open_receipt('sqlite:' + token, token, sha)['result']
    for completed in done:
        assert encoded(by_token[completed.receipt]) == encoded(completed.row)
    ledger.db.close()
