"""A lease from the device to the shared ledger.

`multi.lease_worker` takes the reference lease's range from a SharedFrontier
whose baseline sits at its first v, computes it with the pipelined lease engine
and commits its compact receipt as a gzip blob.  The committed result must be
the golden lease (`tests/golden/lease_575222931200000000.json`: counts, values,
hits and the transcript), its noncube hit must stop the campaign and land in
the candidates table, and the audit must replay the receipt to the same totals.
"""
from __future__ import annotations

import gzip
import json
from pathlib import Path

import pytest

from psuedocubes_amd.accounting import encoded
from psuedocubes_amd.batch_experiment import MaskedMatcher
from psuedocubes_amd.multi import SharedFrontier, lease_worker
from psuedocubes_amd.optimized import LEASE_WIDTH, PROFILE

pytestmark = pytest.mark.gpu

LEASE = json.loads((Path(__file__).parent / 'golden' / 'lease_575222931200000000.json').read_text())
FIRST = int(LEASE['first'])


def test_lease_worker_commits_the_reference_lease(runtime, config619, focuses619, tmp_path) -> None:
    assert int(LEASE['end']) == FIRST + LEASE_WIDTH
    baseline = dict(cursor=str(FIRST), commits='0', candidates='0', cubes='0', subtiles='0',
                    campaign_state='searching')
    db = SharedFrontier(tmp_path, {'p': 619, 'lo': str(config619.lo), 'hi': str(config619.hi)}, 'test-build',
                        FIRST + LEASE_WIDTH, LEASE_WIDTH, ('gpu0',), baseline, PROFILE)
    done = lease_worker(db, 'gpu0', MaskedMatcher(runtime, config619), *focuses619, leases=2)
    assert len(done) == 1
    result, golden = done[0], LEASE['result']
    for key in ('valid_focused_candidates', 'genuine_cubes', 'subtiles_executed', 'group_survivors', 'all_values',
                'cpu_confirmed_hits', 'subtile_scopes'):
        assert result[key] == golden[key], key
    state = db.state()
    assert (state['cursor'], state['campaign_state']) == (LEASE['end'], 'candidate-found')
    hits = [row[0] for row in db.db.execute('SELECT n FROM candidates')]
    assert hits == [h['n'] for h in golden['cpu_confirmed_hits']]
    audit = db.audit_shared()
    assert audit['totals'] == dict(commits=1, candidates=golden['valid_focused_candidates'],
                                   cubes=golden['genuine_cubes'], subtiles=golden['subtiles_executed'])
    token, receipt = db.db.execute("SELECT token,receipt FROM leases WHERE state='committed'").fetchone()
    assert receipt == 'sqlite:' + token
    blob = db.db.execute('SELECT data FROM receipt_blobs WHERE token=?', (token,)).fetchone()[0]
    assert encoded(json.loads(gzip.decompress(blob))['result']) == encoded(result)
    assert db.next_batch('gpu0') is None
    db.db.close()
