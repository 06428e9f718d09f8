"""The production smoke canary on the device (`tools/run_production.py:93-122`).

Two kernels agree on the first two subtiles, the high-magnitude canaries return
exactly their value, the two-batch canary ledger resumes after a reopen with
partitioned batches equal to the whole subtiles, and a full production batch
runs through the per-subtile path.  The record lands in canary.json.
"""
from __future__ import annotations

import json

import pytest

from psuedocubes_amd.production import Frontier
from psuedocubes_amd.run_production import smoke_canary

pytestmark = pytest.mark.gpu


def test_smoke_canary_passes_and_records_its_ledger(runtime, config619, focuses619, tmp_path) -> None:
    record = smoke_canary(runtime, tmp_path, config619, *focuses619, 'test-build')
    assert record['status'] == 'PASS'
    assert record['ledger']['commits'] == 2 and record['ledger']['committed_v_end'] == '200000000'
    assert len(record['full_batch']['subtile_scopes']) == 64
    assert [c['canary'] for c in record['cases'][2:]] == ['674441580981249129037406633', str(1000000007**3),
                                                          str(1000000009**3)]
    assert json.loads((tmp_path / 'canary.json').read_text())['status'] == 'PASS'
    reopened = Frontier(tmp_path / 'canary-ledger', {'purpose': 'production-canary'}, 'test-build',
                        200_000_000, 100_000_000)
    assert reopened.next_batch() is None and reopened.audit_production()['full_interval_complete']
    reopened.db.close()
