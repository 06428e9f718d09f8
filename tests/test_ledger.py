"""The coverage ledger on the CPU: ownership, durable and compact receipts, recovery, audits.

The cases follow the reference's ledger tests (`tests/test_pseudocube_recovery.py`,
`tests/test_pseudocube_production.py:19-55`, `tests/test_pseudocube_multi.py:33-153`,
`tests/test_pseudocube_optimized.py:38-122`), with crashes forced from the host:
a receipt write torn by a failing `os.replace`, a process that exits from a SQLite
TEMP TRIGGER inside `complete` after its receipt blob is inserted, and four spawned
owners contending on one WAL database.  Plus the served form's batched commit
(`Ledger.commit_batch`), which has no counterpart in the reference.
"""
from __future__ import annotations

from concurrent.futures import ProcessPoolExecutor
import gzip
import json
import multiprocessing
import os
from pathlib import Path
from typing import Any
from unittest.mock import MagicMock, patch

import pytest

from psuedocubes_amd.accounting import Ledger
from psuedocubes_amd.multi import SharedFrontier
from psuedocubes_amd.optimized import PROFILE, allocated_bytes, ledger_settings
from psuedocubes_amd.production import Frontier, save_compressed

OWNERS = ('a', 'b', 'c', 'd')
BASE = dict(cursor='0', commits='0', candidates='0', cubes='0', subtiles='0', campaign_state='searching')
CONFIG = {'search_type': 'pseudocube', 'p': 619}
PASS = dict(status='PASS', host_verified=True)


def lease_row(first: int, end: int, hit: bool = False) -> dict[str, Any]:
    return dict(status='PASS', host_verified=True, valid_focused_candidates=11, genuine_cubes=1,
                cpu_confirmed_hits=[dict(n='test-only-witness', valid_pseudocube=True)] if hit else [],
                subtile_scopes=[dict(first=str(first), end=str(end))])


def shared(path: Path, baseline: dict[str, Any] | None = None, end: int = 10000, build: str = 'build',
           width: int = 10, profile: dict[str, Any] | None = None) -> SharedFrontier:
    return SharedFrontier(path, {'test': True}, build, end, width, OWNERS, baseline, profile)


# --- Ledger ------------------------------------------------------------------

@pytest.fixture
def ledger(tmp_path: Path):
    db = Ledger(tmp_path, CONFIG, 'build-v1')
    db.work = db.plan([dict(kind='v-window', first='0', end='100')])[0]
    yield db
    db.db.close()


def test_receipt_is_written_before_coverage_and_a_torn_write_retries(ledger: Ledger) -> None:
    token = ledger.claim(ledger.work)
    with patch('psuedocubes_amd.accounting.os.replace', side_effect=OSError('torn receipt write')):
        with pytest.raises(OSError):
            ledger.commit(ledger.work, token, dict(PASS))
    assert ledger.audit()['states'] == {'running': 1}
    ledger.reclaim(ledger.work, token, 'the test process stopped its owner')
    retry = ledger.claim(ledger.work)
    with pytest.raises(ValueError):
        ledger.commit(ledger.work, token, dict(PASS))
    ledger.commit(ledger.work, retry, dict(PASS, pairs='12'))
    with pytest.raises(ValueError):
        ledger.claim(ledger.work)
    assert [r['pairs'] for r in ledger.audit()['committed_results']] == ['12']
    with pytest.raises(ValueError):
        ledger.reclaim(ledger.work, retry, '')


def test_corrupt_receipt_and_changed_configuration_fail_closed(ledger: Ledger, tmp_path: Path) -> None:
    token = ledger.claim(ledger.work)
    ledger.commit(ledger.work, token, dict(PASS))
    (tmp_path / 'outputs' / f'{token}.json').write_text('{}')
    with pytest.raises(ValueError):
        ledger.audit()
    with pytest.raises(ValueError):
        Ledger(tmp_path, {'search_type': 'pseudosquare', 'p': 619}, 'build-v1')
    with pytest.raises(ValueError):
        ledger.plan([dict(first='1', end='100')])


def test_overflow_and_kernel_error_split_without_double_count(ledger: Ledger) -> None:
    for flag in ('overflow', 'kernel_error'):
        token = ledger.claim(ledger.work)
        with pytest.raises(ValueError):
            ledger.commit(ledger.work, token, dict(PASS, **{flag: True}))
    halves = ledger.split(ledger.work)
    assert len(halves) == 2
    with pytest.raises(ValueError):
        ledger.claim(ledger.work)
    for child in halves:
        ledger.commit(child, ledger.claim(child), dict(PASS))
    assert ledger.audit()['states'] == {'split': 1, 'committed': 2}
    scopes = [json.loads(r[0]) for r in ledger.db.execute('SELECT scope FROM work WHERE parent=?', (ledger.work,))]
    assert sorted((int(s['first']), int(s['end'])) for s in scopes) == [(0, 50), (50, 100)]


def test_budget_persists_across_connections_and_binds_the_build(ledger: Ledger, tmp_path: Path) -> None:
    held = ledger.reserve('gpu', 600, 600)
    other = Ledger(tmp_path, CONFIG, 'build-v1')
    try:
        with pytest.raises(RuntimeError, match='BUDGET_EXHAUSTED'):
            other.reserve('gpu', 1, 600)
    finally:
        other.db.close()
    ledger.settle(held, 20)
    with pytest.raises(ValueError):
        ledger.settle(held, 20)
    ledger.reserve('gpu', 580, 600)
    token = ledger.claim(ledger.work)
    ledger.build = 'another-build'
    with pytest.raises(ValueError):
        ledger.commit(ledger.work, token, dict(PASS))


def test_replays_never_add_coverage_and_quarantine_withdraws(ledger: Ledger) -> None:
    ledger.commit(ledger.work, ledger.claim(ledger.work), dict(PASS, pairs='10'))
    replay = ledger.begin_validation(ledger.work, 'packed table A/B')
    ledger.finish_validation(replay, dict(PASS, pairs='10'))
    audit = ledger.audit()
    assert len(audit['committed_results']) == 1 and len(audit['validations']) == 1
    assert audit['validations'][0]['state'] == 'verified'
    with pytest.raises(ValueError):
        ledger.plan([dict(first='0', end='100')] * 2)
    ledger.quarantine(ledger.work, 'injected independent-replay disagreement')
    audit = ledger.audit()
    assert audit['states'] == {'quarantined': 1} and audit['committed_results'] == []


def test_batched_commit_stores_compact_receipts_and_records_repeats(tmp_path: Path) -> None:
    db = Ledger(tmp_path, CONFIG, 'build-v1')
    scopes = [dict(kind='served-range', first=str(k * 10), end=str(k * 10 + 10)) for k in range(3)]
    tokens = db.commit_batch([(s, dict(PASS, n=k)) for k, s in enumerate(scopes)])
    assert len(set(tokens)) == 3
    # a repeated scope is a replay, never a second unit of coverage
    again = db.commit_batch([(scopes[1], dict(PASS, n=1))])
    audit = db.audit()
    assert audit['states'] == {'committed': 3}
    assert sorted(r['n'] for r in audit['committed_results']) == [0, 1, 2]
    assert [v['purpose'] for v in audit['validations']] == ['served repeat']
    blob = db.db.execute('SELECT data FROM receipt_blobs WHERE token=?', (tokens[0],)).fetchone()[0]
    assert json.loads(gzip.decompress(blob))['result'] == dict(PASS, n=0)
    # an unverified result, or a scope another owner runs, is refused and nothing is written
    running = db.plan([dict(kind='served-range', first='90', end='100')])[0]
    db.claim(running)
    for entries in ([(dict(kind='served-range', first='50', end='60'), dict(status='FAIL'))],
                    [(dict(kind='served-range', first='60', end='70'), dict(PASS)),
                     (dict(kind='served-range', first='90', end='100'), dict(PASS))]):
        with pytest.raises(ValueError):
            db.commit_batch(entries)
    assert db.db.execute('SELECT count(*) FROM receipt_blobs').fetchone()[0] == 4
    db.db.execute("UPDATE receipt_blobs SET data=x'00' WHERE token=?", (again[0],))
    with pytest.raises(ValueError):
        db.audit()
    db.db.close()


# --- Frontier (one owner) ----------------------------------------------------

def production_row(first: int, end: int, hits: list[str] | None = None) -> dict[str, Any]:
    middle = (first + end) // 2
    return dict(PASS, valid_focused_candidates=11, genuine_cubes=0, cpu_confirmed_hits=hits or [],
                subtile_scopes=[dict(first=str(first), end=str(middle)), dict(first=str(middle), end=str(end))])


def test_frontier_resumes_after_a_torn_receipt_with_an_exact_cursor(tmp_path: Path) -> None:
    db = Frontier(tmp_path, {'test': True}, 'build', 20, 10)
    work, token, first, end = db.next_batch()
    with patch('psuedocubes_amd.production.os.replace', side_effect=OSError('torn receipt')):
        with pytest.raises(OSError):
            db.complete(work, token, first, end, production_row(first, end))
    assert db.state()['cursor'] == '0'
    db.db.close()
    db = Frontier(tmp_path, {'test': True}, 'build', 20, 10)
    assert db.recover_stopped_owner('new process holds the GPU and frontier locks') == 1
    retry = db.next_batch()
    assert retry[0] == work and retry[1] != token
    with pytest.raises(ValueError):
        db.complete(work, token, first, end, production_row(first, end))
    db.complete(*retry, production_row(retry[2], retry[3]))
    db.db.close()
    db = Frontier(tmp_path, {'test': True}, 'build', 20, 10)
    db.validate_last_receipt()
    final = db.next_batch()
    assert final[2:] == (10, 20)
    db.complete(*final, production_row(10, 20))
    assert db.next_batch() is None
    audit = db.audit_production()
    assert audit['full_interval_complete'] and audit['commits'] == 2
    db.db.close()


def test_frontier_stops_on_a_hit_and_detects_corruption(tmp_path: Path) -> None:
    db = Frontier(tmp_path, {'test': True}, 'build', 20, 10)
    job = db.next_batch()
    db.complete(*job, production_row(job[2], job[3], ['test-witness']))
    assert db.state()['campaign_state'] == 'candidate-found' and db.next_batch() is None
    (db.directory / db.state()['last_receipt']).write_bytes(b'corrupt')
    with pytest.raises(ValueError):
        db.validate_last_receipt()
    with pytest.raises(ValueError):
        db.audit_production()
    db.db.close()


def test_frontier_refuses_gaps_and_changed_builds(tmp_path: Path) -> None:
    db = Frontier(tmp_path, {'test': True}, 'build', 20, 10)
    job = db.next_batch()
    bad = production_row(job[2], job[3])
    bad['subtile_scopes'][1]['first'] = '6'
    with pytest.raises(ValueError):
        db.complete(*job, bad)
    assert db.state()['cursor'] == '0'
    for build, end in (('changed-build', 20), ('build', 30)):
        with pytest.raises(ValueError):
            Frontier(tmp_path, {'test': True}, build, end, 10)
    db.db.close()


# --- SharedFrontier (several owners) ----------------------------------------

def owner_loop(path: str, owner: str) -> list[tuple[int, int]]:
    """One spawned owner: twenty leases, each committed after a short owner-specific pause."""
    import time
    db = shared(Path(path))
    done = []
    try:
        for _ in range(20):
            job = db.next_batch(owner)
            assert job is not None
            time.sleep(0.002 * (OWNERS.index(owner) + 1))
            db.complete(owner, job, lease_row(job[2], job[3]))
            done.append((job[2], job[3]))
    finally:
        db.db.close()
    return done


def test_four_spawned_owners_share_one_wal_frontier(tmp_path: Path) -> None:
    shared(tmp_path, BASE).db.close()
    with ProcessPoolExecutor(max_workers=4, mp_context=multiprocessing.get_context('spawn')) as pool:
        per_owner = list(pool.map(owner_loop, [str(tmp_path)] * 4, OWNERS))
    assert sorted(s for done in per_owner for s in done) == [(n, n + 10) for n in range(0, 800, 10)]
    db = shared(tmp_path)
    assert db.db.execute('PRAGMA journal_mode').fetchone()[0] == 'wal'
    audit = db.audit_shared()
    assert audit['totals']['commits'] == 80 and audit['committed_v_end'] == '800'
    assert audit['worker_commits'] == dict.fromkeys(OWNERS, 20)
    db.db.close()


def test_out_of_order_commits_and_a_sticky_candidate_stop(tmp_path: Path) -> None:
    db = shared(tmp_path, BASE)
    a, b, c = (db.next_batch(owner) for owner in OWNERS[:3])
    db.complete('b', b, lease_row(b[2], b[3]))
    assert db.state()['cursor'] == '0' and db.audit_shared()['totals']['commits'] == 1
    db.complete('a', a, lease_row(a[2], a[3], hit=True))
    assert db.state()['cursor'] == '20'
    assert db.next_batch('d') is None
    db.complete('c', c, lease_row(c[2], c[3]))
    state = db.state()
    assert (state['campaign_state'], state['cursor']) == ('candidate-found', '30')
    assert db.db.execute('SELECT count(*) FROM candidates').fetchone()[0] == 1
    db.audit_shared()
    db.db.close()


def test_recovery_needs_evidence_and_never_steals_a_live_lease(tmp_path: Path) -> None:
    db = shared(tmp_path, BASE)
    a, b = db.next_batch('a'), db.next_batch('b')
    with pytest.raises(ValueError):
        db.next_batch('a')
    with pytest.raises(ValueError):
        db.complete('c', a, lease_row(a[2], a[3]))
    with pytest.raises(ValueError):
        db.recover('a', '')
    db.db.close()
    db = shared(tmp_path)
    assert db.recover('a', 'the test stopped owner a') == 1
    retry = db.next_batch('a')
    assert (retry[0], retry[2], retry[3]) == (a[0], a[2], a[3]) and retry[1] != a[1]
    with pytest.raises(ValueError):
        db.complete('a', a, lease_row(a[2], a[3]))
    db.complete('b', b, lease_row(b[2], b[3]))
    db.complete('a', retry, lease_row(retry[2], retry[3]))
    assert db.audit_shared()['totals']['commits'] == 2
    db.db.close()


def test_an_orphan_receipt_is_not_coverage_and_corruption_stops(tmp_path: Path) -> None:
    db = shared(tmp_path, BASE)
    job = db.next_batch('a')

    def crash_after_write(path: Path, value: dict[str, Any]) -> str:
        save_compressed(path, value)
        raise OSError('crash after the durable receipt, before the commit')

    with patch('psuedocubes_amd.multi.save_compressed', side_effect=crash_after_write):
        with pytest.raises(OSError):
            db.complete('a', job, lease_row(job[2], job[3]))
    assert len(list((tmp_path / 'outputs').glob('*.json.gz'))) == 1
    assert db.state()['cursor'] == '0' and db.audit_shared()['totals']['commits'] == 0
    db.fail_lease('a', job, 'isolated injected failure')
    retry = db.next_batch('a')
    db.complete('a', retry, lease_row(retry[2], retry[3]))
    row = db.db.execute("SELECT receipt FROM leases WHERE state='committed'").fetchone()
    (tmp_path / row[0]).write_bytes(b'corrupt')
    with pytest.raises(ValueError):
        db.validate_owner_receipt('a')
    db.db.close()


def test_the_domain_tail_completes_and_bindings_fail_closed(tmp_path: Path) -> None:
    db = shared(tmp_path, BASE, end=23)
    jobs = [db.next_batch(owner) for owner in OWNERS]
    assert jobs[3] is None and jobs[2][2:] == (20, 23)
    for owner, job in reversed(list(zip(OWNERS[:3], jobs[:3]))):
        db.complete(owner, job, lease_row(job[2], job[3]))
    assert db.state()['campaign_state'] == 'complete'
    assert db.audit_shared()['full_interval_complete']
    with pytest.raises(ValueError):
        SharedFrontier(tmp_path, {'test': True}, 'different', 23, 10, OWNERS)
    with pytest.raises(ValueError):
        shared(tmp_path, dict(BASE, cursor='10'), end=23)
    with pytest.raises(ValueError):
        shared(tmp_path / 'fresh', None)   # no baseline: a new frontier never starts from zero by default
    db.db.close()


def upgraded(path: Path, width: int = 10) -> SharedFrontier:
    return SharedFrontier(path, {'test': 'upgrade'}, 'new', 1000, width, OWNERS, None, PROFILE)


def test_an_upgrade_keeps_old_receipts_and_retires_the_old_runtime(tmp_path: Path) -> None:
    old = SharedFrontier(tmp_path, {'test': 'upgrade'}, 'old', 1000, 10, OWNERS, BASE)
    job = old.next_batch('a')
    with pytest.raises(ValueError):
        old.upgrade('new', 100, PROFILE, {'tests': 'passed'})   # a lease is still running
    old.complete('a', job, lease_row(0, 10))
    old_file = tmp_path / old.db.execute('SELECT receipt FROM leases').fetchone()[0]
    original = old_file.read_bytes()
    certificate = old.upgrade('new', 100, PROFILE, {'tests': 'passed'})
    assert certificate['boundary'] == '10'
    with pytest.raises(ValueError):
        old.next_batch('b')
    old.db.close()
    with pytest.raises(ValueError):
        SharedFrontier(tmp_path, {'test': 'upgrade'}, 'old', 1000, 10, OWNERS)
    db = upgraded(tmp_path, 100)
    assert ledger_settings(tmp_path) == ('new', 100, PROFILE)
    job = db.next_batch('b')
    assert job[2:] == (10, 110)
    db.complete('b', job, lease_row(10, 110))
    # the new profile keeps receipts as gzip blobs in the database
    assert db.db.execute('SELECT receipt FROM leases WHERE first=10').fetchone()[0] == 'sqlite:' + job[1]
    assert db.audit_shared()['totals']['commits'] == 2
    assert old_file.read_bytes() == original
    db.db.execute("UPDATE receipt_blobs SET data=x'00'")
    with pytest.raises(ValueError):
        db.audit_shared()
    db.db.close()


def exit_inside_complete(path: str) -> None:
    """A spawned owner that exits (status 91) from a TEMP TRIGGER fired when complete()
    inserts its receipt blob, i.e. inside the committing transaction."""
    db = upgraded(Path(path), 100)
    job = db.next_batch('a')
    db.db.create_function('isolated_exit', 0, lambda: os._exit(91))
    db.db.execute('CREATE TEMP TRIGGER isolated_exit AFTER INSERT ON receipt_blobs BEGIN SELECT isolated_exit(); END')
    db.complete('a', job, lease_row(job[2], job[3]))


def test_an_owner_that_dies_inside_complete_keeps_its_range(tmp_path: Path) -> None:
    old = SharedFrontier(tmp_path, {'test': 'upgrade'}, 'old', 1000, 10, OWNERS, BASE)
    old.upgrade('new', 100, PROFILE, {'tests': 'isolated fault probe'})
    old.db.close()
    child = multiprocessing.get_context('spawn').Process(target=exit_inside_complete, args=(str(tmp_path),))
    child.start()
    child.join(60)
    assert not child.is_alive() and child.exitcode == 91
    db = upgraded(tmp_path, 100)
    assert db.state()['cursor'] == '0'
    assert db.db.execute('SELECT count(*) FROM receipt_blobs').fetchone()[0] == 0
    stale = db.db.execute('SELECT token FROM leases').fetchone()[0]
    assert db.recover('a', 'owned isolated child exited with 91') == 1
    job = db.next_batch('a')
    assert job[2:] == (0, 100) and job[1] != stale
    with pytest.raises(ValueError):
        db.complete('a', (0, stale, 0, 100), lease_row(0, 100))
    db.complete('a', job, lease_row(0, 100))
    assert db.audit_shared()['totals']['commits'] == 1
    db.db.close()


def test_a_tampered_upgrade_certificate_is_refused(tmp_path: Path) -> None:
    old = SharedFrontier(tmp_path, {'test': 'upgrade'}, 'old', 1000, 10, OWNERS, BASE)
    old.upgrade('new', 100, PROFILE, {'tests': 'passed'})
    old.db.execute("UPDATE upgrades SET certificate='{}'")
    old.db.close()
    with pytest.raises(ValueError):
        upgraded(tmp_path, 100)


def test_storage_scan_tolerates_only_vanishing_temporary_files(tmp_path: Path) -> None:
    entry = MagicMock()
    entry.is_dir.return_value = False
    entry.stat.side_effect = FileNotFoundError('renamed meanwhile')
    entry.name = 'status.json.abc.tmp'
    scan = MagicMock()
    scan.__enter__.return_value = [entry]
    with patch('psuedocubes_amd.optimized.os.scandir', return_value=scan):
        assert allocated_bytes(tmp_path) == tmp_path.stat().st_blocks * 512
        entry.name = 'committed.json.gz'
        with pytest.raises(FileNotFoundError):
            allocated_bytes(tmp_path)


def test_lease_worker_commits_and_returns_failed_leases(tmp_path: Path) -> None:
    """The owner loop over a stand-in computation: two leases committed; a computation
    that raises sends its lease back to pending, and the next call retries the same range."""
    from psuedocubes_amd.multi import lease_worker
    db = shared(tmp_path, BASE, end=40)
    calls = []

    def compute(matcher, af, bf, first, end):
        calls.append((first, end))
        if len(calls) == 3:
            raise ArithmeticError('injected verification failure')
        return lease_row(first, end)

    assert len(lease_worker(db, 'a', None, None, None, leases=2, compute=compute)) == 2
    with pytest.raises(ArithmeticError):
        lease_worker(db, 'a', None, None, None, leases=1, compute=compute)
    reason = db.db.execute("SELECT reason FROM lease_attempts WHERE state='failed'").fetchone()[0]
    assert 'injected' in reason
    assert len(lease_worker(db, 'a', None, None, None, leases=5, compute=compute)) == 2
    assert calls == [(0, 10), (10, 20), (20, 30), (20, 30), (30, 40)]
    audit = db.audit_shared()
    assert audit['full_interval_complete'] and audit['totals']['commits'] == 4
    db.db.close()
