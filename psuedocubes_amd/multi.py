"""Shared lease frontier for several owners on one SQLite ledger (`pseudocube/multi.py:74-348`).

Owners (one per GPU process) take consecutive leases of batch_width v from an
allocation cursor and commit them in any order; the coverage cursor advances
over the contiguous committed prefix.  Each owner holds at most one unfinished
lease (a partial unique index), ownership never expires (`recover` needs
evidence that the owner stopped), a lease's result becomes coverage only in
the transaction that stores its compact receipt, and a confirmed hit stops
allocation for every owner.  The database runs in WAL mode with a 30 s busy
timeout, so owners in separate processes serialise on short write
transactions.

Receipts are gzip_receipt bytes, kept in the database (`receipt_blobs`) when
the runtime profile's receipt_storage is 'sqlite-gzip' and as durable
`.json.gz` files otherwise.  A build change is an offline `upgrade` with a
certificate chained to the previous one; receipts are checked against the
build of the epoch their lease lies in.  (`load_multi`, the reference's
reviewed-allocation gate with its fixed device list, is deployment tooling and
has no counterpart here.)
"""
from __future__ import annotations

import gzip
import hashlib
import json
from pathlib import Path
import time
from typing import Any
import uuid

from .accounting import Ledger, digest, encoded, gzip_receipt
from .optimized import validate_result
from .production import save_compressed

# (lease first, attempt token, first, end), as next_batch returns it
Job = tuple[int, str, int, int]

COUNTERS = ('commits', 'candidates', 'cubes', 'subtiles')

FRONTIER_SCHEMA = """
CREATE TABLE IF NOT EXISTS leases(
  first INTEGER PRIMARY KEY, end INTEGER NOT NULL, owner TEXT NOT NULL,
  state TEXT NOT NULL, token TEXT, receipt TEXT, sha TEXT,
  CHECK(first>=0 AND end>first), CHECK(state IN ('pending','running','committed')));
CREATE UNIQUE INDEX IF NOT EXISTS one_unfinished_per_owner ON leases(owner) WHERE state!='committed';
CREATE INDEX IF NOT EXISTS owner_commits ON leases(owner,first);
CREATE TABLE IF NOT EXISTS lease_attempts(
  token TEXT PRIMARY KEY, first INTEGER NOT NULL, owner TEXT NOT NULL,
  state TEXT NOT NULL, reason TEXT, build TEXT NOT NULL);
CREATE TABLE IF NOT EXISTS candidates(n TEXT PRIMARY KEY, witness TEXT NOT NULL, first INTEGER NOT NULL);
CREATE TABLE IF NOT EXISTS worker_totals(owner TEXT PRIMARY KEY, commits INTEGER NOT NULL,
  candidates TEXT NOT NULL, cubes TEXT NOT NULL, subtiles INTEGER NOT NULL);
CREATE TABLE IF NOT EXISTS receipt_blobs(token TEXT PRIMARY KEY, data BLOB NOT NULL);
CREATE TABLE IF NOT EXISTS upgrades(id INTEGER PRIMARY KEY, certificate TEXT NOT NULL, sha TEXT NOT NULL);
"""


def subtile_count(result: dict[str, Any]) -> int:
    """Subtiles a result covers: its compact count, else its scope records."""
    return int(result.get('subtiles_executed', len(result['subtile_scopes'])))


class SharedFrontier(Ledger):
    def __init__(self, directory: Path, config: dict[str, Any], build: str, end: int, batch_width: int,
                 owners: tuple[str, ...], baseline: dict[str, Any] | None = None,
                 profile: dict[str, Any] | None = None) -> None:
        super().__init__(directory, config, build)
        self.end, self.batch_width, self.owners = end, batch_width, tuple(owners)
        self.db.execute('PRAGMA busy_timeout=30000')
        self.db.execute('PRAGMA journal_mode=WAL')
        self.db.executescript(FRONTIER_SCHEMA)
        with self.transaction():
            state = self.state()
            running = state.get('active_build', state.get('frozen_build', build))
            if running != build:
                raise ValueError('shared frontier runs another build (active_build)')
            self.db.execute("INSERT OR IGNORE INTO meta VALUES('frozen_build',?)", (build,))
            fixed = {'active_build': build, 'runtime_profile': encoded(profile or {}).decode(),
                     'domain_end': str(end), 'batch_width': str(batch_width), 'owners': encoded(list(owners)).decode()}
            for key, value in fixed.items():
                bound = self.meta(key)
                if bound is not None and bound != value:
                    raise ValueError('shared frontier was created with another ' + key)
                self.db.execute('INSERT OR IGNORE INTO meta VALUES(?,?)', (key, value))
            imported = self.meta('baseline')
            if imported is not None:
                if baseline is not None and imported != encoded(baseline).decode():
                    raise ValueError('imported coverage differs from the bound baseline')
            else:
                if baseline is None:
                    raise ValueError('a new shared frontier needs an explicit audited baseline')
                prefix = int(baseline['cursor'])
                if not 0 <= prefix <= end or baseline['campaign_state'] != 'searching':
                    raise ValueError('baseline outside the domain, or its campaign already stopped')
                start = {'baseline': encoded(baseline).decode(), 'cursor': str(prefix),
                         'allocation_cursor': str(prefix), 'campaign_state': 'searching'}
                start.update({key: str(baseline[key]) for key in COUNTERS})
                self._set(start)
            for owner in self.owners:
                self.db.execute("INSERT OR IGNORE INTO worker_totals VALUES(?,0,'0','0',0)", (owner,))
        if self.db.execute('PRAGMA quick_check').fetchone()[0] != 'ok':
            raise ValueError('corrupt shared ledger')
        self.epochs = self._epochs()

    # --- state ---------------------------------------------------------------

    def state(self) -> dict[str, str]:
        return {row['key']: row['value'] for row in self.db.execute('SELECT key,value FROM meta')}

    def _set(self, values: dict[str, str]) -> None:
        self.db.executemany('INSERT OR REPLACE INTO meta VALUES(?,?)', values.items())

    def _epochs(self) -> list[tuple[int, str]]:
        """(first v, build) of each build epoch, from the baseline and the upgrade
        certificate chain (each certificate names its predecessor's SHA-256)."""
        state = self.state()
        epochs = [(int(json.loads(state['baseline'])['cursor']), state['frozen_build'])]
        previous: str | None = None
        last: dict[str, Any] | None = None
        for row in self.db.execute('SELECT certificate,sha FROM upgrades ORDER BY id'):
            certificate = json.loads(row['certificate'])
            if digest(certificate) != row['sha'] or certificate.get('previous_upgrade_sha256') != previous:
                raise ValueError('upgrade certificate chain broken')
            if certificate.get('config') != self.identity or certificate.get('old_build') != epochs[-1][1]:
                raise ValueError('upgrade certificate of another configuration or build')
            boundary = int(certificate['boundary'])
            if not epochs[-1][0] <= boundary <= self.end:
                raise ValueError('upgrade boundary outside the frontier')
            epochs.append((boundary, certificate['new_build']))
            previous, last = row['sha'], certificate
        if epochs[-1][1] != state['active_build']:
            raise ValueError('active build has no upgrade certificate')
        if last is not None and (last['new_batch_width'] != state['batch_width']
                                 or encoded(last['new_profile']).decode() != state['runtime_profile']):
            raise ValueError('active batch width or runtime profile differs from the last certificate')
        return epochs

    def _check_active(self) -> None:
        if self.meta('active_build') != self.build:
            raise ValueError('this runtime was replaced by a qualified upgrade')

    # --- leases --------------------------------------------------------------

    def upgrade(self, build: str, batch_width: int, profile: dict[str, Any], evidence: dict[str, Any]) -> dict[str, Any]:
        """Move the frontier to a new build, batch width and profile at its cursor.
        Offline only: the caller holds every owner's lock; old receipts stay as they are."""
        if not evidence or batch_width <= 0 or not build or build == self.build:
            raise ValueError('an upgrade needs evidence and a new build')
        if self.audit_shared()['unfinished_leases'] or self.meta('campaign_state') != 'searching':
            raise ValueError('an upgrade needs a stopped, fully committed frontier')
        with self.transaction():
            state = self.state()
            if state['active_build'] != self.build or state['cursor'] != state['allocation_cursor']:
                raise ValueError('frontier changed during the upgrade')
            last = self.db.execute('SELECT sha FROM upgrades ORDER BY id DESC LIMIT 1').fetchone()
            certificate = {'config': self.identity, 'old_build': self.build, 'new_build': build,
                           'boundary': state['cursor'], 'old_batch_width': state['batch_width'],
                           'new_batch_width': str(batch_width), 'old_profile': json.loads(state['runtime_profile']),
                           'new_profile': profile, 'before': state,
                           'previous_upgrade_sha256': last[0] if last else None, 'evidence': evidence,
                           'unix_time': time.time(), 'authorization': 'Deploy to production'}
            self.db.execute('INSERT INTO upgrades(certificate,sha) VALUES(?,?)',
                            (encoded(certificate).decode(), digest(certificate)))
            self._set({'active_build': build, 'batch_width': str(batch_width),
                       'runtime_profile': encoded(profile).decode()})
        return certificate

    def recover(self, owner: str, stopped_owner_evidence: str) -> int:
        """Return an owner's running lease to pending (its attempt abandoned)."""
        if owner not in self.owners or not stopped_owner_evidence:
            raise ValueError('recovery needs an authorised owner and evidence that it stopped')
        with self.transaction():
            self._check_active()
            row = self.db.execute("SELECT first,token FROM leases WHERE owner=? AND state='running'",
                                  (owner,)).fetchone()
            if row is None:
                return 0
            self.db.execute("UPDATE lease_attempts SET state='abandoned',reason=? WHERE token=?",
                            (stopped_owner_evidence, row['token']))
            self.db.execute("UPDATE leases SET state='pending',token=NULL WHERE first=?", (row['first'],))
            return 1

    def next_batch(self, owner: str) -> Job | None:
        """The owner's pending lease again, or a new one at the allocation cursor,
        under a fresh token; None once the campaign stopped or the domain is allocated."""
        if owner not in self.owners:
            raise ValueError('owner outside the authorised allocation')
        with self.transaction():
            self._check_active()
            state = self.state()
            if state['campaign_state'] != 'searching':
                return None
            row = self.db.execute("SELECT first,end,state FROM leases WHERE owner=? AND state!='committed'",
                                  (owner,)).fetchone()
            if row is not None:
                if row['state'] != 'pending':
                    raise ValueError('owner still runs a lease; recover it first')
                first, end = row['first'], row['end']
            else:
                first = int(state['allocation_cursor'])
                if first >= self.end:
                    return None
                end = min(first + self.batch_width, self.end)
                self.db.execute("INSERT INTO leases(first,end,owner,state) VALUES(?,?,?,'pending')",
                                (first, end, owner))
                self._set({'allocation_cursor': str(end)})
            token = uuid.uuid4().hex
            self.db.execute("UPDATE leases SET state='running',token=? WHERE first=?", (token, first))
            self.db.execute("INSERT INTO lease_attempts VALUES(?,?,?,'running',NULL,?)",
                            (token, first, owner, self.build))
            return first, token, first, end

    def _check_lease(self, owner: str, job: Job) -> None:
        self._check_active()
        first, token, start, end = job
        row = self.db.execute('SELECT owner,state,token,end FROM leases WHERE first=?', (first,)).fetchone()
        attempt = self.db.execute('SELECT owner,state,build FROM lease_attempts WHERE token=?', (token,)).fetchone()
        if row is None or (row['owner'], row['state'], row['token'], row['end']) != (owner, 'running', token, end) \
                or start != first:
            raise ValueError('lease is stale, foreign or changed')
        if attempt is None or (attempt['owner'], attempt['state'], attempt['build']) != (owner, 'running', self.build):
            raise ValueError('attempt is stale or of another build')

    def fail_lease(self, owner: str, job: Job, reason: str) -> None:
        with self.transaction():
            self._check_lease(owner, job)
            self.db.execute("UPDATE lease_attempts SET state='failed',reason=? WHERE token=?", (reason, job[1]))
            self.db.execute("UPDATE leases SET state='pending',token=NULL WHERE first=?", (job[0],))

    def complete(self, owner: str, job: Job, result: dict[str, Any]) -> None:
        """Commit a lease's verified result: receipt first (blob bytes, or a durable
        file), then one transaction that stores the blob, commits the lease, records
        its hits, advances the cursor over the contiguous committed prefix and adds
        the result's counts to the frontier's and the owner's totals."""
        first, token, _, end = job
        validate_result(first, end, result)
        self._check_lease(owner, job)
        value = {'config': self.identity, 'build': self.build, 'owner': owner, 'attempt': token,
                 'first': str(first), 'end': str(end), 'result': result}
        in_database = json.loads(self.meta('runtime_profile') or '{}').get('receipt_storage') == 'sqlite-gzip'
        if in_database:
            blob = gzip_receipt(value)
            name, sha = 'sqlite:' + token, hashlib.sha256(blob).hexdigest()
        else:
            path = self.directory / 'outputs' / (token + '.json.gz')
            name, sha = str(path.relative_to(self.directory)), save_compressed(path, value)
        added = {'commits': 1, 'candidates': int(result['valid_focused_candidates']),
                 'cubes': int(result['genuine_cubes']), 'subtiles': subtile_count(result)}
        with self.transaction():
            self._check_lease(owner, job)
            if in_database:
                self.db.execute('INSERT INTO receipt_blobs VALUES(?,?)', (token, blob))
            self.db.execute("UPDATE lease_attempts SET state='committed' WHERE token=?", (token,))
            self.db.execute("UPDATE leases SET state='committed',receipt=?,sha=? WHERE first=?", (name, sha, first))
            self.db.executemany('INSERT OR IGNORE INTO candidates VALUES(?,?,?)',
                                [(str(hit['n']), encoded(hit).decode(), first) for hit in result['cpu_confirmed_hits']])
            state = self.state()
            cursor = int(state['cursor'])
            while True:
                row = self.db.execute("SELECT end FROM leases WHERE first=? AND state='committed'", (cursor,)).fetchone()
                if row is None:
                    break
                cursor = row['end']
            if result['cpu_confirmed_hits'] or state['campaign_state'] == 'candidate-found':
                phase = 'candidate-found'
            else:
                phase = 'complete' if cursor == self.end else 'searching'
            update = {key: str(int(state[key]) + added[key]) for key in COUNTERS}
            update.update(cursor=str(cursor), campaign_state=phase)
            self._set(update)
            totals = self.db.execute('SELECT * FROM worker_totals WHERE owner=?', (owner,)).fetchone()
            self.db.execute('UPDATE worker_totals SET commits=?,candidates=?,cubes=?,subtiles=? WHERE owner=?',
                            tuple(str(int(totals[key]) + added[key]) for key in COUNTERS) + (owner,))

    # --- audit ---------------------------------------------------------------

    def validate_owner_receipt(self, owner: str) -> None:
        """The owner's newest committed receipt is intact."""
        row = self.db.execute("SELECT * FROM leases WHERE owner=? AND state='committed' ORDER BY first DESC LIMIT 1",
                              (owner,)).fetchone()
        if row is not None:
            self._receipt(dict(row))

    def _build_of(self, first: int, end: int) -> str:
        build = self.epochs[0][1]
        for boundary, later in self.epochs[1:]:
            if first < boundary < end:
                raise ValueError('lease straddles a build boundary')
            if first >= boundary:
                build = later
        return build

    def _receipt(self, row: dict[str, Any]) -> dict[str, Any]:
        """A committed lease's result, after checking its receipt's hash, identity,
        attempt and epoch build and the result itself."""
        data = self.receipt_bytes(row['receipt'], row['token'])
        if hashlib.sha256(data).hexdigest() != row['sha']:
            raise ValueError('committed receipt changed')
        try:
            value = json.loads(gzip.decompress(data))
        except (OSError, ValueError, EOFError) as error:
            raise ValueError('committed receipt unreadable') from error
        build = self._build_of(row['first'], row['end'])
        identity = (value.get('config'), value.get('build'), value.get('owner'), value.get('attempt'),
                    int(value.get('first', -1)), int(value.get('end', -1)))
        if identity != (self.identity, build, row['owner'], row['token'], row['first'], row['end']):
            raise ValueError('committed receipt names another lease, owner, attempt or build')
        attempt = self.db.execute('SELECT state,build FROM lease_attempts WHERE token=?', (row['token'],)).fetchone()
        if attempt is None or (attempt['state'], attempt['build']) != ('committed', build):
            raise ValueError('committed lease has no committed attempt of its build')
        validate_result(row['first'], row['end'], value['result'])
        return value['result']

    def audit_shared(self) -> dict[str, Any]:
        """Replay the whole frontier from the database: leases tile the domain from the
        baseline without overlap, every committed receipt checks, and the cursors and
        totals equal what the receipts add up to.  Run it on a frozen copy while
        receipts are immutable."""
        state = self.state()
        self.epochs = self._epochs()
        baseline = json.loads(state['baseline'])
        allocated = contiguous = int(baseline['cursor'])
        totals = {key: int(baseline[key]) for key in COUNTERS}
        workers = dict.fromkeys(self.owners, 0)
        unfinished: list[dict[str, Any]] = []
        for row in self.db.execute('SELECT * FROM leases ORDER BY first').fetchall():
            if row['first'] != allocated or row['end'] > self.end or row['owner'] not in self.owners:
                raise ValueError('leases overlap, leave a gap or name an unknown owner')
            allocated = row['end']
            if row['state'] != 'committed':
                unfinished.append({'first': str(row['first']), 'end': str(row['end']), 'owner': row['owner'],
                                   'state': row['state']})
                continue
            result = self._receipt(dict(row))
            if row['first'] == contiguous:
                contiguous = row['end']
            totals['commits'] += 1
            totals['candidates'] += int(result['valid_focused_candidates'])
            totals['cubes'] += int(result['genuine_cubes'])
            totals['subtiles'] += subtile_count(result)
            workers[row['owner']] += 1
        if (str(allocated), str(contiguous)) != (state['allocation_cursor'], state['cursor']) \
                or any(str(value) != state[key] for key, value in totals.items()):
            raise ValueError('frontier cursors or totals differ from the committed receipts')
        return dict(status='PASS', imported_prefix=baseline['cursor'], committed_v_end=str(contiguous),
                    allocated_v_end=str(allocated), totals=totals, worker_commits=workers,
                    unfinished_leases=unfinished, no_overlap=True, full_interval_complete=contiguous == self.end,
                    independent_arithmetic_replay=False)


def lease_worker(frontier: SharedFrontier, owner: str, matcher: Any, af: Any, bf: Any, leases: int = 1,
                 compute: Any = None) -> list[dict[str, Any]]:
    """One owner's loop (the reference's GPU worker): take a lease, compute it
    (`optimized.compute_lease`, pipelined on the device), commit its compact
    receipt; a lease whose computation fails goes back to pending with the
    error as its attempt's reason, and the error propagates.  Stops after
    `leases` leases or when the frontier has nothing left; returns the
    committed results."""
    if compute is None:
        from .optimized import compute_lease as compute
    done = []
    for _ in range(leases):
        job = frontier.next_batch(owner)
        if job is None:
            break
        try:
            result = compute(matcher, af, bf, job[2], job[3])
        except BaseException as error:
            frontier.fail_lease(owner, job, f'{type(error).__name__}: {error}')
            raise
        frontier.complete(owner, job, result)
        done.append(result)
    return done
