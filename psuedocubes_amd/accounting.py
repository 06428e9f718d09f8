"""Canonical JSON encoding, SHA-256 digests, durable writes and the coverage ledger.

The encoding is part of the bit-exact contract: per-subtile output hashes and
the lease transcript are SHA-256 over exactly these bytes
(`pseudocube/accounting.py:17-49`).  Integers of magnitude >= 2^53 become
decimal strings, keys are sorted, separators are compact, and a trailing
newline is appended.

`Ledger` (`pseudocube/accounting.py:52-233`) binds a SQLite database to one
search configuration and build: work items move pending -> running (under a
fresh token) -> committed, failed, split or quarantined, every committed
result has a receipt whose SHA-256 the database holds, validation replays are
recorded apart from coverage, and budgets persist across processes.  Local
POSIX filesystems only: a receipt is written to a temporary file, fsynced,
renamed into place and its directory fsynced before the SQLite transaction
(synchronous=FULL) that makes it coverage; a receipt without that transaction
is an orphan, never coverage.
"""
from __future__ import annotations

from contextlib import contextmanager
import gzip
import hashlib
import json
import os
from pathlib import Path
import sqlite3
from typing import Any, Iterator
import uuid
import zlib


_EXACT = 1 << 53
_PLAIN_INTS = {int}


def interoperable(value: Any) -> Any:
    """Replace integers that JSON readers could round with strings (`pseudocube/accounting.py:17-24`).

    A list whose elements are all plain ints inside (-2^53, 2^53) is copied
    as it is, checked with C-level min/max instead of one call per element:
    receipts carry many such count lists.
    """
    kind = type(value)
    if kind is int:
        return value if -_EXACT < value < _EXACT else str(value)
    if kind is str:
        return value
    if isinstance(value, dict):
        return {str(k): interoperable(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        if value and set(map(type, value)) == _PLAIN_INTS and -_EXACT < min(value) and max(value) < _EXACT:
            return list(value)
        return [interoperable(v) for v in value]
    return value


def encoded(value: Any) -> bytes:
    text = json.dumps(interoperable(value), sort_keys=True, separators=(',', ':'), allow_nan=False)
    return (text + '\n').encode()


def digest(value: Any) -> str:
    return hashlib.sha256(encoded(value)).hexdigest()


# Every digit mapped to 0: an integer token of 16 or more digits right after a
# separator of the compact encoding is then one of six byte strings, found with
# plain substring searches.  That is the only place an integer of magnitude
# >= 2^53 (16 digits or more) can stand outside a string.
_ZERO_DIGITS = bytes.maketrans(b'123456789', b'000000000')
_LONG_NUMBERS = tuple(sep + sign + b'0' * 16 for sep in (b':', b',', b'[') for sign in (b'', b'-'))
_SIXTEEN = b'0' * 16


def encoded_str_keys(value: Any) -> bytes:
    """encoded(value) for a value whose dict keys are all str (every result row,
    scope record and witness record the engine builds): the C encoder's output
    when it holds no integer of 16 or more digits, otherwise encoded()'s.  It
    skips interoperable()'s copy of the whole value, which dominates the cost of
    a receipt."""
    text = json.dumps(value, sort_keys=True, separators=(',', ':'), allow_nan=False).encode()
    zeros = text.translate(_ZERO_DIGITS)
    if zeros[:1] in (b'0', b'-') or (_SIXTEEN in zeros and any(pattern in zeros for pattern in _LONG_NUMBERS)):
        return encoded(value)
    return text + b'\n'


def durable(path: Path, value: Any) -> str:
    """fsync a temporary file, rename it over ``path``, fsync the directory (`pseudocube/accounting.py:35-49`)."""
    path.parent.mkdir(parents=True, exist_ok=True)
    data = encoded(value)
    temporary = path.with_name(f'{path.name}.{uuid.uuid4().hex}.tmp')
    with temporary.open('xb') as stream:
        stream.write(data)
        stream.flush()
        os.fsync(stream.fileno())
    os.replace(temporary, path)
    fd = os.open(path.parent, os.O_RDONLY)
    try:
        os.fsync(fd)
    finally:
        os.close(fd)
    return hashlib.sha256(data).hexdigest()


def passing(result: dict[str, Any]) -> bool:
    """Whether a result may become coverage: PASS, host verified, no overflow and
    no kernel error (`pseudocube/accounting.py:127-129`)."""
    return (result.get('status') == 'PASS' and bool(result.get('host_verified')) and not result.get('overflow')
            and not result.get('kernel_error'))


def gzip_receipt(value: Any, level: int = 6) -> bytes:
    """The compact receipt bytes of a value whose dict keys are all str: its
    canonical encoding in a gzip member (no name, modification time 0), so equal
    values give equal bytes."""
    deflate = zlib.compressobj(level, zlib.DEFLATED, 31)   # wbits 31: the gzip container
    return deflate.compress(encoded_str_keys(value)) + deflate.flush()


# The ledger's tables (the reference's schema, so its databases open here):
# meta holds the configuration identity and frontier counters as text; work the
# logical scopes; attempts every ownership token; validations the replays;
# budget the persistent reservations; receipt_blobs compact receipts kept in
# the database itself (receipt 'sqlite:<token>').
LEDGER_SCHEMA = """
CREATE TABLE IF NOT EXISTS meta(key TEXT PRIMARY KEY,value TEXT NOT NULL);
CREATE TABLE IF NOT EXISTS work(id TEXT PRIMARY KEY,scope TEXT NOT NULL,state TEXT NOT NULL,token TEXT,parent TEXT,
                                receipt TEXT,sha TEXT);
CREATE TABLE IF NOT EXISTS attempts(token TEXT PRIMARY KEY,work_id TEXT NOT NULL,build TEXT NOT NULL,state TEXT NOT NULL,
                                    reason TEXT);
CREATE TABLE IF NOT EXISTS validations(token TEXT PRIMARY KEY,work_id TEXT NOT NULL,build TEXT NOT NULL,
                                       purpose TEXT NOT NULL,state TEXT NOT NULL,receipt TEXT,sha TEXT);
CREATE TABLE IF NOT EXISTS budget(id TEXT PRIMARY KEY,kind TEXT NOT NULL,reserved REAL NOT NULL,charged REAL,
                                  state TEXT NOT NULL);
CREATE TABLE IF NOT EXISTS receipt_blobs(token TEXT PRIMARY KEY,data BLOB NOT NULL);
"""


class Ledger:
    """Configuration-bound work ownership with durable receipts (`pseudocube/accounting.py:52-233`).

    Ownership never expires: a running item returns to the pool only through
    fail() or reclaim() with explicit evidence that its owner has stopped.
    """

    database = 'coverage.sqlite3'

    def __init__(self, directory: Path, config: dict[str, Any], build: str) -> None:
        directory.mkdir(parents=True, exist_ok=True)
        self.directory = directory
        self.identity = digest(config)
        self.build = build
        # autocommit mode: every write below runs in an explicit BEGIN IMMEDIATE
        self.db = sqlite3.connect(directory / self.database, isolation_level=None, timeout=30.0)
        self.db.row_factory = sqlite3.Row
        self.db.execute('PRAGMA synchronous=FULL')
        self.db.executescript(LEDGER_SCHEMA)
        with self.transaction():
            bound = self.meta('config')
            if bound is not None and bound != self.identity:
                raise ValueError('ledger belongs to another search configuration')
            self.db.execute("INSERT OR IGNORE INTO meta VALUES('config',?)", (self.identity,))

    @contextmanager
    def transaction(self) -> Iterator[None]:
        """One write transaction, taken at once (BEGIN IMMEDIATE) so that concurrent
        owners serialise on it instead of failing at their first write."""
        self.db.execute('BEGIN IMMEDIATE')
        try:
            yield
        except BaseException:
            self.db.execute('ROLLBACK')
            raise
        self.db.execute('COMMIT')

    def meta(self, key: str) -> str | None:
        row = self.db.execute('SELECT value FROM meta WHERE key=?', (key,)).fetchone()
        return None if row is None else row[0]

    def register(self, scope: dict[str, Any], parent: str | None = None) -> str:
        """The work item of a scope under this configuration (created pending once)."""
        work_id = digest({'config': self.identity, 'scope': scope})
        self.db.execute('INSERT OR IGNORE INTO work(id,scope,state,parent) VALUES(?,?,?,?)',
                        (work_id, encoded(scope).decode(), 'pending', parent))
        return work_id

    def plan(self, scopes: list[dict[str, Any]]) -> list[str]:
        """Register the expected work list; a ledger accepts one plan for good."""
        if len(set(map(digest, scopes))) != len(scopes):
            raise ValueError('the planned work list repeats a scope')
        plan = digest(scopes)
        with self.transaction():
            bound = self.meta('plan')
            if bound is not None and bound != plan:
                raise ValueError('the planned work list differs from the bound one')
            self.db.execute("INSERT OR IGNORE INTO meta VALUES('plan',?)", (plan,))
            return [self.register(scope) for scope in scopes]

    def claim(self, work_id: str) -> str:
        """Own a pending or failed item under a fresh token."""
        token = uuid.uuid4().hex
        with self.transaction():
            row = self.db.execute('SELECT state FROM work WHERE id=?', (work_id,)).fetchone()
            if row is None or row['state'] not in ('pending', 'failed'):
                raise ValueError('work item is owned, split, committed or unknown')
            self.db.execute("UPDATE work SET state='running',token=? WHERE id=?", (token, work_id))
            self.db.execute("INSERT INTO attempts VALUES(?,?,?,'running',NULL)", (token, work_id, self.build))
        return token

    def _check_owner(self, work_id: str, token: str) -> None:
        row = self.db.execute('SELECT state,token FROM work WHERE id=?', (work_id,)).fetchone()
        attempt = self.db.execute('SELECT build FROM attempts WHERE token=?', (token,)).fetchone()
        if (row is None or row['state'] != 'running' or row['token'] != token or attempt is None
                or attempt['build'] != self.build):
            raise ValueError('token no longer owns the item, or the build changed')

    def fail(self, work_id: str, token: str, reason: str) -> None:
        with self.transaction():
            self._check_owner(work_id, token)
            self.db.execute("UPDATE attempts SET state='failed',reason=? WHERE token=?", (reason, token))
            self.db.execute("UPDATE work SET state='failed',token=NULL WHERE id=?", (work_id,))

    def reclaim(self, work_id: str, token: str, stopped_owner_evidence: str) -> None:
        """Release a running item whose owner is known to have stopped (no expiry)."""
        if not stopped_owner_evidence:
            raise ValueError('reclaiming needs evidence that the owner has stopped')
        self.fail(work_id, token, 'owner stopped: ' + stopped_owner_evidence)

    def commit(self, work_id: str, token: str, result: dict[str, Any]) -> None:
        """Write the result's receipt durably, then make it coverage in one transaction."""
        if not passing(result):
            self.fail(work_id, token, 'invalid execution receipt')
            raise ValueError('a failed, overflowed or unverified result cannot become coverage')
        path = self.directory / 'outputs' / (token + '.json')
        sha = durable(path, {'config': self.identity, 'build': self.build, 'work_id': work_id, 'attempt': token,
                             'result': result})
        with self.transaction():
            self._check_owner(work_id, token)
            self.db.execute("UPDATE attempts SET state='committed' WHERE token=?", (token,))
            self.db.execute("UPDATE work SET state='committed',receipt=?,sha=? WHERE id=?",
                            (str(path.relative_to(self.directory)), sha, work_id))

    def commit_batch(self, entries: list[tuple[dict[str, Any], dict[str, Any]]]) -> list[str]:
        """Record many verified results in one transaction, each as coverage of its
        scope, with compact receipts (gzip_receipt, level 1) kept in the database.

        A scope already committed, or committed earlier in the same group, is
        not counted twice: that result becomes a validation replay ('served
        repeat').  A scope owned by a running claim, or split or quarantined, is
        refused and nothing of the group is written.  Returns the attempt tokens
        in entry order.  For callers that finish many small results at once (the
        served form): one synchronous commit per group instead of a file and its
        fsyncs each.
        """
        staged = []
        for scope, result in entries:
            if not passing(result):
                raise ValueError('a failed, overflowed or unverified result cannot become coverage')
            token = uuid.uuid4().hex
            work_id = digest({'config': self.identity, 'scope': scope})
            # level 1: a tenth larger than level 6 and several times faster
            blob = gzip_receipt({'config': self.identity, 'build': self.build, 'work_id': work_id, 'attempt': token,
                                 'result': result}, level=1)
            staged.append((encoded(scope).decode(), work_id, token, blob, hashlib.sha256(blob).hexdigest()))
        if not staged:
            return []
        with self.transaction():
            self.db.executemany("INSERT OR IGNORE INTO work(id,scope,state,parent) VALUES(?,?,'pending',NULL)",
                                [(work_id, scope) for scope, work_id, *_ in staged])
            ids = list({entry[1] for entry in staged})
            states: dict[str, str] = {}
            for k in range(0, len(ids), 500):
                part = ids[k:k + 500]
                query = 'SELECT id,state FROM work WHERE id IN (%s)' % ','.join('?' * len(part))
                states.update((row[0], row[1]) for row in self.db.execute(query, part))
            primary, repeats = [], []
            for _, work_id, token, _, sha in staged:
                state = states[work_id]
                if state == 'committed':
                    repeats.append((token, work_id, self.build, 'sqlite:' + token, sha))
                elif state in ('pending', 'failed'):
                    primary.append((token, work_id, sha))
                    states[work_id] = 'committed'
                else:
                    raise ValueError(f'scope is {state}; it cannot take a result')
            self.db.executemany('INSERT INTO receipt_blobs VALUES(?,?)', [(e[2], e[3]) for e in staged])
            self.db.executemany("INSERT INTO validations VALUES(?,?,?,'served repeat','verified',?,?)", repeats)
            self.db.executemany("INSERT INTO attempts VALUES(?,?,?,'committed',NULL)",
                                [(token, work_id, self.build) for token, work_id, _ in primary])
            self.db.executemany("UPDATE work SET state='committed',token=?,receipt=?,sha=? WHERE id=?",
                                [(token, 'sqlite:' + token, sha, work_id) for token, work_id, sha in primary])
        return [entry[2] for entry in staged]

    def split(self, work_id: str) -> list[str]:
        """Replace an unowned item by its two halves (by v), registered as its children."""
        with self.transaction():
            row = self.db.execute('SELECT state,scope FROM work WHERE id=?', (work_id,)).fetchone()
            if row is None or row['state'] not in ('pending', 'failed'):
                raise ValueError('only an unowned, uncommitted item can be split')
            scope = json.loads(row['scope'])
            first, end = int(scope['first']), int(scope['end'])
            if end - first < 2:
                raise ValueError('scope too small to split')
            middle = (first + end) // 2
            halves = []
            for lo, hi in ((first, middle), (middle, end)):
                child = {k: v for k, v in scope.items() if k != 'envelope'}   # envelopes are the caller's to recompute
                child.update(first=str(lo), end=str(hi))
                halves.append(child)
            self.db.execute("UPDATE work SET state='split',token=NULL WHERE id=?", (work_id,))
            return [self.register(child, work_id) for child in halves]

    def begin_validation(self, work_id: str, purpose: str) -> str:
        """Open a replay of committed work; replays never add coverage."""
        token = uuid.uuid4().hex
        with self.transaction():
            row = self.db.execute('SELECT state FROM work WHERE id=?', (work_id,)).fetchone()
            if row is None or row['state'] != 'committed':
                raise ValueError('a replay needs committed primary work')
            self.db.execute("INSERT INTO validations VALUES(?,?,?,?,'running',NULL,NULL)",
                            (token, work_id, self.build, purpose))
        return token

    def finish_validation(self, token: str, result: dict[str, Any]) -> None:
        row = self.db.execute('SELECT * FROM validations WHERE token=?', (token,)).fetchone()
        if row is None or row['state'] != 'running' or row['build'] != self.build:
            raise ValueError('no running replay under this token and build')
        path = self.directory / 'validations' / (token + '.json')
        sha = durable(path, {'config': self.identity, 'build': self.build, 'attempt': token,
                             'work_id': row['work_id'], 'purpose': row['purpose'], 'result': result})
        verdict = 'verified' if result.get('status') == 'PASS' and result.get('host_verified') else 'failed'
        with self.transaction():
            self.db.execute('UPDATE validations SET state=?,receipt=?,sha=? WHERE token=?',
                            (verdict, str(path.relative_to(self.directory)), sha, token))

    def quarantine(self, work_id: str, reason: str) -> None:
        """Withdraw an item (and its attempts) from coverage, e.g. after a replay disagreed."""
        with self.transaction():
            self.db.execute("UPDATE work SET state='quarantined' WHERE id=?", (work_id,))
            self.db.execute("UPDATE attempts SET state='quarantined',reason=? WHERE work_id=?", (reason, work_id))

    def reserve(self, kind: str, seconds: float, maximum: float) -> str:
        """Hold `seconds` of a persistent budget of `kind`; RuntimeError('BUDGET_EXHAUSTED')
        when held plus charged time would pass `maximum`."""
        if not 0 < seconds <= maximum:
            raise ValueError('reservation outside (0, maximum]')
        token = uuid.uuid4().hex
        with self.transaction():
            used = self.db.execute("SELECT coalesce(sum(CASE WHEN state='held' THEN reserved ELSE charged END),0) "
                                   'FROM budget WHERE kind=?', (kind,)).fetchone()[0]
            if used + seconds > maximum:
                raise RuntimeError('BUDGET_EXHAUSTED')
            self.db.execute("INSERT INTO budget VALUES(?,?,?,NULL,'held')", (token, kind, seconds))
        return token

    def settle(self, token: str, elapsed: float) -> None:
        """Close a reservation at the time actually used."""
        if elapsed < 0:
            raise ValueError('negative elapsed time')
        with self.transaction():
            row = self.db.execute('SELECT state FROM budget WHERE id=?', (token,)).fetchone()
            if row is None or row['state'] != 'held':
                raise ValueError('reservation unknown or already closed')
            self.db.execute("UPDATE budget SET state='closed',charged=? WHERE id=?", (elapsed, token))

    def receipt_bytes(self, name: str, token: str) -> bytes:
        """The stored bytes of a receipt: a file under the ledger's directory, or a
        compact blob ('sqlite:<token>')."""
        if name.startswith('sqlite:'):
            if name != 'sqlite:' + token:
                raise ValueError('receipt blob names another attempt')
            row = self.db.execute('SELECT data FROM receipt_blobs WHERE token=?', (token,)).fetchone()
            if row is None:
                raise ValueError('receipt blob missing')
            return row[0]
        return (self.directory / name).read_bytes()

    def open_receipt(self, name: str, token: str, sha: str) -> dict[str, Any]:
        """A receipt checked against its recorded SHA-256 and decoded."""
        data = self.receipt_bytes(name, token)
        if hashlib.sha256(data).hexdigest() != sha:
            raise ValueError('receipt changed since it was committed')
        try:
            return json.loads(gzip.decompress(data) if name.startswith('sqlite:') or name.endswith('.gz') else data)
        except (OSError, ValueError, EOFError) as error:
            raise ValueError('receipt unreadable') from error

    def audit(self) -> dict[str, Any]:
        """Check the database and every committed receipt (hash, configuration, item,
        attempt, build, verdict); return the item states, committed results,
        replays and budget rows."""
        if self.db.execute('PRAGMA quick_check').fetchone()[0] != 'ok':
            raise ValueError('corrupt SQLite ledger')
        states: dict[str, int] = {}
        results = []
        for row in self.db.execute('SELECT * FROM work'):
            states[row['state']] = states.get(row['state'], 0) + 1
            if row['state'] != 'committed':
                continue
            value = self.open_receipt(row['receipt'], row['token'], row['sha'])
            if (value.get('config'), value.get('work_id'), value.get('attempt')) != (self.identity, row['id'],
                                                                                    row['token']):
                raise ValueError('receipt names another configuration, item or attempt')
            attempt = self.db.execute('SELECT build,state FROM attempts WHERE token=?', (row['token'],)).fetchone()
            if attempt is None or attempt['build'] != value.get('build') or attempt['state'] != 'committed' \
                    or not passing(value.get('result', {})):
                raise ValueError('committed attempt has no valid provenance')
            results.append(value['result'])
        validations = []
        for row in self.db.execute('SELECT * FROM validations'):
            if row['receipt']:
                value = self.open_receipt(row['receipt'], row['token'], row['sha'])
                if (value.get('config'), value.get('work_id'), value.get('build')) != (self.identity, row['work_id'],
                                                                                      row['build']):
                    raise ValueError('replay receipt names another configuration, item or build')
            validations.append(dict(row))
        budget = [dict(row) for row in self.db.execute('SELECT * FROM budget')]
        return dict(status='PASS', states=states, committed_results=results, validations=validations, budget=budget,
                    scope='Only explicitly registered pilot tiles; enclosing numerical interval is not complete.')


def _writer_loop(connection: Any, directory: str, config: dict[str, Any], build: str) -> None:
    """The LedgerWriter child: open the ledger (WAL), then commit each group it is
    sent; groups that are waiting together share one transaction."""
    import pickle
    try:
        ledger = Ledger(Path(directory), config, build)
        ledger.db.execute('PRAGMA journal_mode=WAL')
    except BaseException as error:   # reported to the parent, which raises it
        connection.send(('error', f'{type(error).__name__}: {error}'))
        return
    connection.send(('ready', None))
    try:
        while True:
            message = connection.recv_bytes()
            if not message:
                break
            groups = [pickle.loads(message)]
            while connection.poll():
                more = connection.recv_bytes()
                if not more:
                    message = b''
                    break
                groups.append(pickle.loads(more))
            entries = [entry for _, group in groups for entry in group]
            try:
                tokens = ledger.commit_batch(entries)
                at = 0
                for ident, group in groups:
                    connection.send((ident, tokens[at:at + len(group)], None))
                    at += len(group)
            except Exception as error:
                for ident, _ in groups:
                    connection.send((ident, None, f'{type(error).__name__}: {error}'))
            if not message:
                break
    finally:
        ledger.db.close()


class LedgerWriter:
    """A Ledger whose batched commits (`commit_batch`) run in a child process.

    The caller sends each group of (scope, result) entries pickled through a
    pipe, which costs it a small part of encoding, compressing and committing
    the receipts itself; the child commits waiting groups together, one
    transaction each, and acknowledges every group with its attempt tokens (or
    the error that refused it).  A result is coverage once its group is
    acknowledged.  The child is started with the 'spawn' method and holds no
    device state.
    """

    def __init__(self, directory: Path, config: dict[str, Any], build: str) -> None:
        import multiprocessing
        context = multiprocessing.get_context('spawn')
        self._pipe, child = context.Pipe()
        self._process = context.Process(target=_writer_loop, args=(child, str(directory), config, build),
                                        name='ledger-writer', daemon=True)
        self._process.start()
        child.close()
        self.directory, self.config, self.build = directory, config, build
        self._next = 0
        self.pending: set[int] = set()
        status, detail = self._pipe.recv()
        if status != 'ready':
            self._process.join(10)
            raise ValueError('ledger writer could not open the ledger: ' + str(detail))

    def submit(self, entries: list[tuple[dict[str, Any], dict[str, Any]]]) -> int:
        """Queue one group; returns its id."""
        import pickle
        ident = self._next
        self._next += 1
        self._pipe.send_bytes(pickle.dumps((ident, entries), protocol=pickle.HIGHEST_PROTOCOL))
        self.pending.add(ident)
        return ident

    def acknowledged(self, wait: bool = False) -> list[tuple[int, list[str]]]:
        """Groups committed since the last call, as (id, tokens); with wait, block
        until at least one arrives (if any is pending).  A refused group raises
        ValueError."""
        out = []
        while self.pending and (self._pipe.poll() or (wait and not out)):
            ident, tokens, error = self._pipe.recv()
            self.pending.discard(ident)
            if error is not None:
                raise ValueError(f'ledger writer refused group {ident}: {error}')
            out.append((ident, tokens))
        return out

    def close(self) -> None:
        """Wait for every pending group, then stop the child."""
        try:
            while self.pending:
                self.acknowledged(wait=True)
        finally:
            if self._process.is_alive():
                self._pipe.send_bytes(b'')
            self._process.join(30)
            self._pipe.close()

    def __enter__(self) -> 'LedgerWriter':
        return self

    def __exit__(self, *exc: Any) -> None:
        self.close()
