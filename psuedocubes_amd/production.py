"""Single-owner production frontier and compressed receipts (`pseudocube/production.py:41-130`).

`Frontier` hands out consecutive compute batches of one v domain [0, end) to a
single sequential owner: the cursor moves only when the batch at the cursor is
committed, every commit's receipt is a gzip file written durably before the
transaction that makes it coverage, and a hit stops the campaign.
(`load_production`, the reference's reviewed-configuration gate, is part of
its deployment tooling and has no counterpart here.)
"""
from __future__ import annotations

import gzip
import hashlib
import json
import os
from pathlib import Path
from typing import Any
import uuid

from .accounting import Ledger, gzip_receipt, passing


def save_compressed(path: Path, value: dict[str, Any]) -> str:
    """Write gzip_receipt(value) to path durably (temporary file, fsync, rename,
    directory fsync); return the SHA-256 of the compressed bytes."""
    data = gzip_receipt(value)
    path.parent.mkdir(parents=True, exist_ok=True)
    temporary = path.with_name(f'{path.name}.{uuid.uuid4().hex}.tmp')
    with temporary.open('xb') as stream:
        stream.write(data)
        stream.flush()
        os.fsync(stream.fileno())
    os.replace(temporary, path)
    directory = os.open(path.parent, os.O_RDONLY)
    try:
        os.fsync(directory)
    finally:
        os.close(directory)
    return hashlib.sha256(data).hexdigest()


def check_subtiles(first: int, end: int, result: dict[str, Any]) -> int:
    """The number of subtile scopes of a result that tile [first, end) exactly, in order;
    ValueError on a gap, an overlap or a short cover."""
    cursor = first
    scopes = result['subtile_scopes']
    for scope in scopes:
        if int(scope['first']) != cursor or not cursor < int(scope['end']) <= end:
            raise ValueError('subtile scopes leave a gap or overlap')
        cursor = int(scope['end'])
    if cursor != end:
        raise ValueError('subtile scopes stop short of the batch end')
    return len(scopes)


class Frontier(Ledger):
    """One owner's contiguous frontier over v in [0, end) in batches of batch_width."""

    COUNTERS = ('commits', 'candidates', 'cubes', 'subtiles')

    def __init__(self, directory: Path, config: dict[str, Any], build: str, end: int, batch_width: int) -> None:
        super().__init__(directory, config, build)
        self.end, self.batch_width = end, batch_width
        with self.transaction():
            for key, value in (('frozen_build', build), ('domain_end', str(end)), ('batch_width', str(batch_width))):
                bound = self.meta(key)
                if bound is not None and bound != value:
                    raise ValueError(f'frontier was created with another {key}')
                self.db.execute('INSERT OR IGNORE INTO meta VALUES(?,?)', (key, value))
            for key, value in [('cursor', '0'), ('campaign_state', 'searching')] + [(k, '0') for k in self.COUNTERS]:
                self.db.execute('INSERT OR IGNORE INTO meta VALUES(?,?)', (key, value))
        if self.db.execute('PRAGMA quick_check').fetchone()[0] != 'ok':
            raise ValueError('corrupt production ledger')

    def state(self) -> dict[str, str]:
        return {row['key']: row['value'] for row in self.db.execute('SELECT key,value FROM meta')}

    def recover_stopped_owner(self, evidence: str) -> int:
        """Release every running batch, given evidence that its owner has stopped."""
        if not evidence:
            raise ValueError('evidence that the owner has stopped is required')
        running = self.db.execute("SELECT id,token FROM work WHERE state='running'").fetchall()
        for row in running:
            self.reclaim(row['id'], row['token'], evidence)
        return len(running)

    def next_batch(self) -> tuple[str, str, int, int] | None:
        """(work id, token, first, end) of the batch at the cursor, claimed; None when
        the campaign has stopped or the domain is done."""
        state = self.state()
        first = int(state['cursor'])
        if state['campaign_state'] != 'searching' or first >= self.end:
            return None
        end = min(first + self.batch_width, self.end)
        work_id = self.register({'kind': 'production-v-window', 'first': str(first), 'end': str(end)})
        return work_id, self.claim(work_id), first, end

    def complete(self, work_id: str, token: str, first: int, end: int, result: dict[str, Any]) -> None:
        """Commit the batch at the cursor and advance it."""
        if not passing(result):
            self.fail(work_id, token, 'invalid production receipt')
            raise ValueError('a failed, overflowed or unverified result cannot become coverage')
        subtiles = check_subtiles(first, end, result)
        path = self.directory / 'outputs' / (token + '.json.gz')
        name = str(path.relative_to(self.directory))
        sha = save_compressed(path, {'config': self.identity, 'build': self.build, 'work_id': work_id,
                                     'attempt': token, 'first': str(first), 'end': str(end), 'result': result})
        with self.transaction():
            self._check_owner(work_id, token)
            state = self.state()
            scope = json.loads(self.db.execute('SELECT scope FROM work WHERE id=?', (work_id,)).fetchone()[0])
            if int(state['cursor']) != first or (int(scope['first']), int(scope['end'])) != (first, end):
                raise ValueError('the frontier moved or the scope changed')
            self.db.execute("UPDATE attempts SET state='committed' WHERE token=?", (token,))
            self.db.execute("UPDATE work SET state='committed',receipt=?,sha=? WHERE id=?", (name, sha, work_id))
            phase = ('candidate-found' if result['cpu_confirmed_hits'] else
                     'complete' if end == self.end else 'searching')
            added = (1, int(result['valid_focused_candidates']), int(result['genuine_cubes']), subtiles)
            updates = {key: str(int(state[key]) + n) for key, n in zip(self.COUNTERS, added)}
            updates.update(cursor=str(end), last_receipt=name, last_sha=sha, campaign_state=phase)
            self.db.executemany('INSERT OR REPLACE INTO meta VALUES(?,?)', updates.items())

    def _open(self, name: str, sha: str) -> dict[str, Any]:
        data = (self.directory / name).read_bytes()
        if hashlib.sha256(data).hexdigest() != sha:
            raise ValueError('production receipt changed since it was committed')
        try:
            return json.loads(gzip.decompress(data))
        except (OSError, ValueError, EOFError) as error:
            raise ValueError('production receipt unreadable') from error

    def validate_last_receipt(self) -> None:
        """The newest receipt is intact and ends at the cursor."""
        state = self.state()
        if not state.get('last_receipt'):
            return
        value = self._open(state['last_receipt'], state['last_sha'])
        if (value['config'], value['build'], value['end']) != (self.identity, self.build, state['cursor']):
            raise ValueError('newest receipt does not end at the frontier')

    def audit_production(self) -> dict[str, Any]:
        """Replay the committed receipts in commit order: contiguous from 0 to the cursor."""
        cursor = commits = 0
        for row in self.db.execute("SELECT * FROM work WHERE state='committed' ORDER BY rowid"):
            value = self._open(row['receipt'], row['sha'])
            if value['config'] != self.identity or value['build'] != self.build:
                raise ValueError('committed receipt of another configuration or build')
            if int(value['first']) != cursor or value['attempt'] != row['token']:
                raise ValueError('committed receipts leave a gap or overlap')
            cursor = int(value['end'])
            commits += 1
        if cursor != int(self.state()['cursor']):
            raise ValueError('frontier differs from the committed receipts')
        return dict(status='PASS', commits=commits, committed_v_end=str(cursor), domain_end=str(self.end),
                    full_interval_complete=cursor == self.end, independent_arithmetic_replay=False)
