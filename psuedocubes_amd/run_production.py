"""Per-subtile batch aggregation with overflow splitting, and the production smoke canary.

Drop-in for `compute_batch` of the reference's tools/run_production.py:51-83:
each subtile runs through `Matcher.tile`; an OverflowError splits the subtile
at its integer middle and replays both halves; every CPU-confirmed hit is
re-verified by the standalone verifier.  The row it returns is the baseline
that `reproduce --compare-baseline` checks the batched engine against.
"""
from __future__ import annotations

from pathlib import Path
import time
from typing import Any, Callable

from .accounting import digest
from .focus import Config, Tile
from .gpu import GPUFocus, Matcher, Runtime
from .predicate import verify


def compute_batch(matcher: Matcher, af: GPUFocus, bf: GPUFocus, first: int, end: int, width: int,
                  notify: Callable[[str], None] = lambda message: None) -> dict[str, Any]:
    started = time.monotonic()
    scopes: list[dict[str, Any]] = []
    hits: list[dict[str, Any]] = []
    values: list[str] = []
    groups = [0] * len(matcher.filter_groups)
    counts = {key: 0 for key in ('focus_u', 'focus_v', 'cartesian_opportunities', 'valid_focused_candidates',
                                 'genuine_cubes', 'local_survivors', 'cube_checks')}
    timings = {key: 0.0 for key in ('kernel_seconds', 'focus_enumeration_seconds', 'transfer_seconds',
                                    'verification_seconds')}
    overflow_splits: list[dict[str, str]] = []

    def execute(tile: Tile) -> None:
        try:
            row = matcher.tile(tile, af, bf)
        except OverflowError:
            if tile.end - tile.first < 2:
                raise
            overflow_splits.append(dict(first=str(tile.first), end=str(tile.end)))
            left, right = tile.split()
            execute(left)
            execute(right)
            return
        if row['status'] != 'PASS' or not row['host_verified'] or row['cpu_rejected_reports']:
            raise ArithmeticError('unverified GPU tile')
        for n in row['cpu_confirmed_hits']:
            witness = verify(int(n), matcher.config.p)
            if not witness['valid_pseudocube']:
                raise ArithmeticError('standalone candidate verification failed')
            hits.append(witness)
        scopes.append(dict(first=str(tile.first), end=str(tile.end),
                           valid_focused_candidates=row['valid_focused_candidates'],
                           group_survivors=row['group_survivors'], local_survivors=row['local_survivors'],
                           output_sha256=digest(row['all_values'])))
        values.extend(row['all_values'])
        for key in counts:
            counts[key] += int(row[key])
        for key in timings:
            timings[key] += float(row[key])
        for i, value in enumerate(row['group_survivors']):
            groups[i] += value
        notify('WATCHDOG=1')

    for start in range(first, end, width):
        execute(Tile(start, min(start + width, end)))
    if len(values) != len(set(values)):
        raise ArithmeticError('duplicate output across disjoint subtiles')
    return dict(status='PASS', host_verified=True, **counts, **timings, subtile_scopes=scopes,
                filter_groups=matcher.filter_groups, group_survivors=groups, all_values=values,
                cpu_confirmed_hits=hits, overflow_splits=overflow_splits,
                elapsed_seconds=time.monotonic() - started,
                claim='Executed primary coverage; independent replay required before any minimum claim.')


def smoke_canary(runtime: Runtime, directory: Path, config: Config, af: GPUFocus, bf: GPUFocus,
                 build: str) -> dict[str, Any]:
    """The production smoke canary (`tools/run_production.py:93-122`), recorded durably in
    directory/canary.json; raises ArithmeticError on the first disagreement.

    1. Subtiles [0, 1e8) and [1e8, 2e8): the legacy per-subtile path (`Matcher.tile`,
       the direct per-position kernel) against the residue-bitset batch engine
       (`MaskedMatcher` on the same one-subtile batch): values, valid positions,
       group survivors and genuine cubes.  (The reference compares two layouts of
       one kernel there; here the two are different kernels.)
    2. High-magnitude canaries: x in (n - 1, n] for n = 674441580981249129037406633
       at p = 613 and the cubes 1000000007^3, 1000000009^3 at p = 619 give exactly [n].
    3. Ledger resume: a two-batch Frontier of 1e8-wide batches; each batch through
       compute_batch at width 5e7 (a partition of the subtile) must equal the
       subtile's row from 1, with the frontier closed and reopened between the two
       batches, its last receipt validated and the whole ledger audited.
    4. One full production batch (64 subtiles of 1e8) through compute_batch.
    """
    from .accounting import durable
    from .batch_experiment import MaskedMatcher, prepare
    from .focus import A619, B619
    from .production import Frontier
    fields = ('all_values', 'valid_focused_candidates', 'group_survivors', 'genuine_cubes')
    direct, masked = Matcher(runtime, config), MaskedMatcher(runtime, config)
    rows: list[dict[str, Any]] = []
    for first in (0, 100_000_000):
        tile = direct.tile(Tile(first, first + 100_000_000), af, bf)
        batch = masked.run(prepare(runtime, config, af, bf, first, first + 100_000_000, 100_000_000))
        scope = batch['subtile_scopes'][0]
        other = dict(all_values=batch['all_values'], valid_focused_candidates=scope['valid_focused_candidates'],
                     group_survivors=scope['group_survivors'], genuine_cubes=batch['genuine_cubes'])
        for key in fields:
            if tile[key] != other[key]:
                raise ArithmeticError('production-geometry canary mismatch: ' + key)
        rows.append(dict(first=str(first), tile=tile, batch=batch))
    for p, n in ((613, 674441580981249129037406633), (619, 1000000007**3), (619, 1000000009**3)):
        narrow = Config(p, n - 1, n, A619, B619)
        v = narrow.canonical(n)[1]
        row = Matcher(runtime, narrow).tile(Tile(max(0, v - 5_000_000), min(B619, v + 5_000_001)), af, bf)
        if row['all_values'] != [str(n)]:
            raise ArithmeticError('production high-magnitude canary failed')
        rows.append(dict(canary=str(n), result=row))
    ledger_dir = directory / 'canary-ledger'
    purpose = {'purpose': 'production-canary'}
    audited: dict[str, Any] = {}
    for k in range(2):
        frontier = Frontier(ledger_dir, purpose, build, 200_000_000, 100_000_000)
        if k:
            frontier.validate_last_receipt()
        job = frontier.next_batch()
        if job is None or job[2] != 100_000_000 * k:
            raise ArithmeticError('canary ledger did not resume at the next batch')
        result = compute_batch(direct, af, bf, job[2], job[3], 50_000_000)
        for key in fields:
            if result[key] != rows[k]['tile'][key]:
                raise ArithmeticError(('resumed ' if k else '') + 'subtile partition mismatch: ' + key)
        frontier.complete(*job, result)
        if k:
            audited = frontier.audit_production()
        frontier.db.close()
    if not audited.get('full_interval_complete'):
        raise ArithmeticError('canary ledger incomplete')
    full = compute_batch(direct, af, bf, 0, 6_400_000_000, 100_000_000)
    if len(full['subtile_scopes']) < 64:
        raise ArithmeticError('full production batch incomplete')
    record = dict(status='PASS', cases=rows, ledger=audited, full_batch=full, novel_production_coverage=0)
    durable(directory / 'canary.json', record)
    return record
