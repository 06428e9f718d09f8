"""Lease execution and compact, reproducible lease receipts.

Drop-in for the reference's `pseudocube/optimized.py:15-90` (`PROFILE`,
`BATCH_WIDTH`, `LEASE_WIDTH`, `Accumulator`, `compute_lease`) and the
receipt check `validate_result` of `pseudocube/multi.py:36-67`.

`compute_lease` hands the lease to libpcube in device passes of consecutive
batches (8, then 64 at a time, the last passes halving down to 8), each one
call on a worker thread: both focus windows of the pass in one paired
launch, every batch's subtile bounds in one launch, one host synchronisation,
then every batch's bitset build and match queued on the device stream and one
copy of the pass's counters at the end.  The caller's thread folds one pass's
batches into the receipt while the device computes the next: each from its
totals, its values checked natively in one call per pass and its scope
records formatted in one call.  The receipt is field for field that of
`prepare` + `MaskedMatcher.run` per batch, so the transcript is unchanged.

Also here: `allocated_bytes` and `ledger_settings`, the storage helpers of the
reference's lease runner (`pseudocube/optimized.py:93-113`).
"""
from __future__ import annotations

from contextlib import closing
import ctypes
import hashlib
import json
import os
from pathlib import Path
import sqlite3
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from typing import Any, Callable, Iterator, NamedTuple

import numpy as np

from . import native
from .accounting import digest, encoded
from .batch_experiment import (EMPTY_DIGEST, OUTPUT_CAP, MaskedMatcher, batch_row, check_values, prepare,
                               tile_geometry, verify_pass_from)
from .gpu import GPUFocus

PROFILE: dict[str, Any] = dict(engine='residue-bitset-batch-v1', tile_width=100_000_000,
                               tiles_per_compute_batch=64, compute_batches_per_lease=256,
                               receipt_storage='sqlite-gzip', receipt_format='compact-batch-v1',
                               authorization='Deploy to production')
BATCH_WIDTH = PROFILE['tile_width'] * PROFILE['tiles_per_compute_batch']
LEASE_WIDTH = BATCH_WIDTH * PROFILE['compute_batches_per_lease']
COUNTS = ('focus_u', 'focus_v', 'cartesian_opportunities', 'valid_focused_candidates',
          'genuine_cubes', 'local_survivors', 'cube_checks')
TIMINGS = ('kernel_seconds', 'focus_enumeration_seconds', 'transfer_seconds',
           'verification_seconds', 'prefilter_build_seconds')


def validate_result(first: int, end: int, result: dict[str, Any]) -> None:
    """Reject rows that cannot become coverage (`pseudocube/multi.py:36-67`)."""
    if (result.get('status') != 'PASS' or result.get('host_verified') is not True or result.get('overflow')
            or result.get('kernel_error')):
        raise ValueError('unverified execution cannot become coverage')
    cursor = first
    for tile in result['subtile_scopes']:
        if int(tile['first']) != cursor or not cursor < int(tile['end']) <= end:
            raise ValueError('gap or overlap in execution subtiles')
        cursor = int(tile['end'])
    if cursor != end:
        raise ValueError('incomplete execution receipt')
    for key in ('valid_focused_candidates', 'genuine_cubes'):
        if int(result[key]) < 0:
            raise ValueError('negative execution count')
    if result.get('receipt_format') == 'compact-batch-v1':
        width = int(result['tile_width'])
        batch = int(result['compute_batch_width'])
        if width <= 0 or batch != 64 * width or len(result['subtile_scopes']) != 1:
            raise ValueError('invalid compact execution geometry')
        if int(result['subtiles_executed']) != (end - first + width - 1) // width:
            raise ValueError('compact subtile count mismatch')
        if int(result['compute_batches']) != (end - first + batch - 1) // batch:
            raise ValueError('compact batch count mismatch')
        transcript = result['subtile_scopes'][0]['transcript_sha256']
        if len(transcript) != 64 or any(c not in '0123456789abcdef' for c in transcript):
            raise ValueError('invalid execution transcript hash')
        values = result['all_values']
        if len(values) != len(set(values)) or len(values) != int(result['genuine_cubes']) + len(result['cpu_confirmed_hits']):
            raise ValueError('compact output count mismatch')
        if result['subtile_scopes'][0]['output_sha256'] != digest(values):
            raise ValueError('compact output hash mismatch')
    elif 'subtiles_executed' in result or 'receipt_format' in result:
        raise ValueError('unknown receipt format')


def allocated_bytes(directory: Path) -> int:
    """Disk blocks allocated under a ledger directory, in bytes (`pseudocube/optimized.py:93-106`).
    A file that vanishes during the scan is tolerated only if it is a temporary
    receipt (`.tmp`), which an atomic rename may remove at any time."""
    total = directory.stat().st_blocks * 512
    with os.scandir(directory) as entries:
        for entry in entries:
            try:
                if entry.is_dir(follow_symlinks=False):
                    total += allocated_bytes(Path(entry.path))
                else:
                    total += entry.stat(follow_symlinks=False).st_blocks * 512
            except FileNotFoundError:
                if not entry.name.endswith('.tmp'):
                    raise
    return total


def ledger_settings(directory: Path) -> tuple[str, int, dict[str, Any]]:
    """(active build, batch width, runtime profile) of a shared ledger, read only
    (`pseudocube/optimized.py:109-113`)."""
    uri = (directory / 'coverage.sqlite3').as_uri() + '?mode=ro'
    with closing(sqlite3.connect(uri, uri=True)) as db:
        state = dict(db.execute('SELECT key,value FROM meta').fetchall())
    return (state.get('active_build', state['frozen_build']), int(state['batch_width']),
            json.loads(state.get('runtime_profile', '{}')))


class BatchTotals(NamedTuple):
    """A batch by the totals its row would hold: focus_u, focus_v and
    cartesian_opportunities from its bounds and window, valid_focused_candidates
    and group_survivors summed over its counters, and its verified values
    (decimal strings, subtile after subtile, each subtile's sorted) with the
    noncubes' witness records; a batch that emitted nothing has neither."""
    subtiles: int
    span: tuple[int, int]        # (first subtile's first, last subtile's end)
    focus_u: int
    focus_v: int
    opportunities: int
    valid: int
    groups: list[int]
    filter_groups: list[list[int]]
    timings: dict[str, float]
    values: list[str] = []
    hits: list[dict[str, Any]] = []


class Accumulator:
    """Folds batch rows into one compact lease row (`pseudocube/optimized.py:26-78`).

    transcript = SHA-256 over the canonical encoding of every batch's
    subtile_scopes, in batch order.
    """

    def __init__(self, first: int, end: int) -> None:
        if not 0 <= first < end or end - first > LEASE_WIDTH:
            raise ValueError('invalid optimized lease')
        self.first, self.end, self.cursor = first, end, first
        self.counts = dict.fromkeys(COUNTS, 0)
        self.timings = dict.fromkeys(TIMINGS, 0.0)
        self.values: list[str] = []
        self.hits: list[dict[str, Any]] = []
        self.groups: list[int] = []
        self.filters: list[list[int]] | None = None
        self.transcript = hashlib.sha256()
        self.batches = self.subtiles = 0

    def add(self, first: int, end: int, row: dict[str, Any], encoded_scopes: bytes | None = None) -> None:
        """Fold one batch row.  encoded_scopes, when given, is the row's
        encoded(row['subtile_scopes']) already formed (the pipelined lease's)."""
        if first != self.cursor or end != min(first + BATCH_WIDTH, self.end):
            raise ValueError('gap, overlap or changed compute batch geometry')
        validate_result(first, end, row)
        width = PROFILE['tile_width']
        expected = [(v, min(v + width, end)) for v in range(first, end, width)]
        actual = [(int(s['first']), int(s['end'])) for s in row['subtile_scopes']]
        if actual != expected:
            raise ValueError('changed arithmetic subtile geometry')
        if self.filters is None:
            self.filters = [list(group) for group in row['filter_groups']]
            self.groups = [0] * len(self.filters)
        if row['filter_groups'] != self.filters or len(row['group_survivors']) != len(self.groups):
            raise ValueError('residue filters changed during execution')
        for key in COUNTS:
            self.counts[key] += int(row[key])
        for key in TIMINGS:
            self.timings[key] += float(row.get(key, 0.0))
        for i, count in enumerate(row['group_survivors']):
            self.groups[i] += int(count)
        self.values.extend(row['all_values'])
        self.hits.extend(row['cpu_confirmed_hits'])
        self.transcript.update(encoded_scopes if encoded_scopes is not None else encoded(row['subtile_scopes']))
        self.batches += 1
        self.subtiles += len(actual)
        self.cursor = end

    def add_totals(self, first: int, end: int, batch: BatchTotals, encoded_scopes: bytes) -> None:
        """Fold a batch from its totals: the checks and sums of add() on the row
        batch_row builds for it, without building the row."""
        if first != self.cursor or end != min(first + BATCH_WIDTH, self.end):
            raise ValueError('gap, overlap or changed compute batch geometry')
        width = PROFILE['tile_width']
        if batch.span != (first, end) or batch.subtiles != (end - first + width - 1) // width:
            raise ValueError('changed arithmetic subtile geometry')
        if self.filters is None:
            self.filters = [list(group) for group in batch.filter_groups]
            self.groups = [0] * len(self.filters)
        if batch.filter_groups != self.filters or len(batch.groups) != len(self.groups):
            raise ValueError('residue filters changed during execution')
        counts = self.counts
        counts['focus_u'] += batch.focus_u
        counts['focus_v'] += batch.focus_v
        counts['cartesian_opportunities'] += batch.opportunities
        counts['valid_focused_candidates'] += batch.valid
        if batch.values:
            n = len(batch.values)
            counts['local_survivors'] += n
            counts['cube_checks'] += n
            counts['genuine_cubes'] += n - len(batch.hits)
            self.values.extend(batch.values)
            self.hits.extend(batch.hits)
        for key in TIMINGS:
            self.timings[key] += batch.timings.get(key, 0.0)
        self.groups = [a + b for a, b in zip(self.groups, batch.groups)]
        self.transcript.update(encoded_scopes)
        self.batches += 1
        self.subtiles += batch.subtiles
        self.cursor = end

    def finish(self, seconds: float) -> dict[str, Any]:
        if self.cursor != self.end:
            raise ValueError('unfinished lease cannot become coverage')
        row = dict(status='PASS', host_verified=True, receipt_format=PROFILE['receipt_format'],
                   tile_width=PROFILE['tile_width'], compute_batch_width=BATCH_WIDTH,
                   compute_batches=self.batches, subtiles_executed=self.subtiles,
                   subtile_scopes=[dict(first=str(self.first), end=str(self.end),
                                        transcript_sha256=self.transcript.hexdigest(),
                                        output_sha256=digest(self.values))],
                   **self.counts, **self.timings, group_survivors=self.groups, filter_groups=self.filters,
                   all_values=self.values, cpu_confirmed_hits=self.hits, overflow_splits=[], elapsed_seconds=seconds,
                   claim='Executed primary coverage; independent arithmetic replay required before any minimum claim.')
        validate_result(self.first, self.end, row)
        return row


_records = threading.local()


def scope_record_bytes(bounds: np.ndarray, counters: np.ndarray, ng: int, local: np.ndarray | None,
                       shas: bytes) -> bytes:
    """encoded(subtile_scopes) of one batch, written by the library (`pcube_scope_records`).

    bounds: [ntiles, 2] (first, end); counters: [ntiles, stride] rows
    [valid, group_1..group_ng, ...]; local: local survivors per subtile (None:
    all 0); shas: 64 hex digits per subtile.  Every array C-contiguous uint64.
    """
    nt, stride = counters.shape
    need = 256 + nt * (400 + 24 * ng)
    buf = getattr(_records, 'buf', None)
    if buf is None or len(buf) < need:
        buf = _records.buf = ctypes.create_string_buffer(need)
        _records.length = ctypes.c_uint64()
    code = native.lib().pcube_scope_records(nt, native.ptr64(bounds), native.ptr64(counters), stride, ng,
                                            native.ptr64(local) if local is not None else None, shas, buf,
                                            len(buf), ctypes.byref(_records.length))
    native.check(code, 'pcube_scope_records')
    return buf.raw[:_records.length.value]


def pass_record_bytes(per_batch: list[int], bounds: np.ndarray, counters: np.ndarray, ng: int, local: np.ndarray,
                      shas: bytes) -> list[bytes]:
    """scope_record_bytes of every batch of a pass in one library call
    (`pcube_scope_records_many`): batch b holds per_batch[b] consecutive subtiles
    of bounds, counters, local and shas (each over the whole pass)."""
    nt, stride = counters.shape
    need = 256 * len(per_batch) + nt * (400 + 24 * ng)
    buf = getattr(_records, 'many', None)
    if buf is None or len(buf) < need:
        buf = _records.many = ctypes.create_string_buffer(need)
    sizes = np.array(per_batch, dtype=np.uint32)
    offsets = np.zeros(len(per_batch) + 1, dtype=np.uint64)
    native.check(native.lib().pcube_scope_records_many(len(per_batch), native.ptr32(sizes), native.ptr64(bounds),
                                                       native.ptr64(counters), stride, ng,
                                                       native.ptr64(np.ascontiguousarray(local, dtype=np.uint64)),
                                                       shas, buf, len(buf), native.ptr64(offsets)),
                 'pcube_scope_records_many')
    ends = offsets.tolist()
    raw = ctypes.string_at(buf, ends[-1])
    return [raw[a:b] for a, b in zip(ends[:-1], ends[1:])]


LEASE_CHUNK = 64   # most batches per device pass of a pipelined lease
LEASE_EDGE = 8     # batches of its first and last passes


def tapered_sizes(n: int, edge: int, chunk: int) -> list[int]:
    """Sizes of consecutive device passes over n items, each at most chunk: a first
    pass of edge items, so the host starts on results early; passes of chunk (few
    pass boundaries, at each of which the device waits for the next call); and a
    tail halving down to edge, so that the host's work on one pass takes less time
    than the device takes for the next and little is left once the device is done."""
    edge = max(1, min(edge, chunk))
    if n <= 2 * edge:
        return [n] if n <= edge else [edge, n - edge]
    rest = n - edge
    tail = []
    size = edge
    while rest > 0 and size < chunk:
        tail.append(min(size, rest))
        rest -= tail[-1]
        size *= 2
    middle = []
    while rest > 0:
        middle.append(min(chunk, rest))
        rest -= middle[-1]
    return [edge] + middle + tail[::-1]


def pass_sizes(n: int) -> list[int]:
    """Batches per device pass for a lease of n batches (tapered_sizes with
    LEASE_EDGE and LEASE_CHUNK)."""
    return tapered_sizes(n, LEASE_EDGE, LEASE_CHUNK)


def pass_totals(per_batch: list[int], bounds: np.ndarray, counters: np.ndarray,
                ng: int) -> tuple[list[bool], list[int], list[int], list[list[int]]]:
    """Per-batch totals of a device pass whose batch b has per_batch[b]
    consecutive subtiles: whether it emitted nothing and saw no underflow,
    focus_u (sum of u_end - u_first), cartesian_opportunities (sum of
    (u_end - u_first) (v_end - v_first)) and the column sums [valid,
    group_1..group_ng], all as batch_row forms them.  bounds: [ntiles, 4]
    index rows; counters: [ntiles, ng + 3].  A pass with a reversed bound pair
    marks no batch quiet (its rows go through batch_row)."""
    heads = np.cumsum([0] + per_batch[:-1])
    du = bounds[:, 1] - bounds[:, 0]
    ordered = bool((bounds[:, 1] >= bounds[:, 0]).all() and (bounds[:, 3] >= bounds[:, 2]).all())
    flagged = np.add.reduceat((counters[:, ng + 1:ng + 3] != 0).any(axis=1).astype(np.uint32), heads)
    quiet = (flagged == 0).tolist() if ordered else [False] * len(per_batch)
    return (quiet, np.add.reduceat(du, heads).tolist(),
            np.add.reduceat(du * (bounds[:, 3] - bounds[:, 2]), heads).tolist(),
            np.add.reduceat(counters[:, :ng + 1], heads, axis=0).tolist())


def _pass_inputs(matcher: MaskedMatcher, af: GPUFocus, bf: GPUFocus, batches: list[tuple[int, int]],
                 geometry: np.ndarray, counts: list[int], window_cap: int | None) -> tuple[Any, ...]:
    """The host arrays of one device pass over consecutive batches and the
    arguments of its `pcube_lease` call (their windows, bounds, matches and
    counters, with one host synchronisation).  geometry: the pass's subtiles as
    [n, 4] (v_first, v_end, u_min, u_end) rows; counts: subtiles per batch."""
    width = PROFILE['tile_width']
    first, end = batches[0][0], batches[-1][1]
    tiles = np.ascontiguousarray(geometry, dtype=np.uint64).reshape(-1)
    per_batch = np.array(counts, dtype=np.uint32)
    ntot = int(per_batch.sum())
    ng = len(matcher.filter_groups)
    if window_cap is None:
        # expected coordinates in each window of the pass, with margin
        u_span = int(geometry[-1, 3]) - int(geometry[0, 2])
        expect = max(u_span * af.focus.count / af.focus.modulus, (end - first) * bf.focus.count / bf.focus.modulus)
        window_cap = int(expect * 1.1) + 1_000_000
    bounds = np.zeros(ntot * 4, dtype=np.uint64)
    sizes = np.zeros(len(batches) * 2, dtype=np.uint64)
    counters = np.zeros(ntot * (ng + 3), dtype=np.uint64)
    values = np.zeros(2 * (1 << 20), dtype=np.uint64)
    seconds = (ctypes.c_double * 4)()
    args = (matcher.engine.handle, af.handle, bf.handle, len(batches), native.ptr32(per_batch),
            native.ptr64(tiles), width, 8192, window_cap, native.ptr64(bounds), native.ptr64(sizes),
            native.ptr64(counters), native.ptr64(values), values.size // 2, seconds)
    arrays = (tiles, per_batch, bounds.reshape(ntot, 4), sizes, counters.reshape(ntot, ng + 3), values, seconds)
    return arrays, args


def _run_pass(check: Callable[[], None], call: Callable[..., int], args: tuple[Any, ...]) -> float:
    """The worker's part of a pass: the library call between two budget checks.
    Both check the library on this thread: its error text is per thread, and
    the folding thread makes no call that waits for the library's lock."""
    check()
    started = time.perf_counter()
    native.check(call(*args), 'pcube_lease')
    seconds = time.perf_counter() - started
    check()
    return seconds


def lease_rows(matcher: MaskedMatcher, af: GPUFocus, bf: GPUFocus, first: int, end: int,
               batch_width: int = BATCH_WIDTH,
               window_cap: int | None = None,
               quiet: bool = False) -> Iterator[tuple[int, int, dict[str, Any] | BatchTotals, bytes]]:
    """Every batch row of [first, end), from pipelined device passes.

    The batches go to the device in groups (pass_sizes), each group in one
    pass (`pcube_lease`) on a worker thread; the caller folds one group's rows
    while the device computes the next.  Each row equals
    `MaskedMatcher.run(prepare(..., start, stop))` for its batch in every
    field except the timings, which are its pass's phases divided evenly over
    the pass's batches.  With each row come the bytes of
    encoded(row['subtile_scopes']), formatted by the library from the same
    counters and bounds.  With quiet, a batch comes as its BatchTotals instead
    of a row (pass_rows).
    """
    config = matcher.config
    width = PROFILE['tile_width']
    starts = list(range(first, end, batch_width))
    groups = []
    at = 0
    for size in pass_sizes(len(starts)):
        groups.append([(s, min(s + batch_width, end)) for s in starts[at:at + size]])
        at += size
    pool = ThreadPoolExecutor(max_workers=1)
    call = native.lib().pcube_lease
    # the worker takes the interpreter lock only between passes, to return from one
    # call and make the next; a short switch interval hands it over at once while
    # this thread folds, where the default (5 ms) would leave the device idle
    interval = sys.getswitchinterval()
    sys.setswitchinterval(min(interval, 1e-4))
    try:
        pending = []
        # the whole lease's subtile geometry in one library call (an array, no Python
        # tuples), so every pass is submitted at once and the worker never waits for
        # the next one's inputs
        geometry, per_batch = tile_geometry(config, [span for batches in groups for span in batches], width)
        at = row = 0
        for batches in groups:
            counts = per_batch[at:at + len(batches)]
            nt = sum(counts)
            arrays, args = _pass_inputs(matcher, af, bf, batches, geometry[row:row + nt], counts, window_cap)
            pending.append((batches, counts, arrays, pool.submit(_run_pass, matcher.runtime.check, call, args)))
            at += len(batches)
            row += nt
        for batches, counts, arrays, future in pending:
            yield from pass_rows(matcher, batches, counts, arrays, future.result(), quiet)
    finally:
        # a caller that stops early (an error in its fold) leaves no pass behind
        pool.shutdown(wait=True, cancel_futures=True)
        sys.setswitchinterval(interval)


def pass_rows(matcher: Any, batches: list[tuple[int, int]], counts: list[int], arrays: tuple[Any, ...],
              device_seconds: float,
              quiet: bool) -> Iterator[tuple[int, int, dict[str, Any] | BatchTotals, bytes]]:
    """The rows of one finished device pass, batch by batch, with each row's
    encoded(subtile_scopes).

    matcher: config, filter_groups, kernel, layout (a MaskedMatcher's);
    arrays: the pass's host arrays (_pass_inputs); counts: subtiles per batch.

    With quiet, a pass that saw no underflow and no overflow has every emitted
    value checked natively in one call (check_values), its records formatted in
    one call (pcube_scope_records_many) and its totals summed over the counters
    (pass_totals), and each batch comes as its BatchTotals.  Otherwise, and for a
    pass whose values fail a check, each batch's row is built by batch_row with
    the Python predicate, so that the first failing batch raises as the
    reference's loop does (a batch that emitted nothing still comes as its
    totals with quiet).
    """
    config = matcher.config
    ng = len(matcher.filter_groups)
    cap = OUTPUT_CAP
    tiles, _, bounds, sizes, counters, values, seconds = arrays
    n = len(batches)
    timings = dict(kernel_seconds=seconds[2] / n, focus_enumeration_seconds=seconds[0] / n,
                   transfer_seconds=max(0.0, device_seconds - seconds[0] - seconds[2]) / n)
    tile_ends = tiles.reshape(-1, 4)[:, :2].copy()   # (v_first, v_end) per subtile, contiguous
    ntot = len(tile_ends)
    listed = counters[:, ng + 1]
    kept = np.minimum(listed, np.uint64(cap))
    ordered = bool((bounds[:, 1] >= bounds[:, 0]).all() and (bounds[:, 3] >= bounds[:, 2]).all())
    clean = ordered and not counters[:, ng + 2].any() and int(listed.max(initial=0)) <= cap
    if quiet:
        quiet_of, focus_of, pairs_of, sums_of = pass_totals(counts, bounds, counters, ng)
        shares = dict(timings, verification_seconds=0.0, prefilter_build_seconds=seconds[1] / n)
        filters = [list(group) for group in matcher.filter_groups]
    checked = None
    if clean and (quiet or kept.any()):
        try:
            checked = check_values(config, tile_ends[:, 0], tile_ends[:, 1], np.full(ntot, config.p, dtype=np.uint32),
                                   kept, values)
        except ArithmeticError:
            checked = None   # the per-batch path below raises at the first failing batch
    if quiet and checked is not None:
        index, strings_of, digests, hits_of = checked
        shas = bytearray(EMPTY_DIGEST.encode() * ntot)
        for s, text_digest in zip(index, digests):
            shas[64 * s:64 * s + 64] = text_digest.encode()
        records = pass_record_bytes(counts, tile_ends, counters, ng, kept, bytes(shas))
        heads = np.cumsum([0] + counts).tolist()
        batch_of = (np.searchsorted(heads, index, side='right') - 1).tolist() if index else []
        found = 0
        t0 = 0
        for b, (s, e) in enumerate(batches):
            nt = counts[b]
            batch_values: list[str] = []
            batch_hits: list[dict[str, Any]] = []
            while found < len(index) and batch_of[found] == b:
                batch_values += strings_of[found]
                batch_hits += hits_of[found]
                found += 1
            totals = BatchTotals(nt, (int(tile_ends[t0, 0]), int(tile_ends[t0 + nt - 1, 1])), focus_of[b],
                                 int(sizes[2 * b + 1]), pairs_of[b], sums_of[b][0], sums_of[b][1:], filters, shares,
                                 batch_values, batch_hits)
            yield s, e, totals, records[b]
            t0 += nt
        return
    verified = None
    if checked is not None:
        verified = verify_pass_from(checked, counts)
    emitted = kept.tolist()
    limbs: list[int] | None = None
    t0 = 0
    offset = 0
    for b, (s, e) in enumerate(batches):
        nt = counts[b]
        block = counters[t0:t0 + nt]
        if quiet and quiet_of[b]:
            totals = BatchTotals(nt, (int(tile_ends[t0, 0]), int(tile_ends[t0 + nt - 1, 1])), focus_of[b],
                                 int(sizes[2 * b + 1]), pairs_of[b], sums_of[b][0], sums_of[b][1:], filters, shares)
            yield s, e, totals, scope_record_bytes(tile_ends[t0:t0 + nt], block, ng, None, EMPTY_DIGEST.encode() * nt)
            t0 += nt
            continue
        per_tile = None
        if verified is None:
            if limbs is None:
                limbs = values[:2 * sum(emitted)].tolist()
            per_tile = []
            for k in emitted[t0:t0 + nt]:
                per_tile.append([limbs[2 * i] + (limbs[2 * i + 1] << 64) for i in range(offset, offset + k)])
                offset += k
        bnd = [tuple(r) for r in bounds[t0:t0 + nt].tolist()]
        row_started = time.perf_counter()
        ranges = [tuple(r) for r in tiles.reshape(-1, 4)[t0:t0 + nt].tolist()]
        row = batch_row(config, matcher.filter_groups, ranges, bnd, int(sizes[2 * b + 1]), block,
                        per_tile, cap, timings, row_started,
                        verified=verified[b] if verified is not None else None)
        row['kernel_registers'] = int(matcher.kernel['registers'])
        row['kernel_local_bytes'] = int(matcher.kernel['local_bytes'])
        row['matching_layout'] = matcher.layout
        row['prefilter_build_seconds'] = seconds[1] / n
        scopes = row['subtile_scopes']
        if row['local_survivors']:
            local = np.array([sc['local_survivors'] for sc in scopes], dtype=np.uint64)
            shas_of = ''.join(sc['output_sha256'] for sc in scopes).encode()
        else:   # nothing emitted: every record holds 0 and the digest of []
            local, shas_of = None, EMPTY_DIGEST.encode() * nt
        yield s, e, row, scope_record_bytes(tile_ends[t0:t0 + nt], block, ng, local, shas_of)
        t0 += nt


def compute_lease(matcher: MaskedMatcher, af: GPUFocus, bf: GPUFocus, first: int, end: int,
                  heartbeat: Callable[[], None] = lambda: None, pipelined: bool = True) -> dict[str, Any]:
    """One lease of up to 256 compute batches (`pseudocube/optimized.py:81-90`).

    pipelined=False runs the reference's loop literally, `prepare` then
    `MaskedMatcher.run` per batch; both give the same receipt.
    """
    started = time.monotonic()
    acc = Accumulator(first, end)
    if pipelined:
        for start, stop, row, records in lease_rows(matcher, af, bf, first, end, quiet=True):
            if isinstance(row, BatchTotals):
                acc.add_totals(start, stop, row, records)
            else:
                acc.add(start, stop, row, records)
            heartbeat()
    else:
        for start in range(first, end, BATCH_WIDTH):
            stop = min(start + BATCH_WIDTH, end)
            batch = prepare(matcher.runtime, matcher.config, af, bf, start, stop, PROFILE['tile_width'])
            acc.add(start, stop, matcher.run(batch))
            heartbeat()
    return acc.finish(time.monotonic() - started)
