"""Batched subtile engines on MI355X: geometry, the residue-bitset matcher, comparison.

Drop-in for the reference's `pseudocube/batch_experiment.py` (`Prepared`
16-24, `tile_ranges` 27-32, `prepare_arrays` 35-48, `prepare` 51-57,
`FusedMatcher` 79-167, `MaskedMatcher` 170-254, `compare` 257-263).
`MaskedMatcher` runs libpcube's bitset kernel; `FusedMatcher` runs the direct
per-position kernel with the same per-subtile counters.  Both produce the same
row, key for key.
"""
from __future__ import annotations

from dataclasses import dataclass
import ctypes
import hashlib
import threading
import time
from typing import Any, NamedTuple

import numpy as np

from . import native
from .accounting import digest
from .focus import Config, Tile
from .gpu import DeviceArray, GPUFocus, MatchEngine, Runtime
from .predicate import is_perfect_cube, local3, verify
from .verification import CLAIM

# output_sha256 of a subtile that emitted nothing (almost every subtile at p=619)
EMPTY_DIGEST = digest([])
# values one subtile may emit per batch (run's default output capacity)
OUTPUT_CAP = 8192
# candidate opportunities one subtile may hold (`pseudocube/batch_experiment.py:45`)
OPPORTUNITY_LIMIT = 50_000_000


@dataclass
class Prepared:
    first: int
    end: int
    width: int
    us: Any
    vs: Any
    bounds: list[tuple[int, int, int, int]]
    prep_seconds: float
    ranges: list[tuple[int, int, int, int]] | None = None   # tile_ranges(first, end, width), when known
    totals: tuple[int, int] | None = None   # (focus_u, cartesian_opportunities) of the bounds, when known
    started: tuple[Any, int, int] | None = None   # (engine, launch token, cap) of a match prepare() started


def tile_ranges(config: Config, first: int, end: int, width: int) -> list[tuple[int, int, int, int]]:
    """(v_first, v_end, u_min, u_end) for each subtile of [first, end) (`pseudocube/batch_experiment.py:27-32`)."""
    Tile(first, end).validate(config)
    if width <= 0 or (end - first + width - 1) // width > 64:
        raise ValueError('experiment allows at most 64 bounded subtiles')
    # u_interval(v)[0] = (lo + v*A) // B + 1 and u_interval(v)[1] = (hi + v*A) // B + 1,
    # stepped by exact quotient/remainder updates instead of one wide division per bound
    A, B = config.A, config.B
    step_q, step_r = divmod(width * A, B)
    q_lo, r_lo = divmod(config.lo + first * A, B)
    q_hi, r_hi = divmod(config.hi + (first + width - 1) * A, B)
    out = []
    for v in range(first, end, width):
        stop = v + width
        if stop >= end:
            stop = end
            q_hi = (config.hi + (end - 1) * A) // B
        out.append((v, stop, q_lo + 1, q_hi + 1))
        q_lo += step_q
        r_lo += step_r
        if r_lo >= B:
            q_lo += 1
            r_lo -= B
        q_hi += step_q
        r_hi += step_r
        if r_hi >= B:
            q_hi += 1
            r_hi -= B
    return out


def tile_geometry(config: Config, spans: list[tuple[int, int]], width: int) -> tuple[np.ndarray, list[int]]:
    """The subtiles of many batches in one library call (pcube_tile_ranges_many, exact
    128-bit arithmetic): an [n, 4] uint64 array of (v_first, v_end, u_min, u_end)
    rows, the batches' subtiles consecutive, and each batch's subtile count."""
    n = len(spans)
    firsts = np.array([s for s, _ in spans], dtype=np.uint64)
    ends = np.array([e for _, e in spans], dtype=np.uint64)
    widths = np.full(n, width, dtype=np.uint64)
    tiles = np.zeros(4 * 64 * max(1, n), dtype=np.uint64)
    counts = np.zeros(max(1, n), dtype=np.uint32)
    mask = (1 << 64) - 1
    native.call('pcube_tile_ranges_many', config.A, config.B, config.lo & mask, config.lo >> 64, config.hi & mask,
                config.hi >> 64, n, native.ptr64(firsts), native.ptr64(ends), native.ptr64(widths),
                native.ptr64(tiles), 64 * max(1, n), native.ptr32(counts))
    per_batch = counts[:n].tolist()
    return tiles[:4 * sum(per_batch)].reshape(-1, 4), per_batch


def tile_ranges_many(config: Config, spans: list[tuple[int, int]], width: int) -> list[list[tuple[int, int, int, int]]]:
    """tile_ranges of many batches at once: equal to [tile_ranges(config, first, end,
    width) for first, end in spans]."""
    if not spans:
        return []
    tiles, per_batch = tile_geometry(config, spans, width)
    flat = tiles.tolist()
    out = []
    at = 0
    for count in per_batch:
        out.append([tuple(r) for r in flat[at:at + count]])
        at += count
    return out


def searchsorted(array: DeviceArray, queries: list[int]) -> list[int]:
    q = np.array(queries, dtype=np.uint64)
    out = np.zeros(q.size, dtype=np.uint64)
    native.call('pcube_searchsorted', array.address, array.size, native.ptr64(q), q.size, native.ptr64(out))
    return [int(x) for x in out]


def prepare_arrays(runtime: Runtime, config: Config, first: int, end: int, width: int,
                   us: DeviceArray, vs: DeviceArray, started: float) -> Prepared:
    ranges = tile_ranges(config, first, end, width)
    ub = searchsorted(us, [bound for _, _, lo, hi in ranges for bound in (lo, hi)])
    vb = searchsorted(vs, [bound for lo, hi, _, _ in ranges for bound in (lo, hi)])
    bounds = [(ub[2 * i], ub[2 * i + 1], vb[2 * i], vb[2 * i + 1]) for i in range(len(ranges))]
    for a, b, c, d in bounds:
        if (b - a) * (d - c) > OPPORTUNITY_LIMIT:
            raise OverflowError('per-subtile candidate opportunity limit exceeded')
    runtime.check()
    return Prepared(first, end, width, us, vs, bounds, time.perf_counter() - started, ranges)


class _PrepareScratch:
    """Host buffers of prepare()'s library call: the geometry, bounds and totals
    array, pointers into it and the scalar results.  One set per thread, reused
    by its calls; every value is copied out before prepare() returns."""

    def __init__(self) -> None:
        self.out = np.empty(8 * 64 + 3, dtype=np.uint64)   # tiles, bounds, totals
        base = self.out.ctypes.data
        self.tiles = native.ptr64_at(base)
        self.bounds = native.ptr64_at(base + 8 * 256)
        self.totals = native.ptr64_at(base + 8 * 512)
        self.ntiles = ctypes.c_uint32()
        self.nu, self.nv, self.launch = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self.r_ntiles, self.r_nu, self.r_nv, self.r_launch = (
            ctypes.byref(x) for x in (self.ntiles, self.nu, self.nv, self.launch))


_scratch = threading.local()


def _prepare_scratch() -> _PrepareScratch:
    scratch = getattr(_scratch, 'value', None)
    if scratch is None:
        scratch = _scratch.value = _PrepareScratch()
    return scratch


def prepare(runtime: Runtime, config: Config, af: GPUFocus, bf: GPUFocus,
            first: int, end: int, width: int = 100000000) -> Prepared:
    """Both focus windows and the subtile bounds of one batch (`pseudocube/batch_experiment.py:51-57`).

    With both wheels present this is one library call (`pcube_prepare_range`,
    which also computes the subtile geometry); otherwise it is the reference's
    sequence of two windows and two searches.  When a MaskedMatcher of this
    configuration exists, the call is `pcube_prepare_match`: the batch's match
    is enqueued as soon as its bounds are known, and that matcher's run()
    collects it (prep_seconds then includes the enqueue).
    """
    started = time.perf_counter()
    if af.wheel is None or bf.wheel is None:
        ranges = tile_ranges(config, first, end, width)
        us = af.window(ranges[0][2], ranges[-1][3])
        vs = bf.window(first, end)
        return prepare_arrays(runtime, config, first, end, width, us, vs, started)
    Tile(first, end).validate(config)
    if width <= 0 or (end - first + width - 1) // width > 64:
        raise ValueError('experiment allows at most 64 bounded subtiles')
    cap = 1000000
    us, vs = DeviceArray.pooled(cap), DeviceArray.pooled(cap)
    sc = _prepare_scratch()
    runtime.check()
    mask = (1 << 64) - 1
    args = (af.handle, bf.handle, config.A, config.B, config.lo & mask, config.lo >> 64, config.hi & mask,
            config.hi >> 64, first, end, width, sc.tiles, sc.r_ntiles, us.address, cap, vs.address, cap,
            sc.r_nu, sc.r_nv, sc.bounds, sc.totals)
    matcher = runtime.matcher_for(config)
    if matcher is None:
        code = native.lib().pcube_prepare_range(*args)
        launch = 0
    else:
        code = native.lib().pcube_prepare_match(*args, matcher.engine.handle, OUTPUT_CAP, matcher.engine_id,
                                                OPPORTUNITY_LIMIT, sc.r_launch)
        launch = sc.launch.value
    if code == native.RANGE:
        raise OverflowError('focus window exceeds bounded period count')
    if code == native.OVERFLOW:
        raise OverflowError('focus buffer overflow; tile invalid')
    native.check(code, 'pcube_prepare_range')
    us.size, vs.size = sc.nu.value, sc.nv.value
    n = 4 * sc.ntiles.value
    flat = sc.out.tolist()
    focus_u, opportunities, largest = flat[512:515]
    if largest > OPPORTUNITY_LIMIT:
        raise OverflowError('per-subtile candidate opportunity limit exceeded')
    quads = iter(flat[:n])
    ranges = list(zip(quads, quads, quads, quads))
    quads = iter(flat[256:256 + n])
    rows = list(zip(quads, quads, quads, quads))
    runtime.check()
    return Prepared(first, end, width, us, vs, rows, time.perf_counter() - started, ranges,  # type: ignore[arg-type]
                    (focus_u, opportunities),
                    (matcher.engine, launch, OUTPUT_CAP) if launch else None)


def python_witnesses(values: list[int], p: int) -> list[dict[str, object]]:
    """Host verification with the pure-Python predicate (the reference's own path)."""
    if any(not local3(n, p) for n in values):
        raise ArithmeticError('GPU/CPU disagreement')
    return [verify(n, p) for n in values if not is_perfect_cube(n)]


def scope_skeleton(ranges: list[tuple[int, int, int, int]]) -> list[dict[str, Any]]:
    """The per-subtile receipt records of a batch before its counts are known
    (keys and key order of `pseudocube/batch_experiment.py:149-158`), so the
    host can build them while the device matches.  Each starts as the record
    of a subtile that emitted nothing; batch_row fills in the counts."""
    firsts = [str(first) for first, _, _, _ in ranges]
    n = len(ranges)
    out = []
    for k, (_, end, _, _) in enumerate(ranges):
        # consecutive subtiles share a bound: a subtile's end is the next one's first
        out.append({'first': firsts[k],
                    'end': firsts[k + 1] if k + 1 < n and ranges[k + 1][0] == end else str(end),
                    'valid_focused_candidates': 0, 'group_survivors': None, 'local_survivors': 0,
                    'output_sha256': EMPTY_DIGEST})
    return out


class Verified(NamedTuple):
    """Values of a batch already verified natively (pcube_batch_values): per subtile its
    sorted values as decimal strings and their digest (None without values), and the
    witness records of the noncubes in value order."""
    values: list[list[str]]
    digests: list[str | None]
    hits: list[dict[str, object]]


def check_values(config: Config, firsts: np.ndarray, ends: np.ndarray, primes: np.ndarray, kept: np.ndarray,
                 values: np.ndarray) -> tuple[list[int], list[list[str]], list[str], list[list[dict[str, object]]]]:
    """Every value a device pass emitted, checked in one native call
    (pcube_batch_values): per subtile sorted, no duplicates, inside (lo, hi] of
    config, canonical v inside [firsts, ends) of its subtile and local3 at that
    subtile's prime; raises ArithmeticError otherwise.

    firsts, ends, primes, kept: per subtile of the pass (kept = values listed);
    values: the listed values as (lo, hi) uint64 limbs, subtile after subtile.
    Returns, for the subtiles with values in ascending order: their indices,
    sorted decimal strings, the digest of each subtile's strings and the
    noncubes' witness records (the form python_witnesses gives)."""
    counts = np.ascontiguousarray(kept, dtype=np.uint64)
    nvalues = int(counts.sum())
    if nvalues == 0:
        return [], [], [], []
    limbs = np.ascontiguousarray(values[:2 * nvalues], dtype=np.uint64)
    firsts = np.ascontiguousarray(firsts, dtype=np.uint64)
    ends = np.ascontiguousarray(ends, dtype=np.uint64)
    primes = np.ascontiguousarray(primes, dtype=np.uint32)
    cube = np.zeros(nvalues, dtype=np.uint8)
    roots = np.zeros(nvalues, dtype=np.uint64)
    text = ctypes.create_string_buffer(nvalues * 44 + len(counts) * 3 + 16)
    offsets = np.zeros(len(counts) + 1, dtype=np.uint64)
    failure = ctypes.c_uint32()
    mask = (1 << 64) - 1
    native.call('pcube_batch_values', nvalues, native.ptr64(limbs), len(counts), native.ptr64(counts),
                native.ptr64(firsts), native.ptr64(ends), native.ptr32(primes), config.A, config.B, config.lo & mask,
                config.lo >> 64, config.hi & mask, config.hi >> 64, cube.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                native.ptr64(roots), text, len(text), native.ptr64(offsets), ctypes.byref(failure))
    raw_text = text.raw
    index = np.flatnonzero(counts).tolist()
    number = counts[index].tolist()
    starts = offsets[index].tolist()
    stops = offsets[np.array(index) + 1].tolist()
    prime_of = primes[index].tolist()
    cubes = cube.tolist()
    root_list = roots.tolist()
    strings_of, digests, hits_of = [], [], []
    v = 0
    sha256 = hashlib.sha256
    for n, a, b, p in zip(number, starts, stops, prime_of):
        chunk = raw_text[a:b]
        # the library's text of a subtile is the canonical encoding ["d1","d2",...]\n of
        # its sorted decimal strings, digits only
        strings = chunk[2:-3].decode('ascii').split('","')
        strings_of.append(strings)
        digests.append(sha256(chunk).hexdigest())
        flags = cubes[v:v + n]
        roots_of = root_list[v:v + n]
        if 1 in flags:
            hits_of.append([{'n': value, 'p': p, 'cube_root_floor': str(root), 'perfect_cube': False,
                             'failing_primes': [], 'valid_pseudocube': True, 'claim': CLAIM}
                            for value, is_cube, root in zip(strings, flags, roots_of) if not is_cube])
        else:
            hits_of.append([{'n': value, 'p': p, 'cube_root_floor': str(root), 'perfect_cube': False,
                             'failing_primes': [], 'valid_pseudocube': True, 'claim': CLAIM}
                            for value, root in zip(strings, roots_of)])
        v += n
    return index, strings_of, digests, hits_of


def verify_pass(config: Config, firsts: np.ndarray, ends: np.ndarray, primes: np.ndarray, kept: np.ndarray,
                groups: list[int], values: np.ndarray) -> list[Verified]:
    """check_values over a pass whose batches hold groups[b] consecutive subtiles:
    per batch its Verified (decimal strings and digest per subtile, the noncubes'
    witness records in value order)."""
    return verify_pass_from(check_values(config, firsts, ends, primes, kept, values), groups)


def verify_pass_from(checked: tuple[list[int], list[list[str]], list[str], list[list[dict[str, object]]]],
                     groups: list[int]) -> list[Verified]:
    """verify_pass's per-batch Verified from check_values's output."""
    out = [Verified([[] for _ in range(n)], [None] * n, []) for n in groups]
    index, strings_of, digests, hits_of = checked
    if not index:
        return out
    heads = np.cumsum([0] + list(groups))
    batch_of = (np.searchsorted(heads, index, side='right') - 1).tolist()
    for s, b, strings, text_digest, hits in zip(index, batch_of, strings_of, digests, hits_of):
        k = s - int(heads[b])
        out[b].values[k] = strings
        out[b].digests[k] = text_digest
        out[b].hits.extend(hits)
    return out


def batch_row(config: Config, filter_groups: list[list[int]], ranges: list[tuple[int, int, int, int]],
              bounds: list[tuple[int, int, int, int]], nv: int, counters: np.ndarray, per_tile: list[list[int]] | None,
              output_cap: int, timings: dict[str, float], started: float,
              witnesses: Any = python_witnesses, skeleton: list[dict[str, Any]] | None = None,
              totals: tuple[int, int] | None = None, verified: Verified | None = None) -> dict[str, Any]:
    """Host collection and verification of one batch (`pseudocube/batch_experiment.py:121-167`).

    Raises ArithmeticError on underflow, duplicates, a value failing the
    exact predicate or landing in the wrong subtile; OverflowError when a
    subtile emitted more than output_cap values.  With `verified`, the values'
    checks have run natively (the same checks) and per_tile is not read.
    """
    ng = len(filter_groups)
    transfer_started = time.perf_counter()
    nrows = len(counters)
    # a full p = 619 batch usually emits nothing: skip the per-subtile scan then
    flagged = bool(counters[:, ng + 1:ng + 3].any())
    emitted = counters[:, ng + 1].tolist() if flagged else [0] * nrows
    if flagged and counters[:, ng + 2].any():
        raise ArithmeticError('unsigned underflow')
    if emitted and max(emitted) > output_cap:
        raise OverflowError('fused tile output overflow; no truncated coverage')
    # every counted value must have come back, and nothing else
    tiles = verified.values if verified is not None else per_tile
    if tiles is None or len(tiles) != nrows or list(map(len, tiles)) != emitted:
        raise ArithmeticError('emitted values disagree with the subtile counters')
    scopes = skeleton if skeleton is not None else scope_skeleton(ranges)
    if verified is not None:
        transfer_seconds = time.perf_counter() - transfer_started + timings.get('transfer_seconds', 0.0)
        verification_started = time.perf_counter()
        all_text = [s for values in verified.values if values for s in values]
        hits = verified.hits
        if all_text:
            for number, text_digest, scope in zip(emitted, verified.digests, scopes):
                if number:
                    scope['local_survivors'] = number
                    scope['output_sha256'] = text_digest
    else:
        all_values: list[int] = []
        for values in tiles:
            if values:
                all_values.extend(sorted(values))
        transfer_seconds = time.perf_counter() - transfer_started + timings.get('transfer_seconds', 0.0)
        verification_started = time.perf_counter()
        if len(all_values) != len(set(all_values)):
            raise ArithmeticError('duplicate canonical output')
        if any(not config.lo < n <= config.hi for n in all_values):
            raise ArithmeticError('GPU/CPU disagreement')
        hits = witnesses(all_values, config.p)
        if any(not row['valid_pseudocube'] for row in hits):
            raise ArithmeticError('independent witness verification failed')
        if all_values:
            position = 0
            for (first, end, _, _), number, scope in zip(ranges, emitted, scopes):
                if number:
                    values = sorted(all_values[position:position + number])
                    if any(not first <= config.canonical(n)[1] < end for n in values):
                        raise ArithmeticError('output assigned to wrong subtile')
                    scope['local_survivors'] = number
                    scope['output_sha256'] = digest([str(n) for n in values])
                    position += number
        all_text = [str(n) for n in all_values]
    # only the columns the records hold: [valid, group_1..group_ng] of each subtile
    for valid, group, scope in zip(counters[:, 0].tolist(), counters[:, 1:ng + 1].tolist(), scopes):
        scope['valid_focused_candidates'] = valid
        scope['group_survivors'] = group
    verification_seconds = time.perf_counter() - verification_started
    sums = counters[:, :ng + 1].sum(axis=0, dtype=np.uint64).tolist() if nrows else [0] * (ng + 1)
    groups = sums[1:]
    if totals is not None:
        focus_u, opportunities = totals
    else:
        focus_u = opportunities = 0
        for a, b, c, d in bounds:
            focus_u += b - a
            opportunities += (b - a) * (d - c)
    return dict(status='PASS', host_verified=True, focus_u=focus_u, focus_v=int(nv),
                cartesian_opportunities=opportunities,
                valid_focused_candidates=sums[0],
                genuine_cubes=len(all_text) - len(hits), local_survivors=len(all_text), cube_checks=len(all_text),
                kernel_seconds=timings.get('kernel_seconds', 0.0), transfer_seconds=transfer_seconds,
                verification_seconds=verification_seconds,
                focus_enumeration_seconds=timings.get('focus_enumeration_seconds', 0.0), subtile_scopes=scopes,
                filter_groups=filter_groups, group_survivors=groups, all_values=all_text,
                cpu_confirmed_hits=hits, overflow_splits=[],
                elapsed_seconds=time.perf_counter() - started + timings.get('focus_enumeration_seconds', 0.0),
                claim='Isolated validation only; no production coverage.')


class FusedMatcher:
    """Per-subtile counters over a whole batch, direct per-position engine (`pseudocube/batch_experiment.py:79-167`)."""
    engine_id = 1
    layout = 'fused-batch'
    kernel_name = 'match_direct'

    def __init__(self, runtime: Runtime, config: Config) -> None:
        self.runtime, self.config = runtime, config
        self.engine = MatchEngine(runtime, config)
        self.filter_groups = self.engine.filter_groups
        self.kernel = runtime.compile(self.kernel_name)
        self.sm_count = runtime.compute_units
        self.prefilter_seconds = 0.0
        self.intersection_seconds = 0.0   # the match kernel alone (intersection and chain) of the last run

    @property
    def table(self) -> Any:
        """The device allowed-residue tables every engine of this matcher reads."""
        return self.engine.table

    def run(self, batch: Prepared, output_cap: int = OUTPUT_CAP) -> dict[str, Any]:
        started = time.perf_counter()
        ranges = batch.ranges if batch.ranges is not None else tile_ranges(self.config, batch.first, batch.end,
                                                                             batch.width)
        if len(ranges) != len(batch.bounds) or not 0 < output_cap <= 8192:
            raise ValueError('invalid prepared geometry or output capacity')
        # the device matches while the host builds the receipt records; prepare()
        # may have started this batch on this matcher already
        eager = batch.started
        if not (eager is not None and eager[0] is self.engine and eager[2] == output_cap
                and self.engine.adopt(eager[1], len(ranges), output_cap)):
            self.engine.start(batch.us, batch.vs, batch.first, batch.width, len(ranges), output_cap, self.engine_id)
        skeleton = scope_skeleton(ranges)
        counters, per_tile, seconds = self.engine.finish(overflow_ok=True)
        self.prefilter_seconds = seconds[0]
        self.intersection_seconds = seconds[2]
        timings = dict(kernel_seconds=seconds[1], focus_enumeration_seconds=batch.prep_seconds)
        row = batch_row(self.config, self.filter_groups, ranges, batch.bounds, int(batch.vs.size), counters,
                        per_tile, output_cap, timings, started, skeleton=skeleton, totals=batch.totals)
        row['kernel_registers'] = int(self.kernel['registers'])
        row['kernel_local_bytes'] = int(self.kernel['local_bytes'])
        row['matching_layout'] = self.layout
        return row


class MaskedMatcher(FusedMatcher):
    """Residue-bitset engine (`pseudocube/batch_experiment.py:170-254`).

    The first residual filters (up to eight whose classes fit one LDS chunk;
    193..277 at p = 619) are intersected as 64-bit residue bitsets staged in
    LDS before any 128-bit candidate is built; group counts are those of the
    reference's four-filter bitset form.
    """
    engine_id = 0
    layout = 'residue-bitset-batch'
    kernel_name = 'match_staged'

    def __init__(self, runtime: Runtime, config: Config) -> None:
        super().__init__(runtime, config)
        if not self.engine.order:
            raise ValueError('bitset experiment requires at least one residual filter')
        runtime.compile('prefilter_staged')
        runtime.register_matcher(config, self)

    def run(self, batch: Prepared, output_cap: int = OUTPUT_CAP) -> dict[str, Any]:
        result = super().run(batch, output_cap)
        result['prefilter_build_seconds'] = self.prefilter_seconds
        return result


def compare(reference: dict[str, Any], actual: dict[str, Any]) -> None:
    """Raise unless two batch rows agree on every compared field (`pseudocube/batch_experiment.py:257-263`)."""
    for key in ('focus_u', 'focus_v', 'cartesian_opportunities', 'valid_focused_candidates', 'genuine_cubes',
                'local_survivors', 'cube_checks', 'filter_groups', 'group_survivors', 'subtile_scopes',
                'cpu_confirmed_hits'):
        if reference[key] != actual[key]:
            raise ArithmeticError('batch/reference mismatch: ' + key)
    if sorted(reference['all_values']) != sorted(actual['all_values']):
        raise ArithmeticError('complete output mismatch')
