"""Batched form: many windows of uneven v-length and mixed prime bound in one device pass.

A window is a canonical v range [first, end) searched at prime bound p with
subtiles of `width` (the reference's 1e8 by default, at most 64 subtiles).
Its result row is field for field what the plain drop-in returns for it,
`MaskedMatcher(runtime, Config(p, ...)).run(prepare(..., first, end, width))`
(`pseudocube/batch_experiment.py:104-167`), so every window's receipt is the
same alone, in the plain path, and in any mix.

One `run` call:
  * plans each window's subtiles (tile_ranges) and picks its engine: the
    residue-bitset engine, or the direct per-position engine for windows with
    so few v rows that building their bitset tables costs more than it saves
    (the threshold comes from the autotuner);
  * groups windows under a device-memory budget;
  * per group, one `pcube_batch_run`: both focus wheels generate every
    window's coordinates in one pass each, the subtile bounds come from one
    search, and a fixed-grid launch sequence (captured once in a hipGraph)
    builds rows, bitsets and residue caches for all windows and runs one
    persistent match kernel that pulls row chunks from a device queue;
  * verifies every emitted value on the host (native predicate) and builds
    the rows.
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
import ctypes
import gc
import json
from pathlib import Path
import sys
import time
from typing import Any, Iterable

import numpy as np

from . import native
from .batch_experiment import EMPTY_DIGEST, Verified, check_values, tile_ranges, verify_pass
from .focus import Config
from .gpu import GPUFocus, MatchEngine, Runtime
from .optimized import tapered_sizes

TUNING = Path(__file__).resolve().parent / 'tuning.json'
DEFAULT_TUNING = dict(rows_per_chunk=8, blocks_per_cu=4, direct_below_rows=96)
GROUP_WINDOWS = 1024   # windows per device pass of run(): one pass runs while the previous is assembled
GROUP_EDGE = 128       # windows of run()'s first pass and of its last, tapered from GROUP_WINDOWS


def load_tuning() -> dict[str, int]:
    """The autotuner's choices (psuedocubes_amd/tuning.json) over the defaults."""
    tuning = dict(DEFAULT_TUNING)
    if TUNING.exists():
        tuning.update({k: int(v) for k, v in json.loads(TUNING.read_text()).get('choice', {}).items()})
    return tuning


@dataclass(frozen=True)
class Window:
    first: int
    end: int
    p: int
    width: int = 100_000_000


class BatchedEngine:
    """Resident tables for a set of prime bounds sharing A, B and (lo, hi]."""

    def __init__(self, runtime: Runtime, configs: Iterable[Config], af: GPUFocus, bf: GPUFocus, cap: int = 8192,
                 tuning: dict[str, int] | None = None, use_graph: bool = True) -> None:
        self.runtime, self.af, self.bf, self.cap = runtime, af, bf, cap
        self.configs = {c.p: c for c in configs}
        if not 1 <= len(self.configs) <= 3:
            raise ValueError('one to three prime bounds per batched engine')
        base = next(iter(self.configs.values()))
        if any((c.A, c.B, c.lo, c.hi) != (base.A, base.B, base.lo, base.hi) for c in self.configs.values()):
            raise ValueError('batched configurations must share A, B and the interval')
        self.base = base
        self.tuning = dict(load_tuning(), **(tuning or {}))
        self.order = list(self.configs)
        self.engines = {p: MatchEngine(runtime, self.configs[p]) for p in self.order}
        handles = (ctypes.c_int * len(self.order))(*[self.engines[p].handle for p in self.order])
        engine = ctypes.c_int()
        native.call('pcube_batch_create', handles, len(self.order), af.handle, bf.handle, cap,
                    int(self.tuning['blocks_per_cu']), int(use_graph), ctypes.byref(engine))
        self.handle = engine.value
        self.graph_launches = 0
        self.launch_seconds: list[float] = []
        self.kernel = runtime.compile('batch_match')

    def plan(self, window: Window) -> tuple[list[tuple[int, int, int, int]], int]:
        """(subtiles, engine mode) of one window."""
        config = self.configs[window.p]
        ranges = tile_ranges(config, window.first, window.end, window.width)
        rows = (window.end - window.first) * self.bf.focus.count / self.bf.focus.modulus
        mode = 1 if rows < self.tuning['direct_below_rows'] else 0
        return ranges, mode

    def plan_many(self, windows: list[Window]) -> tuple[list[tuple[list[list[int]], int]], np.ndarray]:
        """plan() of every window of a pass, the subtile geometry in one library call
        (pcube_tile_ranges_many, exact 128-bit arithmetic), and the flat tiles array
        the device pass takes."""
        n = len(windows)
        base = self.base
        firsts = np.array([w.first for w in windows], dtype=np.uint64)
        ends = np.array([w.end for w in windows], dtype=np.uint64)
        widths = np.array([w.width for w in windows], dtype=np.uint64)
        tiles = np.zeros(4 * 64 * max(1, n), dtype=np.uint64)
        counts = np.zeros(max(1, n), dtype=np.uint32)
        mask = (1 << 64) - 1
        native.call('pcube_tile_ranges_many', base.A, base.B, base.lo & mask, base.lo >> 64, base.hi & mask,
                    base.hi >> 64, n, native.ptr64(firsts), native.ptr64(ends), native.ptr64(widths),
                    native.ptr64(tiles), 64 * max(1, n), native.ptr32(counts))
        total = int(counts[:n].sum())
        tiles = tiles[:4 * total]
        flat = tiles.reshape(total, 4).tolist()
        density = self.bf.focus.count / self.bf.focus.modulus
        threshold = self.tuning['direct_below_rows']
        plans = []
        at = 0
        for w, count in zip(windows, counts[:n].tolist()):
            plans.append((flat[at:at + count], 1 if (w.end - w.first) * density < threshold else 0))
            at += count
        return plans, tiles

    def footprint(self, window: Window) -> int:
        """Estimated device bytes of one window in a group."""
        return self.footprints([window])[0]

    def footprints(self, windows: list[Window]) -> list[int]:
        """footprint() of many windows at once: u and v coordinates, rows, residue
        rows and the bitset words of the window's mask rows, from its v-length."""
        length = np.array([w.end - w.first for w in windows], dtype=np.float64)
        u_span = length * (self.base.A / self.base.B) + (self.base.hi - self.base.lo) / self.base.B
        nu = u_span * (self.af.focus.count / self.af.focus.modulus)
        nv = length * (self.bf.focus.count / self.bf.focus.modulus)
        words = nu / 64 + 1
        return (np.floor(24 * nu + 40 * nv + 8 * 1300 * words).astype(np.int64) + 4096).tolist()

    def run(self, windows: list[Window], budget: int | None = None) -> list[dict[str, Any]]:
        """One result row per window, in input order.

        Windows go to the device in groups (tapered_sizes: a first group of
        GROUP_EDGE, groups of GROUP_WINDOWS, a tail halving down to GROUP_EDGE)
        within the memory budget, each one device pass (launch) on a worker
        thread, in order; this thread assembles a group's rows while the device
        runs the next.  The cyclic garbage collector is paused meanwhile (the rows hold
        no reference cycles, and full collections over thousands of fresh rows
        only add pauses), and Python's switch interval is 0.1 ms, so the
        worker gets the interpreter lock between passes at once.  On success the
        fresh objects (hundreds of thousands of row, scope and witness containers)
        go straight to the collector's oldest generation before it resumes, unless
        the process has frozen objects of its own: the young collection that would
        otherwise traverse them all at once, to find no garbage, is skipped."""
        if budget is None:
            budget = max(256 << 20, self.runtime.cap // 4)
        rows: list[dict[str, Any] | None] = [None] * len(windows)
        groups: list[list[int]] = []
        group: list[int] = []
        used = 0
        limits = tapered_sizes(len(windows), GROUP_EDGE, GROUP_WINDOWS)
        for index, need in enumerate(self.footprints(windows) if windows else []):
            limit = limits[len(groups)] if len(groups) < len(limits) else GROUP_WINDOWS
            if group and (used + need > budget or len(group) >= limit):
                groups.append(group)
                group, used = [], 0
            group.append(index)
            used += need
        if group:
            groups.append(group)
        collecting = gc.isenabled()
        gc.disable()
        interval = sys.getswitchinterval()
        sys.setswitchinterval(min(interval, 1e-4))
        pool = ThreadPoolExecutor(max_workers=1)
        finished = False
        try:
            passes = [pool.submit(self.launch, [windows[i] for i in g]) for g in groups]
            for g, launched in zip(groups, passes):
                for index, row in zip(g, self.assemble(launched.result())):
                    rows[index] = row
            finished = True
        finally:
            # an error leaves no pass behind
            pool.shutdown(wait=True, cancel_futures=True)
            sys.setswitchinterval(interval)
            if collecting:
                if finished and gc.get_freeze_count() == 0:
                    gc.freeze()
                    gc.unfreeze()
                gc.enable()
        return rows  # type: ignore[return-value]

    def launch(self, windows: list[Window]) -> dict[str, Any]:
        """Device pass over one group of windows; returns the raw results for assemble().

        The library call releases the GIL, so another thread can assemble the
        previous group's rows meanwhile."""
        group = list(range(len(windows)))
        started = time.perf_counter()
        plans, tiles = self.plan_many(windows)
        cfg = np.array([self.order.index(windows[i].p) for i in group], dtype=np.uint32)
        ntiles = np.array([len(r) for r, _ in plans], dtype=np.uint32)
        widths = np.array([windows[i].width for i in group], dtype=np.uint64)
        modes = np.array([m for _, m in plans], dtype=np.uint32)
        ntot = int(ntiles.sum())
        ngs = [self.engines[windows[i].p].ng for i in group]
        bounds = np.zeros(4 * ntot, dtype=np.uint64)
        sizes = np.zeros(2 * len(group), dtype=np.uint64)
        counters = np.zeros(sum(int(n) * (g + 3) for n, g in zip(ntiles, ngs)), dtype=np.uint64)
        values = np.empty(2 * min(max(1, ntot * self.cap), 1 << 22), dtype=np.uint64)
        seconds = (ctypes.c_double * 4)()
        launches = ctypes.c_uint64()
        self.runtime.check()
        native.call('pcube_batch_run', self.handle, len(group), native.ptr32(cfg), native.ptr32(ntiles),
                    native.ptr64(tiles), native.ptr64(widths), native.ptr32(modes),
                    int(self.tuning['rows_per_chunk']), native.ptr64(bounds), native.ptr64(sizes),
                    native.ptr64(counters), native.ptr64(values), values.size // 2, seconds, ctypes.byref(launches))
        self.runtime.check()
        self.graph_launches = launches.value
        self.launch_seconds.append(seconds[2])
        device = time.perf_counter() - started
        n = len(group)
        timings = dict(kernel_seconds=seconds[2] / n, focus_enumeration_seconds=seconds[0] / n,
                       transfer_seconds=max(0.0, device - seconds[0] - seconds[2]) / n)
        return dict(windows=windows, plans=plans, tiles=tiles.reshape(ntot, 4), bounds=bounds.reshape(ntot, 4),
                    sizes=sizes, counters=counters, values=values, timings=timings)

    def assemble(self, raw: dict[str, Any]) -> list[dict[str, Any]]:
        """Verified result rows of a launched group, in its window order.

        Each row is field for field batch_row's for its window (the checks and
        their order too: the pass's values first, then per window an underflow
        and an output overflow), built from pass-wide arrays: every counter
        column gathered once per prime bound, every emitted value checked and
        formatted in one native call (check_values)."""
        started = time.perf_counter()
        windows, plans, bounds, sizes = raw['windows'], raw['plans'], raw['bounds'], raw['sizes']
        counters, values, timings = raw['counters'], raw['values'], raw['timings']
        nw = len(windows)
        if nw == 0:
            return []
        cap = self.cap
        nts = [len(ranges) for ranges, _ in plans]
        ngs = [self.engines[w.p].ng for w in windows]
        nt_of = np.array(nts, dtype=np.int64)
        heads = np.zeros(nw, dtype=np.int64)
        np.cumsum(nt_of[:-1], out=heads[1:])
        stride = np.repeat(np.array(ngs, dtype=np.int64) + 3, nt_of)
        row_at = np.cumsum(stride) - stride            # each subtile's counter row
        ng_of = stride - 3
        emitted = counters[row_at + ng_of + 1]
        kept = np.minimum(emitted, np.uint64(cap))
        tiles = raw.get('tiles')
        if tiles is None:
            tiles = np.array([r for ranges, _ in plans for r in ranges], dtype=np.uint64).reshape(-1, 4)
        primes = np.repeat(np.array([w.p for w in windows], dtype=np.uint32), nt_of)
        index, strings_of, digests, hits_of = check_values(self.base, tiles[:, 0], tiles[:, 1], primes, kept, values)
        underflow = np.add.reduceat(counters[row_at + ng_of + 2] != 0, heads)
        overflow = np.add.reduceat(emitted > np.uint64(cap), heads)
        if underflow.any() or overflow.any():
            first = int(np.flatnonzero((underflow != 0) | (overflow != 0))[0])
            if underflow[first]:
                raise ArithmeticError('unsigned underflow')
            raise OverflowError('fused tile output overflow; no truncated coverage')
        valid_t = counters[row_at].tolist()
        valid_w = np.add.reduceat(counters[row_at], heads).tolist()
        # per prime bound: every subtile's group counts, and their sums per window
        groups_t: dict[int, list[list[int]]] = {}
        groups_w: dict[int, list[list[int]]] = {}
        window_p = np.array([w.p for w in windows])
        for p in set(window_p.tolist()):
            ng = self.engines[p].ng
            mine = window_p == p
            rows_p = row_at[np.repeat(mine, nt_of)]
            block = counters[rows_p[:, None] + 1 + np.arange(ng)] if ng else np.zeros((len(rows_p), 0), np.uint64)
            groups_t[p] = block.tolist()
            starts = np.zeros(int(mine.sum()), dtype=np.int64)
            np.cumsum(nt_of[mine][:-1], out=starts[1:])
            groups_w[p] = np.add.reduceat(block, starts, axis=0).tolist() if ng else [[] for _ in starts]
        at_t = dict.fromkeys(groups_t, 0)
        at_w = dict.fromkeys(groups_w, 0)
        du = (bounds[:, 1] - bounds[:, 0]).astype(np.int64)
        dv = (bounds[:, 3] - bounds[:, 2]).astype(np.int64)
        focus_u = np.add.reduceat(du, heads).tolist()
        pairs = np.add.reduceat(du * dv, heads).tolist()
        nvs = sizes[1::2].tolist()
        kept_t = kept.tolist()
        window_of = np.repeat(np.arange(nw), nt_of)[index].tolist() if index else []
        verification = time.perf_counter() - started
        share = verification / nw
        rows: list[dict[str, Any]] = []
        found = 0
        t0 = 0
        for k in range(nw):
            window = windows[k]
            ranges, mode = plans[k]
            p = window.p
            nt = nts[k]
            lists = groups_t[p]
            g0 = at_t[p]
            at_t[p] = g0 + nt
            scopes = []
            n = len(ranges)
            for j in range(n):
                first, end = ranges[j][0], ranges[j][1]
                scopes.append({'first': str(first), 'end': str(end), 'valid_focused_candidates': valid_t[t0 + j],
                               'group_survivors': lists[g0 + j], 'local_survivors': 0,
                               'output_sha256': EMPTY_DIGEST})
            all_text: list[str] = []
            hits: list[dict[str, Any]] = []
            while found < len(index) and window_of[found] == k:
                scope = scopes[index[found] - t0]
                scope['local_survivors'] = kept_t[index[found]]
                scope['output_sha256'] = digests[found]
                all_text += strings_of[found]
                hits += hits_of[found]
                found += 1
            engine = self.engines[p]
            wg = at_w[p]
            at_w[p] = wg + 1
            nvalues = len(all_text)
            rows.append(dict(status='PASS', host_verified=True, focus_u=focus_u[k], focus_v=nvs[k],
                             cartesian_opportunities=pairs[k], valid_focused_candidates=valid_w[k],
                             genuine_cubes=nvalues - len(hits), local_survivors=nvalues, cube_checks=nvalues,
                             kernel_seconds=timings.get('kernel_seconds', 0.0),
                             transfer_seconds=timings.get('transfer_seconds', 0.0), verification_seconds=share,
                             focus_enumeration_seconds=timings.get('focus_enumeration_seconds', 0.0),
                             subtile_scopes=scopes, filter_groups=engine.filter_groups,
                             group_survivors=groups_w[p][wg], all_values=all_text, cpu_confirmed_hits=hits,
                             overflow_splits=[],
                             elapsed_seconds=share + timings.get('focus_enumeration_seconds', 0.0),
                             claim='Isolated validation only; no production coverage.',
                             matching_layout='batched-direct' if mode else 'batched-residue-bitset',
                             prefilter_build_seconds=0.0))
            t0 += nt
        return rows

    def verify_values(self, windows: list[Window], plans: list, kept: list[np.ndarray],
                      values: np.ndarray) -> list[Verified]:
        """Every emitted value of a pass checked in one native call (pcube_batch_values:
        sorted per subtile, no duplicates, inside (lo, hi], canonical v inside its
        subtile, local3 at its window's p), then per window its subtiles' decimal
        strings, digests and the noncubes' witness records."""
        kept_all = np.concatenate(kept) if kept else np.zeros(0, dtype=np.uint64)
        firsts = np.array([r[0] for ranges, _ in plans for r in ranges], dtype=np.uint64)
        ends = np.array([r[1] for ranges, _ in plans for r in ranges], dtype=np.uint64)
        primes = np.repeat(np.array([w.p for w in windows], dtype=np.uint32), [len(k) for k in kept])
        return verify_pass(self.base, firsts, ends, primes, kept_all, [len(k) for k in kept], values)

    def __del__(self) -> None:
        if native._lib is not None and hasattr(self, 'handle'):
            try:
                native.lib().pcube_batch_destroy(self.handle)
            except Exception:
                pass


def window_stream(config: Config, count: int, seed: int, length_range: tuple[float, float],
                  primes: tuple[int, ...]) -> list[Window]:
    """Seeded uneven windows: v-lengths log-uniform over length_range, p uniform over
    primes, start = canonical v of an x drawn uniformly from (10^27, 3*10^27], clipped to [0, B)."""
    import math
    import random
    rng = random.Random(seed)
    lo, hi = math.log(length_range[0]), math.log(length_range[1])
    out = []
    for _ in range(count):
        length = int(math.exp(rng.uniform(lo, hi)))
        p = rng.choice(primes)
        x = rng.randrange(config.lo + 1, config.hi + 1)
        v = config.canonical(x)[1]
        first = min(v, config.B - 1)
        end = min(config.B, first + max(1, length))
        out.append(Window(first, end, p))
    return out
