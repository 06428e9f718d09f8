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
import hashlib
import json
from pathlib import Path
import sys
import time
from typing import Any, Iterable

import numpy as np

from . import native
from .batch_experiment import Verified, batch_row, tile_ranges, verify_pass
from .focus import Config
from .gpu import GPUFocus, MatchEngine, Runtime
from .verification import CLAIM

TUNING = Path(__file__).resolve().parent / 'tuning.json'
DEFAULT_TUNING = dict(rows_per_chunk=8, blocks_per_cu=4, direct_below_rows=96)
GROUP_WINDOWS = 1024   # windows per device pass of run(): one pass runs while the previous is assembled


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
        u_span = (window.end - window.first) * self.base.A / self.base.B + (self.base.hi - self.base.lo) / self.base.B
        nu = u_span * self.af.focus.count / self.af.focus.modulus
        nv = (window.end - window.first) * self.bf.focus.count / self.bf.focus.modulus
        words = nu / 64 + 1
        return int(8 * nu + 8 * nv + 32 * nv + 16 * nu + 8 * 1300 * words) + 4096

    def run(self, windows: list[Window], budget: int | None = None) -> list[dict[str, Any]]:
        """One result row per window, in input order.

        Windows go to the device in groups of at most GROUP_WINDOWS within the
        memory budget, each one device pass (launch) on a worker thread, in
        order; this thread assembles a group's rows while the device runs the
        next.  The cyclic garbage collector is paused meanwhile (the rows hold
        no reference cycles, and full collections over thousands of fresh rows
        only add pauses), and Python's switch interval is 0.1 ms, so the
        worker gets the interpreter lock between passes at once."""
        if budget is None:
            budget = max(256 << 20, self.runtime.cap // 4)
        rows: list[dict[str, Any] | None] = [None] * len(windows)
        groups: list[list[int]] = []
        group: list[int] = []
        used = 0
        for index, window in enumerate(windows):
            need = self.footprint(window)
            if group and (used + need > budget or len(group) >= GROUP_WINDOWS):
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
        try:
            passes = [pool.submit(self.launch, [windows[i] for i in g]) for g in groups]
            for g, launched in zip(groups, passes):
                for index, row in zip(g, self.assemble(launched.result())):
                    rows[index] = row
        finally:
            # an error leaves no pass behind
            pool.shutdown(wait=True, cancel_futures=True)
            sys.setswitchinterval(interval)
            if collecting:
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
        return dict(windows=windows, plans=plans, bounds=bounds.reshape(ntot, 4), sizes=sizes, counters=counters,
                    values=values, timings=timings)

    def assemble(self, raw: dict[str, Any]) -> list[dict[str, Any]]:
        """Verified result rows of a launched group, in its window order."""
        windows, plans, bounds, sizes = raw['windows'], raw['plans'], raw['bounds'], raw['sizes']
        counters, values, timings = raw['counters'], raw['values'], raw['timings']
        blocks, kept = [], []
        c0 = 0
        for window, (ranges, _) in zip(windows, plans):
            ng = self.engines[window.p].ng
            nt = len(ranges)
            block = counters[c0:c0 + nt * (ng + 3)].reshape(nt, ng + 3)
            blocks.append(block)
            kept.append(np.minimum(block[:, ng + 1], self.cap))
            c0 += nt * (ng + 3)
        verified = self.verify_values(windows, plans, kept, values)
        # each window's focus_u and cartesian opportunities, summed over its subtiles' bounds at once
        starts = np.cumsum([0] + [len(ranges) for ranges, _ in plans[:-1]])
        du = (bounds[:, 1] - bounds[:, 0]).astype(np.int64)
        dv = (bounds[:, 3] - bounds[:, 2]).astype(np.int64)
        focus_u = np.add.reduceat(du, starts).tolist() if len(du) else []
        pairs = np.add.reduceat(du * dv, starts).tolist() if len(du) else []
        nvs = sizes[1::2].tolist()
        rows: list[dict[str, Any]] = []
        for k, (window, (ranges, mode), block, checked) in enumerate(zip(windows, plans, blocks, verified)):
            engine = self.engines[window.p]
            row = batch_row(self.configs[window.p], engine.filter_groups, ranges, [], nvs[k], block, None, self.cap,
                            timings, time.perf_counter(), verified=checked, totals=(focus_u[k], pairs[k]))
            row['matching_layout'] = 'batched-direct' if mode else 'batched-residue-bitset'
            row['prefilter_build_seconds'] = 0.0
            rows.append(row)
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
