"""Served form: continuous batching of coverage-range requests.

Requests are canonical v ranges [first, end) of up to 64 subtiles at one prime
bound; each arrives at its own time.  The server keeps the focus wheels and
filter tables resident and loops:

  * arrivals join a waiting queue;
  * a waiting request is admitted when a slot is free and its memory
    reservation fits the budget.  It reserves the footprint of one piece,
    since at most one of its pieces is on the device at a time;
  * an admitted request is cut into pieces of at most `piece_tiles` subtiles;
  * each device step takes the next piece of every admitted request (round
    robin, so short requests are not stuck behind long ones) and runs them
    as one batched pass.  A device thread runs step k while the host thread
    verifies and assembles step k-1's rows and admits new arrivals, so
    requests that arrive while others are on the GPU overlap them;
  * a request completes when all its pieces are back.  Its row merges its
    pieces' rows and equals the plain drop-in's row for the whole range
    (`prepare` + `MaskedMatcher.run`, pseudocube/batch_experiment.py:104-167):
    per-subtile scopes and values do not depend on what shares a launch;
  * with a ledger, the requests that complete in one step are committed as
    coverage of their scopes in one transaction (`Ledger.commit_batch`:
    compact gzip receipts in the database, a repeated scope recorded as a
    replay) before any of them is reported complete.  A `LedgerWriter` does
    that in a child process: the loop hands each step's rows over and goes on,
    and reports the requests when the writer acknowledges their commit.

Latency is completion time minus arrival time.

The cyclic garbage collector is paused while a stream is served: result rows
hold no reference cycles, so reference counting frees them, and a full
collection over the retained rows of a long stream stalls the loop for tens of
milliseconds at a time.
"""
from __future__ import annotations

from collections import deque
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field
import gc
import math
import random
import statistics
import sys
import time
from typing import Any, Callable

from .accounting import Ledger, LedgerWriter
from .batched import BatchedEngine, Window
from .focus import Config, Tile

MERGED_COUNTS = ('focus_u', 'focus_v', 'cartesian_opportunities', 'valid_focused_candidates', 'genuine_cubes',
                 'local_survivors', 'cube_checks')
MERGED_TIMES = ('kernel_seconds', 'transfer_seconds', 'verification_seconds', 'focus_enumeration_seconds',
                'prefilter_build_seconds', 'elapsed_seconds')


@dataclass
class Request:
    id: int
    first: int
    end: int
    arrival: float = 0.0
    p: int = 619


@dataclass
class Completed:
    request: Request
    row: dict[str, Any]
    admitted: float
    finished: float
    error: str | None = None   # set for a request rejected at admission; row is then its rejection record
    receipt: str | None = None   # the ledger attempt token of the committed row, with a ledger

    @property
    def latency(self) -> float:
        return self.finished - self.request.arrival


@dataclass
class _Active:
    request: Request
    pieces: deque
    rows: list = field(default_factory=list)
    admitted: float = 0.0
    reserved: int = 0
    outstanding: int = 0


def merge_rows(rows: list[dict[str, Any]]) -> dict[str, Any]:
    """One row for consecutive pieces of a range: sums and concatenations in subtile order."""
    out = dict(rows[0])
    for key in MERGED_COUNTS + MERGED_TIMES:
        out[key] = sum(r[key] for r in rows)
    out['subtile_scopes'] = [s for r in rows for s in r['subtile_scopes']]
    out['all_values'] = [v for r in rows for v in r['all_values']]
    out['cpu_confirmed_hits'] = [h for r in rows for h in r['cpu_confirmed_hits']]
    out['overflow_splits'] = [s for r in rows for s in r['overflow_splits']]
    out['group_survivors'] = [sum(g) for g in zip(*(r['group_survivors'] for r in rows))]
    return out


class Server:
    def __init__(self, engine: BatchedEngine, budget: int | None = None, max_slots: int = 256,
                 piece_tiles: int = 8, ledger: Ledger | LedgerWriter | None = None) -> None:
        self.engine = engine
        self.ledger = ledger
        self.budget = budget if budget is not None else max(256 << 20, engine.runtime.cap // 4)
        self.max_slots = max_slots
        self.piece_tiles = piece_tiles
        self.steps = 0
        self.step_sizes: list[int] = []
        self.overlap_seconds = 0.0

    def validate(self, request: Request) -> str | None:
        """Why a request cannot be served (a prime bound the engine lacks, or not a
        nonempty canonical v range), or None."""
        config = self.engine.configs.get(request.p)
        if config is None:
            return f'no resident tables for p={request.p}'
        try:
            Tile(request.first, request.end).validate(config)
        except ValueError as error:
            return str(error)
        return None

    @staticmethod
    def scope(request: Request, width: int = 100_000_000) -> dict[str, str]:
        """The ledger scope of a request: its prime bound, v range and subtile width."""
        return {'kind': 'served-range', 'p': str(request.p), 'first': str(request.first), 'end': str(request.end),
                'tile_width': str(width)}

    def pieces(self, request: Request, width: int = 100_000_000) -> deque:
        span = self.piece_tiles * width
        return deque(Window(s, min(s + span, request.end), request.p, width)
                     for s in range(request.first, request.end, span))

    def serve(self, requests: list[Request], realtime: bool = True,
              on_complete: Callable[[Completed], None] | None = None) -> list[Completed]:
        """Serve requests; with realtime, each becomes visible at its arrival time on the wall
        clock, otherwise all are visible at once (saturated load)."""
        clock = time.perf_counter
        start = clock()
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting: deque = deque()
        active: dict[int, _Active] = {}
        order: deque = deque()           # round-robin order of active request ids
        used = 0
        done: list[Completed] = []
        inflight: tuple[Future, list[tuple[int, Window]]] | None = None
        committing: dict[int, list] = {}   # writer group id -> its finished requests
        writer = isinstance(self.ledger, LedgerWriter)
        pool = ThreadPoolExecutor(max_workers=1)
        collecting = gc.isenabled()
        gc.disable()
        # the device worker needs the interpreter lock to start and end each step; a
        # short switch interval hands it over at once while this thread assembles rows,
        # where the default (5 ms) would delay every step that long
        interval = sys.getswitchinterval()
        sys.setswitchinterval(min(interval, 1e-4))

        def now() -> float:
            return clock() - start

        def admit() -> None:
            nonlocal used
            t = now()
            while pending and (pending[0].arrival <= t or not realtime):
                waiting.append(pending.popleft())
            while waiting and len(active) < self.max_slots:
                request = waiting[0]
                error = self.validate(request)
                if error is not None:
                    # one malformed request is answered on its own; the loop and the
                    # other requests go on
                    waiting.popleft()
                    t = now()
                    rejected = Completed(request, dict(status='REJECTED', error=error), t, t, error)
                    done.append(rejected)
                    if on_complete:
                        on_complete(rejected)
                    continue
                pieces = self.pieces(request)
                reserve = max(self.engine.footprint(w) for w in pieces)
                if active and used + reserve > self.budget:
                    break
                waiting.popleft()
                active[request.id] = _Active(request, pieces, admitted=now(), reserved=reserve,
                                             outstanding=len(pieces))
                order.append(request.id)
                used += reserve

        def next_step() -> list[tuple[int, Window]] | None:
            step = []
            for _ in range(len(order)):
                rid = order[0]
                order.rotate(-1)
                state = active[rid]
                if state.pieces:
                    step.append((rid, state.pieces.popleft()))
            return step or None

        def finish(raw: dict[str, Any], step: list[tuple[int, Window]]) -> None:
            nonlocal used
            rows = self.engine.assemble(raw)
            finished = []
            for (rid, _), row in zip(step, rows):
                state = active[rid]
                state.rows.append(row)
                state.outstanding -= 1
                if state.outstanding == 0:
                    finished.append((state.request, merge_rows(state.rows), state.admitted))
                    used -= state.reserved
                    del active[rid]
                    order.remove(rid)
            if not finished:
                return
            if self.ledger is None:
                report(finished, [None] * len(finished))
                return
            entries = [(self.scope(request), row) for request, row, _ in finished]
            if writer:
                committing[self.ledger.submit(entries)] = finished
            else:
                report(finished, self.ledger.commit_batch(entries))

        def report(finished: list, tokens: list) -> None:
            t = now()
            for (request, row, admitted), token in zip(finished, tokens):
                completed = Completed(request, row, admitted, t, receipt=token)
                done.append(completed)
                if on_complete:
                    on_complete(completed)

        def acknowledged(wait: bool = False) -> None:
            for ident, tokens in self.ledger.acknowledged(wait):
                report(committing.pop(ident), tokens)

        try:
            while pending or waiting or active or inflight:
                if writer and committing:
                    acknowledged()
                admit()
                if inflight is None:
                    step = next_step()
                    if step is not None:
                        inflight = (pool.submit(self.engine.launch, [w for _, w in step]), step)
                        self.steps += 1
                        self.step_sizes.append(len(step))
                    elif pending and realtime:
                        time.sleep(max(0.0, min(0.002, pending[0].arrival - now())))
                    continue
                future, step = inflight
                raw = future.result()
                inflight = None
                # keep the device busy: launch the next step before assembling this one
                admit()
                following = next_step()
                if following is not None:
                    inflight = (pool.submit(self.engine.launch, [w for _, w in following]), following)
                    self.steps += 1
                    self.step_sizes.append(len(following))
                t0 = now()
                finish(raw, step)
                if inflight is not None:
                    self.overlap_seconds += now() - t0
            while committing:
                acknowledged(wait=True)
        finally:
            pool.shutdown(wait=True)
            sys.setswitchinterval(interval)
            if collecting:
                gc.enable()
        return done


def request_stream(config: Config, count: int, seed: int, rate: float, length_range: tuple[float, float] = (1e7, 6.4e9),
                   p: int = 619) -> list[Request]:
    """Poisson arrivals at `rate` per second; v-lengths log-uniform over length_range;
    start = canonical v of x uniform over (10^27, 3*10^27], clipped to [0, B)."""
    rng = random.Random(seed)
    lo, hi = math.log(length_range[0]), math.log(length_range[1])
    t = 0.0
    out = []
    for i in range(count):
        t += rng.expovariate(rate) if rate > 0 else 0.0
        length = int(math.exp(rng.uniform(lo, hi)))
        x = rng.randrange(config.lo + 1, config.hi + 1)
        first = min(config.canonical(x)[1], config.B - 1)
        out.append(Request(i, first, min(config.B, first + max(1, length)), t, p))
    return out


def latency_summary(done: list[Completed]) -> dict[str, float]:
    lat = sorted(c.latency for c in done)
    if not lat:
        return dict(p50=0.0, p99=0.0, mean=0.0, max=0.0)
    def q(f: float) -> float:
        return lat[min(len(lat) - 1, int(math.ceil(f * len(lat))) - 1)]
    return dict(p50=q(0.50), p99=q(0.99), mean=statistics.fmean(lat), max=lat[-1])
