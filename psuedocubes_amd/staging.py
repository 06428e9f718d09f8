"""Producer/consumer staging of focus windows (tools/batch_window_experiment.py:52-82, 256-336).

A window producer computes the u and v focus windows of whole batches and stages each batch's
coordinates in a page-locked host slot; a matcher consumer takes the slot, checks its SHA-256, copies
it to the device, hands the slot back and matches the batch.  The reference ran this across GPUs
(one producer process feeding consumer processes through shared-memory slots) to test whether
separating window generation from matching pays; it found self-sufficient GPU workers faster
(912e9 against 567e9 and 130e9 positions/s, results/benchmarks-2026-09-14.json).

Here one process drives one GPU, so the producer and the consumers are threads of that process
sharing the library's stream; what overlaps is the host work (window slicing, copies and checksums
of one batch against the device work of another).  Producers:

- ``gpu``: the device's CRT wheel (GPUFocus.window) for eight batches at a time, downloaded into
  page-locked buffers and sliced per batch (the reference's super-window of eight jobs,
  tools/batch_window_experiment.py:275-283);
- ``cpu``: the host producer pcube_host_focus_window over host copies of the focus halves
  (CPUFocus, the reference's review/focus_window_cpu.cpp path).

Consumers run MaskedMatcher (``masked-*`` modes) or FusedMatcher, as the reference's
pipeline_consumer does.  Rows are identical to the plain path's: each batch's windows are the same
sorted coordinate lists whichever producer made them, and the consumer's prepare_arrays computes the
same subtile bounds.
"""
from __future__ import annotations

import ctypes
import hashlib
import queue
import threading
import time
from typing import Any

import numpy as np

from . import native
from .accounting import digest
from .batch_experiment import FusedMatcher, MaskedMatcher, prepare_arrays, tile_ranges
from .focus import Config
from .gpu import DeviceArray, GPUFocus, Runtime

WIDTH = 100_000_000
BATCH = WIDTH * 64
CAP = 1_000_000          # coordinates per window in a slot (the reference's CAP)
SUPER = 8                # batches per GPU producer window pass
MODES = ('gpu-producer', 'cpu-producer', 'masked-gpu-producer', 'masked-cpu-producer')


def signature(result: dict[str, Any]) -> str:
    """Digest of a batch row's results (tools/batch_window_experiment.py:47-50)."""
    keys = ('focus_u', 'focus_v', 'cartesian_opportunities', 'valid_focused_candidates', 'genuine_cubes',
            'group_survivors', 'subtile_scopes', 'cpu_confirmed_hits')
    return digest(dict(**{key: result[key] for key in keys}, all_values=sorted(result['all_values'])))


class PinnedBuffer:
    """A page-locked host array of uint64 (pcube_pinned_alloc)."""

    def __init__(self, size: int) -> None:
        self.size = int(size)
        pointer = ctypes.c_void_p()
        native.call('pcube_pinned_alloc', self.size * 8, ctypes.byref(pointer))
        self.pointer = pointer.value
        self.array = np.ctypeslib.as_array((ctypes.c_uint64 * self.size).from_address(self.pointer))

    def close(self) -> None:
        if self.pointer is not None:
            self.array = None
            native.call('pcube_pinned_free', ctypes.c_void_p(self.pointer))
            self.pointer = None

    def __del__(self) -> None:
        if getattr(self, 'pointer', None) is not None and native._lib is not None:
            try:
                self.close()
            except Exception:
                pass


class CPUFocus:
    """Host copy of a focus's sorted halves with the host window producer
    (the reference's CPUFocus, tools/batch_window_experiment.py:52-82)."""

    def __init__(self, focus: GPUFocus) -> None:
        self.C = np.ascontiguousarray(focus.C.get())
        self.D = np.ascontiguousarray(focus.D.get())   # sorted on the device
        self.modulus = focus.focus.modulus

    def window(self, lo: int, hi: int, cap: int = CAP) -> np.ndarray:
        """Sorted focus coordinates in [lo, hi), period by period."""
        if not 0 <= lo <= hi < 1 << 64:
            raise ValueError('invalid CPU focus range')
        out = np.empty(max(1, cap), dtype=np.uint64)
        total = 0
        count = ctypes.c_uint64()
        if hi > lo:
            for period in range(lo // self.modulus, (hi - 1) // self.modulus + 1):
                base = period * self.modulus
                code = native.lib().pcube_host_focus_window(
                    native.ptr64(self.C), self.C.size, native.ptr64(self.D), self.D.size, self.modulus,
                    max(0, lo - base), min(self.modulus, hi - base), base, native.ptr64_at(out.ctypes.data + 8 * total),
                    cap - total, ctypes.byref(count))
                if code == native.OVERFLOW:
                    raise OverflowError('CPU focus output overflow')
                native.check(code, 'pcube_host_focus_window')
                total += count.value
        return out[:total]


class _GPUProducer:
    """Windows of SUPER batches at a time from the device wheel, downloaded into pinned buffers."""

    def __init__(self, config: Config, af: GPUFocus, bf: GPUFocus) -> None:
        self.config, self.af, self.bf = config, af, bf
        self.us_host, self.vs_host = PinnedBuffer(CAP), PinnedBuffer(CAP)

    def windows(self, first: int, end: int) -> tuple[np.ndarray, np.ndarray]:
        us = self.af.window(self.config.u_interval(first)[0], self.config.u_interval(end - 1)[1])
        vs = self.bf.window(first, end)
        for device, host in ((us, self.us_host), (vs, self.vs_host)):
            if device.size:
                native.call('pcube_download', host.pointer, device.address, device.nbytes)
        return self.us_host.array[:us.size], self.vs_host.array[:vs.size]

    def close(self) -> None:
        self.us_host.close()
        self.vs_host.close()


class _CPUProducer:
    """Windows of SUPER batches at a time from the host producer."""

    def __init__(self, config: Config, af: GPUFocus, bf: GPUFocus) -> None:
        self.config = config
        self.af, self.bf = CPUFocus(af), CPUFocus(bf)

    def windows(self, first: int, end: int) -> tuple[np.ndarray, np.ndarray]:
        return (self.af.window(self.config.u_interval(first)[0], self.config.u_interval(end - 1)[1]),
                self.bf.window(first, end))

    def close(self) -> None:
        pass


def _checksum(us: np.ndarray, vs: np.ndarray) -> str:
    h = hashlib.sha256(memoryview(us).cast('B'))
    h.update(memoryview(vs).cast('B'))
    return h.hexdigest()


def pipeline_producer(producer: Any, config: Config, jobs: list[int], slots: list[PinnedBuffer],
                      free: queue.Queue, tasks: list[queue.Queue], stop: threading.Event) -> dict[str, Any]:
    """Stage the windows of each job (batch index) in a free slot and post it to a consumer
    (tools/batch_window_experiment.py:256-298): (index, slot, nu, nv, sha256)."""
    begun = time.perf_counter()
    generation = transport = waiting = 0.0
    super_us = super_vs = None
    covered = range(0)   # the batch indices the current window pass covers
    for n, index in enumerate(jobs):
        first = index * BATCH
        ranges = tile_ranges(config, first, first + BATCH, WIDTH)
        t = time.perf_counter()
        while True:
            if stop.is_set():
                raise RuntimeError('pipeline stopped')
            try:
                owner, slot = free.get(timeout=1.0)
                break
            except queue.Empty:
                continue
        waiting += time.perf_counter() - t
        t = time.perf_counter()
        # one window pass covers this job and up to SUPER - 1 consecutive ones after it
        if index not in covered:
            run = 1
            while run < SUPER and n + run < len(jobs) and jobs[n + run] == index + run:
                run += 1
            covered = range(index, index + run)
            super_us, super_vs = producer.windows(first, first + run * BATCH)
        ua, ub = np.searchsorted(super_us, [ranges[0][2], ranges[-1][3]])
        va, vb = np.searchsorted(super_vs, [first, first + BATCH])
        us, vs = super_us[ua:ub], super_vs[va:vb]
        if us.size > CAP or vs.size > CAP:
            raise OverflowError('staged window exceeds its slot')
        generation += time.perf_counter() - t
        t = time.perf_counter()
        target = slots[slot].array
        target[:us.size] = us
        target[CAP:CAP + vs.size] = vs
        tasks[owner].put((index, slot, int(us.size), int(vs.size), _checksum(us, vs)))
        transport += time.perf_counter() - t
    return dict(role='producer', status='PASS', seconds=time.perf_counter() - begun, generation_seconds=generation,
                staging_seconds=transport, backpressure_seconds=waiting, jobs=len(jobs))


def pipeline_consumer(runtime: Runtime, config: Config, matcher: Any, owner: int, slots: list[PinnedBuffer],
                      free: queue.Queue, tasks: queue.Queue, stop: threading.Event) -> dict[str, Any]:
    """Match staged batches (tools/batch_window_experiment.py:301-336): check the slot's checksum,
    copy it to the device, hand the slot back, match."""
    begun = time.perf_counter()
    rows: list[dict[str, Any]] = []
    waiting = upload = filtering = 0.0
    while True:
        t = time.perf_counter()
        while True:
            if stop.is_set():
                raise RuntimeError('pipeline stopped')
            try:
                task = tasks.get(timeout=1.0)
                break
            except queue.Empty:
                continue
        waiting += time.perf_counter() - t
        if task is None:
            break
        index, slot, nu, nv, sha = task
        source = slots[slot].array
        t = time.perf_counter()
        if _checksum(source[:nu], source[CAP:CAP + nv]) != sha:
            raise ArithmeticError('producer/consumer coordinate checksum mismatch')
        us, vs = DeviceArray.empty(nu), DeviceArray.empty(nv)
        if nu:
            native.call('pcube_upload', us.address, slots[slot].pointer, nu * 8)
        if nv:
            native.call('pcube_upload', vs.address, slots[slot].pointer + 8 * CAP, nv * 8)
        first = index * BATCH
        prepared = prepare_arrays(runtime, config, first, first + BATCH, WIDTH, us, vs, t)
        upload += time.perf_counter() - t
        # the device holds a complete copy (the uploads return once it has arrived)
        free.put((owner, slot))
        t = time.perf_counter()
        row = matcher.run(prepared)
        filtering += time.perf_counter() - t
        rows.append(dict(job=index, signature=signature(row), candidates=row['valid_focused_candidates'], row=row))
    return dict(role='consumer', status='PASS', seconds=time.perf_counter() - begun, waiting_seconds=waiting,
                upload_seconds=upload, filter_seconds=filtering, rows=rows)


def fleet(runtime: Runtime, config: Config, af: GPUFocus, bf: GPUFocus, mode: str, jobs: list[int],
          consumers: int = 1, slots_per_consumer: int = 2) -> dict[str, Any]:
    """Run one producer and `consumers` matcher consumers over the batch indices `jobs`
    (tools/batch_window_experiment.py:339-404, the producer/consumer modes).  Every row
    comes back, ordered by job; a missing or repeated job raises ArithmeticError."""
    if mode not in MODES:
        raise ValueError(f'unknown staging mode {mode!r}')
    if not jobs or consumers < 1 or slots_per_consumer < 1:
        raise ValueError('staging needs jobs, consumers and slots')
    masked = mode.startswith('masked-')
    producer = _GPUProducer(config, af, bf) if mode.endswith('gpu-producer') else _CPUProducer(config, af, bf)
    matchers = [(MaskedMatcher if masked else FusedMatcher)(runtime, config) for _ in range(consumers)]
    slots = [PinnedBuffer(2 * CAP) for _ in range(consumers * slots_per_consumer)]
    free: queue.Queue = queue.Queue()
    for i in range(consumers):
        for k in range(slots_per_consumer):
            free.put((i, i * slots_per_consumer + k))
    tasks = [queue.Queue(maxsize=slots_per_consumer) for _ in range(consumers)]
    stop = threading.Event()
    results: dict[str, Any] = {}
    errors: list[BaseException] = []

    def guarded(name: str, fn: Any, *args: Any) -> None:
        try:
            results[name] = fn(*args)
        except BaseException as exc:   # recorded and re-raised by the caller
            errors.append(exc)
            stop.set()

    def produce() -> None:
        guarded('producer', pipeline_producer, producer, config, list(jobs), slots, free, tasks, stop)
        for q in tasks:
            q.put(None)

    threads = [threading.Thread(target=produce, name='staging-producer')]
    threads += [threading.Thread(target=guarded, name=f'staging-consumer-{i}',
                                 args=(f'consumer-{i}', pipeline_consumer, runtime, config, matchers[i], i, slots,
                                       free, tasks[i], stop)) for i in range(consumers)]
    begun = time.perf_counter()
    try:
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    finally:
        producer.close()
        for s in slots:
            s.close()
    elapsed = time.perf_counter() - begun
    if errors:
        raise errors[0]
    rows = sorted((r for i in range(consumers) for r in results[f'consumer-{i}']['rows']), key=lambda r: r['job'])
    if [r['job'] for r in rows] != sorted(jobs) or len(set(jobs)) != len(jobs):
        raise ArithmeticError('missing or duplicate staging jobs')
    workers = [results['producer']] + [{k: v for k, v in results[f'consumer-{i}'].items() if k != 'rows'}
                                       for i in range(consumers)]
    candidates = sum(r['candidates'] for r in rows)
    return dict(status='PASS', mode=mode, jobs=len(jobs), consumers=consumers, elapsed_seconds=elapsed,
                workers=workers, rows=rows, candidates=candidates, candidates_per_second=candidates / elapsed)
