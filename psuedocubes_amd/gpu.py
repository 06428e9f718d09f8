"""Device runtime, focus windows and the per-subtile matcher on MI355X.

Drop-in for the reference's `pseudocube/gpu.py` (`Runtime` 38-67, `GPUFocus`
70-114, `launch_plan` 117-126, `Matcher` 185-252).  Names, arguments, result
dictionaries and exceptions follow the reference; the work runs in the HIP
kernels of libpcube.so through the C ABI of include/pcube.h.
"""
from __future__ import annotations

import ctypes
import hashlib
import threading
import time
from math import prod
import weakref
from typing import Any, Callable

import numpy as np

from . import native
from .focus import Config, Focus, Tile
from .predicate import is_perfect_cube, local3

MASK = (1 << 64) - 1


# Free lists of DeviceArray.pooled: at most _POOL_DEPTH buffers per size.
_POOL: dict[int, list[int]] = {}
_POOL_LOCK = threading.Lock()
_POOL_DEPTH = 4


class DeviceArray:
    """A uint64 array in device memory (the role CuPy arrays play in the reference)."""

    def __init__(self, address: int, size: int, owner: bool = True, capacity: int | None = None) -> None:
        self.address = int(address)
        self.size = int(size)
        self.capacity = int(capacity if capacity is not None else size)
        self.owner = owner
        self.dtype = np.uint64

    @classmethod
    def empty(cls, size: int) -> 'DeviceArray':
        address = ctypes.c_uint64()
        native.call('pcube_alloc', max(1, int(size)) * 8, ctypes.byref(address))
        return cls(address.value, size, True, size)

    @classmethod
    def pooled(cls, size: int) -> 'DeviceArray':
        """An owning array of size words whose memory, once the array is dropped, goes back to a
        small per-size free list instead of the library (prepare()'s per-batch windows).  Reuse
        is safe because every device use of a buffer is ordered on the library's one stream."""
        with _POOL_LOCK:
            free = _POOL.get(size)
            address = free.pop() if free else None
        if address is None:
            out = cls.empty(size)
        else:
            out = cls(address, size, True, size)
        out.pooled_size = size
        return out

    @classmethod
    def from_host(cls, values: Any) -> 'DeviceArray':
        host = np.ascontiguousarray(np.asarray(values, dtype=np.uint64))
        out = cls.empty(host.size)
        if host.size:
            native.call('pcube_upload', out.address, host.ctypes.data, host.nbytes)
        return out

    @property
    def nbytes(self) -> int:
        return self.size * 8

    def __len__(self) -> int:
        return self.size

    def get(self) -> np.ndarray:
        host = np.empty(self.size, dtype=np.uint64)
        if self.size:
            native.call('pcube_download', host.ctypes.data, self.address, host.nbytes)
        return host

    def __getitem__(self, key: slice) -> 'DeviceArray':
        if not isinstance(key, slice):
            raise TypeError('device arrays support slices only')
        start, stop, step = key.indices(self.size)
        if step != 1:
            raise ValueError('strided device slices are not supported')
        return DeviceArray(self.address + 8 * start, max(0, stop - start), owner=False)

    def __del__(self) -> None:
        if getattr(self, 'owner', False) and native._lib is not None:
            size = getattr(self, 'pooled_size', None)
            if size is not None:
                with _POOL_LOCK:
                    free = _POOL.setdefault(size, [])
                    if len(free) < _POOL_DEPTH:
                        free.append(self.address)
                        self.owner = False
                        return
            try:
                native.lib().pcube_free(self.address)
            except Exception:
                pass
            self.owner = False


class Runtime:
    """One device, one stream, one capped memory pool (`pseudocube/gpu.py:38-67`).

    cap = min(memory_cap, total * memory_fraction, free - 256 MiB).
    """

    def __init__(self, deadline: float, memory_cap: int, memory_fraction: float = 0.5, device: int = 0) -> None:
        lib = native.lib()
        count = ctypes.c_int()
        native.check(lib.pcube_device_count(ctypes.byref(count)))
        if count.value < 1:
            raise RuntimeError('no HIP device is visible')
        cap, free, total = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        native.check(lib.pcube_init(device, int(memory_cap), float(memory_fraction), ctypes.byref(cap),
                                    ctypes.byref(free), ctypes.byref(total)), 'pcube_init')
        self.deadline = deadline
        self.cap = cap.value
        self.total_device_bytes = total.value
        self.compiles: list[dict[str, Any]] = []
        self.peak_device = 0
        name = ctypes.create_string_buffer(256)
        cus = ctypes.c_int()
        native.check(lib.pcube_device_info(name, 256, ctypes.byref(cus)))
        self.device_name = name.value.decode()
        self.compute_units = cus.value
        self._library_sha = hashlib.sha256(native.LIBRARY.read_bytes()).hexdigest()
        # per configuration, the matcher whose match prepare() starts with each batch
        self._batch_matchers: dict[tuple[int, ...], weakref.ref] = {}

    @staticmethod
    def _config_key(config: Config) -> tuple[int, ...]:
        return (config.p, config.lo, config.hi, config.A, config.B)

    def register_matcher(self, config: Config, matcher: Any) -> None:
        """prepare() of a batch of config starts its match on matcher (the latest registered)."""
        self._batch_matchers[self._config_key(config)] = weakref.ref(matcher)

    def matcher_for(self, config: Config) -> Any:
        ref = self._batch_matchers.get(self._config_key(config))
        return ref() if ref is not None else None

    def device_count(self) -> int:
        """Visible devices (the reference asks CuPy, `tools/reproduce.py:36`)."""
        count = ctypes.c_int()
        native.check(native.lib().pcube_device_count(ctypes.byref(count)))
        return count.value

    def memory(self) -> tuple[int, int]:
        reserved, peak = ctypes.c_uint64(), ctypes.c_uint64()
        native.check(native.lib().pcube_memory(ctypes.byref(reserved), ctypes.byref(peak), None))
        return reserved.value, peak.value

    def check(self) -> None:
        if time.monotonic() >= self.deadline - 3:
            raise TimeoutError('BUDGET_EXHAUSTED')
        # called several times per batch: the pool's peak read in place (the
        # library keeps it in one word), with no call
        try:
            peak = self._peak_word
        except AttributeError:
            peak = self._peak_word = ctypes.c_uint64.from_address(native.lib().pcube_memory_peak())
        if peak.value > self.peak_device:
            self.peak_device = peak.value
        if self.peak_device > self.cap:
            raise MemoryError('device memory cap exceeded')

    def compile(self, name: str) -> dict[str, Any]:
        """Record the attributes of a prebuilt kernel (`pseudocube/gpu.py:54-61`).

        Kernels are compiled ahead of time for gfx950; the record keeps the
        reference's keys with the library hash as the source hash.
        """
        self.check()
        started = time.perf_counter()
        lib = native.lib()
        index = 0
        buf = ctypes.create_string_buffer(64)
        regs, local, shared = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        while lib.pcube_kernel_info(index, buf, 64, ctypes.byref(regs), ctypes.byref(local), ctypes.byref(shared)) == 0:
            if buf.value.decode() == name:
                record = dict(name=name, source_sha256=self._library_sha, wall_seconds=time.perf_counter() - started,
                              registers=regs.value, local_bytes=local.value, shared_bytes=shared.value)
                self.compiles.append(record)
                return record
            index += 1
        raise ValueError(f'unknown kernel {name}')

    def measured(self, fn: Callable[..., Any], *args: Any) -> float:
        """Run one synchronous library call and return its wall seconds (`pseudocube/gpu.py:63-67`)."""
        self.check()
        started = time.perf_counter()
        fn(*args)
        elapsed = time.perf_counter() - started
        self.check()
        return elapsed


def focus_terms(halves: tuple[tuple[tuple[int, ...], ...], tuple[tuple[int, ...], ...]]) -> tuple[np.ndarray, np.ndarray]:
    first, second = halves
    terms = np.array([t for ring in first + second for t in ring] or [0], dtype=np.uint64)
    sizes = np.array([len(ring) for ring in first + second] or [1], dtype=np.uint32)
    return terms, sizes


class GPUFocus:
    """Sorted CRT half arrays resident on the device (`pseudocube/gpu.py:70-114`)."""

    def __init__(self, runtime: Runtime, focus: Focus, wheel: bool = True) -> None:
        self.runtime, self.focus = runtime, focus
        first, second = focus.halves()
        self.half_sizes = [prod(map(len, first)), prod(map(len, second))]
        # two uint64 arrays plus their sort scratch, as the reference budgets
        if sum(self.half_sizes) * 32 > runtime.cap // 2:
            raise MemoryError('focus halves exceed pilot budget')
        terms, sizes = focus_terms((first, second))
        handle, nc, nd = ctypes.c_int(), ctypes.c_uint64(), ctypes.c_uint64()
        runtime.check()
        native.call('pcube_focus_create', focus.modulus, native.ptr64(terms), native.ptr32(sizes), len(first),
                    len(second), ctypes.byref(handle), ctypes.byref(nc), ctypes.byref(nd))
        self.handle = handle.value
        c_addr, d_addr = ctypes.c_uint64(), ctypes.c_uint64()
        native.call('pcube_focus_arrays', self.handle, ctypes.byref(c_addr), ctypes.byref(d_addr))
        # D is sorted here too (the reference leaves it unsorted); C is sorted in both
        self.C = DeviceArray(c_addr.value, nc.value, owner=False)
        self.D = DeviceArray(d_addr.value, nd.value, owner=False)
        self.wheel = None
        if wheel:
            # windows from the CRT wheel: inner ring prefix residues plus outer-ring tests
            moduli = np.array([q for q, _ in focus.rings] or [1], dtype=np.uint32)
            counts = np.array([len(r) for _, r in focus.rings] or [1], dtype=np.uint32)
            residues = np.array([a for _, r in focus.rings for a in r] or [0], dtype=np.uint32)
            inner, n1 = ctypes.c_uint64(), ctypes.c_uint32()
            native.call('pcube_focus_wheel', self.handle, native.ptr32(moduli), native.ptr32(counts),
                        native.ptr32(residues), len(focus.rings), ctypes.byref(inner), ctypes.byref(n1))
            self.wheel = dict(inner_modulus=inner.value, inner_residues=n1.value)
            runtime.compile('wheel_slots')
        else:
            runtime.compile('window_count')
        runtime.check()

    def window(self, low: int, high: int, cap: int = 1000000) -> DeviceArray:
        """Sorted focus coordinates in [low, high) on the device (`pseudocube/gpu.py:94-114`)."""
        if not 0 <= low <= high < 1 << 64:
            raise ValueError('invalid bounded focus window')
        if high == low:
            return DeviceArray.empty(0)
        modulus = self.focus.modulus
        if (high - 1) // modulus - low // modulus + 1 > 4096:
            raise OverflowError('focus window exceeds bounded period count')
        out = DeviceArray.empty(max(1, cap))
        count = ctypes.c_uint64()
        self.runtime.check()
        code = native.lib().pcube_focus_window(self.handle, low, high, out.address, cap, ctypes.byref(count))
        if code == native.OVERFLOW:
            raise OverflowError('focus buffer overflow; tile invalid')
        native.check(code, 'pcube_focus_window')
        out.size = count.value
        self.runtime.check()
        return out

    def windows(self, lows: list[int], highs: list[int], cap: int | None = None) -> tuple[DeviceArray, list[int]]:
        """Many windows in one pass: (concatenated sorted coordinates, offsets[0..n])."""
        if self.wheel is None:
            raise ValueError('multi-window generation needs the focus wheel')
        n = len(lows)
        lo = np.array(lows, dtype=np.uint64)
        hi = np.array(highs, dtype=np.uint64)
        offsets = np.zeros(n + 1, dtype=np.uint64)
        if cap is None:
            expect = sum(h - l for l, h in zip(lows, highs)) * self.focus.count / self.focus.modulus
            cap = int(expect * 1.05) + 64 * n + 4096
        out = DeviceArray.empty(max(1, cap))
        self.runtime.check()
        code = native.lib().pcube_focus_windows(self.handle, n, native.ptr64(lo), native.ptr64(hi), out.address,
                                                cap, native.ptr64(offsets))
        if code == native.OVERFLOW:
            return self.windows(lows, highs, int(offsets[n]) + 1)
        native.check(code, 'pcube_focus_windows')
        out.size = int(offsets[n])
        return out, [int(x) for x in offsets]

    def __del__(self) -> None:
        if native._lib is not None and hasattr(self, 'handle'):
            try:
                native.lib().pcube_focus_destroy(self.handle)
            except Exception:
                pass


def launch_plan(nu: int, nv: int, cu_count: int, layout: str, blocks_per_cu: int = 2) -> tuple[int, int]:
    """(blocks, waves per v row) of a row-striped launch (`pseudocube/gpu.py:117-126`).

    Restated for wave64: a 256-thread block holds four 64-lane waves, and a
    row of nu positions is striped over at most ceil(nu / 64) waves.
    """
    if layout not in ('serial', 'warp') or type(blocks_per_cu) is not int or blocks_per_cu not in (1, 2, 4):
        raise ValueError('invalid matching layout')
    if min(nu, nv) < 0 or cu_count < 1:
        raise ValueError('invalid launch dimensions')
    if not nu or not nv:
        return 0, 1
    if layout == 'serial':
        return (nv + 255) // 256, 1
    stripes = min((nu + 63) // 64, max(1, (cu_count * blocks_per_cu * 4 + nv - 1) // nv))
    return (nv * stripes + 3) // 4, stripes


class Matcher:
    """Per-subtile engine without bitsets (`pseudocube/gpu.py:185-252`).

    Runs libpcube's direct kernel: every focused position goes through the
    whole filter chain with x = uB - vA built in 128 bits.  It is the
    baseline the batched engine is compared against.
    """

    def __init__(self, runtime: Runtime, config: Config, variant: str = 'byte',
                 layout: str = 'warp', blocks_per_sm: int = 2) -> None:
        if variant not in ('byte', 'packed', 'constant'):
            raise ValueError('unknown lookup variant')
        if layout not in ('serial', 'warp'):
            raise ValueError('unknown matching layout')
        self.runtime, self.config, self.variant = runtime, config, variant
        self.layout, self.blocks_per_sm = layout, blocks_per_sm
        self.sm_count = runtime.compute_units
        launch_plan(0, 0, self.sm_count, layout, blocks_per_sm)
        self.engine = MatchEngine(runtime, config)
        self.filter_groups = self.engine.filter_groups
        self.kernel = runtime.compile('match_direct')

    @property
    def table(self) -> DeviceTable:
        """The device allowed-residue tables (`pseudocube/gpu.py:199`, the reference's `self.table`)."""
        return self.engine.table

    def arrays(self, us: DeviceArray, vs: DeviceArray, cap: int = 8192) -> dict[str, Any]:
        started = time.perf_counter()
        ng = len(self.filter_groups)
        if int(us.size) * int(vs.size) > 50_000_000:
            raise OverflowError('bounded pair-opportunity limit exceeded')
        if cap < 0:
            raise ValueError('negative output capacity')
        blocks, stripes = launch_plan(int(us.size), int(vs.size), self.sm_count, self.layout, self.blocks_per_sm)
        populations, per_tile, seconds = self.engine.run(us, vs, 0, 1 << 63, 1, max(cap, 0), engine=1,
                                                         overflow_ok=True)
        elapsed = seconds[1]
        transfer = time.perf_counter()
        populations = [int(x) for x in populations[0]]
        number = populations[ng + 1]
        if populations[ng + 2]:
            raise ArithmeticError('unsigned underflow flagged')
        if number > cap:
            raise OverflowError('output overflow; attempt cannot commit')
        values = sorted(per_tile[0])
        transfer = time.perf_counter() - transfer
        verification = time.perf_counter()
        if len(values) != len(set(values)):
            raise ArithmeticError('duplicate canonical output')
        if any(not self.config.lo < n <= self.config.hi or not local3(n, self.config.p) for n in values):
            raise ArithmeticError('GPU/CPU disagreement: attempt quarantined')
        cube_flags = [is_perfect_cube(n) for n in values]
        cubes = [n for n, cube in zip(values, cube_flags) if cube]
        hits = [n for n, cube in zip(values, cube_flags) if not cube]
        verification = time.perf_counter() - verification
        return dict(status='PASS', host_verified=True, variant=self.variant, focus_u=int(us.size), focus_v=int(vs.size),
                    layout=self.layout, blocks_per_sm=self.blocks_per_sm, warps_per_row=stripes,
                    launch_blocks=blocks, threads_per_block=256, launched_warps=blocks * 4,
                    cartesian_opportunities=int(us.size) * int(vs.size), pairs_visited=populations[0],
                    valid_focused_candidates=populations[0], wide_constructions=populations[0],
                    filter_groups=self.filter_groups, group_survivors=populations[1:ng + 1], local_survivors=number,
                    cube_checks=number, genuine_cubes=len(cubes), noncube_reports=len(hits),
                    cpu_confirmed_hits=list(map(str, hits)), cpu_rejected_reports=0,
                    all_values=list(map(str, values)), kernel_seconds=elapsed, transfer_seconds=transfer,
                    verification_seconds=verification, end_to_end_seconds=time.perf_counter() - started)

    def coordinates(self, us: list[int], vs: list[int], cap: int = 8192) -> dict[str, Any]:
        started = time.perf_counter()
        u = DeviceArray.from_host(np.array(sorted(us), dtype=np.uint64))
        v = DeviceArray.from_host(np.array(vs, dtype=np.uint64))
        transfer = time.perf_counter() - started
        result = self.arrays(u, v, cap)
        result['input_transfer_seconds'] = transfer
        result['end_to_end_seconds'] = time.perf_counter() - started
        return result

    def tile(self, tile: Tile, af: GPUFocus, bf: GPUFocus, cap: int = 1000000,
             output_cap: int = 8192) -> dict[str, Any]:
        tile.validate(self.config)
        started = time.perf_counter()
        lower = self.config.u_interval(tile.first)[0]
        upper = self.config.u_interval(tile.end - 1)[1]
        us = af.window(lower, upper, cap)
        vs = bf.window(tile.first, tile.end, cap)
        prep = time.perf_counter() - started
        result = self.arrays(us, vs, output_cap)
        result['focus_enumeration_seconds'] = prep
        result['tile'] = tile.manifest(self.config)
        result['end_to_end_seconds'] = time.perf_counter() - started
        return result


class DeviceTable:
    """A matcher's device allowed-residue tables, with the one operation the
    reference applies to its device byte table, `fill` (`REL!tools/gpu_worker.py:96-103`):
    a nonzero value marks every residue of every filter allowed, zero none.
    Every later batch of the matcher reads the filled tables."""

    def __init__(self, engine: 'MatchEngine') -> None:
        self.engine = engine

    def fill(self, value: int) -> None:
        native.call('pcube_matcher_table_fill', self.engine.handle, 1 if value else 0)


class MatchEngine:
    """One configuration's filter tables resident on the device (a libpcube matcher)."""

    def __init__(self, runtime: Runtime, config: Config) -> None:
        self.runtime, self.config = runtime, config
        order = config.filter_order()
        if not order:
            raise ValueError('no residual filters')
        self.order = order
        self.filter_groups = [order[i:i + 4] for i in range(0, len(order), 4)]
        self.ng = len(self.filter_groups)
        moduli = np.array(order, dtype=np.uint32)
        handle = ctypes.c_int()
        native.call('pcube_matcher_create', config.A, config.B, config.lo & MASK, config.lo >> 64,
                    config.hi & MASK, config.hi >> 64, native.ptr32(moduli), len(order), ctypes.byref(handle))
        self.handle = handle.value
        self.table = DeviceTable(self)
        self._values: np.ndarray | None = None
        self._values_ptr: Any = None
        self._counters = np.empty(0, dtype=np.uint64)   # host copy of the counters, reused
        self._counters_ptr: Any = None
        self._seconds = (ctypes.c_double * 3)()

    def adopt(self, launch: int, ntiles: int, cap: int) -> bool:
        """Take over the batch that pcube_prepare_match started with token launch, if it is still
        this matcher's pending batch; finish() then collects it."""
        if not launch:
            return False
        self.runtime.check()
        ok = ctypes.c_int()
        native.check(native.lib().pcube_match_adoptable(self.handle, launch, ctypes.byref(ok)), 'pcube_match_adoptable')
        if ok.value:
            self._pending = (ntiles, cap)
        return bool(ok.value)

    def run(self, us: DeviceArray, vs: DeviceArray, v_first: int, tile_width: int, ntiles: int, cap: int,
            engine: int, overflow_ok: bool = False) -> tuple[np.ndarray, list[list[int]], list[float]]:
        """Counters [ntiles, ng+3], emitted values per subtile, (prefilter, match, staged match kernel) seconds.

        With overflow_ok an output overflow returns normally so the caller can
        raise with the reference's wording; other failures raise.
        """
        self.start(us, vs, v_first, tile_width, ntiles, cap, engine)
        return self.finish(overflow_ok)

    def start(self, us: DeviceArray, vs: DeviceArray, v_first: int, tile_width: int, ntiles: int, cap: int,
              engine: int) -> None:
        """Enqueue one batch on the device and return at once; finish() collects it."""
        self.runtime.check()
        native.check(native.lib().pcube_match_start(self.handle, us.address, us.size, vs.address, vs.size, v_first,
                                                    tile_width, ntiles, cap, engine), 'pcube_match_start')
        self._pending = (ntiles, cap)

    def finish(self, overflow_ok: bool = False) -> tuple[np.ndarray, list[list[int]], list[float]]:
        counters, limbs, kept, seconds = self.finish_raw(overflow_ok)
        if kept is None or not any(kept):
            return counters, [[] for _ in range(len(counters))], seconds
        words = limbs.tolist()
        per_tile: list[list[int]] = []
        offset = 0
        for n in kept:
            per_tile.append([words[2 * i] + (words[2 * i + 1] << 64) for i in range(offset, offset + n)])
            offset += n
        return counters, per_tile, seconds

    def finish_raw(self, overflow_ok: bool = False) -> tuple[np.ndarray, np.ndarray, list[int] | None, list[float]]:
        """finish() without Python integers: counters, the emitted values as (low, high) uint64
        limbs subtile after subtile, the number listed per subtile (None when a subtile
        overflowed and overflow_ok holds), and the timings."""
        ntiles, cap = self._pending
        stride = self.ng + 3
        if self._counters.size < ntiles * stride:
            self._counters = np.zeros(ntiles * stride, dtype=np.uint64)
            self._counters_ptr = native.ptr64(self._counters)
        values_cap = min(max(1, ntiles * cap), 1 << 22)
        if self._values is None or self._values.size < 2 * values_cap:
            self._values = np.empty(2 * values_cap, dtype=np.uint64)
            self._values_ptr = native.ptr64(self._values)
        values = self._values
        seconds = self._seconds
        code = native.lib().pcube_match_finish(self.handle, self._counters_ptr, self._values_ptr,
                                               values.size // 2, seconds)
        # the caller owns its counters: a copy of the reused host buffer
        counters = self._counters[:ntiles * stride].reshape(ntiles, stride).copy()
        # overflow_ok covers a subtile that emitted more than cap, which the caller
        # reports itself; any other overflow (or a list/counter mismatch) raises here
        timings = [seconds[0], seconds[1], seconds[2]]
        if code == native.OVERFLOW and overflow_ok and bool((counters[:, self.ng + 1] > cap).any()):
            return counters, values[:0], None, timings
        native.check(code, 'pcube_match')
        self.runtime.check()
        emitted = counters[:, self.ng + 1]
        if not emitted.any():   # the usual full p = 619 batch
            return counters, values[:0], [0] * ntiles, timings
        kept = np.minimum(emitted, cap).tolist()
        return counters, values[:2 * sum(kept)].copy(), kept, timings

    def __del__(self) -> None:
        if native._lib is not None and hasattr(self, 'handle'):
            try:
                native.lib().pcube_matcher_destroy(self.handle)
            except Exception:
                pass
