"""GPU parity: every HIP kernel path against the CPU oracle and the golden fixtures.

Recipes follow the reference's GPU validation scripts:
tools/batch_window_experiment.py:100-205 (`single`: four batch scopes,
complete small sets, forced overflow, high-magnitude canaries, CPU producer
arrays equal GPU arrays) and the release's tools/qualify_optimized.py:36-61
(full-lease equivalence).  All comparisons are bit-exact.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
from pathlib import Path
import random

import numpy as np
import pytest

from oracle import oracle as O
from psuedocubes_amd import native
from psuedocubes_amd.accounting import encoded
from psuedocubes_amd.batch_experiment import FusedMatcher, MaskedMatcher, compare, prepare, tile_ranges
from psuedocubes_amd.focus import A619, B619, Config, Tile
from psuedocubes_amd.gpu import DeviceArray, GPUFocus, Matcher
from psuedocubes_amd.optimized import BATCH_WIDTH, compute_lease
from psuedocubes_amd.predicate import local3
from psuedocubes_amd.run_production import compute_batch

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).parent / 'golden'
LEASE = json.loads((GOLDEN / 'lease_575222931200000000.json').read_text())
LEASE_BATCHES = json.loads((GOLDEN / 'lease_575222931200000000_batches.json').read_text())['batches']
LEASE_FIRST = int(LEASE['first'])


def small_config(p: int, lo: int = 0, hi: int = 100000) -> Config:
    return Config(p, lo, hi, 14 * (19 if p >= 19 else 1), 9 * (13 if p >= 13 else 1) * (31 if p >= 31 else 1))


def scopes_hash(row: dict) -> str:
    return hashlib.sha256(encoded(row['subtile_scopes'])).hexdigest()


def oracle_row_fields(row: dict) -> dict:
    keys = ('focus_u', 'focus_v', 'cartesian_opportunities', 'valid_focused_candidates', 'local_survivors',
            'subtile_scopes', 'group_survivors', 'filter_groups', 'all_values')
    return {k: row[k] for k in keys}


@pytest.fixture(scope='module')
def masked(runtime, config619):
    return MaskedMatcher(runtime, config619)


def test_lds_probe_patterns(runtime) -> None:
    """The LDS probe behind the roofline's LDS peak: every pattern reads the bytes
    it reports, and the staged engine's lane-rotated reads beat random per-lane
    units by more than 2x (no bank conflicts against 16-lane conflicts)."""
    seconds, nbytes = ctypes.c_double(), ctypes.c_double()
    rates = []
    for pattern in range(3):
        native.call('pcube_lds_probe', pattern, 200, ctypes.byref(seconds), ctypes.byref(nbytes))
        assert nbytes.value == runtime.compute_units * 1024 * 200 * 16 * 16
        assert seconds.value > 0
        rates.append(nbytes.value / seconds.value)
    assert rates[1] > 2 * rates[2]
    with pytest.raises(ValueError):
        native.call('pcube_lds_probe', 3, 1, ctypes.byref(seconds), ctypes.byref(nbytes))


def test_wide_limb_kernel(runtime) -> None:
    rng = random.Random(619003)
    cases = [(0, 0, 0, 0), (2**64 - 1, 2**64 - 1, 0, 0), (0, 0, 2**64 - 1, 2**64 - 1),
             (2**32, 2**32, 1, 1), (2**64 - 1, 1, 0, 0)]
    cases += [tuple(rng.getrandbits(64) for _ in range(4)) for _ in range(2000)]
    flat = np.array([x for c in cases for x in c], dtype=np.uint64)
    out = np.zeros(3 * len(cases), dtype=np.uint64)
    native.call('pcube_arithmetic', native.ptr64(flat), native.ptr64(out), len(cases))
    for i, (a, b, c, d) in enumerate(cases):
        assert int(out[3 * i]) + (int(out[3 * i + 1]) << 64) == (a * b - c * d) % (1 << 128)
        assert int(out[3 * i + 2]) == int(a * b < c * d)


def test_floor_div_kernel(runtime) -> None:
    """The exact 128/64 division behind every row's u bounds (u_interval,
    pseudocube/focus.py:120-121): Python's integer division on both coefficients,
    small and maximal divisors, exact multiples and their neighbours."""
    rng = random.Random(619121)
    divisors = [B619, A619, 1, 2, 3, 9 * 19 * 37, 2**32 - 1, 2**32 + 1, 2**63, 2**63 + 1, 2**64 - 1]
    cases = []
    for B in divisors:
        cases += [(0, B), (1, B), (B - 1, B), (B, B), (B + 1, B), (2**128 - 1, B), (2**127, B)]
        cases += [(rng.getrandbits(rng.choice([64, 96, 124, 128])), B) for _ in range(300)]
        multiples = [B * rng.getrandbits(rng.choice([30, 64, 66])) for _ in range(100)]
        cases += [(min(2**128 - 1, max(0, k + d)), B) for k in multiples for d in (-1, 0, 1)]
    flat = np.array([w for x, B in cases for w in (x & (2**64 - 1), x >> 64, B)], dtype=np.uint64)
    out = np.zeros(2 * len(cases), dtype=np.uint64)
    native.call('pcube_floor_div', native.ptr64(flat), native.ptr64(out), len(cases))
    for i, (x, B) in enumerate(cases):
        assert int(out[2 * i]) + (int(out[2 * i + 1]) << 64) == x // B, (x, B)
    bad = np.array([5, 0, 0], dtype=np.uint64)
    with pytest.raises(ValueError):
        native.call('pcube_floor_div', native.ptr64(bad), native.ptr64(out), 1)


def test_bit_transpose_kernel(runtime) -> None:
    """The 64 x 64 bit wave transpose the residue-bitset builders apply
    (pseudocube/batch_experiment.py:184-194 builds the same rows one test at a
    time): row r of the result holds bit r of every column, against numpy's
    bit unpacking on zero, full, identity, single-bit and seeded matrices."""
    rng = np.random.default_rng(619184)
    mats = [np.zeros(64, dtype=np.uint64), np.full(64, 2**64 - 1, dtype=np.uint64),
            np.array([1 << t for t in range(64)], dtype=np.uint64),
            np.array([1 << (63 - t) for t in range(64)], dtype=np.uint64)]
    for k in range(64):   # one set bit at (k, 63 - k) and at (k, k)
        m = np.zeros(64, dtype=np.uint64)
        m[63 - k] = np.uint64(1) << np.uint64(k)
        m[k] |= np.uint64(1) << np.uint64(k)
        mats.append(m)
    mats += list(rng.integers(0, 2**64, size=(700, 64), dtype=np.uint64, endpoint=False))
    cols = np.ascontiguousarray(np.stack(mats))
    out = np.zeros_like(cols)
    native.call('pcube_bit_transpose', native.ptr64(cols.ravel()), native.ptr64(out.ravel()), cols.shape[0])
    bits = np.unpackbits(cols.view(np.uint8).reshape(-1, 64, 8), axis=2, bitorder='little')   # [m, t, k]
    expect = np.packbits(bits.transpose(0, 2, 1), axis=2, bitorder='little').reshape(-1, 64, 8)
    assert np.array_equal(out.view(np.uint8).reshape(-1, 64, 8), expect)
    assert np.array_equal(out[2], cols[2]) and np.array_equal(out[0], cols[0])   # identity and zero are fixed


def test_focus_halves_equal_oracle(focuses619, oracle_focuses619) -> None:
    for gpu, cpu in zip(focuses619, oracle_focuses619):
        assert np.array_equal(gpu.C.get(), cpu.C)
        assert np.array_equal(gpu.D.get(), cpu.D)
    assert focuses619[0].half_sizes == [2096640, 2009280]
    assert focuses619[1].half_sizes == [2592000, 2585088]


def test_windows_equal_oracle(focuses619, oracle_focuses619) -> None:
    rng = random.Random(619003)
    for gpu, cpu in zip(focuses619, oracle_focuses619):
        M = gpu.focus.modulus
        cases = [(0, 1), (0, 9_370_000_000), (M - 5_000_000_000, M + 5_000_000_000), (M - 1, M + 1),
                 (2 * M - 10**9, 2 * M), (12345, 12345)]
        cases += [(s, s + rng.choice([10**6, 10**8, 6_400_000_000, 9_400_000_000]))
                  for s in (rng.randrange(2 * M) for _ in range(12))]
        for low, high in cases:
            got = gpu.window(low, high).get()
            assert np.array_equal(got, cpu.window(low, high)), (low, high)
        with pytest.raises(OverflowError):
            gpu.window(0, 10**12, cap=10)
        with pytest.raises(ValueError):
            gpu.window(5, 4)


def test_small_config_windows_all_periods(runtime) -> None:
    for p in (7, 13, 19, 31):
        config = small_config(p)
        for focus in config.focuses():
            cpu = O.Focus(focus)
            for gpu in (GPUFocus(runtime, focus), GPUFocus(runtime, focus, wheel=False)):
                M = focus.modulus
                for low, high in [(0, M), (1, 3 * M - 1), (M - 1, M + 2), (0, 4000 * M), (17, 17)]:
                    assert np.array_equal(gpu.window(low, high).get(), cpu.window(low, high, cap=10**7)), (p, low, high)
                with pytest.raises(OverflowError):
                    gpu.window(0, 4097 * M)


def test_directory_walk_windows_equal_oracle(runtime, config619, oracle_focuses619) -> None:
    # the exact directory-walk generator (used when a focus has no wheel) on p=619 windows
    rng = random.Random(77)
    for focus, cpu in zip(config619.focuses(), oracle_focuses619):
        gpu = GPUFocus(runtime, focus, wheel=False)
        M = focus.modulus
        for low, high in [(0, 9_370_000_000), (M - 10**9, M + 10**9)] + \
                [(s, s + 6_400_000_000) for s in (rng.randrange(2 * M) for _ in range(4))]:
            assert np.array_equal(gpu.window(low, high).get(), cpu.window(low, high)), (low, high)


def test_multi_window_generation(focuses619, oracle_focuses619) -> None:
    rng = random.Random(9)
    for gpu, cpu in zip(focuses619, oracle_focuses619):
        M = gpu.focus.modulus
        lows = [rng.randrange(2 * M) for _ in range(300)]
        highs = [lo + int(10 ** rng.uniform(0, 10)) for lo in lows]
        highs[3] = lows[3]
        lows[7], highs[7] = M - 5, M + 5
        values, offsets = gpu.windows(lows, highs)
        host = values.get()
        for w, (lo, hi) in enumerate(zip(lows, highs)):
            assert np.array_equal(host[offsets[w]:offsets[w + 1]], cpu.window(lo, hi)), w
        small, small_offsets = gpu.windows(lows[:5], highs[:5], cap=1)   # grows past an undersized cap
        assert small_offsets == [x - offsets[0] for x in offsets[:6]]


def test_device_searchsorted(runtime) -> None:
    from psuedocubes_amd.batch_experiment import searchsorted
    rng = np.random.default_rng(5)
    data = np.sort(rng.integers(0, 2**62, 100_000, dtype=np.uint64))
    arr = DeviceArray.from_host(data)
    queries = [0, int(data[0]), int(data[-1]), int(data[-1]) + 1, 2**63] + [int(x) for x in rng.integers(0, 2**62, 1000, dtype=np.uint64)]
    assert searchsorted(arr, queries) == np.searchsorted(data, np.array(queries, dtype=np.uint64)).tolist()


@pytest.mark.parametrize('index', [0, 1, 127, 255])
def test_masked_batches_of_reference_lease(runtime, config619, focuses619, masked, index) -> None:
    golden = LEASE_BATCHES[index]
    first = int(golden['first'])
    batch = prepare(runtime, config619, *focuses619, first, int(golden['end']))
    row = masked.run(batch)
    assert scopes_hash(row) == golden['scopes_sha256']
    assert row['valid_focused_candidates'] == golden['valid_focused_candidates']
    assert row['group_survivors'] == golden['group_survivors']
    assert row['all_values'] == golden['all_values']
    assert (row['focus_u'], row['focus_v']) == (golden['focus_u'], golden['focus_v'])


def seeded_scopes() -> list[tuple[int, int]]:
    """Batch 0, the seeded full batch, B/2 and the tail ending at B (tools/batch_window_experiment.py:117-118)."""
    rng = random.Random(619003)
    k = rng.randrange(B619 // BATCH_WIDTH)
    return [(0, BATCH_WIDTH), (k * BATCH_WIDTH, (k + 1) * BATCH_WIDTH),
            (246112000000000 - BATCH_WIDTH, 246112000000000), (B619 // 2, B619 // 2 + BATCH_WIDTH),
            (B619 - BATCH_WIDTH, B619), (B619 - 100000001, B619)]


@pytest.mark.parametrize('scope', seeded_scopes())
def test_masked_equals_oracle_on_seeded_batches(runtime, config619, focuses619, oracle_focuses619, masked, scope) -> None:
    first, end = scope
    row = masked.run(prepare(runtime, config619, *focuses619, first, end))
    expected = O.batch(config619, *oracle_focuses619, first, end)
    assert oracle_row_fields(row) == oracle_row_fields(expected)


def test_fused_direct_engine_equals_masked(runtime, config619, focuses619, masked) -> None:
    fused = FusedMatcher(runtime, config619)
    batch = prepare(runtime, config619, *focuses619, LEASE_FIRST, LEASE_FIRST + BATCH_WIDTH)
    a, b = fused.run(batch), masked.run(batch)
    compare(a, b)
    assert a['matching_layout'] == 'fused-batch' and b['matching_layout'] == 'residue-bitset-batch'


def test_complete_small_sets_through_gpu(runtime) -> None:
    for p in (7, 13, 19, 31):
        config = small_config(p)
        af, bf = (GPUFocus(runtime, f) for f in config.focuses())
        width = max(1, (config.B + 6) // 7)
        batch = prepare(runtime, config, af, bf, 0, config.B, width)
        expected = [str(n) for n in range(1, 100001) if local3(n, p)]
        for matcher in (FusedMatcher(runtime, config), MaskedMatcher(runtime, config)):
            row = matcher.run(batch)
            assert sorted(row['all_values'], key=int) == expected
        fa, fb = (O.Focus(f) for f in config.focuses())
        assert oracle_row_fields(MaskedMatcher(runtime, config).run(batch)) == \
            oracle_row_fields(O.batch(config, fa, fb, 0, config.B, width))
        if p == 7:
            for matcher in (FusedMatcher(runtime, config), MaskedMatcher(runtime, config)):
                with pytest.raises(OverflowError):
                    matcher.run(batch, output_cap=1)


@pytest.mark.parametrize('p,n', [(613, 674441580981249129037406633), (619, 1000000007**3), (619, 1000000009**3)])
def test_high_magnitude_canaries(runtime, focuses619, p, n) -> None:
    narrow = Config(p, n - 1, n, A619, B619)
    v = narrow.canonical(n)[1]
    first, end = max(0, v - 5000000), min(B619, v + 5000001)
    batch = prepare(runtime, narrow, *focuses619, first, end, 5000000)
    rows = [FusedMatcher(runtime, narrow).run(batch), MaskedMatcher(runtime, narrow).run(batch)]
    compare(rows[0], rows[1])
    assert rows[1]['all_values'] == [str(n)]
    legacy = compute_batch(Matcher(runtime, narrow), *focuses619, first, end, 5000000)
    compare(legacy, rows[1])


def test_full_reference_lease_pipelined(runtime, config619, focuses619, masked) -> None:
    result = compute_lease(masked, *focuses619, LEASE_FIRST, int(LEASE['end']))
    ref = LEASE['result']
    for key in ('focus_u', 'focus_v', 'cartesian_opportunities', 'valid_focused_candidates', 'genuine_cubes',
                'local_survivors', 'cube_checks', 'compute_batches', 'subtiles_executed', 'group_survivors',
                'filter_groups', 'all_values', 'cpu_confirmed_hits'):
        assert result[key] == ref[key], key
    assert result['subtile_scopes'] == ref['subtile_scopes']


def test_pipelined_lease_equals_per_batch_loop(runtime, config619, focuses619, masked) -> None:
    first = LEASE_FIRST + 7 * BATCH_WIDTH
    end = first + 2 * BATCH_WIDTH + 300_000_001
    a = compute_lease(masked, *focuses619, first, end, pipelined=True)
    b = compute_lease(masked, *focuses619, first, end, pipelined=False)
    strip = lambda r: {k: v for k, v in r.items() if not k.endswith('_seconds')}  # noqa: E731
    assert strip(a) == strip(b)
    assert a['subtiles_executed'] == 64 + 64 + 4 and a['compute_batches'] == 3


def test_pipelined_lease_over_several_passes(runtime, config619, focuses619, masked, monkeypatch) -> None:
    """Passes of at most three batches: a lease of seven (the last ragged) spans
    three device passes (3, 3, 1) and folds to the per-batch loop's receipt; a
    fold stopped after its first row leaves no pass behind."""
    from psuedocubes_amd import optimized
    monkeypatch.setattr(optimized, 'LEASE_CHUNK', 3)
    first = LEASE_FIRST + 40 * BATCH_WIDTH
    end = first + 6 * BATCH_WIDTH + 250_000_001
    rows = optimized.lease_rows(masked, *focuses619, first, end)
    s, e, row, _ = next(rows)
    assert (s, e) == (first, first + BATCH_WIDTH)
    rows.close()
    a = compute_lease(masked, *focuses619, first, end, pipelined=True)
    b = compute_lease(masked, *focuses619, first, end, pipelined=False)
    strip = lambda r: {k: v for k, v in r.items() if not k.endswith('_seconds')}  # noqa: E731
    assert strip(a) == strip(b)
    assert a['compute_batches'] == 7 and a['subtiles_executed'] == 6 * 64 + 3


def test_matcher_coordinates(runtime) -> None:
    """Matcher.coordinates (`pseudocube/gpu.py:236-244`) on host coordinate lists:
    every (u, v) pair with x = uB - vA in (lo, hi] is visited and the values
    passing the filters are exactly the local3 ones among them; the reference's
    complete case (every u and v of Config(13, 0, 100000, 14, 117)) gives the
    whole local3 set; a forced output overflow raises."""
    from psuedocubes_amd.focus import targeted_coordinates
    config = Config(7, 0, 100000, 14, 9)
    matcher = Matcher(runtime, config)
    for n in (1, 125, 2197, 4913, 6859):
        us, vs = targeted_coordinates(config, n, radius=3)
        xs = [u * config.B - v * config.A for u in us for v in vs]
        inside = [x for x in xs if config.lo < x <= config.hi]
        row = matcher.coordinates(us, vs)
        assert row['valid_focused_candidates'] == len(inside)
        assert row['all_values'] == [str(x) for x in sorted(x for x in inside if local3(x, 7))]
        assert row['focus_u'] == len(us) and row['focus_v'] == len(vs)
    full = Config(13, 0, 100000, 14, 117)
    fa, fb = full.focuses()
    us = list(fa.window(full.u_interval(0)[0], full.u_interval(full.B - 1)[1]))
    vs = list(fb.window(0, fb.modulus))
    row = Matcher(runtime, full).coordinates(us, vs)
    assert row['all_values'] == [str(n) for n in range(1, 100001) if local3(n, 13)]
    with pytest.raises(OverflowError):
        Matcher(runtime, full).coordinates(us, vs, cap=0)


def test_legacy_matcher_tile_and_split_replay(runtime) -> None:
    config = small_config(7)
    af, bf = (GPUFocus(runtime, f) for f in config.focuses())
    legacy = Matcher(runtime, config)

    class Tiny(Matcher):
        def tile(self, tile, af, bf, cap=1000000, output_cap=8192):
            return super().tile(tile, af, bf, cap, output_cap=1500)

    tiny = Tiny(runtime, config)
    full = compute_batch(legacy, af, bf, 0, config.B, config.B)
    split = compute_batch(tiny, af, bf, 0, config.B, config.B)
    assert full['overflow_splits'] == [] and split['overflow_splits']
    assert sorted(full['all_values'], key=int) == sorted(split['all_values'], key=int)
    assert full['valid_focused_candidates'] == split['valid_focused_candidates']
    assert sorted(full['all_values'], key=int) == [str(n) for n in range(1, 100001) if local3(n, 7)]
    with pytest.raises(ValueError):
        Matcher(runtime, config, variant='float')
    row = legacy.tile(Tile(0, config.B), af, bf)
    assert row['status'] == 'PASS' and row['valid_focused_candidates'] == full['valid_focused_candidates']


def test_endpoint_and_carry_windows(runtime, focuses619, oracle_focuses619) -> None:
    """Narrow x intervals at the interval ends and at 2^64 / 2^90 carry boundaries."""
    rng = random.Random(619005)
    centres = [10**27 + 1, 3 * 10**27, 2**90, 2**90 + 1, 2**90 - 1, (10**27 // 2**64 + 1) * 2**64,
               (3 * 10**27 // 2**64) * 2**64 - 1]
    centres += [rng.randrange(10**27, 3 * 10**27) for _ in range(3)]
    for n in centres:
        config = Config(619, n - 10**9, min(3 * 10**27, n + 10**9), A619, B619)
        v = config.canonical(n)[1]
        first, end = max(0, v - 5000000), min(B619, v + 5000000)
        row = MaskedMatcher(runtime, config).run(prepare(runtime, config, *focuses619, first, end, 1000000))
        expected = O.batch(config, *oracle_focuses619, first, end, 1000000)
        assert oracle_row_fields(row) == oracle_row_fields(expected), n


def test_tile_ranges_match_oracle_geometry(config619) -> None:
    for first, end in [(0, BATCH_WIDTH), (B619 - BATCH_WIDTH, B619), (0, BATCH_WIDTH + 1)]:
        width = 100000001 if end - first > BATCH_WIDTH else 100000000
        assert tile_ranges(config619, first, end, width) == O.tile_ranges(config619, first, end, width)


def test_reproduce_compare_baseline(tmp_path) -> None:
    """tools/reproduce.py --compare-baseline: bitset lease receipt == per-subtile direct engine receipt."""
    from psuedocubes_amd import reproduce
    out = tmp_path / 'record.json'
    summary = reproduce.main(['--first', str(LEASE_FIRST), '--batches', '2', '--compare-baseline',
                              '--output', str(out)])
    assert summary['status'] == 'PASS' and out.exists()
    record = json.loads(out.read_text())
    assert record['baseline_compared'] is True
    assert record['result']['compute_batches'] == 2 and record['result']['subtiles_executed'] == 128
    assert record['result']['valid_focused_candidates'] == sum(b['valid_focused_candidates'] for b in LEASE_BATCHES[:2])
    # the transcript chains the two batches' encoded scope lists: recompute it from the
    # oracle's scopes, each of which must hash to the golden per-batch value first
    config = Config(619, 10**27, 3 * 10**27, A619, B619)
    fa, fb = (O.Focus(f) for f in config.focuses())
    chained = hashlib.sha256()
    values = []
    for golden in LEASE_BATCHES[:2]:
        row = O.batch(config, fa, fb, int(golden['first']), int(golden['end']))
        assert scopes_hash(row) == golden['scopes_sha256']
        chained.update(encoded(row['subtile_scopes']))
        values += row['all_values']
    scope = record['result']['subtile_scopes'][0]
    assert scope['transcript_sha256'] == chained.hexdigest()
    assert scope['output_sha256'] == hashlib.sha256(encoded(values)).hexdigest()
    assert record['result']['all_values'] == values
    with pytest.raises(ValueError):
        reproduce.main(['--first', '0', '--batches', '1', '--output', str(out)])
