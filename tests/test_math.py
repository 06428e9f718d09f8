"""Host mathematics: tables, predicate, focus wheels, canonical coordinates.

Known answers follow the reference's tests/test_pseudocube_math.py
(table counts 24-54, coefficient adjustment 56-72, canonical endpoints 90-101,
targeted values 103-117, wide limbs 119-131); every expectation below is a
closed form or a brute-force recomputation, never the formula under test.
"""
from __future__ import annotations

import ctypes
from itertools import product
from math import gcd, prod
import random

import pytest

from oracle import oracle as O
from psuedocubes_amd.focus import A619, B619, Config, Tile, targeted_coordinates
from psuedocubes_amd.predicate import icbrt, is_perfect_cube, local3, primes_upto, qualifies3, verify
from psuedocubes_amd.tables import allowed, byte_rows, definition, packed, sieve_primes

N613 = 674441580981249129037406633


def small_config(p: int, lo: int = 0, hi: int = 100000) -> Config:
    return Config(p, lo, hi, 14 * (19 if p >= 19 else 1), 9 * (13 if p >= 13 else 1) * (31 if p >= 31 else 1))


def test_roots_and_invalid_inputs() -> None:
    assert icbrt(3 * 10**27) == 1442249570
    for root in [0, 1, 2, 3, 2**21, 2**32 - 1, 2**42, 10**27]:
        n = root**3
        assert icbrt(n) == root and is_perfect_cube(n)
        if root:
            assert icbrt(n - 1) == root - 1
        assert icbrt(n + 1) == (root if root else 1)
    for bad in (-1, 1.5, True):
        with pytest.raises(ValueError):
            icbrt(bad)  # type: ignore[arg-type]
    for p in (2, 3, 5, 9, 15, 25, -7):
        with pytest.raises(ValueError):
            local3(1, p)
    assert local3(1, 619) and not qualifies3(1, 619)
    assert not local3(0, 619) and not local3(-1, 619)


def test_tables_exhaustively() -> None:
    primes = sieve_primes(619)
    assert primes == list(primes_upto(619))
    cubic = [q for q in primes if q % 3 == 1]
    assert (len(primes), len(cubic), sum(q % 3 == 2 for q in primes)) == (114, 54, 59)
    assert sum(cubic) == 15744
    assert sum(4 * len(packed(q)) for q in cubic) == 2076
    for q in cubic:
        residues = allowed(q)
        assert len(residues) == (q - 1) // 3 and 0 not in residues
        for a in range(q):
            assert (a in residues) == (pow(a, (q - 1) // 3, q) == 1)
            assert bool(packed(q)[a // 32] >> (a % 32) & 1) == (a in residues)
    table, offsets = byte_rows([9, 5, 7])
    assert offsets == [0, 9, 14] and list(table[:9]) == [0, 1, 0, 0, 0, 0, 0, 0, 1]
    assert len(definition(619)['sha256']) == 64


def test_adjustment_and_nondivisibility(config619: Config) -> None:
    assert (A619 % 31, B619 % 31, pow(21, 10, 31)) == (0, 21, 5)
    for focus, coefficient in zip(config619.focuses(), (B619, -A619)):
        for q, residues in focus.rings:
            for a in range(q):
                assert (a in residues) == ((coefficient * a) % q in allowed(q))
    assert set(dict(config619.focuses()[0].rings)[31]) != set(allowed(31))
    for q in (5, 11, 617):
        n = q**3
        assert n % 9 in (1, 8)
        assert all(pow(n, (r - 1) // 3, r) == 1 for r in primes_upto(619) if r % 3 == 1)
        assert not local3(n, 619)
        assert config619.coverage()[str(q)] == 'residual'
    assert config619.coverage()['3'] == 'mod9'
    with pytest.raises(ValueError):
        Config(7, 0, 100000, 14, 9 * 13)
    with pytest.raises(ValueError):
        Config(7, 0, 100000, 3 * 7, 3)


def test_filter_order_and_groups(config619: Config) -> None:
    order = config619.filter_order()
    assert len(order) == 93 and order[:4] == [193, 199, 211, 223]
    cubic = [q for q in order if q % 3 == 1]
    assert order[:35] == cubic and len(cubic) == 35 and order[35:] == sorted(order[35:])
    groups = config619.filter_groups()
    assert len(groups) == 24 and groups[-1] == [617]
    for p, n, ncubic in ((313, 44, 11), (433, 63, 21)):
        c = Config(p, 10**27, 3 * 10**27, A619, B619)
        assert len(c.filter_order()) == n and sum(q % 3 == 1 for q in c.filter_order()) == ncubic
        assert c.filter_order()[:4] == [193, 199, 211, 223]


def test_focus_counts_and_halves(config619: Config) -> None:
    af, bf = config619.focuses()
    assert [len(r) for _, r in af.rings] == [1, 2, 4, 10, 14, 24, 26, 42, 46, 52, 60]
    assert [len(r) for _, r in bf.rings] == [2, 6, 12, 20, 22, 32, 34, 36, 50, 54]
    assert (af.count, bf.count) == (4212736819200, 6700548096000)
    sizes = [[prod(map(len, half)) for half in f.halves()] for f in (af, bf)]
    assert sizes == [[2096640, 2009280], [2592000, 2585088]]
    rng = random.Random(619003)
    for f in (af, bf):
        for index in [0, f.count - 1] + [rng.randrange(f.count) for _ in range(50)]:
            y = f.at(index)
            assert f.rank(y) == (0, index)
            assert all(y % q in residues for q, residues in f.rings)


def test_canonical_boundaries_and_normalized_residues(config619: Config) -> None:
    rng = random.Random(619003)
    for config in (small_config(31), config619):
        for n in [config.lo, config.lo + 1, config.hi, config.hi + 1,
                  *[rng.randrange(config.lo, config.hi + 1) for _ in range(300)]]:
            u, v = config.canonical(n)
            assert u * config.B - v * config.A == n and 0 <= v < config.B
            low, high = config.u_interval(v)
            assert (low <= u < high) == (config.lo < n <= config.hi)
            for q in (5, 7, 9, 31, 617, 619):
                assert (u * (config.B % q) - v * (config.A % q)) % q == n % q
            envelope = Tile(v, min(v + 2, config.B)).envelope(config)
            if config.lo < n <= config.hi:
                assert int(envelope['x_min']) <= n <= int(envelope['x_max'])


def test_targeted_known_value_and_cube_emission() -> None:
    assert qualifies3(N613, 613) and not local3(N613, 619) and pow(N613, 206, 619) == 366
    for p, n in [(613, N613), (619, 1000000007**3), (619, 1000000009**3)]:
        assert local3(n, p)
        config = Config(p, max(0, n - 100), min(3 * 10**27, n + 100), A619, B619)
        us, vs = targeted_coordinates(config, n)
        emitted = {u * config.B - v * config.A for u, v in product(us, vs)
                   if config.lo < u * config.B - v * config.A <= config.hi and local3(u * config.B - v * config.A, p)}
        assert n in emitted
    assert 1000000007**3 % 9 == 8 and 1000000009**3 % 9 == 1


def test_witness_record_matches_fixture() -> None:
    import json
    from pathlib import Path
    fixture = json.loads((Path(__file__).parent / 'golden' / 'witness.json').read_text())
    row = verify(int(fixture['n']), 619)
    for key in ('n', 'p', 'cube_root_floor', 'perfect_cube', 'failing_primes', 'valid_pseudocube', 'claim'):
        assert row[key] == fixture[key]
    for check in fixture['residue_checks']:
        q, residue = check['q'], check['residue']
        assert int(fixture['n']) % q == residue
        if q > 3 and q % 3 == 1:
            assert check['cubic_residue_by_enumeration'] == (residue in allowed(q))


def test_complete_small_sets_by_brute_force() -> None:
    for p in (7, 13, 19, 31):
        config = small_config(p)
        expected = [n for n in range(1, 100001) if local3(n, p)]
        assert O.enumerate_small(config, 0, config.B) == expected
        cubes = {n for n in expected if is_perfect_cube(n)}
        assert cubes == {r**3 for r in range(1, icbrt(100000) + 1) if gcd(r, prod(primes_upto(p))) == 1}


def wide_cases() -> list[tuple[int, int, int, int]]:
    rng = random.Random(619003)
    cases = [(0, 0, 0, 0), (2**64 - 1, 2**64 - 1, 0, 0), (0, 0, 2**64 - 1, 2**64 - 1),
             (2**32, 2**32, 1, 1), (2**64 - 1, 1, 0, 0)]
    return cases + [tuple(rng.getrandbits(64) for _ in range(4)) for _ in range(2000)]  # type: ignore[misc]


def test_oracle_wide_limbs() -> None:
    fn = O.lib().oracle_products
    ref = None
    path = O.HERE / '_ref' / 'wide.so'
    if path.exists():
        ref = ctypes.CDLL(str(path)).products
        ref.argtypes = [ctypes.c_uint64] * 4 + [ctypes.POINTER(ctypes.c_uint64)] * 2
        ref.restype = ctypes.c_int
    for a, b, c, d in wide_cases():
        low, high = ctypes.c_uint64(), ctypes.c_uint64()
        borrow = fn(a, b, c, d, ctypes.byref(low), ctypes.byref(high))
        assert borrow == int(a * b < c * d)
        assert low.value + (high.value << 64) == (a * b - c * d) % (1 << 128)
        if ref is not None:
            l2, h2 = ctypes.c_uint64(), ctypes.c_uint64()
            assert ref(a, b, c, d, ctypes.byref(l2), ctypes.byref(h2)) == borrow
            assert (l2.value, h2.value) == (low.value, high.value)


def test_tile_ranges_against_one_division_per_bound() -> None:
    """Subtile geometry stepped by quotient/remainder updates equals the direct
    u_interval of every subtile's first and last v (the reference's tile_ranges,
    pseudocube/batch_experiment.py:27-32), at interior, ragged and final tiles."""
    from psuedocubes_amd.batch_experiment import tile_ranges
    rng = random.Random(27)
    for p in (313, 433, 619):
        config = Config(p, 10**27, 3 * 10**27, A619, B619)
        for _ in range(400):
            width = rng.choice([1, 7, 10**8, rng.randrange(1, 10**10)])
            first = rng.randrange(0, config.B - 1)
            if rng.random() < 0.15:
                first = max(0, config.B - rng.randrange(1, 64 * width + 1))
            end = min(config.B, first + rng.randrange(1, 64 * width + 1))
            expect = [(v, min(v + width, end), config.u_interval(v)[0], config.u_interval(min(v + width, end) - 1)[1])
                      for v in range(first, end, width)]
            assert tile_ranges(config, first, end, width) == expect


def test_library_tile_ranges_equal_python() -> None:
    """The library's 128-bit subtile geometry (used by the one-call prepare)
    equals tile_ranges on seeded ranges, ragged last subtiles and the range's
    end at B; invalid geometry is refused."""
    import numpy as np
    from psuedocubes_amd import native
    from psuedocubes_amd.batch_experiment import tile_ranges
    rng = random.Random(3301)
    mask = (1 << 64) - 1
    for p in (313, 619):
        config = Config(p, 10**27, 3 * 10**27, A619, B619)
        for _ in range(300):
            width = rng.choice([1, 3, 10**8, rng.randrange(1, 10**9)])
            first = rng.randrange(0, config.B - 1)
            if rng.random() < 0.2:
                first = max(0, config.B - rng.randrange(1, 64 * width + 1))
            end = min(config.B, first + rng.randrange(1, 64 * width + 1))
            tiles = np.zeros(256, dtype=np.uint64)
            n = ctypes.c_uint32()
            native.call('pcube_tile_ranges', config.A, config.B, config.lo & mask, config.lo >> 64, config.hi & mask,
                        config.hi >> 64, first, end, width, native.ptr64(tiles), ctypes.byref(n))
            flat = tiles[:4 * n.value].tolist()
            assert [tuple(flat[i:i + 4]) for i in range(0, len(flat), 4)] == tile_ranges(config, first, end, width)
    tiles = np.zeros(256, dtype=np.uint64)
    n = ctypes.c_uint32()
    for first, end, width in ((5, 5, 1), (0, B619 + 1, 10**8), (0, 65, 1)):
        assert native.lib().pcube_tile_ranges(A619, B619, 0, 0, 1, 0, first, end, width, native.ptr64(tiles),
                                              ctypes.byref(n)) == native.INVALID


def test_focus_window_equals_brute_force() -> None:
    """Focus.window (`pseudocube/focus.py:72-79`) against the brute-force set of
    coordinates: windows inside one period, across periods, empty and at 0."""
    from psuedocubes_amd.focus import Focus  # noqa: F401  (the class the drop-in exports)
    for p in (7, 13, 31):
        config = small_config(p)
        for focus in config.focuses():
            members = {focus.at(i) for i in range(focus.count)}
            M = focus.modulus
            for low, high in [(0, M), (0, 3 * M + 5), (M - 3, 2 * M + 3), (5, 5), (7, 8), (2 * M, 2 * M + M // 2)]:
                got = list(focus.window(low, high))
                expected = [y for y in range(low, high) if y % M in members]
                assert got == expected, (p, M, low, high)
            with pytest.raises(ValueError):
                list(focus.window(5, 4))


def test_enumerate_cpu_complete_small_sets_and_partitions() -> None:
    """enumerate_cpu (`pseudocube/focus.py:170-179`) as the reference's own test uses it
    (tests/test_pseudocube_math.py:74-88): the complete set of x <= 1e5 passing
    local3, no duplicates, the same set over a 7-way partition and over the two
    halves of Tile.split, and the cubes among them the cubes of roots coprime to p#."""
    from psuedocubes_amd.focus import enumerate_cpu
    for p in (7, 13, 19, 31):
        config = small_config(p)
        expected = {n for n in range(1, 100001) if local3(n, p)}
        actual = list(enumerate_cpu(config, Tile(0, config.B)))
        assert len(actual) == len(set(actual)) and set(actual) == expected
        pieces: list[int] = []
        step = max(1, config.B // 7)
        for first in range(0, config.B, step):
            pieces.extend(enumerate_cpu(config, Tile(first, min(config.B, first + step))))
        assert len(pieces) == len(set(pieces)) and set(pieces) == expected
        a, b = Tile(0, config.B).split()
        assert set(enumerate_cpu(config, a)) | set(enumerate_cpu(config, b)) == expected
        assert {n for n in actual if is_perfect_cube(n)} == {
            r**3 for r in range(1, icbrt(100000) + 1) if gcd(r, prod(primes_upto(p))) == 1}
    # endpoints: lo excluded, hi included (tests/test_pseudocube_math.py:116-117)
    for n in (1, 125, 343, 1001):
        if local3(n, 7):
            assert n in list(enumerate_cpu(small_config(7, 0, n), Tile(0, 9)))
            assert n not in list(enumerate_cpu(small_config(7, n, n + 100), Tile(0, 9)))


def test_library_tile_ranges_many_equal_python(config619: Config) -> None:
    """The lease's one-call geometry (pcube_tile_ranges_many) equals tile_ranges batch by batch:
    full batches, a ragged last batch, the tail ending at B, single-v batches."""
    from psuedocubes_amd.batch_experiment import tile_ranges, tile_ranges_many
    first = 575222931200000000
    spans = [(first + k * 6_400_000_000, first + (k + 1) * 6_400_000_000) for k in range(40)]
    spans += [(first, first + 250_000_001), (B619 - 6_400_000_000, B619), (B619 - 1, B619), (0, 1), (5, 7)]
    got = tile_ranges_many(config619, spans, 100_000_000)
    assert got == [tile_ranges(config619, s, e, 100_000_000) for s, e in spans]
    with pytest.raises(ValueError):
        tile_ranges_many(config619, [(0, 6_500_000_000)], 100_000_000)
