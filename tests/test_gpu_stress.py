"""Endpoint and carry stress at scale (p = 619).

* 1024 seeded windows straddling the interval ends (v near 0 and near B; the
  canonical v of x just above 10^27 and at 3*10^27) and the 2^64 / 2^90
  carry boundaries of x, all in one batched launch, each bit-exact against
  the CPU oracle;
* narrow intervals (n - 1, n] and (n - 2^20, n + 2^20] at those boundaries through
  the plain drop-in, bit-exact against the oracle.
The full 256-batch lease is covered by test_full_reference_lease_pipelined.
"""
from __future__ import annotations

import random

import pytest

from oracle import oracle as O
from psuedocubes_amd.batch_experiment import MaskedMatcher, prepare
from psuedocubes_amd.batched import BatchedEngine, Window
from psuedocubes_amd.focus import A619, B619, Config

pytestmark = pytest.mark.gpu
LO, HI = 10**27, 3 * 10**27
FIELDS = ('focus_u', 'focus_v', 'cartesian_opportunities', 'valid_focused_candidates', 'local_survivors',
          'subtile_scopes', 'group_survivors', 'filter_groups', 'all_values')


def boundary_points(rng: random.Random, count: int) -> list[int]:
    """x values at the interval ends and the 2^64 / 2^90 carry boundaries."""
    k_lo, k_hi = LO // 2**64 + 1, HI // 2**64
    points = []
    for i in range(count):
        kind = i % 5
        delta = rng.randrange(-2**40, 2**40)
        if kind == 0:
            x = LO + 1 + abs(delta)
        elif kind == 1:
            x = HI - abs(delta)
        elif kind == 2:
            x = rng.randrange(k_lo, k_hi + 1) * 2**64 + delta
        elif kind == 3:
            x = 2**90 + delta
        else:
            x = rng.choice([LO + 1, HI, 2**90, 2**90 - 1, k_lo * 2**64, k_hi * 2**64 - 1]) + rng.randrange(-8, 9)
        points.append(min(HI, max(LO + 1, x)))
    return points


def stress_windows(config: Config) -> list[Window]:
    rng = random.Random(619005)
    windows = []
    for x in boundary_points(rng, 1000):
        v = config.canonical(x)[1]
        first = max(0, v - 5_000_000)
        windows.append(Window(first, min(B619, first + 10_000_000), 619))
    for i in range(12):
        windows.append(Window(i * 3_000_001, i * 3_000_001 + 7_000_000, 619))
        windows.append(Window(B619 - 7_000_000 - i * 3_000_001, B619 - i * 3_000_001, 619))
    return windows


def test_stress_windows_batched_vs_oracle(runtime, config619, focuses619, oracle_focuses619) -> None:
    windows = stress_windows(config619)
    assert len(windows) == 1024
    rows = BatchedEngine(runtime, [config619], *focuses619).run(windows)
    fa, fb = oracle_focuses619
    for window, row in zip(windows, rows):
        expected = O.batch(config619, fa, fb, window.first, window.end, window.width)
        assert {k: row[k] for k in FIELDS} == {k: expected[k] for k in FIELDS}, window


def test_narrow_intervals_at_carry_boundaries(runtime, focuses619, oracle_focuses619) -> None:
    rng = random.Random(619006)
    fa, fb = oracle_focuses619
    for x in boundary_points(rng, 64):
        half = 1 if x % 2 else 2**20
        config = Config(619, x - half, min(HI, x + half - 1), A619, B619)
        v = config.canonical(x)[1]
        first, end = max(0, v - 2_000_000), min(B619, v + 2_000_000)
        row = MaskedMatcher(runtime, config).run(prepare(runtime, config, *focuses619, first, end, 1_000_000))
        expected = O.batch(config, fa, fb, first, end, 1_000_000)
        assert {k: row[k] for k in FIELDS} == {k: expected[k] for k in FIELDS}, x
