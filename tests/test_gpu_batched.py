"""Batched form: uneven windows, mixed prime bounds, one device pass.

Each window's row must equal the plain drop-in's row for the same window
(prepare + MaskedMatcher.run at its p), whatever else shares its launch, and
every replay of the captured launch sequence must equal the eager launch.
"""
from __future__ import annotations

import random

import pytest

from psuedocubes_amd.batch_experiment import MaskedMatcher, compare, prepare
from psuedocubes_amd.batched import BatchedEngine, Window, window_stream
from psuedocubes_amd.focus import A619, B619, Config

pytestmark = pytest.mark.gpu
PRIMES = (313, 433, 619)


def configs() -> dict[int, Config]:
    return {p: Config(p, 10**27, 3 * 10**27, A619, B619) for p in PRIMES}


def strip(row: dict) -> dict:
    return {k: v for k, v in row.items() if not k.endswith('_seconds') and k not in
            ('matching_layout', 'kernel_registers', 'kernel_local_bytes', 'claim')}


@pytest.fixture(scope='module')
def plain(runtime):
    return {p: MaskedMatcher(runtime, c) for p, c in configs().items()}


def mixed_windows() -> list[Window]:
    rng = random.Random(3)
    out = window_stream(configs()[619], 40, 619007, (1e6, 1e9), PRIMES)
    out += [Window(0, 6_400_000_000, 619), Window(B619 - 250_000_001, B619, 313), Window(5, 6, 433),
            Window(10**12, 10**12 + 2_000_000, 619, 1_000_000)]
    rng.shuffle(out)
    return out


def test_batched_rows_equal_plain_rows(runtime, focuses619, plain) -> None:
    # a threshold that sends the short windows to the direct engine, so both engines run
    engine = BatchedEngine(runtime, configs().values(), *focuses619, tuning=dict(direct_below_rows=50))
    windows = mixed_windows()
    rows = engine.run(windows)
    modes = set()
    for window, row in zip(windows, rows):
        expected = plain[window.p].run(prepare(runtime, configs()[window.p], *focuses619, window.first, window.end,
                                               window.width))
        compare(expected, row)
        assert strip(expected) == strip(row), window
        modes.add(row['matching_layout'])
    assert modes == {'batched-direct', 'batched-residue-bitset'}


def test_batched_memory_groups_do_not_change_results(runtime, focuses619) -> None:
    engine = BatchedEngine(runtime, configs().values(), *focuses619)
    windows = mixed_windows()[:20]
    whole = engine.run(windows)
    pieces = engine.run(windows, budget=64 << 20)
    assert [strip(r) for r in whole] == [strip(r) for r in pieces]


def test_graph_replays_equal_eager(runtime, focuses619, plain) -> None:
    graph = BatchedEngine(runtime, configs().values(), *focuses619, use_graph=True)
    eager = BatchedEngine(runtime, configs().values(), *focuses619, use_graph=False)
    windows = mixed_windows()[:24]
    reference = [strip(r) for r in eager.run(windows)]
    before = graph.graph_launches
    for replay in range(3):
        assert [strip(r) for r in graph.run(windows)] == reference, replay
        # eager work and synchronisation between replays
        plain[619].run(prepare(runtime, configs()[619], *focuses619, 0, 6_400_000_000))
    assert graph.graph_launches >= before + 3
    # a different window set through the same captured sequence
    other = window_stream(configs()[619], 16, 5, (1e6, 2e8), PRIMES)
    assert [strip(r) for r in graph.run(other)] == [strip(r) for r in eager.run(other)]


def test_rejects_windows_beyond_64_subtiles(runtime, focuses619) -> None:
    engine = BatchedEngine(runtime, [configs()[619]], *focuses619)
    with pytest.raises(ValueError):
        engine.run([Window(0, 6_500_000_000, 619)])
