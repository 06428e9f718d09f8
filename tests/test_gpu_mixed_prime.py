"""Mixed prime bounds (p = 313, 433, 619) on the device against the CPU oracle.

The batched workload draws p from {313, 433, 619}; p = 313 and 433 use their own
filter orders (44 and 63 residuals, 11 and 16 groups), so their device results
are compared with the oracle directly, field for field:

* the plain drop-in (prepare + MaskedMatcher.run) on the seeded full batch, the
  reference lease's batch 34, and the tails ending at B;
* the batched engine on the first 256 windows of the seeded mixed stream
  (4096 windows, v-lengths log-uniform 1e6..1e9, seed 619003);
* the whole 4096-window stream in the batched engine, checked by what one
  window's rows at different p must satisfy (equal focused positions and first
  two groups, survivors at a larger p exactly those at 313 that pass the extra
  residual filters, the last group equal to the survivors) and on a seeded
  sample of 48 windows against the oracle.
"""
from __future__ import annotations

import json
from pathlib import Path
import random

import pytest

from oracle import oracle as O
from psuedocubes_amd.batch_experiment import MaskedMatcher, prepare
from psuedocubes_amd.batched import BatchedEngine, Window, window_stream
from psuedocubes_amd.focus import A619, B619, Config
from psuedocubes_amd.optimized import BATCH_WIDTH
from psuedocubes_amd.predicate import local3
from psuedocubes_amd.tables import allowed

pytestmark = pytest.mark.gpu
PRIMES = (313, 433, 619)
FIELDS = ('focus_u', 'focus_v', 'cartesian_opportunities', 'valid_focused_candidates', 'local_survivors',
          'subtile_scopes', 'group_survivors', 'filter_groups', 'all_values')
GOLDEN = Path(__file__).parent / 'golden'
LEASE_BATCHES = json.loads((GOLDEN / 'lease_575222931200000000_batches.json').read_text())['batches']


def configs() -> dict[int, Config]:
    return {p: Config(p, 10**27, 3 * 10**27, A619, B619) for p in PRIMES}


def fields(row: dict) -> dict:
    return {k: row[k] for k in FIELDS}


def stream() -> list[Window]:
    return window_stream(configs()[619], 4096, 619003, (1e6, 1e9), PRIMES)


@pytest.fixture(scope='module')
def batched(runtime, focuses619):
    return BatchedEngine(runtime, configs().values(), *focuses619)


def plain_scopes() -> list[tuple[int, int]]:
    rng = random.Random(619003)
    k = rng.randrange(B619 // BATCH_WIDTH)
    batch34 = LEASE_BATCHES[34]
    return [(k * BATCH_WIDTH, (k + 1) * BATCH_WIDTH), (int(batch34['first']), int(batch34['end'])),
            (B619 - BATCH_WIDTH, B619), (B619 - 100000001, B619)]


@pytest.mark.parametrize('p', [313, 433])
def test_plain_drop_in_equals_oracle_at_313_and_433(runtime, focuses619, oracle_focuses619, p) -> None:
    config = configs()[p]
    matcher = MaskedMatcher(runtime, config)
    assert len(matcher.filter_groups) == {313: 11, 433: 16}[p]
    survivors = 0
    for first, end in plain_scopes():
        print(f'p={p} batch [{first}, {end})', flush=True)
        row = matcher.run(prepare(runtime, config, *focuses619, first, end))
        expected = O.batch(config, *oracle_focuses619, first, end)
        assert fields(row) == fields(expected), (p, first, end)
        survivors += row['local_survivors']
    assert survivors > (5000 if p == 313 else 2)


def test_batched_mixed_stream_equals_oracle(runtime, batched, oracle_focuses619) -> None:
    windows = stream()[:256]
    assert {w.p for w in windows} == set(PRIMES)
    rows = batched.run(windows)
    cfg = configs()
    for window, row in zip(windows, rows):
        expected = O.batch(cfg[window.p], *oracle_focuses619, window.first, window.end, window.width)
        assert fields(row) == fields(expected), window


def test_full_mixed_stream_launch(runtime, batched, oracle_focuses619) -> None:
    windows = stream()
    cfg = configs()
    print('4096 windows at their own p', flush=True)
    rows = batched.run(windows)
    print('the same windows at p = 313', flush=True)
    base = batched.run([Window(w.first, w.end, 313, w.width) for w in windows])
    extra = {p: [q for q in cfg[p].residuals() if q > 313] for p in PRIMES}
    for window, row, low in zip(windows, rows, base):
        config = cfg[window.p]
        assert row['filter_groups'] == config.filter_groups()
        assert sum(s['valid_focused_candidates'] for s in row['subtile_scopes']) == row['valid_focused_candidates']
        groups = row['group_survivors']
        assert all(a >= b for a, b in zip(groups, groups[1:])) and groups[-1] == row['local_survivors']
        for key in ('focus_u', 'focus_v', 'cartesian_opportunities', 'valid_focused_candidates'):
            assert row[key] == low[key], (window, key)
        assert groups[:2] == low['group_survivors'][:2]
        values = [int(x) for x in row['all_values']]
        passing = [x for x in map(int, low['all_values']) if all(x % q in allowed(q) for q in extra[window.p])]
        assert sorted(values) == sorted(passing), window
        assert all(config.lo < x <= config.hi and local3(x, window.p) for x in values[:50])
    rng = random.Random(619013)
    for index in rng.sample(range(len(windows)), 48):
        window = windows[index]
        expected = O.batch(cfg[window.p], *oracle_focuses619, window.first, window.end, window.width)
        assert fields(rows[index]) == fields(expected), window
