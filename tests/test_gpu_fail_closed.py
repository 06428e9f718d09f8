"""A corrupted device filter table cannot reach a receipt.

The reference fills its matcher's device byte table with ones and requires the
host verification to raise ArithmeticError (`REL!tools/gpu_worker.py:96-103`:
Config(13, 0, 100000, 14, 117), every u and v of the focus, `bad.table.fill(1)`,
`bad.coordinates(u, v)`).  The same check runs here through every entry point
that emits values: the per-subtile Matcher (coordinates, tile, compute_batch),
the batch drop-in (MaskedMatcher.run), the pipelined lease (compute_lease) and
the batched engine.  With every residue allowed the kernels still run normally
and emit values that fail the exact local predicate; each path must refuse them.
The p = 619 cases use a narrow x interval, so a full batch emits a few thousand
values, well under the output cap, and the failure is the predicate's.
"""
from __future__ import annotations

import pytest

from psuedocubes_amd.batch_experiment import MaskedMatcher, prepare
from psuedocubes_amd.batched import BatchedEngine, Window
from psuedocubes_amd.focus import A619, B619, Config, Tile
from psuedocubes_amd.gpu import GPUFocus, Matcher
from psuedocubes_amd.optimized import BATCH_WIDTH, compute_lease
from psuedocubes_amd.predicate import local3
from psuedocubes_amd.run_production import compute_batch

pytestmark = pytest.mark.gpu
# x in (2e27, 2e27 + 1.2e22]: about 0.1 focused u per v row, some 6,000 positions per batch
NARROW = Config(619, 2 * 10**27, 2 * 10**27 + 12 * 10**21, A619, B619)


def narrow_batch() -> tuple[int, int]:
    v = NARROW.canonical(2 * 10**27 + 6 * 10**21)[1]
    first = min(v // BATCH_WIDTH * BATCH_WIDTH, B619 - BATCH_WIDTH)
    return first, first + BATCH_WIDTH


def test_reference_table_mutation_case(runtime) -> None:
    config = Config(13, 0, 100000, 14, 117)
    fa, fb = config.focuses()
    u = list(fa.window(config.u_interval(0)[0], config.u_interval(config.B - 1)[1]))
    v = list(fb.window(0, fb.modulus))
    good = Matcher(runtime, config, 'byte').coordinates(u, v)
    assert good['all_values'] == [str(n) for n in range(1, 100001) if local3(n, 13)]
    bad = Matcher(runtime, config, 'byte')
    bad.table.fill(1)
    with pytest.raises(ArithmeticError):
        bad.coordinates(u, v)


def test_mutated_tables_fail_every_small_path(runtime) -> None:
    config = Config(13, 1, 100000, 14, 117)   # a key of its own: prepare() must not start batches on it later
    af, bf = (GPUFocus(runtime, f) for f in config.focuses())
    legacy = Matcher(runtime, config)
    legacy.table.fill(1)
    with pytest.raises(ArithmeticError):
        legacy.tile(Tile(0, config.B), af, bf)
    with pytest.raises(ArithmeticError):
        compute_batch(legacy, af, bf, 0, config.B, config.B)
    masked = MaskedMatcher(runtime, config)
    batch = prepare(runtime, config, af, bf, 0, config.B, max(1, (config.B + 6) // 7))
    assert masked.run(batch)['status'] == 'PASS'
    masked.table.fill(1)
    with pytest.raises(ArithmeticError):
        masked.run(prepare(runtime, config, af, bf, 0, config.B, max(1, (config.B + 6) // 7)))
    # with no residue allowed nothing is emitted, and the counts show the empty chain
    masked.table.fill(0)
    row = masked.run(prepare(runtime, config, af, bf, 0, config.B, max(1, (config.B + 6) // 7)))
    assert row['local_survivors'] == 0 and row['group_survivors'][0] == 0


def test_mutated_tables_fail_the_p619_paths(runtime, focuses619) -> None:
    first, end = narrow_batch()
    masked = MaskedMatcher(runtime, NARROW)
    clean = masked.run(prepare(runtime, NARROW, *focuses619, first, end))
    assert clean['status'] == 'PASS' and clean['valid_focused_candidates'] > 1000
    masked.table.fill(1)
    with pytest.raises(ArithmeticError):
        masked.run(prepare(runtime, NARROW, *focuses619, first, end))
    with pytest.raises(ArithmeticError):
        compute_lease(masked, *focuses619, first, end)
    engine = BatchedEngine(runtime, [NARROW], *focuses619)
    assert engine.run([Window(first, end, 619)])[0]['status'] == 'PASS'
    engine.engines[619].table.fill(1)
    with pytest.raises(ArithmeticError):
        engine.run([Window(first, end, 619)])
