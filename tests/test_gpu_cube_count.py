"""Full-domain control on the device: the GPU cube count against the oracle."""
from __future__ import annotations

import random

import pytest

from oracle import oracle as O
from psuedocubes_amd.cube_count import cube_count

pytestmark = pytest.mark.gpu


def test_full_domain_known_answer(runtime) -> None:
    row = cube_count(1000000001, 1442249570, 619)
    assert row['count'] == '38322775'
    assert list(row)[:4] == ['algorithm', 'root_first', 'root_last', 'p']


def test_listed_ranges_equal_oracle(runtime) -> None:
    rng = random.Random(1442249570)
    cases = [(1, 100000, 619), (1, 1, 2), (2, 2, 2), (3, 3, 2), (3, 3, 3), (2**32 - 100001, 2**32 - 1, 2000)]
    cases += [(a, a + rng.randrange(0, 100000), rng.choice([2, 3, 5, 313, 433, 619, 2000]))
              for a in (rng.randrange(1, 2**32 - 200000) for _ in range(24))]
    for first, last, p in cases:
        row = cube_count(first, last, p, list_roots=True)
        count, roots = O.cube_count(first, last, p, list_roots=True)
        assert int(row['count']) == count and [int(r) for r in row['roots']] == roots, (first, last, p)


def test_long_ranges_equal_oracle(runtime) -> None:
    # ranges of many 2^18-odd segments, with ragged ends on both sides
    rng = random.Random(619)
    for first, last, p in [(1000000001, 1100000000, 619), (1, 20000000, 2), (4294000000, 2**32 - 1, 1999)] + \
            [(a, a + rng.randrange(1, 30000000), rng.choice([433, 619])) for a in (rng.randrange(1, 2**32 - 40000000)
                                                                                    for _ in range(4))]:
        assert int(cube_count(first, last, p)['count']) == O.cube_count(first, last, p)[0], (first, last, p)
