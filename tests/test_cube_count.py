"""Full-domain control (cube counts): the oracle pinned on the CPU.

The oracle's count of roots coprime to every prime <= p is checked against a
gcd brute force on seeded short ranges, against the reference's own two
programs (review/cube_count.c, review/cube_count_independent.c, compiled
unmodified under oracle/_ref when the reference is present) on counts and
root lists, and against the reference's known answer for the whole domain.
"""
from __future__ import annotations

import json
from math import gcd, prod
from pathlib import Path
import random
import subprocess

import pytest

from oracle import oracle as O
from psuedocubes_amd.cube_count import cube_count

REF = Path(__file__).resolve().parents[1] / 'oracle' / '_ref'
FULL = (1000000001, 1442249570, 619)
FULL_COUNT = 38322775   # results/current.json:25 of the reference


def primorial(p: int) -> int:
    return prod(q for q in range(2, p + 1) if all(q % d for d in range(2, int(q ** 0.5) + 1)))


def test_oracle_cube_count_equals_gcd_brute_force() -> None:
    rng = random.Random(38322775)
    cases = [(1, 1, 2), (1, 2000, 2), (1, 5000, 2000), (2**32 - 3000, 2**32 - 1, 619), (6, 6, 3), (7, 7, 5)]
    cases += [(a, a + rng.randrange(0, 3000), rng.choice([2, 3, 5, 7, 619, 1999, 2000]))
              for a in (rng.randrange(1, 2**32 - 4000) for _ in range(40))]
    for first, last, p in cases:
        m = primorial(p)
        expected = [r for r in range(first, last + 1) if gcd(r, m) == 1]
        count, roots = O.cube_count(first, last, p, list_roots=True)
        assert count == len(expected) and roots == expected, (first, last, p)


def test_oracle_full_domain_known_answer() -> None:
    assert O.cube_count(*FULL)[0] == FULL_COUNT


@pytest.mark.skipif(not (REF / 'cube_count').exists(), reason='reference programs not built (no reference checkout)')
def test_oracle_cube_count_equals_reference_programs() -> None:
    rng = random.Random(619)
    ranges = [(1, 100000, 619), (1000000001, 1000100001, 619), (2**32 - 100001, 2**32 - 1, 2000)]
    ranges += [(a, a + 100000, rng.choice([3, 433, 619])) for a in (rng.randrange(1, 2**32 - 200000) for _ in range(3))]
    for first, last, p in ranges:
        count, roots = O.cube_count(first, last, p, list_roots=True)
        for program in ('cube_count', 'cube_count_independent'):
            row = json.loads(subprocess.check_output([str(REF / program), str(first), str(last), str(p), '--list'],
                                                     text=True))
            assert int(row['count']) == count and [int(r) for r in row['roots']] == roots, (program, first, p)
    # a long range crossing many segments, counts only
    first, last = 1000000001, 1010000001
    count, _ = O.cube_count(first, last, 619)
    for program in ('cube_count', 'cube_count_independent'):
        row = json.loads(subprocess.check_output([str(REF / program), str(first), str(last), '619'], text=True))
        assert int(row['count']) == count


def test_cube_count_rejects_the_reference_invalid_arguments() -> None:
    for first, last, p, listed in ((0, 5, 619, False), (6, 5, 619, False), (1, 2**32, 619, False),
                                   (1, 5, 1, False), (1, 5, 2001, False), (1, 100002, 619, True)):
        with pytest.raises(ValueError):
            cube_count(first, last, p, list_roots=listed)
