"""The native host verifier (pcube_host_verify, used on every value the batched
and served forms emit) against the Python predicate.

The Python side is the restatement of the reference's `pseudocube/oracle.py:19-45`
(local3, icbrt) and `review/verify_cube.py:8-20` (verify), itself pinned by the
witness fixture in test_math.py.  Values: random integers of every size up to
2^128, values built to pass local3 (so the accepting path is exercised, not only
rejections), perfect cubes and their neighbours, and the 2^64, 2^90 and 2^128
boundaries, where a 128-bit cube root or residue would wrap if computed wrongly.
"""
from __future__ import annotations

import json
from math import gcd, prod
from pathlib import Path
import random

import pytest

from psuedocubes_amd.predicate import icbrt, is_perfect_cube, local3, primes_upto, verify
from psuedocubes_amd.verification import native_check, native_witnesses

GOLDEN = Path(__file__).parent / 'golden'


def local3_values(rng: random.Random, p: int, count: int, bits: int) -> list[int]:
    """Values meant to pass local3(n, p), built by CRT from one random admissible
    residue per modulus (1 or 8 mod 9; a nonzero cube mod q = 1 mod 3; any nonzero
    residue mod the other primes) plus a random multiple of the product.  The
    tests decide membership with local3 itself; only p up to 61 keeps the
    product far enough below 2^128."""
    out: list[int] = []
    while len(out) < count:
        n, m = rng.choice([1, 8]), 9
        for q in primes_upto(p):
            if q == 3:
                continue
            a = rng.choice([r for r in range(1, q) if q % 3 != 1 or pow(r, (q - 1) // 3, q) == 1])
            n += m * ((a - n) * pow(m, -1, q) % q)
            m *= q
        n += m * rng.getrandbits(max(1, bits - m.bit_length()))
        if n < 2**128:
            out.append(n)
    return out


def sample(p: int, seed: int) -> list[int]:
    rng = random.Random(seed)
    values = [rng.getrandbits(rng.choice([8, 32, 63, 64, 65, 90, 91, 96, 127, 128])) for _ in range(6000)]
    if p <= 61:
        values += local3_values(rng, p, 1500, rng.choice([64, 92, 127]))
    coprime = prod(primes_upto(p))
    roots = [rng.getrandbits(rng.choice([10, 21, 22, 30, 31, 42])) | 1 for _ in range(1500)]
    roots += [r for r in (rng.getrandbits(30) for _ in range(20000)) if gcd(r, coprime) == 1][:500]
    roots += [6981463658331, 6981463658330, 2642245, 2642246, 1276971475, 1442249570]
    values += [r ** 3 + d for r in roots for d in (-1, 0, 1) if 0 <= r ** 3 + d < 2**128]
    for edge in (2**64, 2**90, 2**127, 2**128):
        values += [edge + d for d in range(-40, 41) if 0 <= edge + d < 2**128]
    return values


@pytest.mark.parametrize('p', [7, 13, 31, 619])
def test_native_predicate_equals_python(p: int) -> None:
    values = sample(p, 619003 + p)
    assert len(values) >= 5000
    ok, cube, root = native_check(values, p)
    ok, cube, root = ok.tolist(), cube.tolist(), root.tolist()
    accepted = 0
    for n, o, c, r in zip(values, ok, cube, root):
        assert bool(o) == local3(n, p), n
        assert bool(c) == is_perfect_cube(n), n
        assert r == icbrt(n), n
        accepted += o
    assert accepted >= (1000 if p <= 61 else 400)   # the accepting path is exercised, not only rejections


def test_native_predicate_covers_twenty_thousand_values() -> None:
    values = [v for seed in (1, 2, 3, 4) for v in sample(13, seed)]
    assert len(values) >= 20000
    ok, cube, root = native_check(values, 13)
    assert ok.tolist() == [int(local3(n, 13)) for n in values]
    assert cube.tolist() == [int(is_perfect_cube(n)) for n in values]
    assert root.tolist() == [icbrt(n) for n in values]


@pytest.mark.parametrize('p', [7, 31, 61, 619])
def test_native_witnesses_equal_verify(p: int) -> None:
    rng = random.Random(p)
    if p <= 61:
        values = local3_values(rng, p, 300, 100)
    else:
        values = [int(json.loads((GOLDEN / 'witness.json').read_text())['n'])]
    assert all(local3(n, p) for n in values)
    # cubes of roots coprime to every prime <= p pass local3 and must be left out as genuine cubes
    coprime = prod(primes_upto(p))
    cubes = [r ** 3 for r in range(1000000001, 1000002000, 2) if gcd(r, coprime) == 1][:10]
    assert all(local3(n, p) for n in cubes)
    mixed = values + cubes
    rng.shuffle(mixed)
    expected = [verify(n, p) for n in mixed if not is_perfect_cube(n)]
    assert native_witnesses(mixed, p) == expected
    assert all(row['valid_pseudocube'] and row['failing_primes'] == [] for row in expected)


def test_native_witness_of_the_reference_hit() -> None:
    fixture = json.loads((GOLDEN / 'witness.json').read_text())
    (row,) = native_witnesses([int(fixture['n'])], 619)
    for key in ('n', 'p', 'cube_root_floor', 'perfect_cube', 'failing_primes', 'valid_pseudocube', 'claim'):
        assert row[key] == fixture[key], key


def test_native_witnesses_fail_closed() -> None:
    rng = random.Random(5)
    good = local3_values(rng, 13, 50, 64)
    bad = next(n for n in (rng.getrandbits(64) for _ in range(1000)) if not local3(n, 13))
    with pytest.raises(ArithmeticError):
        native_witnesses(good + [bad], 13)
    assert native_witnesses([], 13) == []
    with pytest.raises(ValueError):
        native_check([1], 11)   # p must be a prime = 1 mod 3
