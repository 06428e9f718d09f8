"""Exact host-side predicate used to verify every value the GPU emits.

Restates the reference's independent predicate (`pseudocube/oracle.py:7-45`)
and its standalone witness verifier (`review/verify_cube.py:8-20`).  Prime
lists here come from trial division, deliberately independent of the sieve in
:mod:`psuedocubes_amd.tables`, and cubic residuosity is tested by modular
exponentiation rather than by the cube tables the kernels use.
"""
from __future__ import annotations

from functools import lru_cache
from math import isqrt


@lru_cache(maxsize=128)
def primes_upto(p: int) -> tuple[int, ...]:
    """Primes <= p by trial division (`pseudocube/oracle.py:7-11`)."""
    if type(p) is not int or p < 0:
        raise ValueError('prime limit must be a nonnegative integer')
    found = []
    for q in range(2, p + 1):
        if all(q % d for d in range(2, isqrt(q) + 1)):
            found.append(q)
    return tuple(found)


def validate_p(p: int) -> None:
    """p must be a prime with p = 1 mod 3 and p >= 7 (`pseudocube/oracle.py:14-16`)."""
    if type(p) is not int or p < 7 or p % 3 != 1 or p not in primes_upto(p):
        raise ValueError('p must be a prime congruent to 1 modulo 3')


def icbrt(n: int) -> int:
    """floor(n ** (1/3)) exactly, by bisection (`pseudocube/oracle.py:19-29`)."""
    if type(n) is not int or n < 0:
        raise ValueError('cube root requires a nonnegative integer')
    low, high = 0, 1 << ((n.bit_length() + 2) // 3)
    while low < high:
        mid = (low + high + 1) >> 1
        if mid * mid * mid <= n:
            low = mid
        else:
            high = mid - 1
    return low


def is_perfect_cube(n: int) -> bool:
    return n >= 0 and icbrt(n) ** 3 == n


def local3(n: int, p: int) -> bool:
    """All local cube conditions up to p (`pseudocube/oracle.py:36-41`).

    n > 0, n mod 9 in {1, 8}, n coprime to every prime <= p, and
    n^((q-1)/3) = 1 mod q for every prime q = 1 mod 3 up to p.
    """
    validate_p(p)
    if type(n) is not int or n <= 0 or n % 9 not in (1, 8):
        return False
    for q in primes_upto(p):
        if n % q == 0:
            return False
        if q % 3 == 1 and pow(n, (q - 1) // 3, q) != 1:
            return False
    return True


def qualifies3(n: int, p: int) -> bool:
    """A pseudocube: locally a cube up to p but not a cube (`pseudocube/oracle.py:44-45`)."""
    return local3(n, p) and not is_perfect_cube(n)


def verify(n: int, p: int) -> dict[str, object]:
    """Witness record with the reference verifier's keys (`review/verify_cube.py:8-20`).

    The key set and value types match the reference exactly because these
    dicts are embedded in ``cpu_confirmed_hits`` and therefore in receipts.
    """
    primes = [q for q in range(2, p + 1) if all(q % d for d in range(2, isqrt(q) + 1))]
    if p not in primes or p % 3 != 1:
        raise ValueError('invalid prime cutoff')
    if n < 0:
        raise ValueError('negative integer')
    # Largest r with r^3 <= n, by bisection on [0, n + 1).
    low, high = 0, n + 1
    while high - low > 1:
        mid = (low + high) >> 1
        if mid * mid * mid <= n:
            low = mid
        else:
            high = mid
    failures = [q for q in primes if n % q == 0 or (q % 3 == 1 and pow(n, (q - 1) // 3, q) != 1)]
    cube = low ** 3 == n
    valid = n > 0 and n % 9 in (1, 8) and not failures and not cube
    return dict(n=str(n), p=p, cube_root_floor=str(low), perfect_cube=cube,
                failing_primes=failures, valid_pseudocube=valid,
                claim='Candidate validity only; no minimality claim.')
