"""Coefficient-adjusted CRT focus wheels, canonical coordinates and v tiles.

Host-side restatement of `pseudocube/focus.py:16-167` (the mathematics is in
`docs/mathematics.md:9-34` of the reference).  Every candidate is written
x = u*B - v*A with 0 <= v < B.  The primes of A constrain u, the primes of B
constrain v, and every other prime up to p is a residual filter applied to x.
"""
from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache
from math import gcd, prod
from typing import Any, Iterator

from .predicate import validate_p
from .tables import allowed, definition, sieve_primes

# The split printed in Sorenson arXiv:1001.3316v2 section 4.1
# (`pseudocube/focus.py:11-12`).
A619 = 701856356111039402
B619 = 693110504329192503
MAX = (1 << 64) - 1


def factors(n: int, p: int) -> list[int]:
    """Prime-power rings of a focus modulus (`pseudocube/focus.py:16-32`).

    Only squarefree primes <= p are accepted, except that 3 must appear as
    the complete factor 9.
    """
    if not 1 <= n < 1 << 63:
        raise ValueError('focus modulus must be positive and below 2^63')
    rings: list[int] = []
    rest = n
    for q in sieve_primes(p):
        if rest % q:
            continue
        power = 1
        while rest % q == 0:
            rest //= q
            power *= q
        expected = 9 if q == 3 else q
        if power != expected:
            raise ValueError('only squarefree primes and one complete factor 9 are supported')
        rings.append(power)
    if rest != 1:
        raise ValueError('focus contains a prime above the active cutoff')
    return rings


@dataclass(frozen=True)
class Focus:
    """A product of rings; coordinates are CRT sums of one residue per ring."""
    modulus: int
    rings: tuple[tuple[int, tuple[int, ...]], ...]

    @property
    def count(self) -> int:
        return prod(len(residues) for _, residues in self.rings)

    @property
    def terms(self) -> tuple[tuple[int, ...], ...]:
        """CRT terms a*(M/q)*((M/q)^-1 mod q) mod M per ring (`pseudocube/focus.py:44-47`)."""
        out = []
        for q, residues in self.rings:
            cofactor = self.modulus // q
            lift = cofactor * pow(cofactor, -1, q)
            out.append(tuple((a * lift) % self.modulus for a in residues))
        return tuple(out)

    def at(self, index: int) -> int:
        """Coordinate of a mixed-radix index, first ring least significant (`pseudocube/focus.py:49-56`)."""
        if not 0 <= index < self.count:
            raise ValueError('focus index out of bounds')
        value = 0
        for ring in self.terms:
            index, digit = divmod(index, len(ring))
            value = (value + ring[digit]) % self.modulus
        return value

    def rank(self, coordinate: int) -> tuple[int, int]:
        """(period, index) of a coordinate; inverse of :meth:`at` (`pseudocube/focus.py:58-64`)."""
        period, value = divmod(coordinate, self.modulus)
        index, radix = 0, 1
        for q, residues in self.rings:
            index += residues.index(value % q) * radix
            radix *= len(residues)
        return period, index

    def indices(self, start: int, end: int, period: int = 0) -> Iterator[int]:
        if not 0 <= start <= end <= self.count or period < 0:
            raise ValueError('invalid focus index interval')
        for index in range(start, end):
            yield period * self.modulus + self.at(index)

    def window(self, low: int, high: int) -> Iterator[int]:
        """Focus coordinates in [low, high), a CPU stream for small focuses (`pseudocube/focus.py:72-79`).

        Same values as the reference's stream, in ascending order: the period's
        coordinates are formed once by CRT composition, ring by ring, and sorted;
        each period of the window then yields its in-range shifts."""
        if low < 0 or high < low:
            raise ValueError('invalid half-open window')
        if high == low:
            return
        period_values = _period_coordinates(self)
        for period in range(low // self.modulus, (high - 1) // self.modulus + 1):
            base = period * self.modulus
            for value in period_values:
                if value + base >= high:
                    break
                if value + base >= low:
                    yield value + base

    def halves(self) -> tuple[tuple[tuple[int, ...], ...], tuple[tuple[int, ...], ...]]:
        """Split the rings into two groups with the smallest larger product.

        Same rule and tie-break as `pseudocube/focus.py:81-87`: the first
        subset mask (in increasing order) that minimises the larger product.
        The first group becomes the sorted array C, the second D.
        """
        terms = self.terms
        n = len(terms)
        best_mask, best_cost = 0, None
        for mask in range(1 << n):
            inside = prod(len(t) for i, t in enumerate(terms) if mask >> i & 1)
            outside = prod(len(t) for i, t in enumerate(terms) if not mask >> i & 1)
            cost = max(inside, outside)
            if best_cost is None or cost < best_cost:
                best_mask, best_cost = mask, cost
        first = tuple(t for i, t in enumerate(terms) if best_mask >> i & 1)
        second = tuple(t for i, t in enumerate(terms) if not best_mask >> i & 1)
        return first, second


@lru_cache(maxsize=16)
def _period_coordinates(focus: Focus) -> tuple[int, ...]:
    """Every coordinate of one period of a focus, ascending: the CRT sums of one
    residue per ring, built ring by ring as {s + t mod M : s so far, t in the ring's terms}."""
    sums = [0]
    for ring in focus.terms:
        sums = [(s + term) % focus.modulus for s in sums for term in ring]
    return tuple(sorted(sums))


@lru_cache(maxsize=64)
def _inverse(a: int, m: int) -> int:
    """a^-1 mod m, once per (a, m): host verification asks for it per emitted value."""
    return pow(a, -1, m)


@dataclass(frozen=True)
class Config:
    """One search definition: prime bound p, interval (lo, hi], and the A/B split."""
    p: int
    lo: int
    hi: int
    A: int
    B: int

    def __post_init__(self) -> None:
        validate_p(self.p)
        if not 0 <= self.lo < self.hi <= 3 * 10**27 or gcd(self.A, self.B) != 1:
            raise ValueError('invalid interval, arithmetic bound, or coprimality')
        factors(self.A, self.p)
        factors(self.B, self.p)
        maximum_u = (self.hi + (self.B - 1) * self.A) // self.B
        if maximum_u > MAX or maximum_u * self.B >= 1 << 128:
            raise ValueError('intermediates exceed implementation bounds')

    def focuses(self) -> tuple[Focus, Focus]:
        """u wheel scaled by B^-1, v wheel scaled by (-A)^-1 (`pseudocube/focus.py:108-114`)."""
        made = []
        for modulus, coefficient in ((self.A, self.B), (self.B, -self.A)):
            rings = []
            for q in factors(modulus, self.p):
                inverse = pow(coefficient, -1, q)
                rings.append((q, tuple(sorted({(a * inverse) % q for a in allowed(q)}))))
            made.append(Focus(modulus, tuple(rings)))
        return made[0], made[1]

    def canonical(self, n: int) -> tuple[int, int]:
        """(u, v) with n = u*B - v*A and 0 <= v < B (`pseudocube/focus.py:116-118`)."""
        v = (-n * _inverse(self.A, self.B)) % self.B if self.B > 1 else 0
        return (n + v * self.A) // self.B, v

    def u_interval(self, v: int) -> tuple[int, int]:
        """Half-open u range with lo < u*B - v*A <= hi (`pseudocube/focus.py:120-121`)."""
        return (self.lo + v * self.A) // self.B + 1, (self.hi + v * self.A) // self.B + 1

    def coverage(self) -> dict[str, str]:
        af, bf = factors(self.A, self.p), factors(self.B, self.p)
        where = {}
        for q in [9] + sieve_primes(self.p):
            if q == 3:
                continue
            where[str(q)] = 'A_focus' if q in af else 'B_focus' if q in bf else 'residual'
        where['3'] = 'mod9'
        return where

    def residuals(self) -> list[int]:
        return [int(q) for q, place in self.coverage().items() if place == 'residual']

    def filter_order(self) -> list[int]:
        """Residual moduli in matching order: cubic ones ascending, then the rest.

        The key (q != 9 and q % 3 != 1, q) is the reference's
        (`pseudocube/gpu.py:134`, `pseudocube/batch_experiment.py:84-85`).
        """
        return sorted(self.residuals(), key=lambda q: (q != 9 and q % 3 != 1, q))

    def filter_groups(self) -> list[list[int]]:
        order = self.filter_order()
        return [order[i:i + 4] for i in range(0, len(order), 4)]

    def manifest(self) -> dict[str, Any]:
        af, bf = self.focuses()
        maximum_u = (self.hi + (self.B - 1) * self.A) // self.B
        return dict(schema_version=1, search_type='pseudocube', p=self.p, lo_exclusive=str(self.lo),
                    hi_inclusive=str(self.hi), A=str(self.A), B=str(self.B), canonical='v-in-[0,B)',
                    coverage=self.coverage(), table_sha256=definition(self.p)['sha256'],
                    focus_counts=[str(af.count), str(bf.count)],
                    arithmetic_bounds=dict(max_u=str(maximum_u), max_v=str(self.B - 1),
                                           max_uB=str(maximum_u * self.B), max_vA=str((self.B - 1) * self.A),
                                           x_bits=self.hi.bit_length(),
                                           product_bits=(maximum_u * self.B).bit_length()))


@dataclass(frozen=True)
class Tile:
    """Half-open canonical v interval [first, end) (`pseudocube/focus.py:143-167`)."""
    first: int
    end: int

    def validate(self, config: Config) -> None:
        if not 0 <= self.first < self.end <= config.B:
            raise ValueError('tile must be a nonempty canonical v interval')

    def envelope(self, config: Config) -> dict[str, str]:
        self.validate(config)
        u_min = config.u_interval(self.first)[0]
        u_max = config.u_interval(self.end - 1)[1] - 1
        return dict(u_min=str(u_min), u_max=str(u_max), v_min=str(self.first), v_max=str(self.end - 1),
                    x_min=str(u_min * config.B - (self.end - 1) * config.A),
                    x_max=str(u_max * config.B - self.first * config.A))

    def split(self) -> tuple['Tile', 'Tile']:
        if self.end - self.first < 2:
            raise ValueError('unsplittable tile')
        middle = (self.first + self.end) // 2
        return Tile(self.first, middle), Tile(middle, self.end)

    def manifest(self, config: Config) -> dict[str, Any]:
        return dict(kind='v-window', first=str(self.first), end=str(self.end), envelope=self.envelope(config))


def enumerate_cpu(config: Config, tile: Tile) -> Iterator[int]:
    """Every x = uB - vA of a tile that passes all residual filters, on the CPU
    (`pseudocube/focus.py:170-179`): v over the B-focus window of the tile, u over
    the A-focus window of v's exact interval, each residual prime's allowed
    residues as a set.  For small configurations (tests and regression checks)."""
    tile.validate(config)
    af, bf = config.focuses()
    filters = [(q, frozenset(allowed(q))) for q in config.residuals()]
    for v in bf.window(tile.first, tile.end):
        low, high = config.u_interval(v)
        shift = v * config.A
        for u in af.window(low, high):
            x = u * config.B - shift
            if all(x % q in residues for q, residues in filters):
                yield x


def targeted_coordinates(config: Config, n: int, radius: int = 1) -> tuple[list[int], list[int]]:
    """CRT index neighbourhoods around n's canonical (u, v) (`pseudocube/focus.py:182-191`)."""
    u, v = config.canonical(n)
    af, bf = config.focuses()
    period, ui = af.rank(u)
    vp, vi = bf.rank(v)
    if vp != 0:
        raise ValueError('noncanonical v')
    return (list(af.indices(max(0, ui - radius), min(af.count, ui + radius + 1), period)),
            list(bf.indices(max(0, vi - radius), min(bf.count, vi + radius + 1))))
