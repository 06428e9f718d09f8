"""Prime lists and allowed-residue tables for the pseudocube filters.

Mirrors the table layer of the reference (`pseudocube/tables.py:9-36`): the
allowed residues are generated from cubes (a^3 mod q), independently of the
exponentiation test used by :mod:`psuedocubes_amd.predicate`.
"""
from __future__ import annotations

import hashlib
import json
from typing import Any


def sieve_primes(p: int) -> list[int]:
    """Primes <= p by an Eratosthenes sieve (`pseudocube/tables.py:9-15`)."""
    if p < 2:
        return []
    composite = bytearray(p + 1)
    result = []
    for q in range(2, p + 1):
        if composite[q]:
            continue
        result.append(q)
        for multiple in range(q * q, p + 1, q):
            composite[multiple] = 1
    return result


def allowed(q: int) -> tuple[int, ...]:
    """Admissible residues of x mod q (`pseudocube/tables.py:18-23`).

    q = 9: the units that are cubes, {1, 8}.  q = 1 mod 3: the nonzero cubes.
    Any other prime: every nonzero residue (coprimality only).
    """
    if q == 9:
        return (1, 8)
    if q % 3 == 1:
        return tuple(sorted({(a * a * a) % q for a in range(1, q)}))
    return tuple(range(1, q))


def packed(q: int) -> list[int]:
    """Allowed residues as a little-endian bitmap of 32-bit words (`pseudocube/tables.py:26-30`)."""
    words = [0] * ((q + 31) // 32)
    for a in allowed(q):
        words[a >> 5] |= 1 << (a & 31)
    return words


def definition(p: int) -> dict[str, Any]:
    """Canonical table dictionary and its SHA-256 (`pseudocube/tables.py:33-36`)."""
    tables = {str(q): list(allowed(q)) for q in [9] + sieve_primes(p) if q != 3}
    raw = json.dumps(tables, sort_keys=True, separators=(',', ':')).encode()
    return {'tables': tables, 'sha256': hashlib.sha256(raw).hexdigest()}


def byte_rows(moduli: list[int]) -> tuple[bytes, list[int]]:
    """Concatenated 0/1 byte rows, one row of length q per modulus, plus row offsets.

    This is the lookup layout the reference's byte matcher uploads
    (`pseudocube/gpu.py:135-142`); the HIP kernels take the same layout.
    """
    table = bytearray()
    offsets = []
    for q in moduli:
        offsets.append(len(table))
        row = bytearray(q)
        for a in allowed(q):
            row[a] = 1
        table += row
    return bytes(table), offsets
