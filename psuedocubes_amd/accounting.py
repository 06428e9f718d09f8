"""Canonical JSON encoding, SHA-256 digests and durable file writes.

The encoding is part of the bit-exact contract: per-subtile output hashes and
the lease transcript are SHA-256 over exactly these bytes
(`pseudocube/accounting.py:17-49`).  Integers of magnitude >= 2^53 become
decimal strings, keys are sorted, separators are compact, and a trailing
newline is appended.
"""
from __future__ import annotations

import hashlib
import json
import os
from pathlib import Path
from typing import Any
import uuid


_EXACT = 1 << 53
_PLAIN_INTS = {int}


def interoperable(value: Any) -> Any:
    """Replace integers that JSON readers could round with strings (`pseudocube/accounting.py:17-24`).

    A list whose elements are all plain ints inside (-2^53, 2^53) is copied
    as it is, checked with C-level min/max instead of one call per element:
    receipts carry many such count lists.
    """
    kind = type(value)
    if kind is int:
        return value if -_EXACT < value < _EXACT else str(value)
    if kind is str:
        return value
    if isinstance(value, dict):
        return {str(k): interoperable(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        if value and set(map(type, value)) == _PLAIN_INTS and -_EXACT < min(value) and max(value) < _EXACT:
            return list(value)
        return [interoperable(v) for v in value]
    return value


def encoded(value: Any) -> bytes:
    text = json.dumps(interoperable(value), sort_keys=True, separators=(',', ':'), allow_nan=False)
    return (text + '\n').encode()


def digest(value: Any) -> str:
    return hashlib.sha256(encoded(value)).hexdigest()


def durable(path: Path, value: Any) -> str:
    """fsync a temporary file, rename it over ``path``, fsync the directory (`pseudocube/accounting.py:35-49`)."""
    path.parent.mkdir(parents=True, exist_ok=True)
    data = encoded(value)
    temporary = path.with_name(f'{path.name}.{uuid.uuid4().hex}.tmp')
    with temporary.open('xb') as stream:
        stream.write(data)
        stream.flush()
        os.fsync(stream.fileno())
    os.replace(temporary, path)
    fd = os.open(path.parent, os.O_RDONLY)
    try:
        os.fsync(fd)
    finally:
        os.close(fd)
    return hashlib.sha256(data).hexdigest()
