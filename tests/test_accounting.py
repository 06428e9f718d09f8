"""Canonical encoding of receipts (`psuedocubes_amd.accounting`).

The per-subtile output hashes and the lease transcript are SHA-256 over these
bytes (`pseudocube/accounting.py:17-49`), so the encoder is part of the
bit-exact contract.  Pinned three ways: against the reference lease's golden
per-batch scope hashes (scope records computed by the oracle, encoded here),
against hand-written bytes at the 2^53 boundary, and the fast path for lists
of small ints against the element-by-element rule on random nested values.
The library's own scope-record formatter, which the pipelined lease hashes,
is then checked byte for byte against this encoder on random records.
"""
from __future__ import annotations

import hashlib
import json
from pathlib import Path
import random

import pytest

from oracle import oracle as O
from psuedocubes_amd.accounting import digest, encoded, encoded_str_keys, interoperable

GOLDEN = Path(__file__).parent / 'golden'
LIMIT = 1 << 53


def element_rule(value):
    """Integers of magnitude >= 2^53 become strings, dict keys become strings,
    tuples become lists, visiting every element (the reference's rule)."""
    if type(value) is int and abs(value) >= LIMIT:
        return str(value)
    if isinstance(value, dict):
        return {str(k): element_rule(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return [element_rule(v) for v in value]
    return value


class Tagged(int):
    pass


def random_value(rng: random.Random, depth: int = 0):
    kind = rng.randrange(10 if depth < 3 else 6)
    if kind == 0:
        return rng.choice([0, 1, -1, LIMIT - 1, LIMIT, -LIMIT, -LIMIT + 1, 10**27, -10**30, Tagged(7), Tagged(LIMIT)])
    if kind == 1:
        return rng.choice([True, False, None, 'x', 2.5])
    if kind == 2:   # lists the fast path takes, and near misses it must not
        return [rng.choice([3, LIMIT - 1, -LIMIT + 1, LIMIT, -LIMIT, True, Tagged(2)]) for _ in range(rng.randrange(5))]
    if kind == 3:
        return [rng.randrange(-10**6, 10**6) for _ in range(rng.randrange(30))]
    if kind in (4, 5):
        return tuple(random_value(rng, depth + 1) for _ in range(rng.randrange(4)))
    if kind in (6, 7):
        return [random_value(rng, depth + 1) for _ in range(rng.randrange(5))]
    return {rng.choice(['a', 'b', 3, True, None]): random_value(rng, depth + 1) for _ in range(rng.randrange(4))}


def test_fast_path_equals_element_rule() -> None:
    rng = random.Random(619)
    for _ in range(20000):
        value = random_value(rng)
        expected = element_rule(value)
        got = interoperable(value)
        assert got == expected
        assert json.dumps(got, sort_keys=True) == json.dumps(expected, sort_keys=True)


def str_keyed(rng: random.Random, depth: int = 0):
    """Random JSON values with str keys only: integers at and around +-2^53 (as
    numbers and inside strings next to separators), digit strings, floats."""
    kind = rng.randrange(9 if depth < 4 else 5)
    if kind == 0:
        return rng.choice([0, 1, -1, LIMIT - 1, LIMIT, -LIMIT, -LIMIT + 1, 10**15, 10**16 - 1, 10**16, -10**17,
                           rng.getrandbits(70) * rng.choice([1, -1]), rng.getrandbits(52)])
    if kind == 1:
        return rng.choice(['', 'x', ':1234567890123456', ',-9007199254740993', '[' + '9' * 20, str(rng.getrandbits(90)),
                           'a"b', 'tab\tnew\nline', '\u00e9'])
    if kind == 2:
        return rng.choice([True, False, None, 0.5, 1e17, -2.5e-300, 123456789012345678.0])
    if kind in (3, 4):
        return rng.getrandbits(rng.choice([10, 52, 53, 54, 64])) * rng.choice([1, -1])
    if kind in (5, 6):
        return [str_keyed(rng, depth + 1) for _ in range(rng.randrange(6))]
    if kind == 7:
        return tuple(str_keyed(rng, depth + 1) for _ in range(rng.randrange(4)))
    return {rng.choice(['a', 'b', 'n', 'first', 'end', '9', '10', ':x']): str_keyed(rng, depth + 1)
            for _ in range(rng.randrange(5))}


def test_str_keyed_encoding_equals_encoded() -> None:
    """encoded_str_keys, the receipts' encoder, gives encoded()'s bytes on str-keyed
    values: big integers become strings whether they sit first in a list, after a
    key or after a comma, and digits inside strings change nothing."""
    rng = random.Random(6190)
    fallbacks = 0
    for _ in range(20000):
        value = str_keyed(rng)
        assert encoded_str_keys(value) == encoded(value), value
        fallbacks += encoded(value) != (json.dumps(value, sort_keys=True, separators=(',', ':')) + '\n').encode()
    assert fallbacks > 1000   # the big-integer rule was exercised
    for value in (LIMIT, -LIMIT, [LIMIT], {'a': -LIMIT}, [1, LIMIT], LIMIT - 1, [-LIMIT + 1]):
        assert encoded_str_keys(value) == encoded(value)


def test_boundary_bytes() -> None:
    value = {'b': [1, LIMIT - 1, LIMIT, -LIMIT, -LIMIT + 1], 'a': True, 'c': (None, 10**20)}
    assert encoded(value) == (b'{"a":true,"b":[1,9007199254740991,"9007199254740992","-9007199254740992",'
                              b'-9007199254740991],"c":[null,"100000000000000000000"]}\n')
    assert digest([]) == hashlib.sha256(b'[]\n').hexdigest()


def test_scope_records_hash_to_the_golden_batch(config619, oracle_focuses619) -> None:
    golden = json.loads((GOLDEN / 'lease_575222931200000000_batches.json').read_text())['batches'][0]
    row = O.batch(config619, *oracle_focuses619, int(golden['first']), int(golden['end']))
    assert hashlib.sha256(encoded(row['subtile_scopes'])).hexdigest() == golden['scopes_sha256']


def test_library_scope_records_equal_the_encoder() -> None:
    """The pipelined lease hashes scope records the library formats
    (`pcube_scope_records`); they must be byte for byte encoded(subtile_scopes),
    large counts (quoted from 2^53) and emitted subtiles included."""
    import numpy as np
    from psuedocubes_amd.batch_experiment import EMPTY_DIGEST
    from psuedocubes_amd.optimized import scope_record_bytes
    rng = random.Random(61903)
    values = [0, 1, 7, 12345, LIMIT - 1, LIMIT, 2**64 - 1]
    for _ in range(300):
        nt, ng = rng.randrange(0, 65), rng.choice([1, 4, 24])
        counters = np.array([[rng.choice(values) if rng.random() < 0.3 else rng.getrandbits(40)
                              for _ in range(ng + 3)] for _ in range(nt)], dtype=np.uint64).reshape(nt, ng + 3)
        firsts = sorted(rng.getrandbits(62) for _ in range(nt))
        bounds = np.array([(f, f + rng.randrange(1, 10**8)) for f in firsts], dtype=np.uint64).reshape(nt, 2)
        emitting = rng.random() < 0.5
        local = [rng.choice([0, 0, 3, LIMIT]) if emitting else 0 for _ in range(nt)]
        shas = [EMPTY_DIGEST if n == 0 else hashlib.sha256(bytes([k])).hexdigest() for k, n in enumerate(local)]
        scopes = [{'first': str(int(b[0])), 'end': str(int(b[1])), 'valid_focused_candidates': int(c[0]),
                   'group_survivors': [int(x) for x in c[1:ng + 1]], 'local_survivors': n, 'output_sha256': h}
                  for b, c, n, h in zip(bounds, counters, local, shas)]
        got = scope_record_bytes(bounds, counters, ng, np.array(local, dtype=np.uint64) if emitting else None,
                                 ''.join(shas).encode())
        assert got == encoded(scopes)
    with pytest.raises(ValueError):
        scope_record_bytes(np.zeros((1, 2), dtype=np.uint64), np.zeros((1, 1), dtype=np.uint64), 4, None,
                           EMPTY_DIGEST.encode())
