"""Derives per-batch golden hashes for the 256-batch positive lease.

Runs the CPU oracle over the whole lease, checks that the chained transcript
and the output hash equal the reference's own fixture
(lease_575222931200000000.json, from
results/continuation-2026-09-15/positive-lease-replay.json), and only then
writes lease_575222931200000000_batches.json: for every compute batch, the
SHA-256 of the canonical encoding of its subtile_scopes, its focused
candidate count, group populations and survivors.  Because the reference
transcript is SHA-256 over the concatenation of exactly those encodings, each
per-batch hash is pinned by the reference fixture.

Takes about a minute on 8 cores.  Usage: python tests/golden/make_lease_batches.py
"""
from __future__ import annotations
import hashlib
import json
from pathlib import Path
import sys

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
from oracle import oracle as O  # noqa: E402
from psuedocubes_amd.focus import A619, B619, Config  # noqa: E402


def main() -> None:
    fixture = json.loads((HERE / 'lease_575222931200000000.json').read_text())
    config = Config(619, 10**27, 3 * 10**27, A619, B619)
    fa, fb = (O.Focus(f) for f in config.focuses())
    rows = []

    def keep(start: int, stop: int, row: dict, enc: bytes) -> None:
        rows.append(dict(first=str(start), end=str(stop), scopes_sha256=hashlib.sha256(enc).hexdigest(),
                         valid_focused_candidates=row['valid_focused_candidates'], focus_u=row['focus_u'],
                         focus_v=row['focus_v'], group_survivors=row['group_survivors'],
                         local_survivors=row['local_survivors'], all_values=row['all_values']))

    result = O.lease(config, fa, fb, int(fixture['first']), int(fixture['end']), on_batch=keep)
    scope = fixture['result']['subtile_scopes'][0]
    if result['transcript_sha256'] != scope['transcript_sha256'] or result['output_sha256'] != scope['output_sha256']:
        raise SystemExit('oracle disagrees with the reference lease fixture; nothing written')
    (HERE / 'lease_575222931200000000_batches.json').write_text(
        json.dumps(dict(first=fixture['first'], end=fixture['end'], transcript_sha256=scope['transcript_sha256'],
                        batches=rows), indent=0, sort_keys=True) + '\n')


if __name__ == '__main__':
    main()
