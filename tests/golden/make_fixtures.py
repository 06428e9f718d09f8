"""Writes the golden fixtures under tests/golden/ from data files the
reference repository holds.  Run in a checkout that has the reference at
/root/reference; the GPU host never runs this.

lease_575222931200000000.json
    Trimmed copy of results/continuation-2026-09-15/positive-lease-replay.json
    (the 256-batch positive lease replayed on the reference engine): every
    count, the 24 group populations, the 82 emitted values, the CPU-confirmed
    witness, the output hash and the transcript hash.
witness.json
    results/candidate-2026-09-15/witness.json (the noncube and its residues).
"""
from __future__ import annotations
import json
from pathlib import Path

REFERENCE = Path('/root/reference')
HERE = Path(__file__).resolve().parent

KEEP = ('focus_u', 'focus_v', 'cartesian_opportunities', 'valid_focused_candidates', 'genuine_cubes',
        'local_survivors', 'cube_checks', 'compute_batches', 'subtiles_executed', 'tile_width',
        'compute_batch_width', 'group_survivors', 'filter_groups', 'all_values', 'cpu_confirmed_hits',
        'subtile_scopes', 'receipt_format', 'status', 'host_verified', 'overflow_splits')


def main() -> None:
    replay = json.loads((REFERENCE / 'results/continuation-2026-09-15/positive-lease-replay.json').read_text())
    row = {k: replay['result'][k] for k in KEEP}
    fixture = dict(source='results/continuation-2026-09-15/positive-lease-replay.json',
                   first=replay['first'], end=replay['end'], result=row)
    (HERE / 'lease_575222931200000000.json').write_text(json.dumps(fixture, indent=1, sort_keys=True) + '\n')
    witness = json.loads((REFERENCE / 'results/candidate-2026-09-15/witness.json').read_text())
    (HERE / 'witness.json').write_text(json.dumps(witness, indent=1, sort_keys=True) + '\n')


if __name__ == '__main__':
    main()
