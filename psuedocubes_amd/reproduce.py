"""Bounded, portable GPU reproduction; produces no production coverage.

Drop-in for the reference's tools/reproduce.py:20-61:

    python -m psuedocubes_amd.reproduce --first F --batches N [--compare-baseline] --output O

runs `compute_lease(MaskedMatcher)` over N compute batches from canonical v = F
and, with --compare-baseline, replays the same range through the per-subtile
direct engine (`compute_batch` with `Matcher`) and requires the two receipts
to have equal digests once the `*_seconds` timings are removed.  The record
is written durably to O, which must not exist.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
from pathlib import Path
import time

from .accounting import digest, durable
from .batch_experiment import MaskedMatcher
from .focus import A619, B619, Config
from .gpu import GPUFocus, Matcher, Runtime
from .optimized import Accumulator, BATCH_WIDTH, compute_lease
from .run_production import compute_batch

PACKAGE = Path(__file__).resolve().parent


def source_manifest() -> dict[str, str]:
    """SHA-256 of every source file of the engine (the reference reads provenance/source.json)."""
    files = sorted(p for p in PACKAGE.rglob('*') if p.suffix in ('.py', '.hip', '.h') and '__pycache__' not in p.parts)
    files.append(PACKAGE.parent / 'include' / 'pcube.h')
    return {str(p.relative_to(PACKAGE.parent)): hashlib.sha256(p.read_bytes()).hexdigest() for p in files if p.exists()}


def main(argv: list[str] | None = None) -> dict:
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument('--first', type=int, default=0)
    parser.add_argument('--batches', type=int, default=1)
    parser.add_argument('--compare-baseline', action='store_true')
    parser.add_argument('--output', type=Path, required=True)
    args = parser.parse_args(argv)
    if not 1 <= args.batches <= 256 or not 0 <= args.first < B619:
        raise ValueError('bounded canonical scope required')
    if args.output.exists():
        raise ValueError('choose a new output path to retain previous evidence')
    visible = os.environ.get('HIP_VISIBLE_DEVICES', os.environ.get('ROCR_VISIBLE_DEVICES'))
    if visible is not None and (not visible or ',' in visible):
        raise ValueError('select one available GPU with HIP_VISIBLE_DEVICES')
    config = Config(619, 10**27, 3 * 10**27, A619, B619)
    end = min(B619, args.first + args.batches * BATCH_WIDTH)
    runtime = Runtime(time.monotonic() + 180, 8 * 1024**3)
    if runtime.device_count() != 1:
        raise ValueError('exactly one visible device required')
    af, bf = (GPUFocus(runtime, f) for f in config.focuses())
    result = compute_lease(MaskedMatcher(runtime, config), af, bf, args.first, end)
    baseline_seconds = None
    if args.compare_baseline:
        reference = Matcher(runtime, config)
        accumulator = Accumulator(args.first, end)
        started = time.monotonic()
        for first in range(args.first, end, BATCH_WIDTH):
            stop = min(first + BATCH_WIDTH, end)
            accumulator.add(first, stop, compute_batch(reference, af, bf, first, stop, 100_000_000))
        baseline = accumulator.finish(time.monotonic() - started)
        expected = {k: v for k, v in baseline.items() if not k.endswith('_seconds')}
        actual = {k: v for k, v in result.items() if not k.endswith('_seconds')}
        if digest(expected) != digest(actual):
            raise ArithmeticError('complete batch/reference mismatch')
        baseline_seconds = baseline['elapsed_seconds']
    record = dict(status='PASS', production_coverage=0, first=str(args.first), end=str(end),
                  baseline_compared=args.compare_baseline, baseline_seconds=baseline_seconds, result=result,
                  source_manifest=source_manifest(), peak_device_bytes=runtime.peak_device,
                  compiled_kernels=runtime.compiles, device=runtime.device_name)
    durable(args.output, record)
    summary = dict(status='PASS', output=str(args.output), production_coverage=0,
                   candidates=result['valid_focused_candidates'], seconds=result['elapsed_seconds'],
                   baseline_seconds=baseline_seconds,
                   transcript_sha256=result['subtile_scopes'][0]['transcript_sha256'])
    print(json.dumps(summary))
    return summary


if __name__ == '__main__':
    main()
