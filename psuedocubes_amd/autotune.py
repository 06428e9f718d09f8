"""Autotuner for the batched form.

Chooses, by measurement on the device it runs on:
  * direct_below_rows - the v-row count below which a window runs the direct
    per-position engine instead of building residue bitsets (the bitset build
    costs sum(q) x words regardless of the window's rows);
  * rows_per_chunk    - v rows per work item of the persistent match kernel;
  * blocks_per_cu     - persistent blocks per compute unit.
Results go to psuedocubes_amd/tuning.json (or --output), which the batched
engine reads at construction.  Any choice gives identical results; only the
speed depends on it.

    python -m psuedocubes_amd.autotune [--output PATH] [--windows N]
"""
from __future__ import annotations

import argparse
import json
import math
import random
import sys
import time
from pathlib import Path

from .batched import TUNING, BatchedEngine, Window, window_stream
from .focus import A619, B619, Config
from .gpu import GPUFocus, Runtime

PRIMES = (313, 433, 619)


def device_seconds(engine: BatchedEngine, windows: list[Window], repeats: int = 2) -> float:
    best = math.inf
    for _ in range(repeats):
        engine.launch_seconds.clear()
        engine.run(windows)
        best = min(best, sum(engine.launch_seconds))
    return best


def main(argv: list[str] | None = None) -> dict:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument('--output', type=Path, default=TUNING)
    ap.add_argument('--windows', type=int, default=512)
    ap.add_argument('--seed', type=int, default=619011)
    args = ap.parse_args(argv)
    configs = [Config(p, 10**27, 3 * 10**27, A619, B619) for p in PRIMES]
    runtime = Runtime(time.monotonic() + 1800, 32 * 1024**3)
    af, bf = (GPUFocus(runtime, f) for f in configs[-1].focuses())
    density = bf.focus.count / bf.focus.modulus
    rng = random.Random(args.seed)
    report: dict = dict(device=runtime.device_name, compute_units=runtime.compute_units, measurements={})

    # 1. engine crossover: same-length windows through each engine
    crossover = []
    for length in (3e5, 1e6, 3e6, 1e7, 3e7, 1e8):
        wins = []
        for _ in range(64):
            first = rng.randrange(0, B619 - int(length))
            wins.append(Window(first, first + int(length), 619))
        times = {}
        for mode, threshold in (('direct', 1 << 40), ('bitset', 0)):
            engine = BatchedEngine(runtime, configs, af, bf, tuning=dict(direct_below_rows=threshold))
            times[mode] = device_seconds(engine, wins) / len(wins)
        rows = length * density
        crossover.append(dict(length=length, rows=rows, **{k + '_us': 1e6 * v for k, v in times.items()}))
        print(json.dumps(crossover[-1]), file=sys.stderr, flush=True)
    report['measurements']['engine_crossover'] = crossover
    threshold = 0
    for entry in crossover:
        if entry['direct_us'] < entry['bitset_us']:
            threshold = max(threshold, int(math.ceil(entry['rows'])) + 1)
    # 2. work granularity and persistent grid on a seeded uneven stream
    stream = window_stream(configs[-1], args.windows, args.seed, (1e6, 1e9), PRIMES)
    grid = []
    best = (math.inf, None)
    for blocks in (2, 4, 8):
        for rows_per_chunk in (4, 8, 16, 32):
            tuning = dict(direct_below_rows=threshold, blocks_per_cu=blocks, rows_per_chunk=rows_per_chunk)
            engine = BatchedEngine(runtime, configs, af, bf, tuning=tuning)
            seconds = device_seconds(engine, stream)
            grid.append(dict(blocks_per_cu=blocks, rows_per_chunk=rows_per_chunk, match_ms=1e3 * seconds))
            print(json.dumps(grid[-1]), file=sys.stderr, flush=True)
            if seconds < best[0]:
                best = (seconds, tuning)
    report['measurements']['grid'] = grid
    report['choice'] = best[1]
    report['stream'] = dict(windows=args.windows, seed=args.seed, lengths='log-uniform 1e6..1e9', primes=PRIMES)
    args.output.parent.mkdir(parents=True, exist_ok=True)
    args.output.write_text(json.dumps(report, indent=1) + '\n')
    print(json.dumps(report['choice']))
    return report


if __name__ == '__main__':
    main()
