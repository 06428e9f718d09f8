#!/usr/bin/env python3
"""Benchmark: focused candidate positions per second on one MI355X.

Workload (the configuration the metric is quoted on): p = 619, 10^27 < x <= 3*10^27,
A/B = Sorenson's split; one step is one full compute batch of 64 subtiles x 1e8
canonical v (about 1.07e9 focused candidate positions), its batch index drawn
uniformly from random.Random(619003) over [0, B / 6.4e9).  Each step runs the
drop-in path end to end: both focus windows, subtile bounds, residue-bitset
build, match kernel, device-to-host copy, exact host verification and the
per-subtile output hashes.  The focus wheels and filter tables are resident
in HBM before timing starts, as in the reference's long-running worker.

Before the warmup steps the device is primed with --prime batches drawn from a
separate seeded stream (random.Random(619011)); the timed batches are the same
with or without it.

Prints one JSON line.  Extra fields: kernel-only rate, the pipelined lease rate
(256 batches in seven device passes overlapped with the host fold), the batched
stream (4096 uneven windows of mixed prime bound), the served form (Poisson range
requests, p50/p99, and the same stream with every row committed to a coverage
ledger), the roofline of the match kernel, and the CPU baseline (this
repository's C restatement on the host cores, reported only after its rows equal
the run's GPU rows).
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import random
import statistics
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# The dominant kernel is the staged bitset match, bound by integer VALU (no
# MFMA, no floating point).  Its algorithmic work is the fixed cost table of
# the batch engine: 0.56 int32 operations per focused position (bitset ANDs,
# the 128-bit survivor construction and its expected filters, per-v bounds,
# windows), so roofline.achieved is 0.56 x positions / kernel time.
OPS_PER_POSITION = 0.56
# Per 64 focused positions (one mask word) it reads one 8-byte word of each of
# the eight bitset filters' masks from LDS: 8 * 8 / 64 = 1 LDS byte per position.
STAGED_FILTERS = 8
LDS_BYTES_PER_POSITION = STAGED_FILTERS * 8 / 64
# MI355X LDS read bandwidth for ds_read_b128: 256 CUs x 256 bytes per clock x
# 2.4 GHz = 157.3e12 B/s (`bench.py --lds-probe` reads about 240 B/clk per CU
# with conflict-free addresses).
LDS_PEAK = 256 * 256 * 2.4e9
# Integer VALU issue on MI355X: the spec-derived figure, 256 CUs x 4 SIMDs x 16
# lanes x 2.4 GHz = 39.3e12 lane operations/s, is exceeded by the build's own
# microbenchmark (`bench.py --valu-probe`: independent 32-bit add/xor chains at
# 32 waves per CU run at 52.3e12), so the measured rate is the peak.
VALU_PEAK = 52.3e12
HBM_PEAK = 8.0e12


def parse() -> argparse.Namespace:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2000)
    ap.add_argument('--warmup', type=int, default=50)
    ap.add_argument('--prime', type=int, default=256,
                    help='untimed batches from a separate seeded stream before the warmup steps')
    ap.add_argument('--cpu-seconds', type=float, default=8.0, help='bounded CPU-baseline sample (config 2)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-lease', action='store_true')
    ap.add_argument('--no-batched', action='store_true')
    ap.add_argument('--batched-windows', type=int, default=4096)
    ap.add_argument('--plain-windows', type=int, default=256, help='prefix of the stream timed on the plain path')
    ap.add_argument('--no-served', action='store_true')
    ap.add_argument('--no-staging', action='store_true')
    ap.add_argument('--staging-jobs', type=int, default=64, help='batches through each producer/consumer mode')
    ap.add_argument('--served-requests', type=int, default=20000)
    ap.add_argument('--served-load', type=float, default=0.8, help='offered load as a fraction of capacity')
    ap.add_argument('--copy-probe', type=int, default=0, metavar='BYTES',
                    help='only run the plain copy kernel used to calibrate the HBM byte counters')
    ap.add_argument('--lds-probe', action='store_true',
                    help='only measure the LDS read bandwidth of the staged access pattern')
    ap.add_argument('--valu-probe', action='store_true', help='only measure the integer-VALU issue rate')
    return ap.parse_args()


def measured(key: str) -> int | None:
    """A per-launch figure of the match kernel from the committed counter summary, if present."""
    path = ROOT / 'profiles' / 'pmc_match_staged_p619_batch.md'
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith(key + ':'):
            return int(line.split(':', 1)[1])
    return None


def measured_traffic() -> int | None:
    """HBM bytes per match launch (FETCH_SIZE and WRITE_SIZE, scaled by the copy calibration)."""
    return measured('traffic_bytes_per_launch')


def valu_probe() -> None:
    """Integer-VALU issue rate (pcube_valu_probe), one JSON line."""
    import ctypes
    from psuedocubes_amd import native
    seconds, ops = ctypes.c_double(), ctypes.c_double()
    native.call('pcube_valu_probe', 20000, ctypes.byref(seconds), ctypes.byref(ops))
    rate = ops.value / seconds.value
    print(json.dumps(dict(valu_probe='8 independent add/xor chains per lane, 32 waves per CU', seconds=seconds.value,
                          lane_ops=ops.value, tops=rate / 1e12, frac_of_peak=rate / VALU_PEAK)), flush=True)


def lds_probe() -> None:
    """LDS read bandwidth of three access patterns (pcube_lds_probe), one JSON line each."""
    import ctypes
    from psuedocubes_amd import native
    names = ('consecutive', 'staged rotated columns', 'random 16-byte units')
    seconds, nbytes = ctypes.c_double(), ctypes.c_double()
    for pattern, name in enumerate(names):
        native.call('pcube_lds_probe', pattern, 2000, ctypes.byref(seconds), ctypes.byref(nbytes))
        rate = nbytes.value / seconds.value
        print(json.dumps(dict(lds_probe=name, seconds=seconds.value, bytes=nbytes.value, tbps=rate / 1e12,
                              bytes_per_clock_per_cu_at_2_4ghz=rate / 256 / 2.4e9)), flush=True)


def row_digest(row: dict) -> dict:
    """The fields of a batch row the CPU must reproduce bit for bit: every subtile's
    scope record (valid count, group populations, local survivors, output hash) as
    one SHA-256, the batch's group populations and its emitted values."""
    import hashlib
    from psuedocubes_amd.accounting import encoded
    return dict(scopes_sha256=hashlib.sha256(encoded(row['subtile_scopes'])).hexdigest(),
                group_survivors=row['group_survivors'], all_values=row['all_values'],
                valid_focused_candidates=row['valid_focused_candidates'])


def cpu_baseline(config, seconds: float, gpu_rows: list, small: tuple | None, lease: dict | None) -> dict:
    """The oracle (C, the host's threads) on the GPU's own batches of this run.

    Config 2 (the metric's workload): the first timed steps' seeded full batches,
    recomputed on the CPU for about `seconds`; every CPU row must equal the GPU
    row of the same batch (scope hashes, group populations, values) before any
    time is reported.  Beside it, on the same terms: one thread on a few of those
    batches, config 1 (the first 1e8-v subtile, v in [0, 1e8)) and config 5 (the
    reference's fixture lease of 256 batches, transcript and output hash checked
    against the GPU lease and the fixture)."""
    from oracle import oracle as O
    from psuedocubes_amd.optimized import BATCH_WIDTH
    fa, fb = (O.Focus(f) for f in config.focuses())
    threads = O.threads()
    mismatches: list[str] = []

    def gated(k: int, gpu_row: dict, nthreads: int) -> int:
        row = O.batch(config, fa, fb, k * BATCH_WIDTH, (k + 1) * BATCH_WIDTH, nthreads=nthreads)
        if row_digest(row) != row_digest(gpu_row):
            mismatches.append(f'batch {k}')
        return row['valid_focused_candidates']

    positions, batches = 0, 0
    started = time.perf_counter()
    for k, gpu_row in gpu_rows:
        if batches and time.perf_counter() - started >= seconds:
            break
        positions += gated(k, gpu_row, threads)
        batches += 1
    elapsed = time.perf_counter() - started
    one_positions, one_batches = 0, 0
    one_started = time.perf_counter()
    for k, gpu_row in gpu_rows[batches:]:
        if one_batches and time.perf_counter() - one_started >= seconds / 5:
            break
        one_positions += gated(k, gpu_row, 1)
        one_batches += 1
    one_elapsed = time.perf_counter() - one_started
    configs = {}
    if small is not None:
        (first, end), gpu_row = small
        t0 = time.perf_counter()
        row = O.batch(config, fa, fb, first, end, nthreads=threads)
        small_seconds = time.perf_counter() - t0
        if row_digest(row) != row_digest(gpu_row):
            mismatches.append('config 1 subtile')
        configs['subtile_1e8_v'] = dict(range=[str(first), str(end)], seconds=small_seconds,
                                          value=row['valid_focused_candidates'] / small_seconds,
                                          positions=row['valid_focused_candidates'])
    if lease is not None:
        t0 = time.perf_counter()
        row = O.lease(config, fa, fb, lease['first'], lease['end'])
        lease_seconds = time.perf_counter() - t0
        fixture = json.loads((ROOT / 'tests' / 'golden' / 'lease_575222931200000000.json').read_text())
        expected = fixture['result']['subtile_scopes'][0]
        if (row['transcript_sha256'], row['output_sha256']) != (lease['transcript_sha256'], lease['output_sha256']) \
                or row['transcript_sha256'] != expected['transcript_sha256'] \
                or row['valid_focused_candidates'] != lease['valid_focused_candidates']:
            mismatches.append('config 5 lease')
        configs['lease_256_batches'] = dict(first=str(lease['first']), batches=row['compute_batches'],
                                        seconds=lease_seconds, value=row['valid_focused_candidates'] / lease_seconds,
                                        transcript_sha256=row['transcript_sha256'])
    try:
        model = next(line.split(':', 1)[1].strip() for line in
                     subprocess.run(['lscpu'], capture_output=True, text=True).stdout.splitlines()
                     if line.startswith('Model name'))
    except Exception:
        model = 'unknown'
    gate = dict(passed=not mismatches, compared=dict(full_batches=batches + one_batches,
                                                     subtile_1e8_v='subtile_1e8_v' in configs,
                                                     lease_256_batches='lease_256_batches' in configs),
                fields='per-subtile scope records (SHA-256), group populations, emitted values; '
                       'the lease transcript and output hash', mismatches=mismatches)
    if mismatches:
        # no time is reported for a CPU run that does not reproduce the GPU's rows
        return dict(value=None, unit='focused candidate positions/s', cores=threads, kind='port', gate=gate)
    return dict(value=positions / elapsed, unit='focused candidate positions/s', cores=threads, kind='port',
                sample=f'{batches} of this run\'s GPU batches (seeded full p=619 batches: windows + bitset matcher '
                       f'+ counters), {elapsed:.1f} s on {threads} threads; the process may run on '
                       f'{O.available_cpus()} CPUs, os.cpu_count() = {os.cpu_count()} ({model}); the oracle uses at '
                       f'most {O.THREAD_LIMIT} threads, the CPU share granted to the process that drives one GPU',
                gate=gate,
                single_thread=dict(value=one_positions / one_elapsed if one_batches else None, cores=1,
                                   sample=f'{one_batches} more of the run\'s GPU batches, {one_elapsed:.1f} s on '
                                          f'one thread'),
                configs=configs)


def strip_timings(row: dict) -> dict:
    return {k: v for k, v in row.items() if not k.endswith('_seconds') and k not in
            ('matching_layout', 'kernel_registers', 'kernel_local_bytes', 'claim')}


def batched_leg(runtime, af, bf, windows: int, plain_windows: int) -> dict:
    """The uneven-window stream: 4096 windows, v-lengths log-uniform 1e6..1e9 from seed
    619003, p uniform over {313, 433, 619}, starts at the canonical v of x uniform over
    (10^27, 3*10^27].  Batched form over the whole stream; plain drop-in and batched form
    on the same prefix for the ratio."""
    from psuedocubes_amd.batch_experiment import MaskedMatcher, prepare
    from psuedocubes_amd.batched import BatchedEngine, window_stream
    from psuedocubes_amd.focus import A619, B619, Config
    configs = {p: Config(p, 10**27, 3 * 10**27, A619, B619) for p in (313, 433, 619)}
    engine = BatchedEngine(runtime, configs.values(), af, bf)
    stream = window_stream(configs[619], windows, 619003, (1e6, 1e9), tuple(configs))
    engine.run(stream[:64])  # capture the launch graph, size the buffers
    t0 = time.perf_counter()
    rows = engine.run(stream)
    whole = time.perf_counter() - t0
    positions = sum(r['valid_focused_candidates'] for r in rows)
    prefix = stream[:plain_windows]
    matchers = {p: MaskedMatcher(runtime, c) for p, c in configs.items()}
    for w in prefix[:8]:
        matchers[w.p].run(prepare(runtime, configs[w.p], af, bf, w.first, w.end, w.width))
    t0 = time.perf_counter()
    plain_rows = [matchers[w.p].run(prepare(runtime, configs[w.p], af, bf, w.first, w.end, w.width)) for w in prefix]
    plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    batched_rows = engine.run(prefix)
    batched_prefix = time.perf_counter() - t0
    prefix_positions = sum(r['valid_focused_candidates'] for r in plain_rows)
    identical = [strip_timings(a) for a in plain_rows] == [strip_timings(b) for b in batched_rows]
    return dict(workload=f'{windows} uneven windows, L log-uniform 1e6..1e9 v, p in {{313,433,619}}, seed 619003',
                value=positions / whole, unit='focused candidate positions/s', seconds=whole, positions=positions,
                survivors=sum(r['local_survivors'] for r in rows), graph_launches=engine.graph_launches,
                tuning=engine.tuning,
                same_stream=dict(windows=len(prefix), plain_value=prefix_positions / plain,
                                 batched_value=prefix_positions / batched_prefix,
                                 ratio=plain / batched_prefix, rows_identical=identical))


def served_leg(runtime, af, bf, requests: int, load: float) -> dict:
    """Poisson stream of range requests (v-lengths log-uniform 1e7..6.4e9, p=619) at
    `load` times the capacity measured on a saturated sample of the same stream."""
    from psuedocubes_amd.batched import BatchedEngine
    from psuedocubes_amd.focus import A619, B619, Config
    from psuedocubes_amd.served import Server, latency_summary, request_stream
    config = Config(619, 10**27, 3 * 10**27, A619, B619)
    engine = BatchedEngine(runtime, [config], af, bf)
    sample = request_stream(config, 2000, 619031, rate=0.0)
    Server(engine).serve(sample[:50], realtime=False)   # warm buffers and the launch graph
    gc.collect()
    t0 = time.perf_counter()
    saturated = Server(engine).serve(sample, realtime=False)
    seconds = time.perf_counter() - t0
    positions = sum(c.row['valid_focused_candidates'] for c in saturated)
    saturated_capacity = positions / seconds
    per_request = positions / len(sample)
    # measured capacity: the highest throughput sustained under Poisson arrivals
    # (served rate within 5 % of the offered rate) among probes below saturation
    # (the server keeps up: it finishes within 5 % of the last arrival)
    probes = []
    capacity = 0.0
    # steps of 0.05 near saturation: a probe that misses by a few milliseconds
    # costs 5 % of the capacity, not 15 %
    for fraction in (0.9, 0.85, 0.8, 0.75, 0.7, 0.65, 0.6, 0.5, 0.4, 0.3, 0.2):
        probe = request_stream(config, 3000, 619041 + len(probes), rate=fraction * saturated_capacity / per_request)
        t0 = time.perf_counter()
        got = Server(engine).serve(probe)
        makespan = time.perf_counter() - t0
        offered = sum(c.row['valid_focused_candidates'] for c in got) / probe[-1].arrival
        kept_up = makespan <= 1.05 * probe[-1].arrival + 0.02
        probes.append(dict(fraction=fraction, offered=offered, makespan=makespan, last_arrival=probe[-1].arrival,
                           kept_up=kept_up))
        if kept_up:
            capacity = offered
            break
    if capacity == 0.0:
        capacity = min(p['offered'] for p in probes)
    rate = load * capacity / per_request
    stream = request_stream(config, requests, 619003, rate=rate)
    gc.collect()
    server = Server(engine)
    t0 = time.perf_counter()
    done = server.serve(stream)
    makespan = time.perf_counter() - t0
    served = sum(c.row['valid_focused_candidates'] for c in done)
    lat = latency_summary(done)
    # the same stream again, every completed row committed to a coverage ledger (SQLite,
    # WAL, synchronous=FULL, compact gzip receipts in the database) by a writer process
    # before the request is reported complete
    import tempfile
    from psuedocubes_amd.accounting import Ledger, LedgerWriter
    with tempfile.TemporaryDirectory(prefix='served-ledger-') as directory:
        ledger_config = dict(served='p=619', lo=str(config.lo), hi=str(config.hi))
        with LedgerWriter(Path(directory), ledger_config, 'bench') as writer:
            gc.collect()
            t0 = time.perf_counter()
            with_ledger = Server(engine, ledger=writer).serve(request_stream(config, requests, 619003, rate=rate))
            ledger_makespan = time.perf_counter() - t0
        ledger = Ledger(Path(directory), ledger_config, 'bench')
        audit = ledger.audit()
        receipt_bytes = ledger.db.execute('SELECT coalesce(sum(length(data)),0) FROM receipt_blobs').fetchone()[0]
        ledger.db.close()
    ledger_leg = dict(value=sum(c.row['valid_focused_candidates'] for c in with_ledger) / ledger_makespan,
                      makespan_seconds=ledger_makespan, latency_seconds=latency_summary(with_ledger),
                      committed=audit['states'].get('committed', 0), replays=len(audit['validations']),
                      receipt_bytes=receipt_bytes)
    return dict(workload=f'{requests} Poisson range requests, L log-uniform 1e7..6.4e9 v, p=619, seed 619003',
                value=served / makespan, unit='focused candidate positions/s', capacity=capacity, ledger=ledger_leg,
                saturated_capacity=saturated_capacity, capacity_probes=probes,
                offered_load=load, rate_per_second=rate, makespan_seconds=makespan, requests=len(done),
                latency_seconds=lat, steps=server.steps, mean_requests_per_step=statistics.fmean(server.step_sizes),
                max_requests_per_step=max(server.step_sizes), host_overlap_seconds=server.overlap_seconds)


def staging_leg(runtime, config, af, bf, jobs: int) -> dict:
    """Producer/consumer staging (tools/batch_window_experiment.py:256-404): one producer (device
    wheel or host producer) and one matcher consumer through page-locked slots, over `jobs`
    consecutive full batches from the reference's positive lease; the direct-engine modes over an
    eighth of them.  Every mode's rows must carry the same signatures."""
    from psuedocubes_amd.staging import fleet
    first_job = 575222931200000000 // 6_400_000_000
    legs, signatures = {}, {}
    for mode in ('masked-gpu-producer', 'masked-cpu-producer', 'gpu-producer', 'cpu-producer'):
        n = jobs if mode.startswith('masked-') else max(2, jobs // 8)
        fleet(runtime, config, af, bf, mode, list(range(first_job + n, first_job + n + 2)))   # warm
        result = fleet(runtime, config, af, bf, mode, list(range(first_job, first_job + n)))
        for r in result['rows']:
            if signatures.setdefault(r['job'], r['signature']) != r['signature']:
                raise ArithmeticError(f'staging mode {mode} disagrees on job {r["job"]}')
        producer, consumer = result['workers'][0], result['workers'][1]
        legs[mode] = dict(jobs=n, value=result['candidates_per_second'], seconds=result['elapsed_seconds'],
                          producer_generation_seconds=producer['generation_seconds'],
                          producer_staging_seconds=producer['staging_seconds'],
                          consumer_waiting_seconds=consumer['waiting_seconds'],
                          consumer_upload_seconds=consumer['upload_seconds'],
                          consumer_filter_seconds=consumer['filter_seconds'])
    return dict(workload=f'{jobs} consecutive full p=619 batches from the positive lease (direct-engine modes: '
                         f'{max(2, jobs // 8)}), one producer and one consumer thread, 2 page-locked slots',
                unit='focused candidate positions/s', modes=legs, rows_agree=True)


def main() -> None:
    args = parse()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py measures one GPU: this engine shards by disjoint canonical-v leases with no '
              'inter-GPU traffic, and only the single-GPU figure is reported here', file=sys.stderr)
        raise SystemExit(2)
    from psuedocubes_amd.batch_experiment import MaskedMatcher, prepare
    from psuedocubes_amd.focus import A619, B619, Config
    from psuedocubes_amd.gpu import GPUFocus, Runtime
    from psuedocubes_amd.optimized import BATCH_WIDTH, compute_lease

    config = Config(619, 10**27, 3 * 10**27, A619, B619)
    runtime = Runtime(time.monotonic() + 3600, 16 * 1024**3)
    if args.lds_probe:
        lds_probe()
        return
    if args.valu_probe:
        valu_probe()
        return
    if args.copy_probe:
        import ctypes
        from psuedocubes_amd import native
        seconds = ctypes.c_double()
        for _ in range(3):
            native.call('pcube_copy_probe', args.copy_probe, ctypes.byref(seconds))
        print(json.dumps(dict(copy_probe_bytes=args.copy_probe, seconds=seconds.value,
                              gbps=2 * args.copy_probe / seconds.value / 1e9)), flush=True)
        return
    af, bf = (GPUFocus(runtime, f) for f in config.focuses())
    matcher = MaskedMatcher(runtime, config)
    rng = random.Random(619003)
    nbatch = B619 // BATCH_WIDTH

    last_batch = [0]

    def step() -> dict:
        k = rng.randrange(nbatch)
        last_batch[0] = k
        batch = prepare(runtime, config, af, bf, k * BATCH_WIDTH, (k + 1) * BATCH_WIDTH)
        return matcher.run(batch)

    # priming, before the warmup: seeded batches from a stream of their own (the timed
    # steps draw the same batches with or without it) bring the device to its steady
    # clocks and the matcher's grow-only buffers to the largest batch sizes
    prime_rng = random.Random(619011)
    primed = time.perf_counter()
    for _ in range(args.prime):
        k = prime_rng.randrange(nbatch)
        matcher.run(prepare(runtime, config, af, bf, k * BATCH_WIDTH, (k + 1) * BATCH_WIDTH))
    primed = time.perf_counter() - primed
    for _ in range(args.warmup):
        step()
    print(f'# priming {args.prime} batches ({primed:.3f} s), warmup {args.warmup} done', file=sys.stderr, flush=True)
    positions, kernel, prefilter, intersection = 0, 0.0, 0.0, 0.0
    # the first timed steps' rows (references only) for the CPU baseline's gate
    kept: list = []
    keep = 0 if args.no_cpu else 200
    started = time.perf_counter()
    for _ in range(args.steps):
        row = step()
        positions += row['valid_focused_candidates']
        kernel += row['kernel_seconds']
        prefilter += row['prefilter_build_seconds']
        intersection += matcher.intersection_seconds
        if len(kept) < keep:
            kept.append((last_batch[0], row))
    elapsed = time.perf_counter() - started
    print(f'# timed {args.steps} steps in {elapsed:.3f} s', file=sys.stderr, flush=True)

    per_launch = positions / args.steps
    match_seconds = kernel / args.steps
    # HIP events around each intersection launch on the library's stream
    intersection_seconds = intersection / args.steps
    achieved = OPS_PER_POSITION * per_launch / intersection_seconds
    lds_rate = LDS_BYTES_PER_POSITION * per_launch / intersection_seconds
    result = dict(metric='focused candidate positions/s on 1 MI355X, plain vs batched; p50/p99 latency',
                  value=positions / elapsed, unit='focused candidate positions/s', n_gpus=1, steps=args.steps,
                  warmup=args.warmup, ms_per_step=1000 * elapsed / args.steps, higher_is_better=True,
                  scaling='weak', vs_baseline=None, dtype='u64/u128 exact integer', data='synthetic',
                  config=dict(workload='plain drop-in: p=619 full batch of 64 x 1e8 v, seeded batch index '
                                       '(random.Random(619003)), prepare + MaskedMatcher.run end to end',
                              p=619, tile_width=100_000_000, subtiles=64, seed=619003),
                  kernel_only=dict(value=positions / kernel, unit='focused candidate positions/s',
                                   match_ms=1000 * match_seconds, prefilter_ms=1000 * prefilter / args.steps),
                  roofline=dict(bound='valu', achieved=achieved / 1e12, peak=VALU_PEAK / 1e12, unit='Tops/s',
                                frac=achieved / VALU_PEAK, traffic=measured_traffic(),
                                kernel='k_match_staged<8>', kernel_ms=1000 * intersection_seconds,
                                ops_per_position=OPS_PER_POSITION, positions_per_launch=per_launch),
                  lds=dict(achieved=lds_rate / 1e12, peak=LDS_PEAK / 1e12, unit='TB/s', frac=lds_rate / LDS_PEAK,
                           bytes_per_position=LDS_BYTES_PER_POSITION),
                  priming=dict(batches=args.prime, seed=619011, seconds=primed))
    # executed integer-VALU lane operations and HBM bytes of the same kernel, per
    # launch from the committed counter summary, over this run's kernel time
    valu = measured('valu_instructions_per_launch')
    if valu:
        ops = valu * 64 / intersection_seconds
        result['valu_executed'] = dict(achieved=ops / 1e12, peak=VALU_PEAK / 1e12, unit='Tops/s',
                                       frac=ops / VALU_PEAK, instructions_per_launch=valu)
    if result['roofline']['traffic']:
        rate = result['roofline']['traffic'] / intersection_seconds
        result['hbm'] = dict(achieved=rate / 1e9, peak=HBM_PEAK / 1e9, unit='GB/s', frac=rate / HBM_PEAK)
    # the full-domain control: genuine cubes r^3 of the interval with r coprime to 619#
    from psuedocubes_amd.cube_count import cube_count
    cube_count(1000000001, 1000100001, 619)   # warm
    control = cube_count(1000000001, 1442249570, 619)
    result['cube_count'] = dict(count=int(control['count']), expected=38322775,
                                kernel_ms=1000 * control['kernel_seconds'],
                                roots_per_second=(1442249570 - 1000000001 + 1) / control['kernel_seconds'])
    if not args.no_lease:
        first = 575222931200000000
        # warm the lease's buffers at full size on the next lease's range, as a worker
        # running lease after lease has them
        compute_lease(matcher, af, bf, first + 256 * BATCH_WIDTH, first + 512 * BATCH_WIDTH)
        t0 = time.perf_counter()
        lease = compute_lease(matcher, af, bf, first, first + 256 * BATCH_WIDTH)
        t1 = time.perf_counter() - t0
        result['lease'] = dict(value=lease['valid_focused_candidates'] / t1, unit='focused candidate positions/s',
                               seconds=t1, batches=256, first=str(first),
                               transcript_sha256=lease['subtile_scopes'][0]['transcript_sha256'])
        lease_check = dict(first=first, end=first + 256 * BATCH_WIDTH,
                           transcript_sha256=lease['subtile_scopes'][0]['transcript_sha256'],
                           output_sha256=lease['subtile_scopes'][0]['output_sha256'],
                           valid_focused_candidates=lease['valid_focused_candidates'])
    if not args.no_batched:
        result['batched'] = batched_leg(runtime, af, bf, args.batched_windows, args.plain_windows)
        print('# batched leg done', file=sys.stderr, flush=True)
    if not args.no_served:
        result['served'] = served_leg(runtime, af, bf, args.served_requests, args.served_load)
        print('# served leg done', file=sys.stderr, flush=True)
    if not args.no_staging:
        result['staging'] = staging_leg(runtime, config, af, bf, args.staging_jobs)
        print('# staging leg done', file=sys.stderr, flush=True)
    if not args.no_cpu:
        # config 1: the first 1e8-v subtile through the drop-in path, for the CPU's gate
        small_range = (0, 100_000_000)
        small_row = matcher.run(prepare(runtime, config, af, bf, *small_range))
        result['cpu_baseline'] = cpu_baseline(config, args.cpu_seconds, kept, (small_range, small_row),
                                              None if args.no_lease else lease_check)
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
