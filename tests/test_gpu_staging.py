"""Producer/consumer staging on the GPU (tools/batch_window_experiment.py:256-404).

Each mode's rows must equal the plain path's rows of the same batches, and, on the reference's
positive lease, the golden per-batch scope hashes and group populations.  The CPU producer's windows
must equal the device wheel's; a slot whose contents no longer match its checksum must fail closed.
"""
from __future__ import annotations

import hashlib
import json
import queue
import threading
from pathlib import Path

import numpy as np
import pytest

from psuedocubes_amd.accounting import encoded

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / 'golden'
LEASE = json.loads((GOLDEN / 'lease_575222931200000000_batches.json').read_text())
BATCH = 6_400_000_000
FIRST_JOB = int(LEASE['first']) // BATCH


def plain_rows(runtime, config, af, bf, jobs):
    from psuedocubes_amd.batch_experiment import MaskedMatcher, prepare
    matcher = MaskedMatcher(runtime, config)
    return {k: matcher.run(prepare(runtime, config, af, bf, k * BATCH, (k + 1) * BATCH)) for k in jobs}


@pytest.mark.parametrize('mode', ['masked-gpu-producer', 'masked-cpu-producer'])
def test_masked_staging_equals_golden_lease_batches(runtime, config619, focuses619, mode) -> None:
    from psuedocubes_amd.staging import fleet, signature
    af, bf = focuses619
    # ten batches, so the producer's eight-batch window pass and a second, shorter one both run
    jobs = list(range(FIRST_JOB, FIRST_JOB + 10))
    print(f'staging mode {mode}, jobs {jobs[0]}..{jobs[-1]}', flush=True)
    result = fleet(runtime, config619, af, bf, mode, jobs)
    assert result['status'] == 'PASS' and [r['job'] for r in result['rows']] == jobs
    for r, golden in zip(result['rows'], LEASE['batches'][:10]):
        row = r['row']
        assert hashlib.sha256(encoded(row['subtile_scopes'])).hexdigest() == golden['scopes_sha256']
        assert row['group_survivors'] == golden['group_survivors']
        assert row['all_values'] == golden['all_values']
        assert row['valid_focused_candidates'] == golden['valid_focused_candidates']
    plain = plain_rows(runtime, config619, af, bf, jobs[:3])
    for r in result['rows'][:3]:
        assert r['signature'] == signature(plain[r['job']])
    producer = result['workers'][0]
    assert producer['role'] == 'producer' and producer['jobs'] == 10
    assert result['candidates'] == sum(g['valid_focused_candidates'] for g in LEASE['batches'][:10])


def test_fused_staging_and_scattered_jobs(runtime, config619, focuses619) -> None:
    """The direct-engine consumer (FusedMatcher) on jobs that are not consecutive and one next to
    the top of the domain, through the GPU producer."""
    from psuedocubes_amd.focus import B619
    from psuedocubes_amd.staging import fleet, signature
    af, bf = focuses619
    last = B619 // BATCH - 1
    jobs = [3, 5, 6, last]
    print(f'staging fused, jobs {jobs}', flush=True)
    result = fleet(runtime, config619, af, bf, 'gpu-producer', jobs)
    plain = plain_rows(runtime, config619, af, bf, jobs)
    assert [r['job'] for r in result['rows']] == jobs
    for r in result['rows']:
        assert r['signature'] == signature(plain[r['job']])


def test_cpu_producer_windows_equal_device_windows(config619, focuses619) -> None:
    from psuedocubes_amd.staging import CPUFocus
    af, bf = focuses619
    for gpu, (lo, hi) in ((af, (0, 10**10)), (bf, (123, 6_400_000_123)),
                          (af, (af.focus.modulus - 5 * 10**9, af.focus.modulus + 5 * 10**9)),   # two periods
                          (bf, (7 * bf.focus.modulus + 11, 7 * bf.focus.modulus + 6_400_000_011))):
        host = CPUFocus(gpu).window(lo, hi)
        assert np.array_equal(host, gpu.window(lo, hi).get()), (lo, hi)
    with pytest.raises(OverflowError):
        CPUFocus(bf).window(0, 10**12, cap=1000)


def test_consumer_rejects_corrupted_slot(runtime, config619) -> None:
    """A slot whose contents differ from the producer's checksum raises ArithmeticError before
    anything is matched (tools/batch_window_experiment.py:319-320)."""
    from psuedocubes_amd.batch_experiment import MaskedMatcher
    from psuedocubes_amd.staging import CAP, PinnedBuffer, _checksum, pipeline_consumer
    slot = PinnedBuffer(2 * CAP)
    try:
        us = np.arange(10, dtype=np.uint64)
        vs = np.arange(5, dtype=np.uint64)
        slot.array[:10] = us
        slot.array[CAP:CAP + 5] = vs
        sha = _checksum(us, vs)
        slot.array[3] += np.uint64(1)   # the slot changes after the producer's checksum
        free, tasks = queue.Queue(), queue.Queue()
        tasks.put((FIRST_JOB, 0, 10, 5, sha))
        tasks.put(None)
        with pytest.raises(ArithmeticError, match='checksum'):
            pipeline_consumer(runtime, config619, MaskedMatcher(runtime, config619), 0, [slot], free, tasks,
                              threading.Event())
        assert free.empty()   # the slot was not handed back
    finally:
        slot.close()


def test_staging_rejects_bad_arguments(runtime, config619, focuses619) -> None:
    from psuedocubes_amd.staging import fleet
    af, bf = focuses619
    with pytest.raises(ValueError):
        fleet(runtime, config619, af, bf, 'fused', [1])
    with pytest.raises(ValueError):
        fleet(runtime, config619, af, bf, 'masked-gpu-producer', [])
