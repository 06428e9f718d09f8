"""Shared pytest configuration.

Tests marked `gpu` need an MI355X and the built libpcube.so; everything else
runs on the CPU (the oracle, host logic, and the C ABI's exported symbols).
"""
from __future__ import annotations

import sys
import time
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line('markers', 'gpu: needs an MI355X GPU and libpcube.so')
    config.addinivalue_line('markers', 'slow: takes more than ten seconds on the CPU')


@pytest.fixture(scope='session')
def runtime():
    from psuedocubes_amd.gpu import Runtime
    return Runtime(time.monotonic() + 3600, 16 * 1024**3)


@pytest.fixture(scope='session')
def config619():
    from psuedocubes_amd.focus import A619, B619, Config
    return Config(619, 10**27, 3 * 10**27, A619, B619)


@pytest.fixture(scope='session')
def focuses619(runtime, config619):
    from psuedocubes_amd.gpu import GPUFocus
    return tuple(GPUFocus(runtime, f) for f in config619.focuses())


@pytest.fixture(scope='session')
def oracle_focuses619(config619):
    from oracle import oracle as O
    return tuple(O.Focus(f) for f in config619.focuses())
