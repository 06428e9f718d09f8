"""Builds libpcube.so for gfx950 in-tree with hipcc.

The library is the only native code of the package; the Python layer loads it
with ctypes and refuses to run without it.
"""
from __future__ import annotations

import os
from pathlib import Path
import subprocess

HERE = Path(__file__).resolve().parent
SOURCES = [HERE / 'csrc' / 'pcube.hip']
HEADERS = [HERE / 'csrc' / name for name in ('pcube_device.h', 'pcube_kernels.h', 'pcube_wheel.h', 'pcube_staged.h', 'pcube_cubecount.h')] + [
    HERE.parent / 'include' / 'pcube.h']
LIBRARY = HERE / 'libpcube.so'
ARCH = os.environ.get('PCUBE_OFFLOAD_ARCH', 'gfx950')


def command() -> list[str]:
    return ['/opt/rocm/bin/hipcc', f'--offload-arch={ARCH}', '-O3', '-std=c++17', '-shared', '-fPIC',
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
            '-Wall', '-Wno-unused-result', '-I', str(HERE.parent / 'include'),
            *map(str, SOURCES), '-o', str(LIBRARY)]


def build(verbose: bool = False) -> Path:
    """Compile unconditionally (seconds) and return the library path."""
    tmp = LIBRARY.with_suffix('.so.tmp')
    cmd = command()
    cmd[-1] = str(tmp)
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError('hipcc failed:\n' + result.stdout + result.stderr)
    if verbose and (result.stdout or result.stderr):
        print(result.stdout + result.stderr)
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == '__main__':
    print(build(verbose=True))
