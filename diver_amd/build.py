"""Build recipe for libdiver.so (gfx950).  Run ``python -m diver_amd.build``.

Every HIP translation unit is compiled with hipcc for gfx950 only, with packed-FP32 code
generation disabled (``-packed-fp32-ops``) and IEEE-rounded FP32 division/sqrt, then linked into
``diver_amd/libdiver.so`` next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, 'libdiver.so')
BUILD = os.path.join(HERE, '_build')
ARCH = os.environ.get('DIVER_OFFLOAD_ARCH', 'gfx950')

SOURCES = ['traverse.hip', 'integrate.hip', 'composite.hip', 'rt_decoder.hip', 'render.hip',
           'occupancy.hip', 'render_tile.hip', 'adjoint.hip', 'capi.cpp']

COMMON = ['-O3', '-fPIC', '-std=c++17', f'--offload-arch={ARCH}',
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
          '-fhip-fp32-correctly-rounded-divide-sqrt', '-Wno-unused-result']


def hipcc() -> str:
    for c in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', shutil.which('hipcc')):
        if c and os.path.exists(c):
            return c
    raise RuntimeError('hipcc not found')


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + '.o')
    cmd = [hipcc(), *COMMON, '-x', 'hip', '-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'compile failed: {src}\n{" ".join(cmd)}\n{r.stdout}\n{r.stderr}')
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + '.tmp'
    cmd = [hipcc(), '-shared', '-fPIC', f'--offload-arch={ARCH}', *objs, '-o', tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed\n{" ".join(cmd)}\n{r.stdout}\n{r.stderr}')
    os.replace(tmp, OUT)
    if verbose:
        print(f'built {OUT}')
    return OUT


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
