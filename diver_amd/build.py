"""Build recipe for libdiver.so (gfx950).  Run ``python -m diver_amd.build``.

Every HIP translation unit is compiled with hipcc for gfx950 only, with packed-FP32 code
generation disabled (``-packed-fp32-ops``) and IEEE-rounded FP32 division/sqrt, then linked into
``diver_amd/libdiver.so`` next to this file.  The registers, spills and scratch of every kernel,
read from the compiled code objects' metadata, go to ``_build/kernel_resources.json``.
"""
from __future__ import annotations

import concurrent.futures as cf
import json
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
           'occupancy.hip', 'render_tile.hip', 'render_tile_wide.hip', 'adjoint.hip', 'adjoint_mlp.hip', 'prune.hip', 'optim.hip', 'capi.cpp', 'server.cpp']

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


def _llvm_tool(name: str) -> str:
    for d in (os.environ.get('ROCM_PATH', '/opt/rocm') + '/llvm/bin', '/opt/rocm/llvm/bin'):
        p = os.path.join(d, name)
        if os.path.exists(p):
            return p
    raise RuntimeError(f'{name} not found')


def kernel_resources(obj: str) -> dict:
    """Per-kernel resource use of one compiled object, read from the AMDGPU metadata of its gfx950
    code object: {kernel symbol: {vgpr, agpr, vgpr_spill, sgpr_spill, scratch_bytes, lds_bytes}}."""
    import re
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        fat = os.path.join(tmp, 'fatbin')
        co = os.path.join(tmp, 'co')
        subprocess.run([_llvm_tool('llvm-objcopy'), f'--dump-section=.hip_fatbin={fat}', obj, os.devnull],
                       check=True, capture_output=True)
        subprocess.run([_llvm_tool('clang-offload-bundler'), '--unbundle', '--type=o', f'--input={fat}',
                        f'--targets=hipv4-amdgcn-amd-amdhsa--{ARCH}', f'--output={co}'],
                       check=True, capture_output=True)
        notes = subprocess.run([_llvm_tool('llvm-readelf'), '--notes', co], check=True,
                               capture_output=True, text=True).stdout
    out = {}
    for blk in notes.split('  - .'):
        m = re.search(r'\.name:\s+(\S+)', blk)
        if not m or '.kernarg_segment_size' not in blk:
            continue

        def field(k):
            f = re.search(r'\.' + k + r':\s+(\d+)', blk)
            return int(f.group(1)) if f else 0

        out[m.group(1)] = {'vgpr': field('vgpr_count'), 'agpr': field('agpr_count'),
                           'vgpr_spill': field('vgpr_spill_count'), 'sgpr_spill': field('sgpr_spill_count'),
                           'scratch_bytes': field('private_segment_fixed_size'),
                           'lds_bytes': field('group_segment_fixed_size')}
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    report = {}
    for src, obj in zip(SOURCES, objs):
        if src.endswith('.hip'):
            report[src] = kernel_resources(obj)
    with open(os.path.join(BUILD, 'kernel_resources.json'), 'w') as f:
        json.dump(report, f, indent=1, sort_keys=True)
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
