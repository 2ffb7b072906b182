"""Register and scratch budgets of the compiled kernels (CPU: reads the gfx950 code-object metadata
that diver_amd.build records for every kernel).

The tile kernel's occupancy and latency depend on staying inside its register budget: a change that
makes the compiler spill (as a restructured shading block once did, to 1.8 KB per lane) slows every
frame without changing a single result, so no numerical test would notice it.
"""
import json
import os

import pytest

from diver_amd import build as B

REPORT = os.path.join(B.BUILD, 'kernel_resources.json')


@pytest.fixture(scope='module')
def report():
    if not os.path.exists(REPORT):
        B.build()
    with open(REPORT) as f:
        return json.load(f)


def _tile_kernels(report):
    ks = {**report['render_tile.hip'], **report['render_tile_wide.hip']}
    return {name: r for name, r in ks.items() if 'render_tile_kernel' in name}


def test_every_source_reported(report):
    assert set(report) == {s for s in B.SOURCES if s.endswith('.hip')}
    assert all(report[s] for s in report), 'a source compiled to no kernels'


def test_tile_kernels_instantiated(report):
    # 3 decoder shapes x 3 precisions x 2 walks x 2 density forms x 2 feature addressings
    # (32-bit byte offsets in render_tile.hip, 64-bit in render_tile_wide.hip), plus the
    # segment-recording training forward: 3 decoder shapes x FP32 x offline walk x channel-0
    # density x 2 feature addressings.  Mangled names end in ...ELb<B32>ELb<SEG>EEEv.
    ks = _tile_kernels(report)
    assert len(ks) == 78
    seg = [name for name in ks if name.endswith('ELb1EEEvNS_9BatchArgsE')]
    assert len(seg) == 6
    assert all('EEELi0ELb0ELb1ELb' in name for name in seg)  # FP32, offline walk, channel-0 density
    assert all(('ELb1ELb0EEEv' in name or 'ELb1ELb1EEEv' in name) for name in report['render_tile.hip']
               if 'render_tile_kernel' in name)
    assert all(('ELb0ELb0EEEv' in name or 'ELb0ELb1EEEv' in name)
               for name in report['render_tile_wide.hip'] if 'render_tile_kernel' in name)


def test_matrix_core_tile_kernels_do_not_spill(report):
    # ...EEELi<PREC>ELb<RT>ELb<DCH0>EEEv: PREC 1 = BF16, 2 = FP16; DCH0 = density from channel 0
    # (every decoder the models ship).  The generic linear-density form is held to the scratch bound.
    # The 64-channel kernels run 3 waves per SIMD with the ray's output slot and request (at most
    # 12 bytes) in scratch, written once before the walk and read once after it; nothing else may
    # spill.
    checked = 0
    for name, r in _tile_kernels(report).items():
        tail = name.split('EEELi')[1]
        prec, dch0 = tail[0], tail.split('ELb')[2][0]
        if prec in '12' and dch0 == '1':
            if 'DecILi64ELi64' in name:
                assert r['vgpr'] <= 168 and r['vgpr_spill'] <= 3 and r['scratch_bytes'] <= 16, (name, r)
            else:
                assert r['vgpr_spill'] == 0 and r['scratch_bytes'] == 0, (name, r)
            checked += 1
    assert checked == 24


def test_benchmark_kernel_fits_four_waves_per_simd(report):
    # BF16, real-time walk, density from channel 0, 32-bit feature offsets: 16 waves per CU need
    # at most 128 registers per lane (and the LDS of 16 waves, 7.75 KB each)
    name = 'render_tile_kernelINS_3DecILi32ELi32ELi0ELb0ELi1ELi1EEELi1ELb1ELb1ELb1ELb0EE'
    (r,) = [v for k, v in _tile_kernels(report).items() if name in k]
    assert r['vgpr'] + r['agpr'] <= 128 and r['vgpr_spill'] == 0, r
    assert r['lds_bytes'] == 0  # dynamic LDS only (the wave's queue, accumulators, view terms)


def test_scratch_bounded_everywhere(report):
    for src, ks in report.items():
        for name, r in ks.items():
            assert r['scratch_bytes'] <= 64, (src, name, r)
