"""Several requests in one persistent launch (diver_render_batch).

Every request's pixels must come out exactly as when that request is rendered alone (same
interval counts, same rgba bits): a lane's ray is walked, composited and shaded by the same code
whichever rays share its wave, and the per-ray colour sum adds its intervals front to back.
One view of the mix is also checked against the FP64 oracle.
"""
import numpy as np
import pytest
import torch

from oracle import fp64
from oracle.render import render as oracle_render
from diver_amd import _lib, calls
from diver_amd import synthetic as S
from diver_amd.decoder import Decoder, from_baseline
from diver_amd.renderer import SceneGrid, camera_rays, focal_for

pytestmark = pytest.mark.gpu


def _requests(cameras, dev):
    gen = torch.Generator().manual_seed(7)
    o = (torch.rand((1000, 3), generator=gen) * 2.0 - 1.0) * 1.4
    tgt = (torch.rand((1000, 3), generator=gen) * 2.0 - 1.0) * 0.3
    d = tgt - o
    d = d / d.norm(dim=1, keepdim=True)
    reqs = [
        dict(c2w=np.asarray(cameras[0], np.float32), width=64, height=64, focal=focal_for(64, 0.69)),
        dict(c2w=np.asarray(cameras[5], np.float32), width=50, height=37, focal=focal_for(50, 0.69)),
        dict(o=o.to(dev), d=d.to(dev)),
        dict(c2w=np.asarray(cameras[11], np.float32), width=96, height=96, focal=focal_for(96, 0.69)),
        dict(o=o[:0].to(dev), d=d[:0].to(dev)),
    ]
    return reqs


@pytest.mark.parametrize('walk', ['realtime', 'offline'])
@pytest.mark.parametrize('precision', ['bf16', 'fp32'])
def test_batch_equals_alone(scene64, cameras, dev, precision, walk):
    sc = scene64
    dec = Decoder(from_baseline(S.baseline_decoder_layers(32, 32), 32))
    grid = SceneGrid.from_synthetic(sc, dev)
    opts = calls.render_opts(precision, walk, 1e-4, 0.0)
    reqs = _requests(cameras, dev)
    mixed = calls.render_batch(grid, dec, reqs, opts)
    torch.cuda.synchronize()
    for i, rq in enumerate(reqs):
        alone = calls.render_batch(grid, dec, [rq], opts)[0]
        if 'c2w' in rq:
            single = calls.render_camera(grid, dec, rq['c2w'], rq['width'], rq['height'],
                                         rq['focal'], opts)
        else:
            single = calls.render_rays(grid, dec, rq['o'], rq['d'], opts)
        for ref in (alone, single):
            assert torch.equal(mixed[i][1], ref[1]), (i, 'count')
            diff = (mixed[i][0] - ref[0]).abs().max().item() if ref[0].numel() else 0.0
            assert torch.equal(mixed[i][0], ref[0]), (i, 'rgba', diff)


def test_batch_view_matches_oracle(scene64, cameras, dev):
    sc = scene64
    layers = S.baseline_decoder_layers(32, 32)
    dec = Decoder(from_baseline(layers, 32))
    grid = SceneGrid.from_synthetic(sc, dev)
    opts = calls.render_opts('bf16', 'realtime', 1e-4, 0.0)
    reqs = _requests(cameras, dev)
    mixed = calls.render_batch(grid, dec, reqs, opts)
    W, H = 50, 37
    o, d = camera_rays(cameras[5], W, H, 0.69, dev)
    ref = oracle_render(sc.dense_voxels(), sc.mask, sc.n, sc.xyzmin, sc.voxel_size,
                        lambda F, dd: fp64.baseline_decoder(layers, F, dd), 1,
                        o.cpu().numpy(), d.cpu().numpy(), 1e-4, walk='realtime')
    rgba = mixed[1][0].cpu().numpy().reshape(-1, 4)
    rgb = rgba[:, :3] + rgba[:, 3:]
    rgb_ref = ref['rgb'] + ref['T'][:, None]
    mse = float(np.mean((rgb.astype(np.float64) - rgb_ref) ** 2))
    assert -10 * np.log10(max(mse, 1e-30)) >= 45.0
    assert np.abs(rgb - rgb_ref).max() <= 2e-2
    # counts equal on every ray whose FP64 transmittance never comes within 1e-3 (relative) of
    # the 1e-4 threshold
    cnt = mixed[1][1].cpu().numpy().reshape(-1)
    assert ref['near'].mean() <= 0.01
    assert np.array_equal(cnt[~ref['near']], ref['used'][~ref['near']])


def test_batch_limits(scene64, cameras, dev):
    sc = scene64
    dec = Decoder(from_baseline(S.baseline_decoder_layers(32, 32), 32))
    grid = SceneGrid.from_synthetic(sc, dev)
    opts = calls.render_opts('bf16', 'realtime', 1e-4, 0.0)
    assert calls.render_batch(grid, dec, [], opts) == []
    one = _requests(cameras, dev)[0]
    with pytest.raises(ValueError):
        calls.render_batch(grid, dec, [one] * (_lib.MAX_REQUESTS + 1), opts)
    outs = calls.render_batch(grid, dec, [one] * _lib.MAX_REQUESTS, opts)
    for rgba, count in outs[1:]:
        assert torch.equal(rgba, outs[0][0]) and torch.equal(count, outs[0][1])


@pytest.mark.parametrize('precision', ['bf16', 'fp32'])
def test_rays_per_wave_does_not_change_results(scene64, cameras, dev, precision):
    """Small launches give each wave fewer rays (64 at >= 131,072 rays, then 32, 16, 8, 4 down to
    about 2,048 waves).  The same 4,096 rays rendered in launches of 4,096, 16,384, 40,000, 80,000
    and 140,000 rays (4, 8, 16, 32 and 64 rays per wave; the rest of each launch is other rays) must
    come out bit for bit the same."""
    from diver_amd.renderer import render_rays
    grid = SceneGrid.from_synthetic(scene64, dev)
    dec = Decoder(from_baseline(S.baseline_decoder_layers(32, 32), 32))
    o, d = camera_rays(cameras[3], 400, 350, 0.69, dev)  # 140,000 rays
    o, d = o.reshape(-1, 3).contiguous(), d.reshape(-1, 3).contiguous()
    idx = torch.arange(0, o.shape[0], o.shape[0] // 4096, device=dev)[:4096]
    rest = torch.ones(o.shape[0], dtype=torch.bool, device=dev)
    rest[idx] = False
    others = torch.nonzero(rest).flatten()
    results = []
    for total in (4096, 16384, 40000, 80000, 140000):
        sel = torch.cat([idx, others[:total - 4096]])
        rgb, T, cnt = render_rays(grid, dec, o[sel].contiguous(), d[sel].contiguous(), precision)
        results.append((rgb[:4096], T[:4096], cnt[:4096]))
    assert int(results[0][2].sum()) > 0  # the rays see the scene
    for rgb, T, cnt in results[1:]:
        assert torch.equal(rgb, results[0][0]) and torch.equal(T, results[0][1])
        assert torch.equal(cnt, results[0][2])
