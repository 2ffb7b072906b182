"""Fused renderer (traversal + integral + decoder + compositing in one launch) against the FP64
oracle on identical FP32 rays.

Stated tolerances (rgb with white background, per image, vs FP64):
  FP32 decoder: PSNR >= 80 dB, max-abs <= 1e-4
  FP16 decoder: PSNR >= 60 dB, max-abs <= 5e-3
  BF16 decoder: PSNR >= 45 dB, max-abs <= 2e-2
Interval counts (intervals composited before T < 1e-4 or exit) must be equal on every ray
except those whose FP64 transmittance after some interval lies within 1e-3 (relative) of the
threshold (the oracle's ``near`` flag), where FP32 may stop one interval earlier or later; such
rays must stay below 1% of a view.
"""
import numpy as np
import pytest
import torch

from oracle import fp64
from oracle.render import render as oracle_render
from diver_amd import synthetic as S
from diver_amd.decoder import Decoder, from_baseline, from_state_dict
from diver_amd.renderer import SceneGrid, camera_rays, render_camera, render_rays

pytestmark = pytest.mark.gpu

TOL = {'fp32': (80.0, 1e-4), 'fp16': (60.0, 5e-3), 'bf16': (45.0, 2e-2)}


def psnr(a, b):
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float('inf') if mse == 0 else -10 * np.log10(mse)


def _oracle(sc, dec_kind, layers, o, d, t_stop=1e-4, walk='realtime'):
    vox = sc.dense_voxels()
    if dec_kind == 'baseline':
        decode = lambda F, dd: fp64.baseline_decoder(layers, F, dd)
        act = 1
    else:
        decode = lambda F, dd: fp64.mlp_chain(layers, F, dd)
        act = 0
    return oracle_render(vox, sc.mask, sc.n, sc.xyzmin, sc.voxel_size, decode, act, o, d, t_stop,
                         walk=walk)


def _check_counts(got, ref):
    """Counts equal on every ray the FP64 oracle does not flag as threshold-adjacent."""
    got = np.asarray(got).reshape(-1)
    near = ref['near']
    assert near.mean() <= 0.01, near.mean()
    bad = np.nonzero((got != ref['used']) & ~near)[0]
    assert bad.size == 0, (bad[:8], got[bad[:8]], ref['used'][bad[:8]])


@pytest.mark.parametrize('walk', ['realtime', 'offline'])
@pytest.mark.parametrize('hidden', [32, 64])
@pytest.mark.parametrize('precision', ['fp32', 'bf16', 'fp16'])
def test_render_camera_baseline_decoder(scene64, cameras, dev, hidden, precision, walk):
    sc = scene64
    layers = S.baseline_decoder_layers(32, hidden)
    dec = Decoder(from_baseline(layers, 32))
    grid = SceneGrid.from_synthetic(sc, dev)
    W = H = 64
    for cam in (cameras[0], cameras[11]):
        rgb, T, cnt = render_camera(grid, dec, cam, W, H, precision=precision, with_count=True,
                                    traversal=walk)
        o, d = camera_rays(cam, W, H, 0.69, dev)
        o, d = o.cpu().numpy(), d.cpu().numpy()
        ref = _oracle(sc, 'baseline', layers, o, d, walk=walk)
        rgb_ref = (ref['rgb'] + ref['T'][:, None]).reshape(H, W, 3)
        p = psnr(rgb.cpu().numpy(), rgb_ref)
        err = np.abs(rgb.cpu().numpy() - rgb_ref).max()
        min_psnr, max_abs = TOL[precision]
        assert p >= min_psnr and err <= max_abs, (precision, p, err)
        _check_counts(cnt.cpu().numpy(), ref)


@pytest.mark.parametrize('hidden', [32, 64])
def test_render_rays_reference_decoder_fp32(scene64, cameras, dev, hidden):
    sc = scene64
    sd = S.reference_decoder_state(32, hidden)
    dec = Decoder(from_state_dict(sd, 32))
    grid = SceneGrid.from_synthetic(sc, dev)
    o, d = S.camera_rays(cameras[3], 64, 64)
    rgb, T, cnt = render_rays(grid, dec, torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev),
                              'fp32', t_stop=1e-4, traversal='offline')
    ref = _oracle(sc, 'reference', sd, o, d, walk='offline')
    rgb_ref = ref['rgb'] + ref['T'][:, None]
    assert np.abs(rgb.cpu().numpy() - rgb_ref).max() < 1e-4
    _check_counts(cnt.cpu().numpy(), ref)


@pytest.mark.parametrize('walk', ['realtime', 'offline'])
def test_render_without_termination_uses_every_interval(scene64, cameras, dev, walk):
    sc = scene64
    layers = S.baseline_decoder_layers(32, 32)
    dec = Decoder(from_baseline(layers, 32))
    grid = SceneGrid.from_synthetic(sc, dev)
    o, d = S.camera_rays(cameras[4], 64, 64)
    rgb, T, cnt = render_rays(grid, dec, torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev),
                              'fp32', t_stop=0.0, traversal=walk)
    ref = _oracle(sc, 'baseline', layers, o, d, t_stop=0.0, walk=walk)
    assert np.array_equal(cnt.cpu().numpy(), ref['num'])
    assert np.abs(T.cpu().numpy() - ref['T']).max() < 1e-5


def test_render_alpha_skip_reference_mode(scene64, cameras, dev):
    """mlp_eval's alpha skip (mlp_eval.cu:153-155) reproduced by the fused kernel."""
    sc = scene64
    sd = S.reference_decoder_state(32, 64)
    dec = Decoder(from_state_dict(sd, 32))
    grid = SceneGrid.from_synthetic(sc, dev)
    o, d = S.camera_rays(cameras[9], 64, 64)
    rgb, T, cnt = render_rays(grid, dec, torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev),
                              'fp32', t_stop=1e-2, alpha_skip=1e-2)
    ref = oracle_render(sc.dense_voxels(), sc.mask, sc.n, sc.xyzmin, sc.voxel_size,
                        lambda F, dd: fp64.mlp_chain(sd, F, dd), 0, o, d, 1e-2, alpha_skip=1e-2,
                        walk='realtime')
    err = np.abs(rgb.cpu().numpy() - (ref['rgb'] + ref['T'][:, None])).max(1)
    # every ray within 1e-4 except those whose FP64 alpha or transmittance sits within 1e-3
    # (relative) of the 1e-2 thresholds, where FP32 may take the other branch
    near = ref['near']
    assert near.mean() <= 0.01
    assert err[~near].max() < 1e-4, err[~near].max()


def test_render_256_subsample(cameras, dev):
    """256^3 scene (12 spheres + 4 shells): strided pixels of an 800x800 view."""
    sc = S.config2_scene()
    layers = S.baseline_decoder_layers(32, 32)
    dec = Decoder(from_baseline(layers, 32))
    grid = SceneGrid.from_synthetic(sc, dev)
    o, d = S.camera_rays(cameras[8], 800, 800)
    sel = np.arange(0, o.shape[0], 37)
    o, d = o[sel], d[sel]
    ref = _oracle(sc, 'baseline', layers, o, d, walk='realtime')
    rgb_ref = ref['rgb'] + ref['T'][:, None]
    for precision in ('fp32', 'bf16'):
        rgb, T, cnt = render_rays(grid, dec, torch.from_numpy(o).to(dev),
                                  torch.from_numpy(d).to(dev), precision)
        min_psnr, max_abs = TOL[precision]
        p = psnr(rgb.cpu().numpy(), rgb_ref)
        err = np.abs(rgb.cpu().numpy() - rgb_ref).max()
        assert p >= min_psnr and err <= max_abs, (precision, p, err)
        _check_counts(cnt.cpu().numpy(), ref)


@pytest.mark.parametrize('hidden', [32, 64])
def test_fp32_operands_follow_device_update(scene64, cameras, dev, hidden):
    """The FP32 decoder runs on the matrix cores from operands packed on the device, and
    diver_decoder_update (how training pushes new weights) repacks them: a decoder updated to new
    parameters must render as one created from those parameters.  Operands left from the old
    weights would be off by ~1e-2; the bound 1e-6 leaves room only for the density's
    full-integral form that an updated decoder takes instead of the channel-0 shortcut."""
    from diver_amd import calls
    sc = scene64
    grid = SceneGrid.from_synthetic(sc, dev)
    layers = S.baseline_decoder_layers(32, hidden)
    dec = Decoder(from_baseline(layers, 32))
    new = from_baseline([(w * 0.8, b + 0.05) for w, b in layers], 32)
    fresh = Decoder(new)
    before = [render_camera(grid, dec, cam, 64, 64, precision='fp32')[0]
              for cam in (cameras[0], cameras[11])]
    calls.decoder_update(dec, torch.from_numpy(np.ascontiguousarray(new.params)).to(dev))
    for cam, rgb0 in zip((cameras[0], cameras[11]), before):
        rgb_u, T_u, _ = render_camera(grid, dec, cam, 64, 64, precision='fp32')
        rgb_f, T_f, _ = render_camera(grid, fresh, cam, 64, 64, precision='fp32')
        assert float((rgb_u - rgb0).abs().max()) > 1e-3  # the update changed the image
        assert float((rgb_u - rgb_f).abs().max()) <= 1e-6
        assert float((T_u - T_f).abs().max()) <= 1e-6


def test_range_update_keeps_channel0_density(scene64, cameras, dev):
    """diver_decoder_update_range over the colour part only (params[C+1:]) leaves the density row
    and its channel-0 shortcut in place: the updated decoder renders bit for bit like one created
    from the new parameters (which detects the same shortcut)."""
    from diver_amd import calls
    sc = scene64
    grid = SceneGrid.from_synthetic(sc, dev)
    layers = S.baseline_decoder_layers(32, 64)
    dec = Decoder(from_baseline(layers, 32))
    new = from_baseline([(w * 0.8, b + 0.05) for w, b in layers], 32)
    fresh = Decoder(new)
    calls.decoder_update(dec, torch.from_numpy(np.ascontiguousarray(new.params)).to(dev), first=33)
    for cam in (cameras[0], cameras[11]):
        rgb_u, T_u, n_u = render_camera(grid, dec, cam, 64, 64, precision='fp32', with_count=True)
        rgb_f, T_f, n_f = render_camera(grid, fresh, cam, 64, 64, precision='fp32', with_count=True)
        assert torch.equal(rgb_u, rgb_f) and torch.equal(T_u, T_f) and torch.equal(n_u, n_f)
