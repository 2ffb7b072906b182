"""Checkpoints through the MI355X path: a synthetic DIVeR32/DIVeR64 model goes through the
product-side converters (weight.pth, serialize.pth, diver_amd.checkpoint) and the weights-only
loaders into the real-time renderer (mlp_eval on the blob, eval_rt.py's structure) and the fused
renderer; both must equal the direct path (the state dict's decoder on the dense grid) and the
FP64 oracle.

Tolerances (rgb with white background, FP32): serialized vs direct path max-abs 1e-4 (the
DIVeR32 blob multiplies the first layer into the features before the integral instead of after);
vs the FP64 oracle max-abs 2e-4, except rays whose FP64 alpha or transmittance lies within 1e-3
(relative) of the reference thresholds (at most 1%).
"""
import numpy as np
import pytest
import torch

from diver_amd import calls, checkpoint as ck
from diver_amd import synthetic as S
from diver_amd.decoder import Decoder, from_state_dict
from diver_amd.realtime import RealTimeRenderer
from diver_amd.renderer import SceneGrid, render_rays
from oracle import fp64
from oracle.render import render as oracle_render

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('hidden', [32, 64])
def test_serialized_model_renders_like_direct(tmp_path, cameras, dev, hidden):
    sc = S.config1_scene()
    sd = S.reference_decoder_state(32, hidden)
    alpha = np.where(sc.mask, 0.5, 1e-3).astype(np.float32)
    state = {f'model.{k}': torch.from_numpy(np.asarray(v, np.float32)) for k, v in sd.items()}
    state['model.voxels'] = torch.from_numpy(sc.dense_voxels())
    torch.save(ck.weight_from_checkpoint(state, alpha), tmp_path / 'weight.pth')
    wc = ck.load_weight(tmp_path / 'weight.pth', dev)
    torch.save(ck.serialize(torch.load(tmp_path / 'weight.pth', weights_only=True)),
               tmp_path / 'serialize.pth')
    ser = ck.load_serialized(tmp_path / 'serialize.pth', sc.n, dev)

    # real-time loop: blob + corner map (eval_rt.py) vs the state dict on the dense grid
    rt_ser = RealTimeRenderer.from_serialized(ser, sc.xyzmin, sc.voxel_size)
    rt_dir = RealTimeRenderer(wc.voxels, wc.mask, sc.xyzmin, sc.voxel_size,
                              Decoder(wc.decoder()))
    cam = np.asarray(cameras[3], np.float32)
    img_s, _ = rt_ser.render(cam, 64, 64)
    img_d, _ = rt_dir.render(cam, 64, 64)
    o, d = calls.camera_rays(cam, 64, 64, S.focal_for(64), dev)
    ref = oracle_render(wc.voxels.cpu().numpy(), sc.mask, sc.n, sc.xyzmin, sc.voxel_size,
                        lambda F, dd: fp64.mlp_chain(sd, F, dd), 0, o.cpu().numpy(),
                        d.cpu().numpy(), 1e-2, alpha_skip=1e-2, walk='realtime')
    near = ref['near']
    assert near.mean() <= 0.01
    rgb_ref = ref['rgb'] + ref['T'][:, None]
    for img in (img_s, img_d):
        err = np.abs(img.cpu().numpy().reshape(-1, 3) - rgb_ref).max(1)
        assert err[~near].max() < 2e-4
    assert float((img_s - img_d).abs().reshape(-1, 3).max(1).values[
        torch.from_numpy(~near).to(dev)].max()) < 1e-4

    # fused renderer: the serialized scene's dense grid and decoder vs the direct ones
    g_ser = SceneGrid(ser.dense_voxels(), ser.mask, sc.xyzmin, sc.voxel_size)
    g_dir = SceneGrid(wc.voxels, wc.mask, sc.xyzmin, sc.voxel_size)
    rgb_s, _, _ = render_rays(g_ser, Decoder(ser.fused_decoder()), o, d, 'fp32', t_stop=1e-4,
                              traversal='offline')
    rgb_d, _, _ = render_rays(g_dir, Decoder(from_state_dict(sd, 32)), o, d, 'fp32', t_stop=1e-4,
                              traversal='offline')
    ref2 = oracle_render(wc.voxels.cpu().numpy(), sc.mask, sc.n, sc.xyzmin, sc.voxel_size,
                         lambda F, dd: fp64.mlp_chain(sd, F, dd), 0, o.cpu().numpy(),
                         d.cpu().numpy(), 1e-4, walk='offline')
    near2 = ref2['near']
    rgb_ref2 = ref2['rgb'] + ref2['T'][:, None]
    for rgb in (rgb_s, rgb_d):
        err = np.abs(rgb.cpu().numpy() - rgb_ref2).max(1)
        assert err[~near2].max() < 2e-4
    assert float((rgb_s - rgb_d).abs().max()) < 1e-4
