"""The deployment formats (diver_amd.checkpoint) on the CPU: weight.pth and serialize.pth built
from a synthetic DIVeR state dict, written with torch.save and read back with the weights-only
loader.

- the blob builder equals the oracle's FP64 restatement of convert.py:63-140 (3168 / 8288 floats);
- the decoder read back from the blob (decoder.from_rt_blob) decodes like the state dict's own
  (FP64 evaluation of both canonical forms on the same features and directions);
- the sparse grid keeps exactly the corners of the occupied voxels, with their features;
- serialize.pth's index, corner flags and pre-multiplied features rebuild the mask, the corner map
  and the features (DIVeR32: F @ W0^T).
"""
import numpy as np
import pytest
import torch

from diver_amd import checkpoint as ck
from diver_amd import synthetic as S
from diver_amd.decoder import from_rt_blob, from_state_dict
from oracle import fp64


def _model_state(hidden, n=16, seed=0):
    sc = S.make_scene(n, 32, 2, 0, seed=seed)
    sd = S.reference_decoder_state(32, hidden)
    vox = sc.dense_voxels()
    rng = np.random.default_rng(seed)
    vox = vox + rng.normal(scale=0.05, size=vox.shape).astype(np.float32)  # features everywhere
    alpha = np.where(sc.mask, 0.5, 1e-3).astype(np.float32)
    state = {f'model.{k}': torch.from_numpy(np.asarray(v, np.float32)) for k, v in sd.items()}
    state['model.voxels'] = torch.from_numpy(vox)
    state['model.mlp_voxel.net.0.weight'] = torch.zeros(4, 4)  # the implicit MLP is dropped
    return sc, sd, vox, alpha, state


def _canon_eval(c, F, d):
    """Evaluate a canonical decoder in FP64 on features F [K, C] and directions d [K, 3]."""
    L = c.layers()
    z = np.asarray(F, np.float64)
    for w, b in L['pre']:
        z = np.maximum((z if w is None else z @ w.T.astype(np.float64)) + b, 0.0)
    s = z @ L['w_s'] + L['b_s']
    y = z @ L['W_c'].T + fp64.positional_encoding(d) @ L['W_v'].T + L['b_c']
    for w, b in L['post']:
        y = np.maximum(y, 0) @ w.T + b
    rgb = 1 / (1 + np.exp(-(np.maximum(y, 0) @ L['W_o'].T + L['b_o'])))
    return s, rgb


@pytest.mark.parametrize('hidden', [32, 64])
def test_blob_matches_oracle_and_decodes_alike(hidden):
    sd = S.reference_decoder_state(32, hidden)
    blob = ck.rt_blob({k: torch.from_numpy(np.asarray(v, np.float32)) for k, v in sd.items()})
    ref = fp64.rt_blob_from_state(sd, hidden)
    assert blob.size == {32: 3168, 64: 8288}[hidden] == ref.size
    np.testing.assert_allclose(blob, ref, rtol=1e-6, atol=1e-6)
    rng = np.random.default_rng(1)
    F = rng.normal(scale=0.3, size=(200, 32))
    d = rng.normal(size=(200, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    s0, c0 = _canon_eval(from_state_dict(sd, 32), F, d)
    W0 = np.asarray(sd['mlp1.0.weight'], np.float64)
    s1, c1 = _canon_eval(from_rt_blob(blob, hidden), F @ W0.T if hidden == 32 else F, d)
    assert np.abs(s0 - s1).max() < 1e-4 and np.abs(c0 - c1).max() < 1e-5


def test_blob_branch_keyed_on_decoder_shape():
    sd = S.reference_decoder_state(32, 32)
    sd['mlp1.4.weight'] = sd['mlp1.4.weight'][:, :16]
    sd['mlp1.0.weight'] = sd['mlp1.0.weight'][:16]
    with pytest.raises(ValueError):
        ck.rt_blob(sd)


@pytest.mark.parametrize('hidden', [32, 64])
def test_weight_and_serialize_round_trip(tmp_path, hidden):
    sc, sd, vox, alpha, state = _model_state(hidden)
    weight = ck.weight_from_checkpoint(state, alpha)
    assert not any('mlp_voxel' in k for k in weight)
    torch.save(weight, tmp_path / 'weight.pth')
    wc = ck.load_weight(tmp_path / 'weight.pth')
    assert torch.equal(wc.mask, torch.from_numpy(sc.mask))
    corners = S.corner_occupancy(sc.mask)
    dense = wc.voxels.numpy()
    assert np.array_equal(dense[corners], vox[corners])
    assert not dense[~corners].any()
    ser = ck.serialize(torch.load(tmp_path / 'weight.pth', weights_only=True))
    torch.save(ser, tmp_path / 'serialize.pth')
    sc_rt = ck.load_serialized(tmp_path / 'serialize.pth', sc.n)
    assert torch.equal(sc_rt.mask, torch.from_numpy(sc.mask))
    np.testing.assert_allclose(sc_rt.blob, fp64.rt_blob_from_state(sd, hidden), rtol=1e-6,
                               atol=1e-6)
    n1 = sc.n + 1
    flat = vox.reshape(-1, 32)
    idx = np.flatnonzero(corners.reshape(-1))
    rows = sc_rt.voxel_map.numpy()[idx]
    want = flat[idx].astype(np.float64)
    if hidden == 32:
        want = want @ np.asarray(sd['mlp1.0.weight'], np.float64).T
    np.testing.assert_allclose(sc_rt.features.numpy()[rows], want, rtol=1e-5, atol=1e-6)
    assert sc_rt.voxel_map.numel() == n1 ** 3
    dv = sc_rt.dense_voxels().numpy().reshape(-1, 32)
    np.testing.assert_allclose(dv[idx], want, rtol=1e-5, atol=1e-6)
