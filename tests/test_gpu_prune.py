"""Occupancy extraction (diver_alpha_map, diver_amd.prune) against the FP64 oracle
(oracle/prune.py) on identical FP32 rays.

Stated tolerances:
  alpha map values: |GPU - FP64| <= 1e-5 (weights lie in [0, 1]; the FP32 integral, decoder and
  transmittance each carry ~1e-6 relative error);
  touched voxels: identical sets (the walk is bit-exact; a voxel counts as touched once its FP64
  weight exceeds 1e-20, below which FP32 may underflow to 0);
  voxel mask at thresh 1e-2: identical except for voxels whose FP64 weight is within 1e-5 of it;
  batch splits and repeated runs: bit-identical maps (the maximum is order-independent);
  fine-ray selection: identical to the oracle's masked_intersect counts > 0.
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import fp64
from oracle.prune import alpha_map as oracle_alpha_map
from diver_amd import prune
from diver_amd import synthetic as S
from diver_amd.decoder import Decoder, from_baseline, from_state_dict
from diver_amd.renderer import SceneGrid

pytestmark = pytest.mark.gpu

ATOL = 1e-5


@pytest.fixture(scope='module')
def rays(cameras):
    o, d = S.random_rays(cameras[:8], 4096, 64, 64, seed=7)
    return np.ascontiguousarray(o, np.float32), np.ascontiguousarray(d, np.float32)


def _decoders():
    sd = S.reference_decoder_state(32, 32)
    layers = S.baseline_decoder_layers(32, 32)
    return {
        'diver32': (Decoder(from_state_dict(sd, 32)), lambda F, d: fp64.mlp_chain(sd, F, d), 0),
        'baseline': (Decoder(from_baseline(layers, 32)),
                     lambda F, d: fp64.baseline_decoder(layers, F, d), 1),
    }


@pytest.mark.parametrize('kind', ['diver32', 'baseline'])
@pytest.mark.parametrize('pad', [0, 3])
def test_alpha_map_matches_oracle(scene64, rays, dev, kind, pad):
    sc = scene64
    dec, decode, act = _decoders()[kind]
    grid = SceneGrid.from_synthetic(sc, dev)
    o, d = rays
    got = prune.alpha_map(grid, dec, torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev),
                          pad=pad).cpu().numpy().astype(np.float64)
    ref = oracle_alpha_map(sc.dense_voxels(), sc.mask, sc.n, sc.xyzmin, sc.voxel_size, decode, act,
                           o, d, pad=pad)
    assert np.abs(got - ref).max() <= ATOL
    touched = ref > 1e-20
    assert touched.sum() > 1000
    assert not np.any((got > 0) & (ref == 0))
    assert not np.any(touched & (got == 0))
    if pad:
        border = np.ones_like(touched)
        border[pad:-pad, pad:-pad, pad:-pad] = False
        assert (got[border] == 0).all()
    near = np.abs(ref - 1e-2) <= ATOL
    gm = prune.voxel_mask(torch.from_numpy(got), 1e-2).numpy()
    assert np.array_equal(gm[~near], (ref > 1e-2)[~near])


def test_alpha_map_batches_and_repeats_bit_identical(scene64, rays, dev):
    dec = _decoders()['diver32'][0]
    grid = SceneGrid.from_synthetic(scene64, dev)
    o, d = (torch.from_numpy(x).to(dev) for x in rays)
    one = prune.alpha_map(grid, dec, o, d)
    split = prune.alpha_map(grid, dec, o, d, batch=1000)
    again = prune.alpha_map(grid, dec, o, d)
    assert torch.equal(one, split) and torch.equal(one, again)


def test_unmasked_walk(scene64, rays, dev):
    sc = scene64
    dec, decode, act = _decoders()['baseline']
    grid = SceneGrid.from_synthetic(sc, dev)
    o, d = rays
    ot, dt = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    got = prune.alpha_map(grid, dec, ot[:1024], dt[:1024], masked=False)
    ones = SceneGrid(grid.voxels, torch.ones_like(grid.mask), sc.xyzmin, sc.voxel_size)
    assert torch.equal(got, prune.alpha_map(ones, dec, ot[:1024], dt[:1024]))
    ref = oracle_alpha_map(sc.dense_voxels(), np.ones_like(sc.mask), sc.n, sc.xyzmin,
                           sc.voxel_size, decode, act, o[:1024], d[:1024])
    assert np.abs(got.cpu().numpy() - ref).max() <= ATOL


def test_rays_hitting_matches_oracle(scene64, rays, dev):
    sc = scene64
    grid = SceneGrid.from_synthetic(sc, dev)
    o, d = rays
    rng = np.random.default_rng(3)
    new_mask = sc.mask & (rng.random(sc.mask.shape) < 0.3)
    hit = prune.rays_hitting(grid, torch.from_numpy(new_mask).to(dev), torch.from_numpy(o).to(dev),
                             torch.from_numpy(d).to(dev)).cpu().numpy()
    _, num, _ = oracle.masked_intersect(o, d, np.float32(sc.xyzmin),
                                        np.float32(sc.xyzmin + sc.n * sc.voxel_size), sc.n,
                                        np.float32(sc.voxel_size), new_mask, 1.0, max_n=3 * sc.n)
    assert np.array_equal(hit, num > 0)
    assert 0 < hit.sum() < len(hit)
