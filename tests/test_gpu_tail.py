"""Tail bits: a walk ends once no occupied voxel can lie ahead on its ray (diver_tail_bits).

The bits are checked against a numpy restatement of their definition (the 8^3-block occupancy
dilated by one block).  Ending the walks early must change nothing:
every kernel that reads the bits gives the same bits with and without them -- the fused render
(real-time and offline walks, BF16 and FP32 decoders, the sparsity sums), batched requests, the
walk's interval lists and the alpha map -- on camera views and on random rays that start inside
the grid, run along axes and lie on cell boundaries.
"""
import copy

import numpy as np
import pytest
import torch

from diver_amd import calls, prune
from diver_amd import synthetic as S
from diver_amd.decoder import Decoder, from_baseline, from_state_dict
from diver_amd.renderer import SceneGrid, focal_for, render_camera

pytestmark = pytest.mark.gpu


def _tail_numpy(mask):
    n = mask.shape[0]
    m = n // 8
    cell = mask.reshape(m, 8, m, 8, m, 8).any(axis=(1, 3, 5))
    pad = np.pad(cell, 1)
    dil = np.zeros_like(cell)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                dil |= pad[dz:dz + m, dy:dy + m, dx:dx + m]
    bits = np.zeros((m ** 3 + 31) // 32, dtype=np.uint32)
    idx = np.nonzero(dil.reshape(-1))[0]
    np.bitwise_or.at(bits, idx // 32, (np.uint32(1) << (idx % 32).astype(np.uint32)))
    return bits


def _without_tail(grid):
    g = copy.copy(grid)
    g.tail = None
    return g


@pytest.fixture(scope='module')
def scene256(dev):
    sc = S.config2_scene(32, 256)
    return sc, SceneGrid.from_synthetic(sc, dev)


def _edge_rays(n_rays, seed):
    """Rays from inside and outside the grid, some along axes, some starting on block corners
    (grid units multiples of 8 -> world coordinates of the [-1, 1]^3 box)."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(-1.6, 1.6, (n_rays, 3)).astype(np.float32)
    d = rng.normal(size=(n_rays, 3)).astype(np.float32)
    q = n_rays // 4
    d[:q, 1:] = 0.0                       # along x
    d[q:2 * q, 0] = 0.0                   # in the y-z plane
    cells = rng.integers(0, 33, (q, 3)).astype(np.float32)
    o[2 * q:3 * q] = (-1.0 + cells * (2.0 / 32)).astype(np.float32)  # on block corners
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def test_tail_bits_match_definition(dev):
    for sc in (S.config1_scene(), S.config2_scene(32, 256), S.make_scene(192, 8, 5, 2, seed=3),
               S.make_scene(40, 8, 3, 1, seed=4)):
        grid = SceneGrid.from_synthetic(sc, dev)
        got = grid.tail.cpu().numpy().view(np.uint32)
        assert np.array_equal(got, _tail_numpy(sc.mask.reshape((grid.n,) * 3).astype(bool)))


@pytest.mark.parametrize('prec', ['bf16', 'fp32'])
def test_camera_views_unchanged(dev, scene256, prec):
    _sc, grid = scene256
    dec = Decoder(from_baseline(S.baseline_decoder_layers(32, 32), 32))
    cams = S.hemisphere_cameras()
    bare = _without_tail(grid)
    for v in (0, 37, 111):
        a = render_camera(grid, dec, cams[v], 400, 400, precision=prec, with_count=True)
        b = render_camera(bare, dec, cams[v], 400, 400, precision=prec, with_count=True)
        for x, y in zip(a, b):
            assert torch.equal(x, y), v


@pytest.mark.parametrize('traversal,t_stop', [('offline', 0.0), ('realtime', 1e-4)])
def test_random_rays_unchanged(dev, scene256, traversal, t_stop):
    _sc, grid = scene256
    dec = Decoder(from_baseline(S.baseline_decoder_layers(32, 64), 32))
    o, d = _edge_rays(16384, 3)
    og, dg = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    opts = calls.render_opts('fp32', traversal, t_stop, 0.0)
    a = calls.render_rays_sparsity(grid, dec, og, dg, opts)
    b = calls.render_rays_sparsity(_without_tail(grid), dec, og, dg, opts)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    assert int(a[1].sum()) > 10000


@pytest.mark.parametrize('n', [192, 40])
def test_other_sizes_unchanged(dev, n):
    """Grids whose size is no power of two (the real-time walk needs n % 64 == 0, so 40 runs the
    offline walk only): the same intervals and renders with and without the tail bits."""
    sc = S.make_scene(n, 32, 5, 2, seed=3)
    grid = SceneGrid.from_synthetic(sc, dev)
    dec = Decoder(from_baseline(S.baseline_decoder_layers(32, 32), 32))
    o, d = _edge_rays(8192, 13)
    og, dg = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    walks = ['offline', 'realtime'] if n % 64 == 0 else ['offline']
    for traversal in walks:
        ia, na = calls.walk_rays(grid, og, dg, traversal, max_n=3 * n)
        ib, nb = calls.walk_rays(_without_tail(grid), og, dg, traversal, max_n=3 * n)
        assert torch.equal(na, nb) and torch.equal(ia, ib), traversal
        assert int(na.sum()) > 1000
        opts = calls.render_opts('fp32', traversal, 1e-4, 0.0)
        a = calls.render_rays(grid, dec, og, dg, opts)
        b = calls.render_rays(_without_tail(grid), dec, og, dg, opts)
        assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]), traversal


@pytest.mark.parametrize('traversal', ['offline', 'realtime'])
def test_walk_intervals_unchanged(dev, scene256, traversal):
    _sc, grid = scene256
    cams = S.hemisphere_cameras()
    o, d = S.random_rays(cams, 8192, 800, 800, 7)
    o2, d2 = _edge_rays(8192, 11)
    og = torch.from_numpy(np.concatenate([o, o2])).to(dev)
    dg = torch.from_numpy(np.concatenate([d, d2])).to(dev)
    ia, na = calls.walk_rays(grid, og, dg, traversal, max_n=768)
    ib, nb = calls.walk_rays(_without_tail(grid), og, dg, traversal, max_n=768)
    assert torch.equal(na, nb) and torch.equal(ia, ib)


def test_batched_requests_unchanged(dev, scene256):
    _sc, grid = scene256
    dec = Decoder(from_baseline(S.baseline_decoder_layers(32, 32), 32))
    cams = S.hemisphere_cameras()
    opts = calls.render_opts('bf16', 'realtime', 1e-4, 0.0)
    reqs = [{'c2w': np.asarray(cams[(13 * i) % len(cams)], np.float32), 'width': w, 'height': h,
             'focal': focal_for(w, 0.69)} for i, (w, h) in enumerate([(200, 200), (480, 270),
                                                                      (800, 800), (64, 48)])]
    a = calls.render_batch(grid, dec, reqs, opts)
    b = calls.render_batch(_without_tail(grid), dec, reqs, opts)
    for (ra, ca), (rb, cb) in zip(a, b):
        assert torch.equal(ra, rb) and torch.equal(ca, cb)


@pytest.mark.parametrize('pre_layers', [False, True])
def test_alpha_map_unchanged(dev, scene256, pre_layers):
    _sc, grid = scene256
    if pre_layers:  # the reference DIVeR32 decoder: two pre-layers on the matrix cores
        dec = Decoder(from_state_dict(S.reference_decoder_state(32, 32), 32))
    else:
        dec = Decoder(from_baseline(S.baseline_decoder_layers(32, 32), 32))
    cams = S.hemisphere_cameras()
    o, d = S.random_rays(cams, 65536, 800, 800, 5)
    og, dg = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    a = prune.alpha_map(grid, dec, og, dg)
    b = prune.alpha_map(_without_tail(grid), dec, og, dg)
    assert torch.equal(a, b)
    assert int((a > 0).sum()) > 1000
