"""The alpha-map oracle (oracle/prune.py) pinned by closed forms it does not compute itself.

- A constant field along a ray through one row of voxels (tilted by ~1e-3 so the walk's
  divisions stay finite, as the reference's recurrence needs): every voxel has the same
  density, so voxel k from the entry gets w_k = alpha (1 - alpha)^k exactly.
- Rays in both directions through the same row: each voxel keeps the larger of its two weights.
- pad: voxels within pad of the border stay 0.
- The unmasked path: masked_intersect over an all-ones mask yields exactly the consecutive
  crossing pairs of ray_voxel_intersect (the reference's unmasked traversal, prune.py:108-113).
"""
import numpy as np
import pytest

import oracle
from oracle import fp64
from oracle.prune import alpha_map


def _constant_scene(n, value):
    vox = np.full((n + 1, n + 1, n + 1, 4), value, np.float32)
    return vox, np.ones((n, n, n), bool)


def _unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def _decode_ch0(F, d):
    return F[:, 0], None


@pytest.mark.parametrize('act', [0, 1])
def test_geometric_weights_along_x(act):
    n, value = 8, 0.7
    vox, mask = _constant_scene(n, value)
    o = np.array([[-2.0, -0.5 + 1.5 * (2.0 / n), 0.25 + 0.5 * (2.0 / n)]], np.float32)
    d = _unit([[1.0, 1e-3, 2e-3]])
    amap = alpha_map(vox, mask, n, -1.0, 2.0 / n, _decode_ch0, act, o, d)
    v = float(np.float32(value))  # the grid stores float32
    sigma = v if act == 0 else np.log1p(np.exp(v))
    alpha = 1.0 - np.exp(-sigma)
    hit = np.argwhere(amap > 0)
    assert len(hit) == n
    z, y = hit[0][0], hit[0][1]
    assert (hit[:, 0] == z).all() and (hit[:, 1] == y).all()
    expect = alpha * (1.0 - alpha) ** np.arange(n)
    np.testing.assert_allclose(amap[z, y, :], expect, rtol=1e-12)


def test_max_over_opposite_rays_and_pad():
    n, value = 8, 0.3
    vox, mask = _constant_scene(n, value)
    y0, z0 = -1.0 + 3.5 * (2.0 / n), -1.0 + 4.5 * (2.0 / n)
    o = np.array([[-2.0, y0, z0], [2.0, y0, z0]], np.float32)
    d = _unit([[1.0, 1e-3, 2e-3], [-1.0, 1e-3, 2e-3]])
    alpha = 1.0 - np.exp(-float(np.float32(value)))
    k = np.arange(n)
    both = np.maximum(alpha * (1 - alpha) ** k, alpha * (1 - alpha) ** (n - 1 - k))
    amap = alpha_map(vox, mask, n, -1.0, 2.0 / n, _decode_ch0, 0, o, d)
    np.testing.assert_allclose(amap[4, 3, :], both, rtol=1e-12)
    padded = alpha_map(vox, mask, n, -1.0, 2.0 / n, _decode_ch0, 0, o, d, pad=2)
    np.testing.assert_allclose(padded[4, 3, 2:n - 2], both[2:n - 2], rtol=1e-12)
    assert (padded[4, 3, :2] == 0).all() and (padded[4, 3, n - 2:] == 0).all()
    assert padded.sum() == padded[4, 3, 2:n - 2].sum()


def test_all_ones_mask_matches_crossing_pairs():
    n = 16
    rng = np.random.default_rng(5)
    R = 64
    o = rng.uniform(-1.8, 1.8, (R, 3)).astype(np.float32)
    o[:, 2] = 2.5
    tgt = rng.uniform(-0.9, 0.9, (R, 3))
    d = tgt - o
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    vs = np.float32(2.0 / n)
    isect, num, _ = oracle.masked_intersect(o, d, np.float32(-1.0), np.float32(1.0), n, vs,
                                            np.ones((n, n, n), bool), 1.0, max_n=3 * n)
    cross, cnum, _ = oracle.ray_voxel_intersect(o, d, np.float32(-1.0), np.float32(1.0), n, vs,
                                                max_n=3 * n)
    for r in range(R):
        pairs = np.concatenate([cross[r, :cnum[r] - 1], cross[r, 1:cnum[r]]], 1) if cnum[r] > 1 \
            else np.zeros((0, 6), np.float32)
        assert num[r] == len(pairs), r
        np.testing.assert_array_equal(isect[r, :num[r]], pairs)
