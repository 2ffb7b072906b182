"""Traversal parity on the GPU: hit lists, counts, entry/exit points and distances are bit-exact
against the CPU restatement of the reference kernels (oracle/traverse.c)."""
import numpy as np
import pytest
import torch

import oracle
from diver_amd import calls
from diver_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _scene256():
    return S.config2_scene()


@pytest.fixture(scope='module')
def scene256():
    return _scene256()


def _sub_rays(cam, w, h, step):
    o, d = S.camera_rays(cam, w, h)
    idx = (np.arange(h)[::step, None] * w + np.arange(w)[None, ::step]).reshape(-1)
    return o[idx], d[idx]


def _check_masked(o, d, sc, max_n, dev):
    n = sc.n
    vs = np.float32(sc.voxel_size)
    xmin = np.float32(sc.xyzmin)
    ri, rn, rt = oracle.masked_intersect(o, d, xmin, xmin + n * vs, n, vs, sc.mask, 1.0,
                                         max_n=max_n)
    assert rn.max() <= max_n
    gi, gn, gt = calls.masked_intersect_raw(torch.from_numpy(o).to(dev),
                                            torch.from_numpy(d).to(dev),
                                            torch.from_numpy(sc.mask).to(dev), xmin, n, vs, 1.0,
                                            max_n)
    gi, gn, gt = gi.cpu().numpy(), gn.cpu().numpy(), gt.cpu().numpy()
    assert np.array_equal(gn, rn)
    assert np.array_equal(gi, ri)
    assert np.array_equal(gt, rt)
    return rn


def test_masked_intersect_config1_bit_exact(scene64, cameras, dev):
    o, d = S.camera_rays(cameras[0], 64, 64)
    rn = _check_masked(o, d, scene64, 3 * 64, dev)
    assert rn.sum() > 0


def test_masked_intersect_config2_bit_exact(scene256, cameras, dev):
    total = 0
    for c in (0, 7):
        o, d = _sub_rays(cameras[c], 800, 800, 5)
        total += _check_masked(o, d, scene256, 256, dev).sum()
    assert total > 10000


def test_masked_intersect_random_rays_and_mask_scale(dev):
    n = 32
    rng = np.random.default_rng(3)
    coarse = rng.uniform(size=(16, 16, 16)) < 0.3
    o = rng.uniform(-3, 3, size=(5000, 3)).astype(np.float32)
    o[:1000] = rng.uniform(-0.9, 0.9, size=(1000, 3))   # origins inside the grid
    tgt = rng.uniform(-1, 1, size=(5000, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    d[10:20, 0] = 0.0      # +0.0 components (reference stops after entry)
    d[20:30, 1] = -0.0
    vs = np.float32(2.0 / n)
    ri, rn, rt = oracle.masked_intersect(o, d, np.float32(-1), np.float32(1), n, vs, coarse, 0.5,
                                         max_n=3 * n)
    gi, gn, gt = calls.masked_intersect_raw(torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev),
                                            torch.from_numpy(coarse).to(dev), -1.0, n, vs, 0.5, 3 * n)
    assert np.array_equal(gn.cpu().numpy(), rn)
    assert np.array_equal(gi.cpu().numpy(), ri)
    assert np.array_equal(gt.cpu().numpy(), rt)


def test_ray_voxel_intersect_bit_exact(scene64, cameras, dev):
    o, d = S.camera_rays(cameras[1], 64, 64)
    n = 64
    vs = np.float32(scene64.voxel_size)
    ri, rn, rt = oracle.ray_voxel_intersect(o, d, np.float32(-1), np.float32(1), n, vs)
    gi, gn, gt = calls.ray_voxel_intersect_raw(torch.from_numpy(o).to(dev),
                                               torch.from_numpy(d).to(dev), -1.0, n, vs, 3 * n)
    assert np.array_equal(gn.cpu().numpy(), rn)
    assert np.array_equal(gi.cpu().numpy(), ri)
    assert np.array_equal(gt.cpu().numpy(), rt)


def test_shims_trim_like_reference(scene64, cameras, dev):
    from diver_amd.utils.ray_voxel_intersection import masked_intersect, ray_voxel_intersect
    o, d = S.camera_rays(cameras[2], 64, 64)
    n, vs = 64, np.float32(scene64.voxel_size)
    ri, rn, rt = oracle.masked_intersect(o, d, np.float32(-1), np.float32(1), n, vs,
                                         scene64.mask, 1.0)
    k = rn.max()
    coord, hit, ts = masked_intersect(torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev),
                                      -1.0, 1.0, n, float(vs),
                                      torch.from_numpy(scene64.mask).to(dev), 1.0)
    assert coord.shape == (o.shape[0], k, 6)
    assert np.array_equal(coord.cpu().numpy(), ri[:, :k])
    assert np.array_equal(ts.cpu().numpy(), rt[:, :k])
    assert np.array_equal(hit.cpu().numpy(), rt[:, :k] >= 0)
    ri, rn, rt = oracle.ray_voxel_intersect(o, d, np.float32(-1), np.float32(1), n, vs)
    k = rn.max()
    coord, hit, ts = ray_voxel_intersect(torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev),
                                         -1.0, 1.0, n, float(vs))
    assert np.array_equal(coord.cpu().numpy(), ri[:, :k])
    assert np.array_equal(ts.cpu().numpy(), rt[:, :k])


def _rt_rounds(sc, o, d, dev, stride=None, poison=None, rounds=80):
    """aabb_intersect then march rounds on both sides; finish advances at the -1 terminator.
    Returns the number of rounds compared."""
    n = sc.n
    B = d.shape[0]
    centre = ((o[0] - np.float32(sc.xyzmin)) / np.float32(sc.voxel_size)).astype(np.float32)
    octrees = oracle.octree_levels(sc.mask)
    mask = sc.mask.astype(np.uint8)
    c_ref = np.zeros((8, B, 6), np.float32)
    f_ref = np.zeros(B, np.uint8)
    rgba_ref = np.zeros((B, 4), np.float32)
    oracle.aabb_intersect(c_ref, d, centre, f_ref, rgba_ref, octrees, n)
    c_gpu = torch.zeros((8, B, 6), dtype=torch.float32, device=dev)
    f_gpu = torch.zeros(B, dtype=torch.bool, device=dev)
    rgba_gpu = torch.zeros((B, 4), dtype=torch.float32, device=dev)
    dg = torch.from_numpy(d).to(dev)
    calls.aabb_intersect(c_gpu, dg, torch.from_numpy(centre).to(dev), f_gpu, rgba_gpu,
                         torch.from_numpy(octrees).to(dev), n)
    assert np.array_equal(c_gpu.cpu().numpy(), c_ref)
    assert np.array_equal(f_gpu.cpu().numpy().astype(np.uint8), f_ref)
    assert np.array_equal(rgba_gpu.cpu().numpy(), rgba_ref)
    if poison is not None:
        c_ref[7, poison, 3] = -1.0
        c_gpu[7, torch.from_numpy(poison).to(dev), 3] = -1.0
    mg = torch.from_numpy(mask).to(dev)
    stride = stride or oracle.REF_RAY_MARCH_STRIDE
    done = 0
    for _ in range(rounds):
        if f_ref.all():
            break
        oracle.ray_march(c_ref, d, mask, f_ref, n, stride)
        calls.ray_march(c_gpu, dg, mg, f_gpu, n, stride)
        assert np.array_equal(c_gpu.cpu().numpy(), c_ref)
        term = (c_ref[:, :, 0] < 0).any(0)
        f_ref = (f_ref.astype(bool) | term).astype(np.uint8)
        f_gpu = torch.from_numpy(f_ref.astype(bool)).to(dev)
        done += 1
    return done


def test_aabb_and_ray_march_bit_exact(scene256, cameras, dev):
    o, d = _sub_rays(cameras[3], 800, 800, 8)
    assert _rt_rounds(scene256, o, d, dev) >= 2


@pytest.mark.parametrize('walk', ['realtime', 'offline'])
def test_fused_walk_bit_exact(scene256, cameras, dev, walk):
    """The fused renderer's walk (reciprocal-corrected divisions, block-bitmap skipping) emits
    exactly the reference's intervals: masked_intersect for the offline walk, aabb_intersect +
    ray_march rounds for the real-time walk (each ray from its own origin)."""
    from diver_amd.renderer import SceneGrid
    sc = scene256
    grid = SceneGrid.from_synthetic(sc, dev)
    rng = np.random.default_rng(17)
    o, d = _sub_rays(cameras[5], 800, 800, 6)
    # plus rays from random origins, inside and outside the grid, and axis-aligned components
    o2 = rng.uniform(-2.5, 2.5, size=(4000, 3)).astype(np.float32)
    o2[:800] = rng.uniform(-0.9, 0.9, size=(800, 3))
    d2 = rng.uniform(-1, 1, size=(4000, 3)) - 0.3 * o2
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    d2 = d2.astype(np.float32)
    d2[:40, 0] = 0.0
    d2[40:80, 2] = -0.0
    o = np.concatenate([o, o2]).astype(np.float32)
    d = np.concatenate([d, d2]).astype(np.float32)
    max_n = 512
    if walk == 'realtime':
        ri, rn = oracle.rt_walk(o, d, sc.xyzmin, sc.voxel_size, sc.n,
                                oracle.octree_levels(sc.mask), sc.mask, max_n)
    else:
        vs = np.float32(sc.voxel_size)
        ri, rn, _ = oracle.masked_intersect(o, d, np.float32(sc.xyzmin), np.float32(1), sc.n, vs,
                                            sc.mask, 1.0, max_n=max_n)
    gi, gn = calls.walk_rays(grid, torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev),
                             walk, max_n)
    assert rn.max() <= max_n and rn.sum() > 10000
    assert np.array_equal(gn.cpu().numpy(), rn)
    assert np.array_equal(gi.cpu().numpy(), ri)


def test_ray_march_grid_stride_quirk(scene64, cameras, dev):
    """A resumable ray whose resume point is negative ends its thread's sweep
    (ray_march.cu:171-173): with a small stride the later rays of that thread are skipped."""
    o, d = S.camera_rays(cameras[4], 64, 64)
    poison = np.arange(5, 4096, 97)
    _rt_rounds(scene64, o, d, dev, stride=300, poison=poison, rounds=3)
