"""CPU ORACLE -- test infrastructure only.

FP64 re-evaluation of a full render: the FP32 traversal restatement (traverse.c) gives the
occupied intervals -- masked_intersect's walk ('offline') or the real-time walk of
aabb_intersect + ray_march ('realtime') -- and the integral, decoder and front-to-back compositing
then run in float64 (fp64.py).  Density activations: 0 = relu (model/diver.py:195),
1 = softplus (benchmark decoder).
"""
from __future__ import annotations

import numpy as np

from . import fp64, masked_intersect, octree_levels, rt_walk


def intervals(o, d, mask, n, xyzmin, voxel_size, max_n=None, walk='offline'):
    """Rows (ray index, p_in, p_out) of every occupied interval, in ray then depth order."""
    max_n = 3 * n if max_n is None else max_n
    if walk == 'offline':
        isect, num, _ = masked_intersect(o, d, np.float32(xyzmin),
                                         np.float32(xyzmin + n * voxel_size), n,
                                         np.float32(voxel_size), mask, 1.0, max_n=max_n)
    elif walk == 'realtime':
        isect, num = rt_walk(o, d, xyzmin, voxel_size, n, octree_levels(mask), mask, max_n)
    else:
        raise ValueError(walk)
    if num.max(initial=0) > max_n:
        raise ValueError('max_n too small for these rays')
    ray = np.repeat(np.arange(o.shape[0]), num)
    sel = np.arange(max_n)[None, :] < num[:, None]
    rows = isect[sel]
    return ray, rows[:, :3], rows[:, 3:], num


def render(voxels, mask, n, xyzmin, voxel_size, decode, density_act, o, d, t_stop=1e-4,
           alpha_skip=0.0, max_n=None, walk='offline'):
    """Returns dict(rgb [R,3] without background, T [R], used [R], num [R], near [R]); near marks
    rays that meet an alpha or transmittance within 1e-3 (relative) of alpha_skip or t_stop."""
    ray, pin, pout, num = intervals(o, d, mask, n, xyzmin, voxel_size, max_n, walk)
    R = o.shape[0]
    if ray.size == 0:
        return dict(rgb=np.zeros((R, 3)), T=np.ones(R), used=np.zeros(R, np.int64), num=num,
                    near=np.zeros(R, bool))
    F = fp64.integrate_einsum(voxels, pin, pout)
    s_raw, rgb = decode(F, np.asarray(d, np.float64)[ray])
    sigma = np.maximum(s_raw, 0.0) if density_act == 0 else fp64.softplus(s_raw)
    alpha = -np.expm1(-sigma)
    acc, T, used, near = fp64.render_front_to_back(ray, alpha, rgb, R, t_stop, alpha_skip,
                                                   return_near=True)
    return dict(rgb=acc, T=T, used=used, num=num, near=near)
