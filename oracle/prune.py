"""CPU ORACLE -- test infrastructure only.

Float64 restatement of the alpha-map extraction of prune.py:83-137: the occupied intervals of
masked_intersect (the FP32 traversal restatement in traverse.c; an all-ones mask stands for the
unmasked ray_voxel_intersect path), the segment integral and decoder density in float64
(fp64.py), the rendering weights of model/diver.py:194-202 (w_i = alpha_i * prod_{j<i} (1 -
alpha_j), no early termination), the voxel corner min(int(p_in + 1e-4), int(p_out + 1e-4))
(prune.py:118), the pad bound (prune.py:121-124), the flat index x + y N + z N^2 (prune.py:128)
and the scatter max (prune.py:132-133).
"""
from __future__ import annotations

import numpy as np

from . import fp64
from .render import intervals


def alpha_map(voxels, mask, n, xyzmin, voxel_size, decode, density_act, o, d, pad=0, max_n=None):
    """float64 [N, N, N] ([z][y][x]) map of the largest weight any ray gives each voxel."""
    amap = np.zeros(n ** 3)
    ray, pin, pout, num = intervals(o, d, mask, n, xyzmin, voxel_size, max_n, 'offline')
    if ray.size == 0:
        return amap.reshape(n, n, n)
    F = fp64.integrate_einsum(voxels, pin, pout)
    s_raw, _ = decode(F, np.asarray(d, np.float64)[ray])
    sigma = np.maximum(s_raw, 0.0) if density_act == 0 else fp64.softplus(s_raw)
    alpha = -np.expm1(-sigma)
    # transmittance in front of each interval: exp(-(sum of sigma over the ray's earlier rows)),
    # which is prod (1 - alpha_j) exactly; rows come in ray order, then depth order
    cs = np.cumsum(sigma)
    counts = np.asarray(num, np.int64)
    counts = counts[counts > 0]
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    before_ray = np.repeat(cs[starts] - sigma[starts], counts)
    w = alpha * np.exp(-(cs - sigma - before_ray))
    c = fp64.voxel_corner(pin, pout)
    keep = np.all((c >= pad) & (c <= n - 1 - pad), axis=1)
    flat = c[:, 0] + c[:, 1] * n + c[:, 2] * n * n
    np.maximum.at(amap, flat[keep], w[keep])
    return amap.reshape(n, n, n)
