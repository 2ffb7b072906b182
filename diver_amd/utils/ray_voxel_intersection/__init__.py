"""Drop-in for the reference's ``utils.ray_voxel_intersection`` module
(utils/ray_voxel_intersection/__init__.py:18-88), backed by libdiver instead of a JIT-built
pybind11 extension.

Same call signatures, argument orders and return values:
``ray_voxel_intersect(o, v, xyzmin, xyzmax, voxel_num, voxel_size)`` and
``masked_intersect(o, v, xyzmin, xyzmax, voxel_num, voxel_size, mask, mask_scale)`` return
``(intersection[:, :K], tns >= 0, tns[:, :K])`` with ``K`` the largest per-ray count.

The reference allocates ``3N``-wide zero-filled buffers and trims them; here a count-only pass
finds ``K`` first and the second pass writes exactly ``K`` entries per ray (hits, then the
reference's padding values 0 / -1), so the result is identical without the wide buffers.
"""
from __future__ import annotations

import torch
from torch.autograd import Function

from ... import calls


class RayVoxelIntersect(Function):
    @staticmethod
    def forward(ctx, o, v, xyzmin, xyzmax, voxel_num, voxel_size):
        _, num, _ = calls.ray_voxel_intersect_raw(o, v, xyzmin, voxel_num, voxel_size, 0,
                                                  counts_only=True)
        k = int(num.max().item()) if num.numel() else 0
        intersection, _, tns = calls.ray_voxel_intersect_raw(o, v, xyzmin, voxel_num, voxel_size,
                                                             k)
        return intersection, tns >= 0, tns

    @staticmethod
    def backward(ctx, a, b, c):
        return None, None, None, None, None, None


class MaskedIntersect(Function):
    @staticmethod
    def forward(ctx, o, v, xyzmin, xyzmax, voxel_num, voxel_size, mask, mask_scale):
        _, num, _ = calls.masked_intersect_raw(o, v, mask, xyzmin, voxel_num, voxel_size,
                                               mask_scale, 0, counts_only=True)
        k = int(num.max().item()) if num.numel() else 0
        intersection, _, tns = calls.masked_intersect_raw(o, v, mask, xyzmin, voxel_num,
                                                          voxel_size, mask_scale, k)
        return intersection, tns >= 0, tns

    @staticmethod
    def backward(ctx, a, b, c):
        return None, None, None, None, None, None, None, None


ray_voxel_intersect = RayVoxelIntersect.apply
masked_intersect = MaskedIntersect.apply


def intersect_counts(o, v, xyzmin, voxel_num, voxel_size, mask=None, mask_scale=1.0):
    """Per-ray interval counts only (no interval buffers)."""
    if mask is None:
        return calls.ray_voxel_intersect_raw(o, v, xyzmin, voxel_num, voxel_size, 0, True)[1]
    return calls.masked_intersect_raw(o, v, mask, xyzmin, voxel_num, voxel_size, mask_scale, 0,
                                      True)[1]
