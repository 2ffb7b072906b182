"""Drop-in for the reference's ``utils.ray_march`` module (utils/ray_march/__init__.py:16-17):
``aabb_intersect(o, v, center, finish, rgba, octrees, voxel_num)`` and
``ray_march(o, v, mask, finish, voxel_num)``, both in place on the caller's buffers
(``o`` = coords ``[8, H, W, 6]``, ``v`` = directions ``[H, W, 3]``, ``center`` = shared origin in grid
units on the device).  ``voxel_num`` may be any positive multiple of 64 (the reference compiles 128
and 256 and silently does nothing otherwise; here other sizes raise).
"""
from __future__ import annotations

import torch
import torch.nn.functional as NF

from ... import calls


def aabb_intersect(o, v, center, finish, rgba, octrees, voxel_num):
    calls.aabb_intersect(o, v, center, finish, rgba, octrees, voxel_num)


def ray_march(o, v, mask, finish, voxel_num):
    calls.ray_march(o, v, mask, finish, voxel_num)


def build_octrees(voxel_masks: torch.Tensor) -> torch.Tensor:
    """Occupancy pyramid for aabb_intersect (eval_rt.py:236-239): max-pooled masks at block
    sizes 64, 16 and 4, flattened and concatenated coarse first."""
    m = voxel_masks[None, None].float()
    levels = [NF.max_pool3d(m, s, stride=s)[0, 0].bool().reshape(-1) for s in (64, 16, 4)]
    return torch.cat(levels, 0).contiguous()
