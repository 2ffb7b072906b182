"""Occupancy extraction on libdiver: the alpha map, the new voxel mask and the fine-ray
selection of prune.py (prune.py:83-180).

The reference loops over batches of training rays: intersect, decode, render weights, then a
torch_scatter max into an N^3 map (prune.py:83-137); thresholds the map into the voxel mask
(prune.py:146); and, for bias sampling, keeps the rays that hit the new mask (prune.py:161-169,
209-211).  Here one kernel per batch does the walk, integral, density, weights and max
(`diver_alpha_map`); the walk runs over the grid's occupancy mask (the reference's
masked_intersect path) or over an all-ones mask when the model has none (its
ray_voxel_intersect path: the same recurrence, every voxel on the ray).
"""
from __future__ import annotations

import torch

from . import calls
from .renderer import SceneGrid


class _WalkGrid:
    """A SceneGrid view whose walk mask is replaced (no octrees, blocks or bricks needed)."""

    def __init__(self, grid: SceneGrid, mask: torch.Tensor):
        self.voxels = grid.voxels
        self.mask = mask
        self.n = grid.n
        self.channels = grid.channels
        self.xyzmin = grid.xyzmin
        self.voxel_size = grid.voxel_size
        self.octrees = None
        self.blocks = None
        self.bricks = None
        self.tail = None


def alpha_map(grid: SceneGrid, decoder, o: torch.Tensor, d: torch.Tensor, pad: int = 0,
              batch: int = 1 << 24, masked: bool = True) -> torch.Tensor:
    """[N, N, N] float32 map (indexed [z][y][x]) of the largest rendering weight any ray gives
    each voxel (prune.py:83-137).  masked=False walks every voxel (a model without a mask).
    The kernel keeps no per-ray state in memory, so a batch only bounds one launch: large batches
    amortise the launch's tail of long rays (the map is the same for any batching)."""
    n = grid.n
    walk = grid if masked else _WalkGrid(grid, torch.ones_like(grid.mask))
    out = torch.zeros(n ** 3, dtype=torch.float32, device=grid.voxels.device)
    for b in range(0, o.shape[0], batch):
        calls.alpha_map(walk, decoder, o[b:b + batch], d[b:b + batch], pad, out)
    return out.view(n, n, n)


def voxel_mask(amap: torch.Tensor, thresh: float = 1e-2) -> torch.Tensor:
    """The pruned occupancy mask: voxels whose alpha exceeds thresh (prune.py:146)."""
    return amap > thresh


def rays_hitting(grid: SceneGrid, mask: torch.Tensor, o: torch.Tensor, d: torch.Tensor,
                 batch: int = 1 << 20) -> torch.Tensor:
    """bool [R]: the rays that cross at least one voxel of `mask` (prune.py:161-169, 209-211)."""
    n = grid.n
    hit = torch.empty(o.shape[0], dtype=torch.bool, device=o.device)
    m = mask.to(torch.uint8).contiguous()
    for b in range(0, o.shape[0], batch):
        _, num, _ = calls.masked_intersect_raw(o[b:b + batch], d[b:b + batch], m, grid.xyzmin, n,
                                               grid.voxel_size, 1.0, 3 * n, counts_only=True)
        hit[b:b + batch] = num > 0
    return hit
