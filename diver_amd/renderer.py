"""Fused render path: resident scene grids, decoders, and single-launch rendering of rays or
pinhole cameras (traversal + segment integral + decoder + compositing per ray)."""
from __future__ import annotations

import math

import numpy as np
import torch

from . import calls
from .decoder import Decoder
from .utils.ray_march import build_octrees


class SceneGrid:
    """A scene resident in device memory: dense corner features [(N+1)^3, C] fp32, the [N^3]
    occupancy mask and its 64/16/4 occupancy pyramid, over [xyzmin, xyzmin + N*voxel_size]^3."""

    def __init__(self, voxels: torch.Tensor, mask: torch.Tensor, xyzmin: float, voxel_size: float):
        if voxels.dim() != 4 or mask.dim() != 3:
            raise ValueError('voxels must be [N+1, N+1, N+1, C] and mask [N, N, N]')
        n = mask.shape[0]
        if tuple(voxels.shape[:3]) != (n + 1,) * 3:
            raise ValueError('voxels and mask sizes disagree')
        self.voxels = voxels.contiguous()
        self.mask = mask.to(torch.uint8).contiguous()
        self.n = n
        self.channels = voxels.shape[-1]
        self.xyzmin = float(xyzmin)
        self.voxel_size = float(voxel_size)
        self.octrees = (build_octrees(self.mask.bool()).to(torch.uint8).contiguous()
                        if n % 64 == 0 else None)
        self.blocks = calls.block_bits(self.mask) if n % 8 == 0 else None
        self.bricks = calls.brick_bits(self.mask) if n % 8 == 0 else None
        # where each walk may end (the block bits dilated by one block); no effect on results
        self.tail = calls.tail_bits(self.blocks, n) if n % 8 == 0 else None

    @classmethod
    def from_synthetic(cls, scene, device) -> 'SceneGrid':
        vox = scene.dense_voxels_torch(device)
        mask = torch.from_numpy(scene.mask).to(device)
        return cls(vox, mask, scene.xyzmin, scene.voxel_size)

    @property
    def nbytes(self) -> int:
        extra = self.octrees.numel() if self.octrees is not None else 0
        return self.voxels.numel() * 4 + self.mask.numel() + extra


def _walk(grid: SceneGrid, traversal: str) -> str:
    if traversal == 'auto':
        return 'realtime' if grid.octrees is not None else 'offline'
    return traversal


def render_rays(grid: SceneGrid, decoder: Decoder, o: torch.Tensor, d: torch.Tensor,
                precision: str = 'fp32', t_stop: float = 1e-4, white_back: bool = True,
                traversal: str = 'auto', alpha_skip: float = 0.0):
    """Returns (rgb [R,3], transmittance [R], intervals composited [R])."""
    opts = calls.render_opts(precision, _walk(grid, traversal), t_stop, alpha_skip)
    rgba, count = calls.render_rays(grid, decoder, o, d, opts)
    rgb = rgba[:, :3] + rgba[:, 3:] if white_back else rgba[:, :3]
    return rgb, rgba[:, 3], count


def focal_for(width: int, fov: float) -> float:
    return 0.5 * width / math.tan(0.5 * fov)


def render_camera(grid: SceneGrid, decoder: Decoder, c2w, width: int, height: int,
                  fov: float = 0.69, precision: str = 'fp32', t_stop: float = 1e-4,
                  white_back: bool = True, with_count: bool = False, traversal: str = 'auto',
                  alpha_skip: float = 0.0):
    """One pinhole view: returns (rgb [H, W, 3], transmittance [H, W], count or None)."""
    c2w = np.asarray(c2w, np.float32).reshape(3, 4)
    opts = calls.render_opts(precision, _walk(grid, traversal), t_stop, alpha_skip)
    rgba, count = calls.render_camera(grid, decoder, c2w, width, height, focal_for(width, fov),
                                      opts, with_count)
    rgb = rgba[..., :3] + rgba[..., 3:] if white_back else rgba[..., :3]
    return rgb, rgba[..., 3], count


def camera_rays(c2w, width: int, height: int, fov: float, device):
    """The exact FP32 rays render_camera traces, row-major pixels: (o, d) [H*W, 3]."""
    c2w = np.asarray(c2w, np.float32).reshape(3, 4)
    return calls.camera_rays(c2w, width, height, focal_for(width, fov), device)
