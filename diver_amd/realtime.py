"""The reference's real-time frame loop (eval_rt.py:52-135) on libdiver's drop-in ops: the plain
FP32 GPU path.

Setup (eval_rt.py:52-89): sparse corner features + (N+1)^3 corner map, dense occupancy, the
64/16/4 octree pyramid, the decoder, and the per-frame buffers coords [8, H, W, 6], rgba [H, W, 4],
finish [H, W].  Per frame (eval_rt.py:106-122): aabb_intersect, then march rounds
(ray_march: <= 8 occupied intervals per ray) alternating with mlp_eval (decode + composite) until
every ray is finished; the white background is rgba[:3] + rgba[3].
"""
from __future__ import annotations

import numpy as np
import torch

from . import calls
from .decoder import CanonicalDecoder, Decoder
from .renderer import focal_for
from .utils.ray_march import build_octrees


class RealTimeRenderer:
    def __init__(self, voxels: torch.Tensor, mask: torch.Tensor, xyzmin: float, voxel_size: float,
                 decoder: Decoder, alpha_skip: float = 1e-2, t_stop: float = 1e-2):
        dev = voxels.device
        n = mask.shape[0]
        if n % 64:
            raise ValueError('the octree descent needs N to be a multiple of 64')
        self.n = n
        self.xyzmin = float(xyzmin)
        self.voxel_size = float(voxel_size)
        self.decoder = decoder
        self.alpha_skip = alpha_skip
        self.t_stop = t_stop
        self.mask = mask.to(torch.uint8).contiguous()
        # corners of occupied voxels -> sparse rows (convert.py:35-44, eval_rt.py:230-233)
        m = mask.bool()
        cocc = torch.zeros((n + 1,) * 3, dtype=torch.bool, device=dev)
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    cocc[dz:dz + n, dy:dy + n, dx:dx + n] |= m
        idx = torch.nonzero(cocc.reshape(-1)).reshape(-1)
        C = voxels.shape[-1]
        self.features = voxels.reshape(-1, C).index_select(0, idx).contiguous()
        vmap = torch.zeros((n + 1) ** 3, dtype=torch.int32, device=dev)
        vmap[idx] = torch.arange(idx.numel(), dtype=torch.int32, device=dev)
        self.voxel_map = vmap.reshape(n + 1, n + 1, n + 1)
        self.octrees = build_octrees(m).to(torch.uint8)
        self._buffers = {}

    @classmethod
    def from_serialized(cls, scene, xyzmin: float, voxel_size: float, alpha_skip: float = 1e-2,
                        t_stop: float = 1e-2) -> 'RealTimeRenderer':
        """From a loaded ``serialize.pth`` (diver_amd.checkpoint.load_serialized): its rows, corner
        map and blob decoder as eval_rt.py:52-89 installs them."""
        self = cls.__new__(cls)
        n = scene.n
        if n % 64:
            raise ValueError('the octree descent needs N to be a multiple of 64')
        self.n = n
        self.xyzmin = float(xyzmin)
        self.voxel_size = float(voxel_size)
        self.decoder = Decoder(scene.rt_decoder())
        self.alpha_skip = alpha_skip
        self.t_stop = t_stop
        self.mask = scene.mask.to(torch.uint8).contiguous()
        self.features = scene.features
        self.voxel_map = scene.voxel_map.reshape(n + 1, n + 1, n + 1)
        self.octrees = build_octrees(scene.mask.bool()).to(torch.uint8)
        self._buffers = {}
        return self

    def _bufs(self, height, width, dev):
        key = (height, width)
        if key not in self._buffers:
            self._buffers[key] = (
                torch.zeros((8, height, width, 6), dtype=torch.float32, device=dev),
                torch.zeros((height, width, 4), dtype=torch.float32, device=dev),
                torch.zeros((height, width), dtype=torch.bool, device=dev))
        return self._buffers[key]

    def render(self, c2w, width: int, height: int, fov: float = 0.69, max_rounds: int = 10000):
        """Returns (rgb [H, W, 3] with white background, rounds)."""
        dev = self.mask.device
        c2w = np.asarray(c2w, np.float32).reshape(3, 4)
        o, v = calls.camera_rays(c2w, width, height, focal_for(width, fov), dev)
        v = v.reshape(height, width, 3)
        centre = ((torch.from_numpy(c2w[:, 3]).to(dev) - self.xyzmin) / self.voxel_size)
        centre = centre.to(torch.float32).contiguous()
        coords, rgba, finish = self._bufs(height, width, dev)
        calls.aabb_intersect(coords, v, centre, finish, rgba, self.octrees, self.n)
        rounds = 0
        while not bool(finish.all()) and rounds < max_rounds:
            calls.ray_march(coords, v, self.mask, finish, self.n)
            calls.mlp_eval(self.decoder, rgba, coords, self.features, self.voxel_map, v, finish,
                           self.alpha_skip, self.t_stop)
            rounds += 1
        return rgba[..., :3] + rgba[..., 3:], rounds
