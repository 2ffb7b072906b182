"""Seeded synthetic scenes, decoders and cameras for the render-and-train path.

Everything here is a pure function of its seeds, computed with numpy on the host so that the
GPU path and the CPU checks see the same inputs.

Conventions follow the reference:
- the feature grid is ``voxels[z][y][x][c]`` with ``(N+1)^3`` corners (model/diver.py:37-39,
  utils/integrator.py:80-85) and spans ``[xyzmin, xyzmin + N*voxel_size]^3`` (model/diver.py:20-30);
- the occupancy mask is ``mask[z][y][x]`` over ``N^3`` voxels (eval_rt.py:224-226);
- cameras are pinhole, looking down ``-z`` of the camera frame with pixel centres at ``+0.5``
  and normalised directions (utils/dataset/ray_utils.py:20-27,48), ``focal = 0.5 W / tan(fov/2)``
  (utils/dataset/blender.py:28-31).

Scene recipe (sizes in world units over the grid ``[-1, 1]^3``):
- spheres: centre ~ U(-0.6, 0.6)^3, radius ~ U(0.1, 0.3); a voxel is occupied when its centre
  lies inside;
- shells: same centre/radius law, a voxel is occupied when its centre lies within
  ``thickness/2`` of the sphere surface, ``thickness = 2`` voxels;
- a corner is occupied when any of the (up to) 8 voxels sharing it is occupied; occupied
  corners get features ~ N(0, 0.1^2) (drawn in z, y, x, c order), all others are 0.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

SEED_SCENE = 0
SEED_GT = 1
SEED_CAMERAS = 2
SEED_WEIGHTS = 3
SEED_ARRIVALS = 4

FOV = 0.69
CAM_RADIUS = 4.0


@dataclasses.dataclass
class Scene:
    n: int                    # voxels per side
    channels: int
    xyzmin: float
    voxel_size: float
    mask: np.ndarray          # bool [N, N, N]  (z, y, x)
    corner_idx: np.ndarray    # int64 [K] flat index of occupied corners in the (N+1)^3 grid
    corner_feat: np.ndarray   # float32 [K, C]
    seed: int

    @property
    def n1(self) -> int:
        return self.n + 1

    @property
    def occupancy(self) -> float:
        return float(self.mask.mean())

    def dense_voxels(self) -> np.ndarray:
        """Dense ``float32 [N+1, N+1, N+1, C]`` grid (host memory)."""
        g = np.zeros((self.n1 ** 3, self.channels), np.float32)
        g[self.corner_idx] = self.corner_feat
        return g.reshape(self.n1, self.n1, self.n1, self.channels)

    def dense_voxels_torch(self, device):
        """Dense grid built directly on ``device`` (avoids a host copy of the full grid)."""
        import torch
        g = torch.zeros((self.n1 ** 3, self.channels), dtype=torch.float32, device=device)
        idx = torch.from_numpy(self.corner_idx).to(device)
        g.index_copy_(0, idx, torch.from_numpy(self.corner_feat).to(device))
        return g.view(self.n1, self.n1, self.n1, self.channels)


def _voxel_centres(n: int, xyzmin: float, vs: float) -> np.ndarray:
    return (xyzmin + (np.arange(n, dtype=np.float64) + 0.5) * vs)


def make_occupancy(n: int, n_spheres: int, n_shells: int, seed: int,
                   xyzmin: float = -1.0, grid_size: float = 2.0,
                   shell_thickness_vox: float = 2.0) -> np.ndarray:
    """Union of seeded spheres and thin shells; returns bool ``[N, N, N]`` (z, y, x)."""
    rng = np.random.default_rng(seed)
    k = n_spheres + n_shells
    centres = rng.uniform(-0.6, 0.6, size=(k, 3))
    radii = rng.uniform(0.1, 0.3, size=k)
    vs = grid_size / n
    half_t = 0.5 * shell_thickness_vox * vs
    c = _voxel_centres(n, xyzmin, vs)
    mask = np.zeros((n, n, n), dtype=bool)
    for i in range(k):
        cx, cy, cz = centres[i]
        r_out = radii[i] + (half_t if i >= n_spheres else 0.0)
        lo = np.clip(np.floor(((np.array([cz, cy, cx]) - r_out) - xyzmin) / vs).astype(int) - 1, 0, n)
        hi = np.clip(np.ceil(((np.array([cz, cy, cx]) + r_out) - xyzmin) / vs).astype(int) + 1, 0, n)
        zz = c[lo[0]:hi[0], None, None]
        yy = c[None, lo[1]:hi[1], None]
        xx = c[None, None, lo[2]:hi[2]]
        dist = np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2 + (zz - cz) ** 2)
        if i < n_spheres:
            m = dist <= radii[i]
        else:
            m = np.abs(dist - radii[i]) <= half_t
        mask[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] |= m
    return mask


def corner_occupancy(mask: np.ndarray) -> np.ndarray:
    """Corners touched by an occupied voxel: bool ``[N+1, N+1, N+1]``."""
    n = mask.shape[0]
    c = np.zeros((n + 1, n + 1, n + 1), dtype=bool)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                c[dz:dz + n, dy:dy + n, dx:dx + n] |= mask
    return c


def make_scene(n: int = 64, channels: int = 32, n_spheres: int = 6, n_shells: int = 0,
               seed: int = SEED_SCENE, feat_std: float = 0.1,
               xyzmin: float = -1.0, grid_size: float = 2.0) -> Scene:
    mask = make_occupancy(n, n_spheres, n_shells, seed, xyzmin, grid_size)
    cocc = corner_occupancy(mask)
    corner_idx = np.flatnonzero(cocc.reshape(-1)).astype(np.int64)
    rng = np.random.default_rng(seed + 1000)
    feat = (rng.standard_normal((corner_idx.size, channels), dtype=np.float32)
            * np.float32(feat_std))
    return Scene(n=n, channels=channels, xyzmin=float(xyzmin), voxel_size=float(grid_size / n),
                 mask=mask, corner_idx=corner_idx, corner_feat=feat, seed=seed)


def config1_scene() -> Scene:
    """64^3, 32 channels, 6 spheres."""
    return make_scene(64, 32, 6, 0, SEED_SCENE)


def config2_scene(channels: int = 32, n: int = 256, seed: int = SEED_SCENE) -> Scene:
    """256^3, 12 spheres + 4 shells (2 voxels thick)."""
    return make_scene(n, channels, 12, 4, seed)


# ---------------------------------------------------------------------------------------------
# cameras
# ---------------------------------------------------------------------------------------------

def focal_for(width: int, fov: float = FOV) -> float:
    return 0.5 * width / math.tan(0.5 * fov)


def look_at_c2w(position: np.ndarray, target=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0)) -> np.ndarray:
    """Camera-to-world ``[3, 4]``; the camera looks down its ``-z`` axis at ``target``."""
    pos = np.asarray(position, np.float64)
    zc = pos - np.asarray(target, np.float64)
    zc /= np.linalg.norm(zc)
    upv = np.asarray(up, np.float64)
    xc = np.cross(upv, zc)
    if np.linalg.norm(xc) < 1e-8:
        xc = np.cross(np.array([0.0, 1.0, 0.0]), zc)
    xc /= np.linalg.norm(xc)
    yc = np.cross(zc, xc)
    return np.stack([xc, yc, zc, pos], axis=1)


def hemisphere_cameras(count: int = 200, seed: int = SEED_CAMERAS,
                       radius: float = CAM_RADIUS) -> np.ndarray:
    """``count`` poses on the upper hemisphere (z up) looking at the origin: ``[count, 3, 4]``."""
    rng = np.random.default_rng(seed)
    z = rng.uniform(0.05, 0.95, size=count)
    phi = rng.uniform(0.0, 2.0 * math.pi, size=count)
    s = np.sqrt(1.0 - z * z)
    pos = radius * np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)
    return np.stack([look_at_c2w(p) for p in pos], axis=0)


def camera_rays(c2w: np.ndarray, width: int, height: int, fov: float = FOV):
    """Rays of one pinhole view: ``(o, d)`` float32 ``[H*W, 3]`` each, row-major pixels."""
    focal = focal_for(width, fov)
    i = (np.arange(width, dtype=np.float64) + 0.5)[None, :].repeat(height, 0)
    j = (np.arange(height, dtype=np.float64) + 0.5)[:, None].repeat(width, 1)
    dirs = np.stack([(i - width / 2) / focal, -(j - height / 2) / focal, -np.ones_like(i)], -1)
    d = dirs @ c2w[:, :3].T
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    o = np.broadcast_to(c2w[:, 3], d.shape)
    return (np.ascontiguousarray(o.reshape(-1, 3), np.float32),
            np.ascontiguousarray(d.reshape(-1, 3), np.float32))


def random_rays(cameras: np.ndarray, count: int, width: int, height: int, seed: int):
    """``count`` random pixels drawn uniformly over the given cameras (training batches)."""
    rng = np.random.default_rng(seed)
    cam = rng.integers(0, cameras.shape[0], size=count)
    px = rng.integers(0, width, size=count)
    py = rng.integers(0, height, size=count)
    focal = focal_for(width)
    dirs = np.stack([(px + 0.5 - width / 2) / focal, -(py + 0.5 - height / 2) / focal,
                     -np.ones(count)], -1)
    rot = cameras[cam, :, :3]
    d = np.einsum('nij,nj->ni', rot, dirs)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    o = cameras[cam, :, 3]
    return o.astype(np.float32), d.astype(np.float32)


# ---------------------------------------------------------------------------------------------
# decoders
# ---------------------------------------------------------------------------------------------

def _kaiming_linear(rng, fan_in: int, fan_out: int):
    w = rng.standard_normal((fan_out, fan_in)) * math.sqrt(2.0 / fan_in)
    bound = 1.0 / math.sqrt(fan_in)
    b = rng.uniform(-bound, bound, size=fan_out)
    return w.astype(np.float32), b.astype(np.float32)


def baseline_decoder_layers(channels: int = 32, hidden: int = 32, seed: int = SEED_WEIGHTS,
                            pe_bands: int = 4):
    """Decoder of the benchmark configs: density = softplus(f[0]); colour =
    sigmoid(L3(relu(L2(relu(L1([f[1:], PE(d)])))))) with L1: (C-1+6L+3)->H, L2: H->H, L3: H->3.

    Returns the three ``(weight [out, in], bias [out])`` pairs, Kaiming-normal weights and
    PyTorch-default uniform biases.
    """
    rng = np.random.default_rng(seed)
    pe_dim = 3 * (2 * pe_bands + 1)
    l1 = _kaiming_linear(rng, channels - 1 + pe_dim, hidden)
    l2 = _kaiming_linear(rng, hidden, hidden)
    l3 = _kaiming_linear(rng, hidden, 3)
    return [l1, l2, l3]


def reference_decoder_state(channels: int = 32, hidden: int = 32, seed: int = SEED_WEIGHTS,
                            pe_bands: int = 4) -> dict:
    """State dict of the reference's decoder MLPs (model/diver.py:45-54, model/mlps.py:21-31).

    hidden 32 -> DIVeR32 (mlp_point [32,2,33], mlp_view [32,1]);
    hidden 64 -> DIVeR64 (mlp_point [64,1,65], mlp_view [64,1]).
    Keys ``mlp1.{0,2,..}.weight/bias`` and ``mlp2.{0,2}.weight/bias``.
    """
    rng = np.random.default_rng(seed)
    pe_dim = 3 * (2 * pe_bands + 1)
    depth = 2 if hidden == 32 else 1
    out1 = hidden + 1
    sd = {}
    dims = [channels] + [hidden] * depth
    for i in range(depth):
        w, b = _kaiming_linear(rng, dims[i], dims[i + 1])
        sd[f'mlp1.{2 * i}.weight'], sd[f'mlp1.{2 * i}.bias'] = w, b
    w, b = _kaiming_linear(rng, dims[-1], out1)
    sd[f'mlp1.{2 * depth}.weight'], sd[f'mlp1.{2 * depth}.bias'] = w, b
    w, b = _kaiming_linear(rng, out1 - 1 + pe_dim, hidden)
    sd['mlp2.0.weight'], sd['mlp2.0.bias'] = w, b
    w, b = _kaiming_linear(rng, hidden, 3)
    sd['mlp2.2.weight'], sd['mlp2.2.bias'] = w, b
    return sd
