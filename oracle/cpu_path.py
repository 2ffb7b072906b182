"""CPU ORACLE -- test infrastructure only (bench.py times it as the CPU baseline; tests use it as
a checker).

The reference's offline render path restated on the CPU in FP32 with torch, stage by stage:
  1. masked traversal: the C restatement of masked_intersect_kernel (traverse.c), trimmed to the
     largest count like utils/ray_voxel_intersection/__init__.py:76-80;
  2. ``integrate``: utils/integrator.py:62-113 (eight gathers, A/B/C terms, einsums);
  3. decoder: nn.Linear chains (model/mlps.py:21-31, model/diver.py:159-171) or the benchmark
     decoder (softplus density, colour MLP on [f[1:], PE(d)]);
  4. scatter to [B, M] and ``render``: model/diver.py:175-178, 183-207 (cumprod, white background).
No early termination (the offline path has none).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import masked_intersect


def integrate(voxels: torch.Tensor, coord_i: torch.Tensor, coord_o: torch.Tensor) -> torch.Tensor:
    N, C = voxels.shape[-2:]
    pmin = torch.min((coord_i + 1e-4).long(), (coord_o + 1e-4).long())
    xmin, ymin, zmin = pmin.split(1, dim=-1)
    xmax, ymax, zmax = (pmin + 1).clamp_max(N - 1).split(1, dim=-1)
    F = torch.stack([voxels[zmin, ymin, xmin], voxels[zmin, ymin, xmax],
                     voxels[zmin, ymax, xmin], voxels[zmin, ymax, xmax],
                     voxels[zmax, ymin, xmin], voxels[zmax, ymin, xmax],
                     voxels[zmax, ymax, xmin], voxels[zmax, ymax, xmax]], 0).reshape(2, 2, 2, -1, C)
    P = torch.stack([coord_i - pmin, coord_o - pmin], 0).permute(0, 2, 1)
    A = P.sum(0).prod(0) / 12 + P.prod(1).sum(0) / 6
    X = F[1] - F[0]
    X = X[0] - X[1]
    X = X[0] - X[1]
    feature = A[:, None] * X
    P_roll = P[:, [1, 2, 0]]
    B = (P * P_roll).sum(0) / 3 + (P * P_roll[[1, 0]]).sum(0) / 6
    Y = (F[0, 0, 0, None] - F[[0, 0, 1], [1, 1, 0], 0] - F[[0, 1, 0], 0, [1, 0, 1]]
         + F[[0, 1, 1], [1, 1, 0], [1, 0, 1]])
    feature = feature + torch.einsum('dk,dkc->kc', B, Y)
    Cc = P.sum(0) / 2
    Z = F[[0, 0, 1], [0, 1, 0], [1, 0, 0]] - F[0, 0, 0, None]
    feature = feature + torch.einsum('dk,dkc->kc', Cc, Z)
    return feature + F[0, 0, 0]


def positional_encoding(x: torch.Tensor, L: int = 4) -> torch.Tensor:
    out = [x]
    for l in range(L):
        out.append(torch.sin((2 ** l * math.pi) * x))
        out.append(torch.cos((2 ** l * math.pi) * x))
    return torch.cat(out, -1)


def baseline_decode(layers):
    ts = [(torch.from_numpy(w), torch.from_numpy(b)) for w, b in layers]

    def decode(feature, d):
        x = torch.cat([feature[:, 1:], positional_encoding(d)], -1)
        for k, (w, b) in enumerate(ts):
            x = torch.nn.functional.linear(x, w, b)
            if k + 1 < len(ts):
                x = torch.relu(x)
        return feature[:, 0], torch.sigmoid(x)
    return decode, 1


def render(o: np.ndarray, d: np.ndarray, voxels: torch.Tensor, mask: np.ndarray, n: int,
           xyzmin: float, voxel_size: float, decode, density_act: int, white_back: bool = True):
    """FP32 CPU render of rays (o, d) [R, 3]; returns rgb [R, 3] (torch)."""
    xmin = np.float32(xyzmin)
    vs = np.float32(voxel_size)
    isect, num, tns = masked_intersect(o, d, xmin, xmin + n * vs, n, vs, mask, 1.0, max_n=3 * n)
    k = int(num.max(initial=0))
    coord = torch.from_numpy(isect[:, :k])
    ts = torch.from_numpy(tns[:, :k])
    hit = ts >= 0
    B = o.shape[0]
    if k == 0:
        return torch.ones(B, 3)
    rows = coord[hit]
    feature = integrate(voxels, rows[:, :3], rows[:, 3:])
    dd = torch.from_numpy(d)[torch.where(hit)[0]]
    s_, c_ = decode(feature, dd)
    sigma = torch.zeros(B, k)
    sigma[hit] = s_
    color = torch.zeros(B, k, 3)
    color[hit] = c_
    s = torch.relu(sigma) if density_act == 0 else torch.nn.functional.softplus(sigma)
    alpha = 1 - torch.exp(-s * hit)
    shifted = torch.nn.functional.pad(1 - alpha[None, :, :-1], (1, 0), value=1)[0]
    weight = alpha * torch.cumprod(shifted, -1)
    rgb = (weight[:, :, None] * color).sum(1)
    if white_back:
        rgb = rgb + (1 - weight.sum(1, keepdim=True))
    return rgb
