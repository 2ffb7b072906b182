"""CPU ORACLE -- test infrastructure only.

Float64 restatement of the implicit voxel grid: the voxel MLP of model/mlps.py:34-81 (positional
encoding with `bands` bands, Linear + ReLU layers with the encoding concatenated in front of the
input of the layers listed in `skips`, a final Linear) and integrate_mlp of
utils/integrator.py:3-60 (the integral of a voxel whose 8 corner features are the MLP at the
corners' (z, y, x) / (N / 2) - 1, upper corner clamped to N - 1).
"""
from __future__ import annotations

import numpy as np

from . import fp64


def implicit_mlp(sd: dict, skips, bands: int, x: np.ndarray) -> np.ndarray:
    """sd: pts_linears.{i}.weight/bias, feature_linear.weight/bias (numpy); x [P, 3] -> [P, C]."""
    points = fp64.positional_encoding(np.asarray(x, np.float64), bands)
    p = points
    i = 0
    while f'pts_linears.{i}.weight' in sd:
        if i in skips:
            p = np.concatenate([points, p], 1)
        p = np.maximum(p @ np.asarray(sd[f'pts_linears.{i}.weight'], np.float64).T
                       + np.asarray(sd[f'pts_linears.{i}.bias'], np.float64), 0.0)
        i += 1
    return p @ np.asarray(sd['feature_linear.weight'], np.float64).T \
        + np.asarray(sd['feature_linear.bias'], np.float64)


def corner_positions(N: int, zyx: np.ndarray) -> np.ndarray:
    """Integer (z, y, x) corners -> the MLP's input, float32 as integrator.py:30 computes it."""
    return (np.asarray(zyx, np.float32) / np.float32(N * 0.5) - np.float32(1.0)).astype(np.float64)


def integrate_mlp(sd: dict, skips, bands: int, N: int, coord_in, coord_out) -> np.ndarray:
    """[K, C]: per row, the corner weights of fp64.corner_weights against the MLP at the row's 8
    corners (corner j = (j&1, (j>>1)&1, j>>2) in x, y, z)."""
    w, pmin = fp64.corner_weights(coord_in, coord_out)
    out = 0.0
    for j in range(8):
        x = np.minimum(pmin[:, 0] + (j & 1), N - 1)
        y = np.minimum(pmin[:, 1] + ((j >> 1) & 1), N - 1)
        z = np.minimum(pmin[:, 2] + (j >> 2), N - 1)
        F = implicit_mlp(sd, skips, bands, corner_positions(N, np.stack([z, y, x], -1)))
        out = out + w[:, j:j + 1] * F
    return out


def dense_grid(sd: dict, skips, bands: int, N: int) -> np.ndarray:
    """The MLP at every corner: [N, N, N, C] indexed [z][y][x] (model/diver.py:65-72)."""
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing='ij')
    F = implicit_mlp(sd, skips, bands, corner_positions(N, np.stack([z, y, x], -1).reshape(-1, 3)))
    return F.reshape(N, N, N, -1)
