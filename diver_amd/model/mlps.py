"""The reference's MLP building blocks (model/mlps.py): positional encoding, the Linear/ReLU chain
constructor of the decoders and the implicit voxel-feature MLP."""
from __future__ import annotations

import math

import torch
import torch.nn as nn


class PositionalEncoding(nn.Module):
    """[x, sin(2^0 pi x), cos(2^0 pi x), ..., sin(2^(L-1) pi x), cos(2^(L-1) pi x)]."""

    def __init__(self, L):
        super().__init__()
        self.L = L

    def forward(self, inputs):
        parts = [inputs]
        for l in range(self.L):
            f = (2 ** l) * math.pi
            parts.append(torch.sin(f * inputs))
            parts.append(torch.cos(f * inputs))
        return torch.cat(parts, -1)


def mlp(dim_in, dims, dim_out):
    """Linear layers dim_in -> dims[0] -> ... -> dims[-1] -> dim_out with ReLU between them."""
    widths = [dim_in] + list(dims)
    layers = []
    for a, b in zip(widths[:-1], widths[1:]):
        layers.append(nn.Linear(a, b))
        layers.append(nn.ReLU(inplace=True))
    layers.append(nn.Linear(widths[-1], dim_out))
    return nn.Sequential(*layers)


class ImplicitMLP(nn.Module):
    """The implicit voxel-feature MLP (model/mlps.py:34-81): PE(x) with dim_enc bands, D Linear +
    ReLU layers of width C with PE(x) concatenated back in before the layers listed in S, then a
    Linear to dim_out.  Same parameter names (pts_linears.*, feature_linear.*)."""

    def __init__(self, D, C, S, dim_out, dim_enc):
        super().__init__()
        self.input_ch = dim_enc * 2 * 3 + 3
        self.D = D
        self.C = C
        self.skips = list(S)
        self.dim_out = dim_out
        self.point_encode = PositionalEncoding(dim_enc)
        self.pts_linears = nn.ModuleList(
            [nn.Linear(self.input_ch, C)] +
            [nn.Linear(C + self.input_ch, C) if i in self.skips else nn.Linear(C, C)
             for i in range(1, D)])
        self.feature_linear = nn.Linear(C, dim_out)

    def forward(self, x):
        points = self.point_encode(x)
        p = points
        for i, layer in enumerate(self.pts_linears):
            if i in self.skips:
                p = torch.cat([points, p], 1)
            p = torch.relu(layer(p))
        return self.feature_linear(p)
