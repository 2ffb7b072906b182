"""The reference's decoder building blocks (model/mlps.py:6-31): positional encoding of the view
direction and the Linear/ReLU chain constructor."""
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
