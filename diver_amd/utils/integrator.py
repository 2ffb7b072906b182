"""Drop-in for the reference's ``utils.integrator`` (utils/integrator.py).

``integrate(voxels, coord_i, coord_o)`` returns the mean of the trilinear feature field over each
segment (utils/integrator.py:62-113), computed by one libdiver kernel instead of ~30 torch ops;
gradients flow to ``voxels`` through the library's scatter-add adjoint.  The coordinates carry no
gradient (they come from the traversal ops, whose backward is None in the reference too).

``integrate_mlp`` (the implicit-grid stage, utils/integrator.py:3-60) evaluates the voxel MLP once
per distinct corner (keys from one kernel, ``torch.unique``) and integrates from that table with the
same corner weights (one kernel, with a scatter-add adjoint), so gradients reach the MLP's
parameters through autograd.
"""
from __future__ import annotations

import torch
from torch.autograd import Function

from .. import calls


class _Integrate(Function):
    @staticmethod
    def forward(ctx, voxels, coord_i, coord_o):
        coord_i = coord_i.detach().contiguous()
        coord_o = coord_o.detach().contiguous()
        ctx.save_for_backward(coord_i, coord_o)
        ctx.shape = tuple(voxels.shape)
        return calls.integrate(voxels.detach(), coord_i, coord_o)

    @staticmethod
    def backward(ctx, grad):
        coord_i, coord_o = ctx.saved_tensors
        gv = None
        if ctx.needs_input_grad[0]:
            gv = calls.integrate_backward(ctx.shape, coord_i, coord_o, grad.contiguous())
        return gv, None, None


def integrate(voxels, coord_i, coord_o):
    """voxels [N+1, N+1, N+1, C], coord_i / coord_o [K, 3] (grid units) -> [K, C]."""
    return _Integrate.apply(voxels, coord_i, coord_o)


class _IntegrateTable(Function):
    @staticmethod
    def forward(ctx, table, corner_row, coord_i, coord_o):
        ctx.save_for_backward(corner_row, coord_i, coord_o)
        ctx.rows = table.shape[0]
        return calls.integrate_table(table.detach(), corner_row, coord_i, coord_o)

    @staticmethod
    def backward(ctx, grad):
        corner_row, coord_i, coord_o = ctx.saved_tensors
        gt = None
        if ctx.needs_input_grad[0]:
            gt = calls.integrate_table_backward(ctx.rows, corner_row, coord_i, coord_o,
                                                grad.contiguous())
        return gt, None, None, None


def corner_table(N, coord_i, coord_o):
    """Distinct corners of the rows' integrals: (positions [U, 3] in (z, y, x) order mapped to
    [-1, 1) as integrator.py:30 does, corner_row [K, 8] into them)."""
    keys = calls.corner_keys(N, coord_i, coord_o)
    uniq, inv = torch.unique(keys.view(-1), return_inverse=True)
    zyx = torch.stack([uniq // (N * N), (uniq // N) % N, uniq % N], -1)
    return zyx.float() / (N * 0.5) - 1.0, inv.view(-1, 8)


def integrate_mlp(voxel_mlp, N, C, coord_i, coord_o):
    """voxel_mlp: (z, y, x) in [-1, 1) -> C features; N: corners per axis (voxel_num + 1);
    coord_i / coord_o [K, 3] (grid units, x y z) -> [K, C] (utils/integrator.py:3-60)."""
    coord_i = coord_i.detach().contiguous()
    coord_o = coord_o.detach().contiguous()
    pos, corner_row = corner_table(N, coord_i, coord_o)
    table = voxel_mlp(pos).float().contiguous()
    return _IntegrateTable.apply(table, corner_row, coord_i, coord_o)
