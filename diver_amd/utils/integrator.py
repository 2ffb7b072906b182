"""Drop-in for the reference's ``utils.integrator`` (utils/integrator.py).

``integrate(voxels, coord_i, coord_o)`` returns the mean of the trilinear feature field over each
segment (utils/integrator.py:62-113), computed by one libdiver kernel instead of ~30 torch ops;
gradients flow to ``voxels`` through the library's scatter-add adjoint.  The coordinates carry no
gradient (they come from the traversal ops, whose backward is None in the reference too).

``integrate_mlp`` (the implicit-grid stage, utils/integrator.py:3-60) is outside this path.
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


def integrate_mlp(voxel_mlp, N, C, coord_i, coord_o):
    raise NotImplementedError('the implicit-grid integral (integrate_mlp) is outside the '
                              'explicit-grid render path built here')
