"""Adam with the per-parameter update in one HIP pass (``diver_adam_step``).

The reference trains with ``torch.optim.Adam(self.parameters(), lr=..., weight_decay=...)``
(train.py:39-47): L2 weight decay folded into the gradient, no amsgrad.  This optimizer keeps that
update rule and torch's state names (``step``, ``exp_avg``, ``exp_avg_sq``); each dense float32
device parameter is updated by one kernel that reads the parameter, gradient and both moments once
and writes them back once, which for the 2.17 GB trainable grid is the whole cost of the step.

With ``clear_grad=True`` every gradient the step consumed is also set to zero in the same pass, so
the caller can drop its ``zero_grad()`` (a separate 2.17 GB memset) for the next backward.

A parameter group may carry ``rows`` (an int64 device tensor) and ``row_len``: the step then
updates only those rows of length ``row_len`` (``diver_adam_step_rows``).  Without weight decay
(the reference's default) an element whose gradient and moments are zero is a fixed point of the
update, so listing every row that can receive a gradient gives exactly the full step.  For the
trainable grid those are the corners of the occupied voxels
(``adjoint.FusedRenderer.trainable_rows``): about a tenth of the grid on the 256^3 scenes.
"""
from __future__ import annotations

import math

import torch

from . import _lib


class Adam(torch.optim.Optimizer):
    def __init__(self, params, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8,
                 weight_decay: float = 0.0, clear_grad: bool = False):
        if not 0.0 <= lr:
            raise ValueError(f'invalid learning rate {lr}')
        if not (0.0 <= betas[0] < 1.0 and 0.0 <= betas[1] < 1.0):
            raise ValueError(f'invalid betas {betas}')
        if not 0.0 <= eps:
            raise ValueError(f'invalid eps {eps}')
        defaults = dict(lr=lr, betas=tuple(betas), eps=eps, weight_decay=weight_decay,
                        clear_grad=clear_grad, rows=None, row_len=0)
        super().__init__(params, defaults)

    @staticmethod
    def _check(p: torch.Tensor):
        if not (p.is_cuda and p.dtype == torch.float32 and p.is_contiguous()):
            raise TypeError('diver_amd.optim.Adam updates contiguous float32 device parameters; '
                            f'got {p.dtype} on {p.device} (contiguous={p.is_contiguous()})')
        if p.grad.is_sparse or not p.grad.is_contiguous() or p.grad.dtype != torch.float32:
            raise TypeError('the gradient must be a dense contiguous float32 tensor')

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        for group in self.param_groups:
            b1, b2 = group['betas']
            for p in group['params']:
                if p.grad is None:
                    continue
                self._check(p)
                state = self.state[p]
                if not state:
                    state['step'] = 0
                    state['exp_avg'] = torch.zeros_like(p, memory_format=torch.contiguous_format)
                    state['exp_avg_sq'] = torch.zeros_like(p, memory_format=torch.contiguous_format)
                state['step'] += 1
                t = state['step']
                step_size = group['lr'] / (1.0 - b1 ** t)
                bc2_sqrt = math.sqrt(1.0 - b2 ** t)
                rows = group['rows']
                if rows is not None:
                    if rows.dtype != torch.int64 or not rows.is_cuda or not rows.is_contiguous():
                        raise TypeError('rows must be a contiguous int64 device tensor')
                    _lib.call('diver_adam_step_rows', p.data_ptr(), p.grad.data_ptr(),
                              state['exp_avg'].data_ptr(), state['exp_avg_sq'].data_ptr(),
                              p.numel(), rows.data_ptr(), rows.numel(), int(group['row_len']), b1,
                              b2, group['eps'], group['weight_decay'], step_size, bc2_sqrt,
                              int(bool(group['clear_grad'])), _lib.stream_of(p))
                    continue
                _lib.call('diver_adam_step', p.data_ptr(), p.grad.data_ptr(),
                          state['exp_avg'].data_ptr(), state['exp_avg_sq'].data_ptr(), p.numel(),
                          b1, b2, group['eps'], group['weight_decay'], step_size, bc2_sqrt,
                          int(bool(group['clear_grad'])), _lib.stream_of(p))
        return loss
