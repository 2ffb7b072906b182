"""Differentiable fused render: a hand-derived adjoint behind ``torch.autograd.Function``.

Forward: one fused kernel per batch of rays (FP32), keeping only (rgb, T) and the interval
count per ray -- no samples are stored for the backward.

With ``sparsity=True`` the forward also returns the reference's sparsity term
mean(log(1 + sigma^2 / 0.5)) over the batch's [B, M] raw densities (M = the largest interval count
of the batch, train.py:84-85,110; the zero padding adds nothing): the kernel sums
log(1 + 2 s^2) per ray, and its adjoint adds g / (B M) * 4 s / (1 + 2 s^2) to each interval's ds.

Backward (for decoders without pre-layers, e.g. the benchmark decoder):
  1. each composited interval's integral, density and transmittance, from the entry and exit the
     forward recorded (``diver_render_rays_segments`` / ``diver_render_rows_segments``) or, for a
     batch too large to record, by re-walking every ray (``diver_render_rows``);
  2. the decoder over the rows on the matrix cores (``diver_adjoint_decoder_forward``: relu(y1),
     relu(y2), colour c and weight w = T a per row), then the compositing adjoint, per interval k
     of a ray with C = sum_i T_i a_i c_i + bg T_n:
        dC/dc_k     = T_k a_k
        dC/dsigma_k = T_{k+1} c_k - (sum_{i>k} T_i a_i c_i + bg T_n)
     (the per-ray suffix sums in FP64, ``diver_ray_suffix``);
  3. the decoder's backward pass on the matrix cores (``diver_adjoint_decoder_backward``):
        do = dc * c (1 - c);  dr2 = (do W_o) * [y2 > 0];  dr1 = (dr2 W_2) * [y1 > 0];
        dF = dr1 W_c + ds w_s
     and the weight gradients as reductions over the rows in one hand-written pair of launches
     (``diver_adjoint_weight_grads``, a fixed summation order):
        dW_o = do^T r2;  dW_2 = dr2^T r1;  dW_c = dr1^T F;  dW_v = dr1^T PE(d)
  4. grid adjoint: ``diver_integrate_backward`` scatters dF through the 8 corner weights.
The reference obtains the same gradient from torch autograd over stored K x C / B x M
activations (model/diver.py:141-207, utils/integrator.py:62-113, train.py:102-113).
"""
from __future__ import annotations

import math

import numpy as np
import torch
import torch.nn as nn
from torch.autograd import Function

from . import calls
from .decoder import CanonicalDecoder, Decoder, PE_DIM


def positional_encoding(d: torch.Tensor, bands: int = 4) -> torch.Tensor:
    parts = [d]
    for l in range(bands):
        parts.append(torch.sin((2 ** l * math.pi) * d))
        parts.append(torch.cos((2 ** l * math.pi) * d))
    return torch.cat(parts, -1)


def split_canonical(p: torch.Tensor, C: int, H: int):
    """Views of the canonical parameter vector of a decoder without pre-layers, one post layer."""
    sizes = [C, 1, H * C, H * PE_DIM, H, H * H, H, 3 * H, 3]
    parts = torch.split(p, sizes)
    ws, bs, wc, wv, bc, w2, b2, wo, bo = parts
    return (ws, bs, wc.view(H, C), wv.view(H, PE_DIM), bc, w2.view(H, H), b2, wo.view(3, H), bo)


# Device memory the forward may spend on recorded segments (24 bytes per ray and step bound); a
# larger batch re-walks its rays in the backward instead.
SEGMENT_BYTES = 2 << 30


class _FusedRender(Function):
    @staticmethod
    def forward(ctx, voxels, params, grid, decoder, o, d, t_stop, white_back, traversal,
                accumulate, sparsity):
        # the density row (w_s = e_0, b_s = 0: softplus(F[0])) is constant, so only the colour
        # part is pushed and the kernel keeps its channel-0 density path
        calls.decoder_update(decoder, params, first=decoder.channels + 1)
        opts = calls.render_opts('fp32', traversal, t_stop, 0.0)
        # With a gradient to come and no early termination (the reference's training), the
        # offline walk records each interval's entry and exit
        # ([R, cap, 6], cap = the walk's step bound 3N + 8, so every interval fits), and the
        # backward builds its rows from them instead of walking every ray again.
        cap = 3 * grid.n + 8
        segments = None
        if (any(ctx.needs_input_grad[:2]) and traversal == 'offline' and t_stop <= 0.0
                and getattr(grid, 'blocks', None) is not None
                and getattr(grid, 'bricks', None) is not None
                and o.shape[0] * cap * 24 <= SEGMENT_BYTES):
            rgba, count, sreg, segments = calls.render_rays_segments(grid, decoder, o, d, opts, cap,
                                                                     sparsity=sparsity)
        elif sparsity:
            rgba, count, sreg = calls.render_rays_sparsity(grid, decoder, o, d, opts)
        else:
            rgba, count = calls.render_rays(grid, decoder, o, d, opts)
        if sparsity:
            # B x M with M the batch's largest interval count, kept on the device (no sync)
            bm = (count.max().to(torch.float32) * o.shape[0]).clamp_min(1.0)
        else:
            bm = torch.ones((), device=o.device)
        ctx.save_for_backward(params, o, d, rgba, count, bm, segments)
        ctx.grid, ctx.decoder, ctx.opts = grid, decoder, opts
        ctx.white_back = white_back
        ctx.vshape = tuple(voxels.shape)
        ctx.accumulate = voxels if accumulate else None
        ctx.sparsity = sparsity
        bg = rgba[:, 3:] if white_back else torch.zeros_like(rgba[:, 3:])
        rgb = rgba[:, :3] + bg
        if sparsity:
            return rgb, sreg.to(torch.float64).sum().to(torch.float32) / bm
        return rgb

    @staticmethod
    def backward(ctx, g, g_reg=None):
        params, o, d, rgba, count, bm, segments = ctx.saved_tensors
        if g is None:
            g = torch.zeros_like(rgba[:, :3])
        grid, decoder = ctx.grid, ctx.decoder
        C, H = decoder.channels, decoder.hidden
        R = o.shape[0]
        cnt64 = count.to(torch.int64)
        ends = torch.cumsum(cnt64, 0)
        row_off = ends - cnt64
        # queued ahead of the one host sync (the row count): the decoder's packed operands
        calls.decoder_update(decoder, params, first=decoder.channels + 1)
        K = int(ends[-1].item())
        dev = o.device
        if K == 0:
            gv0 = None if ctx.accumulate is not None else torch.zeros(ctx.vshape, device=dev)
            return gv0, torch.zeros_like(params), *([None] * 9)
        if segments is not None:
            rows, ray = calls.render_rows_segments(grid, decoder, ctx.opts, count, row_off, K,
                                                   segments, with_ray=True)
        else:
            rows = calls.render_rows(grid, decoder, o, d, ctx.opts, count, row_off, K)
            ray = torch.repeat_interleave(torch.arange(R, device=dev), cnt64, output_size=K)
        pin, pout = rows[:, 0:3], rows[:, 3:6]
        # the decoder over the rows on the matrix cores: relu(y1), relu(y2), colour, weight
        r1, r2, c, w = calls.adjoint_decoder_forward(decoder, rows, ray, d)
        # suffix sums within each ray: sum over the later rows of w c, plus the background
        # (one lane per ray, FP64 sums)
        suffix = calls.ray_suffix(row_off, count, w, c,
                                  rgba[:, 3].contiguous() if ctx.white_back else None)
        # sparsity term: d/ds of g_reg * sum log(1 + s^2 / 0.5) / (B M), applied in the kernel
        sscale = (g_reg / bm) if (ctx.sparsity and g_reg is not None) else None
        dF, do, ds, dr1, dr2 = calls.adjoint_decoder_backward(decoder, rows, ray, K, g.contiguous(),
                                                              suffix, r1, r2, c, w, sscale)
        # weight gradients: reductions over every row, straight into the canonical order
        # (dw_s, db_s, dW_c, dW_v, db_c, dW_2, db_2, dW_o, db_o)
        gp = calls.adjoint_weight_grads(decoder, K, rows, ray, d, r1, r2, do, ds, dr1, dr2)
        if ctx.accumulate is not None:
            # scatter straight into the parameter's gradient: no dense temporary of the grid
            leaf = ctx.accumulate
            if leaf.grad is None:
                leaf.grad = torch.zeros_like(leaf)
            calls.integrate_backward(ctx.vshape, pin, pout, dF, leaf.grad)
            gv = None
        else:
            gv = calls.integrate_backward(ctx.vshape, pin, pout, dF)
        return gv, gp, None, None, None, None, None, None, None, None, None


def corner_rows(mask: torch.Tensor) -> torch.Tensor:
    """int64 [K], sorted: the corners z (N+1)^2 + y (N+1) + x of the occupied voxels of an [N]^3
    mask, i.e. the rows of the [(N+1)^3, C] grid that the integrals of occupied voxels read."""
    m = mask.bool()
    n = m.shape[0]
    c = torch.zeros((n + 1,) * 3, dtype=torch.bool, device=m.device)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                c[dz:dz + n, dy:dy + n, dx:dx + n] |= m
    return torch.nonzero(c.reshape(-1)).reshape(-1).contiguous()


class BenchmarkDecoder(nn.Module):
    """Trainable benchmark decoder: density softplus(f[0]); colour MLP on [f[1:], PE(d)]
    (Linear (C-1+27)->H, ReLU, Linear H->H, ReLU, Linear H->3, sigmoid)."""

    def __init__(self, layers, channels: int):
        super().__init__()
        (w1, b1), (w2, b2), (w3, b3) = layers
        t = lambda x: nn.Parameter(torch.as_tensor(np.asarray(x, np.float32)).clone())
        self.w1, self.b1, self.w2, self.b2, self.w3, self.b3 = (t(w1), t(b1), t(w2), t(b2),
                                                                 t(w3), t(b3))
        self.channels = channels
        self.hidden = w1.shape[0]

    def canonical(self) -> torch.Tensor:
        C, H = self.channels, self.hidden
        dev = self.w1.device
        ws = torch.zeros(C, device=dev)
        ws[0] = 1.0
        wc = torch.cat([torch.zeros(H, 1, device=dev), self.w1[:, :C - 1]], 1)
        wv = self.w1[:, C - 1:]
        return torch.cat([ws, torch.zeros(1, device=dev), wc.reshape(-1), wv.reshape(-1), self.b1,
                          self.w2.reshape(-1), self.b2, self.w3.reshape(-1), self.b3])

    def layers(self):
        f = lambda x: x.detach().cpu().numpy()
        return [(f(self.w1), f(self.b1)), (f(self.w2), f(self.b2)), (f(self.w3), f(self.b3))]


class FusedRenderer(nn.Module):
    """Trainable grid + benchmark decoder rendered by the fused kernel with the hand-written
    adjoint.  ``voxels`` [(N+1)^3, C] and the decoder are parameters; the occupancy mask (and its
    acceleration structures) is fixed."""

    def __init__(self, grid, decoder_module: BenchmarkDecoder, traversal: str = 'offline',
                 accumulate_grad: bool = True):
        super().__init__()
        self.voxels = nn.Parameter(grid.voxels)
        grid.voxels = self.voxels.data  # the kernels read the parameter's storage
        self.grid = grid
        self.decoder = decoder_module
        canon = CanonicalDecoder(decoder_module.channels, decoder_module.hidden, 0, False, 1, 1,
                                 decoder_module.canonical().detach().cpu().numpy())
        self.handle = Decoder(canon)
        self.traversal = traversal
        self.accumulate_grad = accumulate_grad

    def trainable_rows(self) -> torch.Tensor:
        """The rows of ``voxels`` a walk can send a gradient to (``corner_rows`` of the mask, which
        is fixed); the ``rows`` of an ``optim.Adam`` group for the grid."""
        return corner_rows(self.grid.mask)

    def forward(self, o, d, t_stop: float = 1e-4, white_back: bool = True,
                sparsity: bool = False):
        """rgb [R, 3]; with ``sparsity`` also the reference's sigma_loss (train.py:84-85), the
        mean of log(1 + sigma^2 / 0.5) over the batch's [R, max count] raw densities, so a
        training step is ``mse(rgb, target) + l_s * sigma_loss`` as in train.py:109-113."""
        self.grid.voxels = self.voxels.data
        return _FusedRender.apply(self.voxels, self.decoder.canonical(), self.grid, self.handle,
                                  o.contiguous(), d.contiguous(), t_stop, white_back,
                                  self.traversal, self.accumulate_grad, bool(sparsity))
