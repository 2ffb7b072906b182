"""The reference's ``model.diver.DIVeR`` calling contract (model/diver.py:17-207) on libdiver.

Same hyper-parameter names (voxel_num, voxel_dim, grid_size, mask_scale, white_back, fine,
mlp_point, mlp_view, dir_encode), same parameter/buffer names (``voxels``, ``voxel_mask``,
``mlp1.*``, ``mlp2.*``), and the same method signatures and return shapes for
``extract_features``, ``decode``, ``forward`` and ``render``.  Traversal, the segment integral and
compositing run as libdiver kernels; the decoder MLPs stay torch ``nn.Linear`` layers as in the
reference.  With ``hparams.implicit`` the corner features come from the implicit voxel MLP
(``mlp_voxel``, model/diver.py:32-34) through ``integrate_mlp`` until ``init_voxels`` bakes them
into an explicit grid (model/diver.py:56-74).
"""
from __future__ import annotations

import torch
import torch.nn as nn
from torch.autograd import Function

from .. import calls
from ..utils.integrator import integrate, integrate_mlp
from ..utils.ray_voxel_intersection import masked_intersect, ray_voxel_intersect
from .mlps import ImplicitMLP, PositionalEncoding, mlp


class _Composite(Function):
    @staticmethod
    def forward(ctx, sigma, color, mask, white_back):
        sigma, color = sigma.contiguous(), color.contiguous()
        mask = mask.contiguous()
        rgb, weight = calls.composite(sigma, color, mask, white_back)
        ctx.save_for_backward(sigma, color, mask)
        ctx.white_back = white_back
        return rgb, weight

    @staticmethod
    def backward(ctx, g_rgb, g_weight):
        sigma, color, mask = ctx.saved_tensors
        if g_rgb is None:
            g_rgb = torch.zeros(sigma.shape[0], 3, device=sigma.device)
        gs, gc = calls.composite_backward(sigma, color, mask, ctx.white_back, g_rgb.contiguous(),
                                          None if g_weight is None else g_weight.contiguous())
        return gs, gc, None, None


def composite(sigma, color, mask, white_back=True):
    return _Composite.apply(sigma, color, mask, bool(white_back))


class DIVeR(nn.Module):
    def __init__(self, hparams):
        super().__init__()
        self.voxel_num = hparams.voxel_num
        self.voxel_dim = hparams.voxel_dim
        self.grid_size = hparams.grid_size
        self.voxel_size = self.grid_size / self.voxel_num
        self.mask_scale = hparams.mask_scale
        self.white_back = hparams.white_back
        extent = self.voxel_num * self.voxel_size
        self.xyzmin = -extent * 0.5
        self.xyzmax = extent * 0.5
        if getattr(hparams, 'implicit', False):
            self.mlp_voxel = ImplicitMLP(hparams.implicit_network_depth, hparams.implicit_channels,
                                         hparams.implicit_skips, self.voxel_dim,
                                         hparams.implicit_point_encode)
        else:
            voxels = torch.randn(*[self.voxel_num + 1] * 3, self.voxel_dim) * 1e-2
            self.register_parameter('voxels', nn.Parameter(voxels, requires_grad=True))
        if getattr(hparams, 'fine', 0) != 0:
            m = int(self.voxel_num * self.mask_scale)
            self.register_parameter('voxel_mask', nn.Parameter(
                torch.zeros(m, m, m, dtype=torch.bool), requires_grad=False))
        mlp_dim, mlp_depth, mlp_out = hparams.mlp_point
        self.mlp1 = mlp(self.voxel_dim, [mlp_dim] * mlp_depth, mlp_out)
        self.view_enc = PositionalEncoding(hparams.dir_encode)
        mlp_dim, mlp_depth = hparams.mlp_view
        view_dim = hparams.dir_encode * 2 * 3 + 3
        self.mlp2 = mlp(view_dim + mlp_out - 1, [mlp_dim] * mlp_depth, 3)

    def init_voxels(self, evaluate=True):
        """Bake the explicit grid [N+1]^3 x C: the implicit MLP at every corner (positions
        (z, y, x) / ((N+1)/2) - 1, several z-slices per evaluation so each GEMM has ~2^18 rows),
        or zeros (model/diver.py:56-74)."""
        n = self.voxel_num + 1
        ref = self.mlp_voxel if hasattr(self, 'mlp_voxel') else self.voxels
        device = next(ref.parameters()).device if isinstance(ref, nn.Module) else ref.device
        if evaluate:
            step = max(1, (1 << 18) // (n * n))
            voxels = torch.empty(n, n, n, self.voxel_dim, device=device)
            with torch.no_grad():
                for z0 in range(0, n, step):
                    zz, y, x = torch.meshgrid(torch.arange(z0, min(n, z0 + step), device=device),
                                              torch.arange(n, device=device),
                                              torch.arange(n, device=device), indexing='ij')
                    p = torch.stack([zz, y, x], -1).float() / (n * 0.5) - 1.0
                    voxels[z0:z0 + zz.shape[0]] = self.mlp_voxel(p.reshape(-1, 3)).reshape(
                        zz.shape[0], n, n, self.voxel_dim)
        else:
            voxels = torch.zeros(n, n, n, self.voxel_dim, device=device)
        self.register_parameter('voxels', nn.Parameter(voxels, requires_grad=True))

    def extract_features(self, os, ds):
        """os, ds [B, 3] -> (mask [B, M], features [K, C] in mask order, ts [B, M])."""
        if hasattr(self, 'voxel_mask'):
            coord, mask, ts = masked_intersect(
                os.contiguous(), ds.contiguous(), self.xyzmin, self.xyzmax, int(self.voxel_num),
                self.voxel_size, self.voxel_mask.contiguous(), self.mask_scale)
            coord = coord[mask]
            coord_in, coord_out = coord[:, :3], coord[:, 3:]
        else:
            coord, mask, ts = ray_voxel_intersect(
                os.contiguous(), ds.contiguous(), self.xyzmin, self.xyzmax, int(self.voxel_num),
                self.voxel_size)
            ts = ts[:, :-1]
            coord = coord.clamp_min(0)
            mask = mask[:, :-1] & mask[:, 1:]
            coord_in = coord[:, :-1][mask]
            coord_out = coord[:, 1:][mask]
        if not mask.any():
            return mask, None, None
        if hasattr(self, 'voxels'):
            features = integrate(self.voxels, coord_in.contiguous(), coord_out.contiguous())
        else:
            features = integrate_mlp(self.mlp_voxel, self.voxel_num + 1, self.voxel_dim, coord_in,
                                     coord_out)
        return mask, features, ts

    def _decode_features(self, feature, ds, mask):
        B, M = mask.shape
        x = self.mlp1(feature)
        sigma_, x = x[:, 0], x[:, 1:]
        x = torch.cat([x, self.view_enc(ds[torch.where(mask)[0]])], dim=-1)
        color_ = torch.sigmoid(self.mlp2(x))
        sigma = torch.zeros(B, M, device=mask.device)
        sigma[mask] = sigma_
        color = torch.zeros(B, M, 3, device=mask.device)
        color[mask] = color_
        return color, sigma

    def decode(self, coord_in, coord_out, ds, mask):
        if hasattr(self, 'voxels'):
            feature = integrate(self.voxels, coord_in.contiguous(), coord_out.contiguous())
        else:
            feature = integrate_mlp(self.mlp_voxel, self.voxel_num + 1, self.voxel_dim,
                                    coord_in, coord_out)
        return self._decode_features(feature, ds, mask)

    def forward(self, os, ds):
        """-> (color [B, M, 3], sigma [B, M], mask [B, M], ts [B, M])."""
        mask, feature, ts = self.extract_features(os, ds)
        if feature is None:
            return None, None, mask, None
        color, sigma = self._decode_features(feature, ds, mask)
        return color, sigma, mask, ts

    def render(self, color, sigma, mask):
        """-> (rgb [B, 3], weight [B, M])."""
        return composite(sigma, color, mask, self.white_back)
