"""Implicit-grid benchmark: integrate_mlp (utils/integrator.py:3-60) and init_voxels
(model/diver.py:56-74) with the reference's implicit MLP configuration (8 layers of 512 with a
skip at layer 4, 10 encoding bands, 32 feature channels) on a 128^3 grid (129^3 corners).

  integrate_mlp, libdiver: corner keys (kernel) -> torch.unique on int64 keys -> MLP (FP32 GEMMs)
      -> table integral (kernel; scatter-add adjoint in the backward).
  integrate_mlp, reference structure: the same steps as integrator.py in torch -- corners.unique
      over (z, y, x) rows, the MLP, the gather and the einsum-form integral.
  init_voxels: DIVeR.init_voxels, the MLP at every corner in batches of whole z-slices.

The MLP costs (63 + 3 x 512 + 512 x 512 x 7 + 63 x 512 + 512 x 32) multiply-adds per point.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from . import calls
from . import synthetic as S
from .model.mlps import ImplicitMLP
from .utils.integrator import integrate_mlp


def _reference_structure(mlp, N, ci, co):
    pmin = torch.min((ci + 1e-4).long(), (co + 1e-4).long())
    xmin, ymin, zmin = pmin.split(1, -1)
    xmax, ymax, zmax = (pmin + 1).clamp_max(N - 1).split(1, -1)
    corners = torch.stack([zmin, ymin, xmin, zmin, ymin, xmax, zmin, ymax, xmin, zmin, ymax, xmax,
                           zmax, ymin, xmin, zmax, ymin, xmax, zmax, ymax, xmin, zmax, ymax, xmax],
                          -1).reshape(-1, 3)
    cu, rev = corners.unique(dim=0, return_inverse=True)
    F = mlp(cu.float() / (N * 0.5) - 1.0)[rev].reshape(-1, 2, 2, 2, F_C).permute(1, 2, 3, 0, 4)
    P = torch.stack([ci - pmin, co - pmin], 0).permute(0, 2, 1)
    A = P.sum(0).prod(0) / 12 + P.prod(1).sum(0) / 6
    X = F[1] - F[0]
    X = X[0] - X[1]
    X = X[0] - X[1]
    feat = A[:, None] * X
    P_roll = P[:, [1, 2, 0]]
    B = (P * P_roll).sum(0) / 3 + (P * P_roll[[1, 0]]).sum(0) / 6
    Y = F[0, 0, 0, None] - F[[0, 0, 1], [1, 1, 0], 0] - F[[0, 1, 0], 0, [1, 0, 1]] \
        + F[[0, 1, 1], [1, 1, 0], [1, 0, 1]]
    feat = feat + torch.einsum('dk,dkc->kc', B, Y)
    Cc = P.sum(0) / 2
    Z = F[[0, 0, 1], [0, 1, 0], [1, 0, 0]] - F[0, 0, 0, None]
    feat = feat + torch.einsum('dk,dkc->kc', Cc, Z)
    return feat + F[0, 0, 0]


F_C = 32


def _timed(fn, reps):
    fn()
    torch.cuda.synchronize()
    best = float('inf')
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--voxels', type=int, default=128)
    ap.add_argument('--rays', type=int, default=1024)
    ap.add_argument('--reps', type=int, default=3)
    args = ap.parse_args(argv)
    dev = torch.device('cuda', 0)
    n = args.voxels
    N = n + 1
    torch.manual_seed(0)
    mlp = ImplicitMLP(8, 512, [4], F_C, 10).to(dev)
    cams = S.hemisphere_cameras()
    o, d = S.random_rays(cams, args.rays, 800, 800, seed=9)
    o, d = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    vs = 2.0 / n
    isect, num, _ = calls.ray_voxel_intersect_raw(o, d, -1.0, n, vs, 3 * n)
    m = int(num.max().item())
    valid = torch.arange(m, device=dev)[None, :] < num[:, None]
    pair = valid[:, :-1] & valid[:, 1:]
    ci = isect[:, :m - 1][pair].contiguous()
    co = isect[:, 1:m][pair].contiguous()
    K = ci.shape[0]
    keys = calls.corner_keys(N, ci, co)
    U = int(torch.unique(keys.view(-1)).numel())
    macs = 63 * 512 + (512 * 512) * 7 + 63 * 512 + 512 * F_C
    res = {'workload': f'implicit MLP 8x512 (skip 4, 10 bands) -> {F_C} channels, {n}^3 grid, '
                       f'{args.rays} random training rays', 'intervals': K, 'distinct_corners': U,
           'macs_per_point': macs}

    def fwd():
        with torch.no_grad():
            integrate_mlp(mlp, N, F_C, ci, co)

    def fwd_bwd():
        mlp.zero_grad(set_to_none=True)
        integrate_mlp(mlp, N, F_C, ci, co).sum().backward()

    def ref_fwd():
        with torch.no_grad():
            _reference_structure(mlp, N, ci, co)

    t_f, t_fb, t_r = _timed(fwd, args.reps), _timed(fwd_bwd, args.reps), _timed(ref_fwd, args.reps)
    with torch.no_grad():
        diff = (integrate_mlp(mlp, N, F_C, ci, co) - _reference_structure(mlp, N, ci, co)).abs().max()
    res['integrate_mlp'] = {'fwd_ms': round(t_f * 1e3, 2), 'fwd_bwd_ms': round(t_fb * 1e3, 2),
                            'mlp_tflops_fwd': round(2 * macs * U / t_f / 1e12, 1)}
    res['reference_structure'] = {'fwd_ms': round(t_r * 1e3, 2), 'speedup_fwd': round(t_r / t_f, 2),
                                  'max_abs_diff': float(diff)}

    from types import SimpleNamespace
    from .model.diver import DIVeR
    hp = SimpleNamespace(voxel_num=n, voxel_dim=F_C, grid_size=2.0, mask_scale=1.0,
                         white_back=True, fine=0, mlp_point=[32, 1, 33], mlp_view=[32, 1],
                         dir_encode=4, implicit=True, implicit_network_depth=8,
                         implicit_channels=512, implicit_skips=[4], implicit_point_encode=10)
    model = DIVeR(hp).to(dev)
    model.mlp_voxel.load_state_dict(mlp.state_dict())

    def init():
        model.init_voxels(True)

    t_i = _timed(init, 1)
    res['init_voxels'] = {'corners': N ** 3, 'ms': round(t_i * 1e3, 1),
                          'tflops': round(2 * macs * N ** 3 / t_i / 1e12, 1),
                          'frac_of_fp32_peak_157': round(2 * macs * N ** 3 / t_i / 157.3e12, 3)}
    print(json.dumps(res))


if __name__ == '__main__':
    main()
