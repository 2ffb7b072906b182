"""Occupancy-extraction benchmark: the alpha map of prune.py:83-137 over the rays of 800x800
views of the 256^3 x 32 synthetic scene with the reference DIVeR32 decoder (FP32, relu density).

  fused: one diver_alpha_map launch per batch (walk, integral, density, weights, max).
  plain: prune.py's loop on the drop-in ops -- masked_intersect, integrate, the decoder's mlp1 in
         torch, the cumprod weights of DIVeR.render and a torch scatter max -- in batches of
         20,480 rays, as the reference chunks its traversal.

Prints one JSON line.  Algorithmic bytes per batch: 24 B per ray (o, d) + 8 x C x 4 B per
evaluated interval (corner features) + 4 B per interval (map update).
"""
from __future__ import annotations

import argparse
import json
import time

import torch

from . import calls, prune
from . import synthetic as S
from .decoder import Decoder, from_state_dict
from .renderer import SceneGrid, camera_rays


def _plain_alpha_map(grid, sd, o, d, chunk=20480):
    n = grid.n
    dev = o.device
    amap = torch.zeros(n ** 3, dtype=torch.float32, device=dev)
    W = {k: torch.as_tensor(v, device=dev) for k, v in sd.items()}
    names = sorted({k.rsplit('.', 1)[0] for k in W if k.startswith('mlp1.')},
                   key=lambda s: int(s.split('.')[1]))
    vox = grid.voxels
    for b in range(0, o.shape[0], chunk):
        ob, db = o[b:b + chunk], d[b:b + chunk]
        isect, num, _ = calls.masked_intersect_raw(ob, db, grid.mask, grid.xyzmin, n,
                                                   grid.voxel_size, 1.0, 3 * n)
        m = int(num.max().item())
        if m == 0:
            continue
        isect = isect[:, :m]
        mask = torch.arange(m, device=dev)[None, :] < num[:, None]
        rows = isect[mask]
        x = calls.integrate(vox, rows[:, :3].contiguous(), rows[:, 3:].contiguous())
        for i, nm in enumerate(names):
            x = torch.nn.functional.linear(x, W[nm + '.weight'], W[nm + '.bias'])
            if i + 1 < len(names):
                x = torch.relu(x)
        sigma = torch.zeros(mask.shape, device=dev)
        sigma[mask] = x[:, 0]
        alpha = 1 - torch.exp(-torch.relu(sigma) * mask)
        shifted = torch.nn.functional.pad(1 - alpha[None, :, :-1], (1, 0), value=1)[0]
        weight = (alpha * torch.cumprod(shifted, -1))[mask]
        c = torch.minimum((rows[:, :3] + 1e-4).long(), (rows[:, 3:] + 1e-4).long())
        flat = c[:, 0] + c[:, 1] * n + c[:, 2] * n * n
        amap.scatter_reduce_(0, flat, weight, reduce='amax')
    return amap.view(n, n, n)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--views', type=int, default=4)
    ap.add_argument('--reps', type=int, default=3)
    ap.add_argument('--no-plain', action='store_true')
    ap.add_argument('--batch', type=int, default=1 << 24, help='rays per alpha-map launch')
    args = ap.parse_args(argv)
    dev = torch.device('cuda', 0)
    sc = S.config2_scene()
    grid = SceneGrid.from_synthetic(sc, dev)
    sd = S.reference_decoder_state(32, 32)
    dec = Decoder(from_state_dict(sd, 32))
    cams = S.hemisphere_cameras()
    od = [camera_rays(cams[i], 800, 800, 0.69, dev) for i in range(args.views)]
    o = torch.cat([x[0] for x in od]).contiguous()
    d = torch.cat([x[1] for x in od]).contiguous()
    R = o.shape[0]

    prune.alpha_map(grid, dec, o[:65536], d[:65536], batch=args.batch)
    torch.cuda.synchronize()
    best = float('inf')
    for _ in range(args.reps):
        t0 = time.perf_counter()
        fused = prune.alpha_map(grid, dec, o, d, batch=args.batch)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    _, num, _ = calls.masked_intersect_raw(o, d, grid.mask, grid.xyzmin, grid.n, grid.voxel_size,
                                           1.0, 3 * grid.n, counts_only=True)
    intervals = int(num.sum().item())
    alg_bytes = R * 24 + intervals * (8 * grid.channels * 4 + 4)
    res = {'workload': f'alpha map (prune.py:83-137), {args.views} views of 800x800 rays, 256^3 x 32 '
                       'grid, DIVeR32 decoder FP32, offline walk over the mask',
           'rays': R, 'intervals': intervals, 'rays_per_launch': min(R, args.batch),
           'fused': {'ms': round(best * 1e3, 3), 'rays_per_s': round(R / best, 1),
                     'achieved_GBps': round(alg_bytes / best / 1e9, 1)},
           'voxels_over_1e-2': int((fused > 1e-2).sum().item())}
    if not args.no_plain:
        _plain_alpha_map(grid, sd, o[:65536], d[:65536])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plain = _plain_alpha_map(grid, sd, o, d)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        res['plain'] = {'ms': round(tp * 1e3, 3), 'rays_per_s': round(R / tp, 1)}
        res['speedup'] = round(tp / best, 2)
        res['max_abs_fused_vs_plain'] = float((fused - plain).abs().max().item())
    print(json.dumps(res))


if __name__ == '__main__':
    main()
