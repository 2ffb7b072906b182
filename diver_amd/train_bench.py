"""Training-step benchmark: the fused renderer's hand-written adjoint (recompute) against the
reference's autograd structure (stored samples), on one GPU.

    python -m diver_amd.train_bench [--steps 20] [--warmup 3] [--rays 8192]

Scene: the 256^3 x 32-channel sphere/shell grid with trainable corner features and a benchmark
decoder of hidden width 64 (density softplus(F[0]); colour MLP (31 + 27) -> 64 -> 64 -> 3).
Each step draws 8192 random rays from the 200 hemisphere cameras (800x800), renders them, takes
the reference's loss -- MSE against colours rendered from a second seeded grid plus
l_s * mean(log(1 + sigma^2 / 0.5)) over the [B, M] raw densities, l_s = 1e-5 (train.py:84-85,
109-113) -- and runs Adam on features and decoder.  Both paths composite every interval (no early termination, as the reference's offline
training forward, model/diver.py:183-207).

  recompute: FusedRenderer -- traversal, integral, decoder and compositing in one kernel; the
             backward re-walks each ray (diver_render_rows) instead of keeping samples.
  stored:    masked_intersect -> integrate -> decoder -> composite with torch autograd over the
             drop-in ops, keeping every interval's coordinates, features and activations, as the
             reference's model/diver.py:141-207 does.

Both paths use the same Adam: by default diver_amd.optim.Adam (one HIP pass per parameter, which
also clears the gradient), or torch's fused Adam with --optimizer torch.  Prints one JSON line per run: steps/s, the
forward + backward and Adam milliseconds per step (HIP events on the compute stream), and the
forward + backward device-memory high-water mark above the resident state (parameters, gradients,
optimizer moments) for each path.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import calls
from . import synthetic as S
from .adjoint import BenchmarkDecoder, FusedRenderer, corner_rows, positional_encoding
from .decoder import Decoder, from_baseline
from .model.diver import composite
from .optim import Adam as DiverAdam
from .renderer import SceneGrid
from .utils.integrator import integrate
from .utils.ray_voxel_intersection import masked_intersect


L_S = 1e-5  # the reference's default sparsity weight (configs/*.py 'l_s')


def stored_forward(voxels, dec, mask, n, o, d):
    """Reference-structured forward with autograd over stored samples: (rgb, sigma_loss)."""
    xyzmin, vs = -1.0, 2.0 / n
    coord, hit, _tns = masked_intersect(o, d, xyzmin, 1.0, float(n), vs, mask, 1.0)
    rows = coord[hit]
    F = integrate(voxels, rows[:, :3].contiguous(), rows[:, 3:].contiguous())
    ray = torch.where(hit)[0]
    x = torch.cat([F[:, 1:], positional_encoding(d[ray])], -1)
    x = torch.relu(torch.nn.functional.linear(x, dec.w1, dec.b1))
    x = torch.relu(torch.nn.functional.linear(x, dec.w2, dec.b2))
    col = torch.sigmoid(torch.nn.functional.linear(x, dec.w3, dec.b3))
    B, M = hit.shape
    sigma = torch.zeros(B, M, device=o.device).masked_scatter(hit, torch.nn.functional.softplus(F[:, 0]))
    color = torch.zeros(B, M, 3, device=o.device).masked_scatter(hit[..., None].expand(B, M, 3),
                                                                 col)
    # composite() applies relu to sigma (model/diver.py:195); softplus output is already >= 0
    rgb, _w = composite(sigma, color, hit, True)
    raw = torch.zeros(B, M, device=o.device).masked_scatter(hit, F[:, 0])
    return rgb, torch.log(1 + raw ** 2 / 0.5).mean()


def run(path, steps, warmup, n_rays, dev, optimizer='diver'):
    sc = S.config2_scene(32, 256)
    gt = S.config2_scene(32, 256, seed=S.SEED_GT)
    cams = S.hemisphere_cameras()
    layers = S.baseline_decoder_layers(32, 64)
    grid = SceneGrid.from_synthetic(sc, dev)
    # targets: the ground-truth grid rendered by the same decoder with a shifted output bias
    gt_grid = SceneGrid.from_synthetic(gt, dev)
    gt_layers = [(w, b.copy()) for w, b in layers]
    gt_layers[-1] = (gt_layers[-1][0], gt_layers[-1][1] + 0.5)
    gt_dec = Decoder(from_baseline(gt_layers, 32))
    batches = []
    for i in range(warmup + steps):
        o, d = S.random_rays(cams, n_rays, 800, 800, 100 + i)
        og, dg = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
        opts = calls.render_opts('fp32', 'offline', 0.0, 0.0)
        rgba, _ = calls.render_rays(gt_grid, gt_dec, og, dg, opts, with_count=False)
        batches.append((og, dg, (rgba[:, :3] + rgba[:, 3:]).clone()))
    del gt_grid, gt_dec
    dec = BenchmarkDecoder(layers, 32).to(dev)
    if path == 'recompute':
        model = FusedRenderer(grid, dec, traversal='offline')
        params = list(model.parameters())
        fwd = lambda o, d: model(o, d, 0.0, sparsity=True)
    else:
        voxels = torch.nn.Parameter(grid.voxels.clone())
        del grid
        mask = torch.from_numpy(sc.mask).to(dev)
        params = [voxels] + list(dec.parameters())
        fwd = lambda o, d: stored_forward(voxels, dec, mask, 256, o, d)
    if optimizer == 'diver':
        # one HIP pass per parameter that also clears its gradient (no zero_grad memset); the grid
        # is updated on the corners of the occupied voxels only, the rows a gradient can reach
        # (exactly the full update without weight decay)
        rows = corner_rows(torch.from_numpy(sc.mask).to(dev))
        opt = DiverAdam([{'params': [params[0]], 'rows': rows, 'row_len': 32},
                         {'params': params[1:]}], lr=1e-3, clear_grad=True)
    elif optimizer == 'diver-dense':
        opt = DiverAdam(params, lr=1e-3, clear_grad=True)
    else:
        # torch's fused update (its default multi-pass foreach Adam takes ~2.3x longer on the grid)
        opt = torch.optim.Adam(params, lr=1e-3, fused=True)
    ev = []

    def step(b, track=None, timed=False):
        og, dg, tg = b
        if optimizer != 'diver':
            opt.zero_grad(set_to_none=False)
        if track is not None:
            torch.cuda.synchronize()
            torch.cuda.reset_peak_memory_stats(dev)
            before = torch.cuda.memory_allocated(dev)
        if timed:
            e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            e[0].record()
        rgb, sigma_loss = fwd(og, dg)
        loss = torch.mean((rgb - tg) ** 2) + L_S * sigma_loss
        loss.backward()
        if timed:
            e[1].record()
        if track is not None:
            torch.cuda.synchronize()
            track.append(torch.cuda.max_memory_allocated(dev) - before)
        opt.step()
        if timed:
            e[2].record()
            ev.append(e)
        return loss

    for i in range(warmup):
        step(batches[i])
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated(dev)
    t0 = time.perf_counter()
    losses = []
    for i in range(steps):
        losses.append(step(batches[warmup + i], timed=True))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    fwd_bwd_ms = sum(e[0].elapsed_time(e[1]) for e in ev) / len(ev)
    opt_ms = sum(e[1].elapsed_time(e[2]) for e in ev) / len(ev)
    # forward + backward memory high-water mark above the resident state (parameters,
    # gradients, optimizer moments), on one more batch outside the timed loop
    track = []
    step(batches[-1], track)
    return {'steps_per_s': steps / dt, 'ms_per_step': 1e3 * dt / steps,
            'rays_per_s': steps * n_rays / dt, 'fwd_bwd_ms': fwd_bwd_ms, 'adam_ms': opt_ms,
            'resident_bytes': base,
            'fwd_bwd_peak_above_resident_bytes': track[0],
            'loss_first': float(losses[0]), 'loss_last': float(losses[-1])}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--rays', type=int, default=8192)
    ap.add_argument('--path', choices=['both', 'recompute', 'stored'], default='both')
    ap.add_argument('--optimizer', choices=['diver', 'diver-dense', 'torch'], default='diver',
                    help="diver: one-pass HIP Adam, the grid on its trainable rows "
                         "(diver_adam_step_rows); diver-dense: the same over the whole grid; "
                         "torch: torch's fused Adam")
    args = ap.parse_args(argv)
    dev = torch.device('cuda', 0)
    out = {'workload': 'training step, 256^3 x 32 trainable grid, hidden-64 benchmark decoder, '
                       f'{args.rays} random rays, MSE + 1e-5 sparsity loss, Adam, no early '
                       'termination',
           'steps': args.steps, 'optimizer': args.optimizer}
    paths = ('recompute', 'stored') if args.path == 'both' else (args.path,)
    for path in paths:
        print(f'[train] {path}', file=sys.stderr, flush=True)
        out[path] = run(path, args.steps, args.warmup, args.rays, dev, args.optimizer)
        torch.cuda.empty_cache()
    if len(paths) == 2:
        out['speedup_recompute_vs_stored'] = (out['recompute']['steps_per_s'] /
                                              out['stored']['steps_per_s'])
        out['fwd_bwd_speedup_recompute_vs_stored'] = (out['stored']['fwd_bwd_ms'] /
                                                      out['recompute']['fwd_bwd_ms'])
    print(json.dumps(out))


if __name__ == '__main__':
    main()
