"""Benchmark: rays/s and 800x800 frame latency of the fused render path (BF16 matrix-core decoder)
against the plain FP32 GPU path (the reference's real-time op sequence on libdiver), on the
256^3 / 32-channel synthetic scene (12 spheres + 4 shells, ~5% occupied), benchmark DIVeR32
decoder, cameras on the upper hemisphere at radius 4, early termination at T < 1e-4.

Prints one JSON line.  A "step" is one 800x800 frame (640,000 rays) rendered from a resident
scene; inputs (grid, mask, decoder) are in HBM before timing starts.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the render kernel on this workload, from the committed rocprofv3
# FETCH_SIZE / WRITE_SIZE passes (scaled by a 2 GiB copy-kernel calibration)
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                            'render_800x800_256cube_traffic.json')


def measured_traffic(n, width, height, channels):
    if (n, width, height, channels) != (256, 800, 800, 32) or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        return json.load(f)['bf16']['traffic_bytes_per_launch']


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--res', type=int, default=800)
    ap.add_argument('--n', type=int, default=256)
    ap.add_argument('--plain-frames', type=int, default=5)
    ap.add_argument('--cpu-rays', type=int, default=16384)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--calib-copy-mb', type=int, default=0,
                    help='copy this many MiB device-to-device first (counter calibration)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({'error': 'this build runs on one GPU; multi-GPU sharding is described '
                                   'in DESIGN.md but not launched', 'n_gpus': args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    from diver_amd import synthetic as S
    from diver_amd.decoder import Decoder, from_baseline
    from diver_amd.realtime import RealTimeRenderer
    from diver_amd.renderer import SceneGrid, camera_rays, focal_for, render_camera
    from diver_amd import calls

    dev = torch.device('cuda', 0)
    W = H = args.res
    R = W * H
    log(f'[bench] building {args.n}^3 scene')
    sc = S.config2_scene(n=args.n)
    grid = SceneGrid.from_synthetic(sc, dev)
    layers = S.baseline_decoder_layers(32, 32)
    dec = Decoder(from_baseline(layers, 32))
    cams = S.hemisphere_cameras()
    t_stop = 1e-4
    if args.calib_copy_mb:
        # a plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE scale the counters
        src = torch.ones(args.calib_copy_mb * 262144, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        del src, dst
    torch.cuda.synchronize()

    def timed(fn, frames):
        evs = []
        out = None
        for i in range(frames):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            out = fn(i)
            b.record()
            evs.append((a, b))
        torch.cuda.synchronize()
        return [a.elapsed_time(b) for a, b in evs], out

    # ---- plain FP32 path: reference real-time op sequence --------------------------------------
    log('[bench] plain FP32 path (aabb_intersect + ray_march/mlp_eval rounds)')
    rt = RealTimeRenderer(grid.voxels, grid.mask, sc.xyzmin, sc.voxel_size, dec,
                          alpha_skip=0.0, t_stop=t_stop)
    rt.render(cams[0], W, H)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_lat = []
    for i in range(args.plain_frames):
        ts = time.perf_counter()
        img_plain, rounds = rt.render(cams[i % len(cams)], W, H)
        torch.cuda.synchronize()
        plain_lat.append((time.perf_counter() - ts) * 1e3)
    plain_rays_s = R * args.plain_frames / (time.perf_counter() - t0)

    # ---- fused paths -----------------------------------------------------------------------
    counts = {}

    def fused(prec):
        def fn(i):
            rgb, T, cnt = render_camera(grid, dec, cams[i % len(cams)], W, H, precision=prec,
                                        t_stop=t_stop, with_count=True)
            counts[i] = cnt
            return rgb
        return fn

    log('[bench] fused FP32')
    for i in range(args.warmup):
        fused('fp32')(i)
    lat32, _ = timed(fused('fp32'), args.steps)
    log('[bench] fused BF16 (timed region)')
    for i in range(args.warmup):
        fused('bf16')(i)
    counts.clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lat16, _ = timed(fused('bf16'), args.steps)
    wall = time.perf_counter() - t0
    i_eval = float(sum(int(c.sum()) for c in counts.values())) / args.steps

    # ---- the same frames with two in flight on two streams (the render server's launch pattern):
    # the central tiles' tail of one frame overlaps the next frame's start.  Reported beside the
    # headline, which times one frame after another on one stream.
    streams = [torch.cuda.Stream(device=dev), torch.cuda.Stream(device=dev)]
    for i in range(max(2, args.warmup)):
        with torch.cuda.stream(streams[i % 2]):
            render_camera(grid, dec, cams[i % len(cams)], W, H, precision='bf16', t_stop=t_stop)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        with torch.cuda.stream(streams[i % 2]):
            render_camera(grid, dec, cams[i % len(cams)], W, H, precision='bf16', t_stop=t_stop)
    torch.cuda.synchronize()
    two_stream_rays_s = R * args.steps / (time.perf_counter() - t0)

    ms_per_step = float(np.mean(lat16))
    value = R * args.steps / wall
    rays_fp32 = R / (np.mean(lat32) / 1e3)
    C = 32
    b_alg = R * 40 + i_eval * 8 * C * 4
    achieved = b_alg / (ms_per_step / 1e3) / 1e9

    # ---- accuracy on a sample of one frame: BF16 / FP32 GPU vs FP64 oracle ---------------------
    result_extra = {}
    cpu = None
    if not args.no_cpu:
        import oracle
        from oracle import cpu_path, fp64
        from oracle.render import render as oracle_render
        cam = cams[0]
        o_all, d_all = camera_rays(cam, W, H, 0.69, dev)
        stride = max(1, R // args.cpu_rays)
        sel = torch.arange(0, R, stride, device=dev)[:args.cpu_rays]
        o_s, d_s = o_all[sel].cpu().numpy(), d_all[sel].cpu().numpy()
        rgb16, _, _ = render_camera(grid, dec, cam, W, H, precision='bf16', t_stop=t_stop)
        rgb32, _, _ = render_camera(grid, dec, cam, W, H, precision='fp32', t_stop=t_stop)
        rgb16 = rgb16.reshape(-1, 3)[sel].cpu().numpy()
        rgb32 = rgb32.reshape(-1, 3)[sel].cpu().numpy()
        img_plain, _ = rt.render(cam, W, H)
        img_p = img_plain.reshape(-1, 3)[sel].cpu().numpy()
        vox_np = sc.dense_voxels()
        ref = oracle_render(vox_np, sc.mask, sc.n, sc.xyzmin, sc.voxel_size,
                            lambda F, dd: fp64.baseline_decoder(layers, F, dd), 1, o_s, d_s,
                            t_stop, walk='realtime')
        rgb_ref = ref['rgb'] + ref['T'][:, None]

        def psnr(a):
            mse = float(np.mean((a.astype(np.float64) - rgb_ref) ** 2))
            return round(-10 * np.log10(max(mse, 1e-30)), 2)
        result_extra['psnr_vs_fp64'] = {'bf16': psnr(rgb16), 'fp32': psnr(rgb32),
                                        'plain_fp32': psnr(img_p)}
        result_extra['max_abs_vs_fp64'] = {
            'bf16': float(np.abs(rgb16 - rgb_ref).max()),
            'fp32': float(np.abs(rgb32 - rgb_ref).max())}
        # CPU baseline: the reference's offline path restated in torch FP32 on host cores
        ncpu = os.cpu_count() or 1
        torch.set_num_threads(ncpu)
        vox_t = torch.from_numpy(vox_np)
        decode, act = cpu_path.baseline_decode(layers)
        cpu_path.render(o_s[:512], d_s[:512], vox_t, sc.mask, sc.n, sc.xyzmin, sc.voxel_size,
                        decode, act)
        t0 = time.perf_counter()
        cpu_rgb = cpu_path.render(o_s, d_s, vox_t, sc.mask, sc.n, sc.xyzmin, sc.voxel_size,
                                  decode, act)
        cpu_s = time.perf_counter() - t0
        cpu = {'value': round(o_s.shape[0] / cpu_s, 1), 'unit': 'rays/s', 'cores': ncpu,
               'kind': 'port',
               'sample': f'{o_s.shape[0]} strided pixels of one {W}x{H} view, offline path '
                         '(traverse, integrate, nn.Linear decoder, cumprod composite), FP32, '
                         'no early termination'}
        result_extra['psnr_vs_fp64']['cpu_fp32'] = psnr(cpu_rgb.numpy())

    line = {
        'metric': 'rays/s + p50/p99 800^2 frame latency, BF16 vs FP32 path at >=45 dB vs FP64',
        'value': round(value, 1),
        'unit': 'rays/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ms_per_step, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'bf16 decoder (fp32 traversal/integral/transmittance)',
        'data': 'synthetic (seeded spheres+shells grid, Kaiming decoder, hemisphere cameras)',
        'config': {'workload': f'fused render, {W}x{H} frames, {args.n}^3 grid, 32 channels, '
                               'DIVeR32 benchmark decoder, T<1e-4 termination',
                   'frames': args.steps, 'rays_per_frame': R,
                   'occupancy': round(sc.occupancy, 4)},
        'p50_ms': round(float(np.percentile(lat16, 50)), 4),
        'p99_ms': round(float(np.percentile(lat16, 99)), 4),
        'fused_fp32_rays_s': round(rays_fp32, 1),
        'plain_fp32_rays_s': round(plain_rays_s, 1),
        'plain_fp32_p50_ms': round(float(np.percentile(plain_lat, 50)), 3),
        'plain_rounds_per_frame': rounds,
        'speedup_vs_plain_fp32': round(value / plain_rays_s, 2),
        'two_frames_in_flight_rays_s': round(two_stream_rays_s, 1),
        'intervals_per_frame': i_eval,
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4),
                     'traffic': measured_traffic(args.n, W, H, C)},
        'cpu_baseline': cpu,
    }
    line.update(result_extra)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
