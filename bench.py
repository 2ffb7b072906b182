"""Benchmark: rays/s and 800x800 frame latency of the BF16 serving path (the native render server
packing frames into fused launches, BF16 matrix-core decoder) against the plain FP32 GPU path (the
reference's real-time op sequence on libdiver, one frame at a time), on the 256^3 / 32-channel
synthetic scene (12 spheres + 4 shells, ~5% occupied), benchmark DIVeR32 decoder, cameras on the
upper hemisphere at radius 4, early termination at T < 1e-4.

Prints one JSON line.  A "step" is one 800x800 frame (640,000 rays) rendered from a resident
scene; inputs (grid, mask, decoder) are in HBM before timing starts.  ``value`` is the server's
throughput with all ``steps`` frames submitted at once (saturated); ``p50_ms`` / ``p99_ms`` are its
submit-to-done frame latencies under Poisson arrivals at ``--rate`` frames/s.  Beside them: the
fused kernel one frame at a time (its launch durations give the roofline), the FP32 decoder, the
plain FP32 path, and the FP64-oracle accuracy of each.
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


COUNTER_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                            'render_800x800_256cube_counters.json')


def counter_fractions(n, width, height, channels):
    """Issue and matrix-core fractions of the render kernel from the committed rocprofv3 counter
    passes (profiles/render_800x800_256cube_counters.json), and what limits the kernel."""
    if (n, width, height, channels) != (256, 800, 800, 32) or not os.path.exists(COUNTER_FILE):
        return {}
    with open(COUNTER_FILE) as f:
        j = json.load(f)
    c = j['bf16']
    out = {k: c[k] for k in ('limiter', 'valu_issue_frac', 'issue_frac', 'mfma_frac',
                             'mfma_tflops', 'instructions_per_voxel_step',
                             'resident_waves_per_simd') if k in c}
    cb = j.get('bf16_batch16', {})
    out.update({'batch16_' + k: cb[k] for k in ('valu_issue_frac', 'issue_frac',
                                                'resident_waves_per_simd') if k in cb})
    return out


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def serving_keys(serve, args, R, lat16):
    """The line's latency keys: the server's at the offered rate, or (without the serving legs)
    the kernel's one frame at a time."""
    import numpy as np
    if serve is None:
        return {'p50_ms': round(float(np.percentile(lat16, 50)), 4),
                'p99_ms': round(float(np.percentile(lat16, 99)), 4)}
    lat, sat = serve['rate_lat'], serve['saturated_lat']
    return {
        'p50_ms': round(float(np.percentile(lat, 50)), 4),
        'p99_ms': round(float(np.percentile(lat, 99)), 4),
        'latency_offered_hz': args.rate,
        'latency_offered_rays_s': round(args.rate * R, 1),
        'latency_frames': serve['offered_frames'],
        'latency_delivered_hz': round(serve['offered_frames'] / serve['rate_wall_s'], 1),
        'saturated_p50_ms': round(float(np.percentile(sat, 50)), 4),
        'saturated_p99_ms': round(float(np.percentile(sat, 99)), 4),
        'serving_launches': serve['launches'],
    }


def serving_legs(args, grid, dec, cams, W, H, t_stop):
    """The render server (three launches in flight) on 800x800 frames of the scene: (1) saturated:
    ``steps`` frames submitted at once into output buffers allocated beforehand, wall clock from the
    first submit to the last result (the timed region; GPU synchronised on both sides); (2) Poisson
    arrivals at ``rate`` frames/s for
    ``rate_seconds`` after 0.3 s of untimed arrivals, a collector thread taking each result as it
    is rendered.  Returns the wall time, per-frame latencies of both legs and launch counts."""
    import collections
    import gc
    import threading

    import numpy as np
    import torch

    from diver_amd import synthetic as S
    from diver_amd.server import RenderServer

    log('[bench] serving: saturated (timed region), then Poisson arrivals')
    # A full collection of Python's cyclic garbage collector walks every object the process holds
    # (tens of ms after importing torch) and stops the submitting thread meanwhile; a serving
    # client moves the start-up objects out of its reach first (gc.freeze), as is usual for
    # latency-bound Python services.
    gc.collect()
    gc.freeze()
    out = {}
    with RenderServer(grid.voxels.device, n_streams=3) as srv:
        srv.add_scene('s', grid, dec, precision='bf16', t_stop=t_stop)
        # the client's output buffers for the timed frames, allocated before timing starts
        outs = [(torch.empty((H, W, 4), dtype=torch.float32, device=grid.voxels.device),
                 torch.empty((H, W), dtype=torch.int32, device=grid.voxels.device))
                for _ in range(args.steps)]
        for r in [srv.submit('s', cams[i % len(cams)], W, H, out=outs[i % args.steps])
                  for i in range(args.warmup)]:
            r.result(timeout=60)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reqs = [srv.submit('s', cams[i % len(cams)], W, H, out=outs[i])
                for i in range(args.steps)]
        for r in reqs:
            r.result(timeout=120)
        torch.cuda.synchronize()
        out['wall_s'] = time.perf_counter() - t0
        out['launches'] = len({r.launch for r in reqs})
        out['saturated_lat'] = [r.latency * 1e3 for r in reqs]
        del reqs, outs
        rng = np.random.default_rng(S.SEED_ARRIVALS)
        q = collections.deque()
        done = threading.Event()
        lat = []

        def collect():
            while not (done.is_set() and not q):
                if not q:
                    time.sleep(50e-6)
                    continue
                r = q.popleft()
                r.result(timeout=120)
                lat.append(r.latency * 1e3)

        def arrivals(seconds):
            t_start = time.perf_counter()
            t_next = t_start
            k = 0
            while True:
                t_next += rng.exponential(1.0 / args.rate)
                if t_next - t_start > seconds:
                    break
                now = time.perf_counter()
                if t_next > now:
                    time.sleep(t_next - now)
                q.append(srv.submit('s', cams[k % len(cams)], W, H))
                k += 1
            return k, time.perf_counter() - t_start

        th = threading.Thread(target=collect)
        th.start()
        warm, _ = arrivals(0.3)  # untimed: the clients' output buffers reach their steady state
        while len(lat) < warm:
            time.sleep(1e-3)
        lat.clear()
        out['offered_frames'], out['rate_wall_s'] = arrivals(args.rate_seconds)
        done.set()
        th.join()
        out['rate_lat'] = lat
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--rate', type=float, default=2400.0,
                    help='offered frames/s for the serving latency leg')
    ap.add_argument('--rate-seconds', type=float, default=1.0)
    ap.add_argument('--res', type=int, default=800)
    ap.add_argument('--n', type=int, default=256)
    ap.add_argument('--plain-frames', type=int, default=5)
    ap.add_argument('--cpu-rays', type=int, default=16384)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-serve', action='store_true',
                    help='skip the serving legs (kernel profiling runs); value is then the '
                         'one-frame-per-launch throughput')
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
    log('[bench] fused BF16, one frame per launch (kernel launch durations)')
    for i in range(args.warmup):
        fused('bf16')(i)
    counts.clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lat16, _ = timed(fused('bf16'), args.steps)
    wall = time.perf_counter() - t0
    i_eval = float(sum(int(c.sum()) for c in counts.values())) / args.steps

    # ---- the same frames with two in flight on two streams, launched from Python: the central
    # tiles' tail of one frame overlaps the next frame's start (what the server does natively)
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

    # ---- 16 frames in one launch (diver_render_batch, as the server packs frames): the steady
    # throughput of the kernel without the frame-end tail
    opts16 = calls.render_opts('bf16', 'realtime', t_stop, 0.0)
    reqs16 = [{'c2w': np.asarray(cams[(7 * i) % len(cams)], np.float32), 'width': W, 'height': H,
               'focal': focal_for(W, 0.69)} for i in range(16)]
    outs16 = calls.render_batch(grid, dec, reqs16, opts16)
    i_b16 = float(sum(int(cnt.sum()) for _rgba, cnt in outs16))  # intervals over the 16 frames
    del outs16
    lat_b16, _ = timed(lambda i: calls.render_batch(grid, dec, reqs16, opts16), 5)
    batch16_rays_s = 16 * R / (float(np.median(lat_b16)) / 1e3)

    # ---- serving path: the native render server, frames of 800x800 -------------------------
    serve = None if args.no_serve else serving_legs(args, grid, dec, cams, W, H, t_stop)
    frame_rays_s = R * args.steps / wall
    if serve is not None:
        ms_per_step = serve['wall_s'] * 1e3 / args.steps
        value = R * args.steps / serve['wall_s']
    else:
        ms_per_step = wall * 1e3 / args.steps
        value = frame_rays_s
    rays_fp32 = R / (np.mean(lat32) / 1e3)
    C = 32
    b_alg = R * 40 + i_eval * 8 * C * 4
    launch_ms = float(np.mean(lat16))
    achieved = b_alg / (launch_ms / 1e3) / 1e9

    # ---- accuracy on a sample of one frame: BF16 / FP32 GPU vs FP64 oracle ---------------------
    result_extra = {}
    cpu = None
    if not args.no_cpu:
        import oracle
        from oracle import cpu_path, fp64
        from oracle.render import render as oracle_render
        # the 4 views of the precision report (0, 50, 100, 150), strided pixels of each
        views = [0, 50, 100, 150]
        per_view = max(1, args.cpu_rays // len(views))
        stride = max(1, R // per_view)
        sel = torch.arange(0, R, stride, device=dev)[:per_view]
        o_l, d_l, g16, g32, gp = [], [], [], [], []
        for v in views:
            cam = cams[v]
            o_all, d_all = camera_rays(cam, W, H, 0.69, dev)
            o_l.append(o_all[sel].cpu().numpy())
            d_l.append(d_all[sel].cpu().numpy())
            g16.append(render_camera(grid, dec, cam, W, H, precision='bf16',
                                     t_stop=t_stop)[0].reshape(-1, 3)[sel].cpu().numpy())
            g32.append(render_camera(grid, dec, cam, W, H, precision='fp32',
                                     t_stop=t_stop)[0].reshape(-1, 3)[sel].cpu().numpy())
            gp.append(rt.render(cam, W, H)[0].reshape(-1, 3)[sel].cpu().numpy())
        o_s, d_s = np.concatenate(o_l), np.concatenate(d_l)
        rgb16, rgb32, img_p = np.concatenate(g16), np.concatenate(g32), np.concatenate(gp)
        vox_np = sc.dense_voxels()
        ref = oracle_render(vox_np, sc.mask, sc.n, sc.xyzmin, sc.voxel_size,
                            lambda F, dd: fp64.baseline_decoder(layers, F, dd), 1, o_s, d_s,
                            t_stop, walk='realtime')
        rgb_ref = ref['rgb'] + ref['T'][:, None]

        def psnr(a, want=rgb_ref):
            mse = float(np.mean((a.astype(np.float64) - want) ** 2))
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
               'sample': f'{o_s.shape[0]} strided pixels of the {W}x{H} views {views} (the '
                         'FP64-oracle views), offline path (traverse, integrate, nn.Linear '
                         'decoder, cumprod composite), FP32, no early termination'}
        # the CPU configuration in full: 64^3 x 32, 6 spheres, one 64x64 view (4096 rays),
        # median of 20 runs after 3 warm-ups
        sc1 = S.config1_scene()
        o1, d1 = S.camera_rays(cams[0], 64, 64)
        vox1 = torch.from_numpy(sc1.dense_voxels())
        runs = []
        for i in range(23):
            t1 = time.perf_counter()
            cpu_path.render(o1, d1, vox1, sc1.mask, sc1.n, sc1.xyzmin, sc1.voxel_size, decode,
                            act)
            runs.append(time.perf_counter() - t1)
        med = float(np.median(runs[3:]))
        result_extra['cpu_config1'] = {'rays_per_s': round(4096 / med, 1),
                                       'ms_per_frame': round(med * 1e3, 2), 'cores': ncpu,
                                       'runs': 20, 'sample': '64^3 x 32, 6 spheres, one 64x64 '
                                                             'view, 4096 rays, FP32'}
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
        'config': {'workload': ('serving' if serve else 'fused render, one launch per frame:')
                               + f' {W}x{H} frames (BF16 decoder), {args.n}^3 grid, 32 channels, '
                               'DIVeR32 benchmark decoder, T<1e-4 termination',
                   'frames': args.steps, 'rays_per_frame': R,
                   'occupancy': round(sc.occupancy, 4)},
        **serving_keys(serve, args, R, lat16),
        'single_frame_rays_s': round(frame_rays_s, 1),
        'single_frame_p50_ms': round(float(np.percentile(lat16, 50)), 4),
        'single_frame_p99_ms': round(float(np.percentile(lat16, 99)), 4),
        'fused_fp32_rays_s': round(rays_fp32, 1),
        'plain_fp32_rays_s': round(plain_rays_s, 1),
        'plain_fp32_p50_ms': round(float(np.percentile(plain_lat, 50)), 3),
        'plain_fp32_p99_ms': round(float(np.percentile(plain_lat, 99)), 3),
        'plain_rounds_per_frame': rounds,
        'speedup_vs_plain_fp32': round(value / plain_rays_s, 2),
        'single_frame_speedup_vs_plain_fp32': round(frame_rays_s / plain_rays_s, 2),
        'two_frames_in_flight_rays_s': round(two_stream_rays_s, 1),
        'batch16_rays_s': round(batch16_rays_s, 1),
        'intervals_per_frame': i_eval,
        # the render kernel, one 800x800 frame per launch: algorithmic bytes over its event-timed
        # launch duration, priced on the HBM roofline; 'limiter' is what bounds it instead
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4),
                     'traffic': measured_traffic(args.n, W, H, C),
                     'kernel': 'render_tile_kernel', 'launch_ms': round(launch_ms, 4),
                     # the same kernel over 16 frames per launch (the serving regime)
                     'batch16_launch_ms': round(float(np.median(lat_b16)), 4),
                     'batch16_achieved': round((16 * R * 40 + i_b16 * 8 * C * 4)
                                               / (float(np.median(lat_b16)) / 1e3) / 1e9, 1),
                     **counter_fractions(args.n, W, H, C)},
        'cpu_baseline': cpu,
    }
    line.update(result_extra)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
