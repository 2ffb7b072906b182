"""Serving benchmark: Poisson frame arrivals at mixed resolutions over three resident scenes.

    python -m diver_amd.serve_bench [--rate 400] [--seconds 4] [--precision bf16]
    python -m diver_amd.serve_bench --baseline      # one request at a time, FP32 decoder

Scenes (synthetic, seeded): 128^3 x 32 channels, 256^3 x 32, 256^3 x 64, each with the benchmark
decoder of its width.  Resolutions 200^2, 400^2, 800^2 and 1920x1080 and the scenes are drawn
uniformly, cameras from the hemisphere set, arrival times from a Poisson process (seed 4).  The
served configuration packs up to 16 pending requests of a scene into one launch with three
launches in flight (``--streams``); ``--baseline`` serves the same arrivals one request per launch, one launch at a time,
with the FP32 decoder.  Prints one JSON line: throughput (rays/s over the run), p50/p99 frame
latency overall and per resolution, the unloaded latency of each resolution alone, launch counts.
"""
from __future__ import annotations

import argparse
import collections
import gc
import json
import sys
import threading
import time

import numpy as np
import torch

from . import synthetic as S
from .decoder import Decoder, from_baseline
from .renderer import SceneGrid
from .server import RenderServer

RESOLUTIONS = [((200, 200), 0.25), ((400, 400), 0.25), ((800, 800), 0.25), ((1920, 1080), 0.25)]
SCENES = [('s128c32', 128, 32), ('s256c32', 256, 32), ('s256c64', 256, 64)]


def build_scenes(dev):
    out = {}
    for name, n, c in SCENES:
        sc = S.config2_scene(channels=c, n=n)
        out[name] = (SceneGrid.from_synthetic(sc, dev),
                     Decoder(from_baseline(S.baseline_decoder_layers(c, c if c > 32 else 32), c)))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--rate', type=float, default=400.0, help='mean frame arrivals per second')
    ap.add_argument('--seconds', type=float, default=4.0)
    ap.add_argument('--precision', default='bf16')
    ap.add_argument('--streams', type=int, default=3)
    ap.add_argument('--baseline', action='store_true',
                    help='one request per launch, one launch in flight, FP32 decoder')
    ap.add_argument('--split-rays', type=int, default=1 << 18,
                    help='frames above this many rays launch apart from smaller ones (0: mix)')
    args = ap.parse_args(argv)
    max_requests = 16
    if args.baseline:
        args.precision, args.streams, max_requests = 'fp32', 1, 1
    dev = torch.device('cuda', 0)
    print('[serve] building scenes', file=sys.stderr, flush=True)
    scenes = build_scenes(dev)
    cams = S.hemisphere_cameras()
    rng = np.random.default_rng(S.SEED_ARRIVALS)
    with RenderServer(dev, n_streams=args.streams, max_requests=max_requests,
                      split_rays=args.split_rays or None) as srv:
        for name, (g, d) in scenes.items():
            srv.add_scene(name, g, d, precision=args.precision)
        # unloaded latency of each resolution (scene 256^3 x 32, one frame at a time)
        alone = {}
        for (w, h), _p in RESOLUTIONS:
            lat = []
            for i in range(6):
                r = srv.submit('s256c32', cams[i], w, h)
                r.result(timeout=60)
                lat.append(r.latency * 1e3)
            alone[f'{w}x{h}'] = float(np.median(lat[1:]))
        print('[serve] arrivals', file=sys.stderr, flush=True)
        # the client takes each frame as it is rendered and drops it (a collector thread), and
        # keeps its start-up objects out of reach of Python's full garbage collections, whose
        # pauses (tens of ms) would otherwise stop the submitting thread
        q = collections.deque()
        done = threading.Event()

        def collect():
            while not (done.is_set() and not q):
                if not q:
                    time.sleep(50e-6)
                    continue
                q.popleft().result(timeout=120)

        gc.collect()
        gc.freeze()
        th = threading.Thread(target=collect)
        th.start()
        t_start = time.perf_counter()
        probs = np.array([p for _r, p in RESOLUTIONS])
        t = t_start
        while True:
            t += rng.exponential(1.0 / args.rate)
            if t - t_start > args.seconds:
                break
            (w, h), _p = RESOLUTIONS[rng.choice(len(RESOLUTIONS), p=probs)]
            name = SCENES[rng.integers(len(SCENES))][0]
            cam = cams[rng.integers(len(cams))]
            now = time.perf_counter()
            if t > now:
                time.sleep(t - now)
            q.append(srv.submit(name, cam, w, h))
        done.set()
        th.join()
        st = srv.stats(since=t_start)
    st.update({'offered_rate_hz': args.rate, 'seconds': args.seconds, 'precision': args.precision,
               'streams': args.streams, 'max_requests_per_launch': max_requests,
               'unloaded_p50_ms': alone,
               'scenes': [f'{n}^3 x {c}' for _name, n, c in SCENES]})
    print(json.dumps(st))


if __name__ == '__main__':
    main()
