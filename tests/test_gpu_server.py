"""Continuous-batching render server: results, admission and statistics.

Frames of two resident scenes at several resolutions are submitted from several threads while
launches are in flight; every frame must equal (bit for bit: rgba and interval counts) the same
frame rendered alone with diver_render_camera.
"""
import threading

import numpy as np
import pytest
import torch

from diver_amd import calls
from diver_amd import synthetic as S
from diver_amd.decoder import Decoder, from_baseline
from diver_amd.renderer import SceneGrid, focal_for
from diver_amd.server import AdmissionError, RenderServer

pytestmark = pytest.mark.gpu


def _scenes(dev):
    a = S.config1_scene()
    b = S.config2_scene(channels=64, n=64)
    ga = SceneGrid.from_synthetic(a, dev)
    gb = SceneGrid.from_synthetic(b, dev)
    da = Decoder(from_baseline(S.baseline_decoder_layers(32, 32), 32))
    db = Decoder(from_baseline(S.baseline_decoder_layers(64, 64), 64))
    return (ga, da), (gb, db)


def test_server_mixed_frames_equal_alone(cameras, dev):
    (ga, da), (gb, db) = _scenes(dev)
    sizes = [(32, 32), (48, 40), (64, 64), (100, 75), (17, 9)]
    jobs = []
    for i in range(20):
        scene = 'a' if i % 3 else 'b'
        w, h = sizes[i % len(sizes)]
        jobs.append((scene, cameras[(7 * i) % len(cameras)], w, h))
    with RenderServer(dev, n_streams=2) as srv:
        srv.add_scene('a', ga, da, precision='bf16')
        srv.add_scene('b', gb, db, precision='bf16')
        reqs = [None] * len(jobs)

        def client(k):
            for i in range(k, len(jobs), 4):
                sc, cam, w, h = jobs[i]
                reqs[i] = srv.submit(sc, cam, w, h)

        threads = [threading.Thread(target=client, args=(k,)) for k in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for r in reqs:
            r.result(timeout=60)
        st = srv.stats()
    assert st['frames'] == len(jobs) and st['p99_ms'] >= st['p50_ms'] > 0
    torch.cuda.synchronize()
    for (sc, cam, w, h), r in zip(jobs, reqs):
        grid, dec = (ga, da) if sc == 'a' else (gb, db)
        opts = calls.render_opts('bf16', 'realtime', 1e-4, 0.0)
        rgba, count = calls.render_camera(grid, dec, np.asarray(cam, np.float32), w, h,
                                          focal_for(w, 0.69), opts)
        assert torch.equal(r.rgba, rgba), (sc, w, h)
        assert torch.equal(r.count, count), (sc, w, h)


def test_server_admission(cameras, dev):
    (ga, da), _ = _scenes(dev)
    need = ga.nbytes + ga.blocks.numel() * 4 + ga.bricks.numel() * 8 + da.nbytes
    with RenderServer(dev, hbm_budget=need + 64 * 64 * 20) as srv:
        srv.add_scene('a', ga, da)
        with pytest.raises(AdmissionError):
            srv.add_scene('a2', ga, da)           # a second copy does not fit
        with pytest.raises(AdmissionError):
            srv.submit('a', cameras[0], 128, 128)  # output larger than the free budget
        with pytest.raises(KeyError):
            srv.submit('nope', cameras[0], 8, 8)
        # frames that fit one at a time still all complete (they wait for memory)
        reqs = [srv.submit('a', cameras[i], 64, 64) for i in range(4)]
        for r in reqs:
            rgb, T, cnt = r.result(timeout=60)
            assert rgb.shape == (64, 64, 3)
        assert srv.launches >= 4
