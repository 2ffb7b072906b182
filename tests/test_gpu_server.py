"""Continuous-batching render server: results, admission and statistics.

Frames of resident scenes at several resolutions are submitted from several threads while
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


def test_server_pinhole_frames_equal_alone(cameras, dev):
    """Tanks and Temples style frames (principal point off the centre, fx != fy) through the
    server equal the same pinhole requests rendered alone, and Blender frames in the same launches
    keep their camera."""
    (ga, da), _ = _scenes(dev)
    intr = [(70.0, 76.0, 30.2, 27.7), (55.5, 50.0, 12.0, 31.4)]
    with RenderServer(dev, n_streams=2) as srv:
        srv.add_scene('a', ga, da, precision='bf16')
        reqs = []
        for i in range(6):
            cam = cameras[(11 * i) % len(cameras)]
            if i % 2:
                reqs.append(srv.submit('a', cam, 64, 48))
            else:
                reqs.append(srv.submit('a', cam, 64, 48, intrinsics=intr[(i // 2) % 2]))
        for r in reqs:
            r.result(timeout=60)
    torch.cuda.synchronize()
    opts = calls.render_opts('bf16', 'realtime', 1e-4, 0.0)
    for i, r in enumerate(reqs):
        c2w = np.asarray(cameras[(11 * i) % len(cameras)], np.float32)
        rq = {'c2w': c2w, 'width': 64, 'height': 48, 'focal': focal_for(64, 0.69)}
        if i % 2 == 0:
            fx, fy, cx, cy = intr[(i // 2) % 2]
            rq.update(focal=fx, fy=fy, cx=cx, cy=cy)
        (rgba, count), = calls.render_batch(ga, da, [rq], opts)
        assert torch.equal(r.rgba, rgba) and torch.equal(r.count, count), i


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


def test_server_scene_built_after_decoder(cameras, dev):
    """A scene whose grid structures are still being built on the caller's stream when it is
    added: the server's launches wait for them (the decoder is created first, so nothing else
    synchronises the device between building the grid and the first launch)."""
    sc = S.config2_scene(channels=32, n=128)
    dec = Decoder(from_baseline(S.baseline_decoder_layers(32, 32), 32))
    grid = SceneGrid.from_synthetic(sc, dev)
    with RenderServer(dev) as srv:
        srv.add_scene('s', grid, dec, precision='bf16')
        reqs = [srv.submit('s', cameras[i], 96, 64) for i in range(3)]
        outs = [r.result(timeout=60) for r in reqs]
    for i, (rgb, T, cnt) in enumerate(outs):
        opts = calls.render_opts('bf16', 'realtime', 1e-4, 0.0)
        rgba, count = calls.render_camera(grid, dec, np.asarray(cameras[i], np.float32), 96, 64,
                                          focal_for(96, 0.69), opts)
        assert torch.equal(reqs[i].rgba, rgba) and torch.equal(cnt, count)


def test_server_unclaimed_frames_and_history(cameras, dev):
    """Frames whose results are never asked for still render and free their memory; close()
    renders everything submitted; the statistics come from the library's bounded history."""
    (ga, da), _ = _scenes(dev)
    need = ga.nbytes + ga.blocks.numel() * 4 + ga.bricks.numel() * 8 + da.nbytes
    with RenderServer(dev, hbm_budget=need + 2 * 64 * 64 * 20) as srv:
        srv.add_scene('a', ga, da)
        for i in range(12):
            srv.submit('a', cameras[i], 64, 64)  # dropped at once
        last = srv.submit('a', cameras[12], 64, 64)
        last.result(timeout=60)
        kept = srv.submit('a', cameras[13], 32, 32)
        st = srv.stats()
    assert st['frames'] >= 13 and st['launches'] >= 7  # one 64x64 frame fits the budget at a time
    rgb, T, cnt = kept.result(timeout=1)  # rendered by close()
    assert rgb.shape == (32, 32, 3)


def _oracle_psnr(sc, layers, cam, stride, rgba_frame, w, h):
    from oracle import fp64
    from oracle.render import render as oracle_render
    o, d = calls.camera_rays(np.asarray(cam, np.float32), w, h, focal_for(w, 0.69),
                             torch.device('cuda', 0))
    sel = np.arange(0, w * h, stride)
    o, d = o.cpu().numpy()[sel], d.cpu().numpy()[sel]
    ref = oracle_render(sc.dense_voxels(), sc.mask, sc.n, sc.xyzmin, sc.voxel_size,
                        lambda F, dd: fp64.baseline_decoder(layers, F, dd), 1, o, d, 1e-4,
                        walk='realtime')
    rgba = rgba_frame.cpu().numpy().reshape(-1, 4)[sel].astype(np.float64)
    rgb = rgba[:, :3] + rgba[:, 3:]
    rgb_ref = ref['rgb'] + ref['T'][:, None]
    mse = float(np.mean((rgb - rgb_ref) ** 2))
    return -10 * np.log10(max(mse, 1e-30)), float(np.abs(rgb - rgb_ref).max())


def test_server_config5_scenes(cameras, dev):
    """The serving configuration's three resident scenes (128^3 x 32, 256^3 x 32, 256^3 x 64
    channels, each with the benchmark decoder of its width, BF16) and its resolutions (200x200,
    800x800, 1920x1080), submitted from three threads: every frame equals the same frame rendered
    alone, bit for bit, and one strided view per scene is within 45 dB of the FP64 oracle."""
    specs = [('s128c32', 128, 32), ('s256c32', 256, 32), ('s256c64', 256, 64)]
    scenes = {}
    for name, n, c in specs:
        sc = S.config2_scene(channels=c, n=n)
        layers = S.baseline_decoder_layers(c, c if c > 32 else 32)
        scenes[name] = (sc, layers, SceneGrid.from_synthetic(sc, dev),
                        Decoder(from_baseline(layers, c)))
    sizes = [(200, 200), (800, 800), (1920, 1080)]
    jobs = [(specs[i % 3][0], cameras[(11 * i) % 200], *sizes[(i // 3) % 3]) for i in range(18)]
    reqs = [None] * len(jobs)
    with RenderServer(dev, n_streams=2) as srv:
        for name, (sc, layers, g, d) in scenes.items():
            srv.add_scene(name, g, d, precision='bf16')

        def client(k):
            for i in range(k, len(jobs), 3):
                name, cam, w, h = jobs[i]
                reqs[i] = srv.submit(name, cam, w, h)

        threads = [threading.Thread(target=client, args=(k,)) for k in range(3)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for r in reqs:
            r.result(timeout=120)
        st = srv.stats()
    assert st['frames'] == len(jobs)
    opts = calls.render_opts('bf16', 'realtime', 1e-4, 0.0)
    checked = set()
    for (name, cam, w, h), r in zip(jobs, reqs):
        sc, layers, g, d = scenes[name]
        rgba, count = calls.render_camera(g, d, np.asarray(cam, np.float32), w, h,
                                          focal_for(w, 0.69), opts)
        assert torch.equal(r.rgba, rgba), (name, w, h)
        assert torch.equal(r.count, count), (name, w, h)
        if name not in checked and (w, h) == (800, 800):
            p, mx = _oracle_psnr(sc, layers, cam, 53, r.rgba, w, h)
            assert p >= 45.0 and mx <= 2e-2, (name, p, mx)
            checked.add(name)
    assert checked == set(scenes)


def test_server_renders_into_caller_buffers(cameras, dev):
    """submit(out=...) renders into the client's own buffers (reused across frames) exactly as
    into fresh ones; mismatched buffers are refused."""
    (ga, da), _ = _scenes(dev)
    rgba = torch.full((40, 48, 4), -1.0, device=dev)
    count = torch.full((40, 48), -1, dtype=torch.int32, device=dev)
    with RenderServer(dev) as srv:
        srv.add_scene('a', ga, da)
        for i in range(3):
            r = srv.submit('a', cameras[i], 48, 40, out=(rgba, count))
            rgb, T, cnt = r.result(timeout=60)
            assert r.rgba.data_ptr() == rgba.data_ptr()
            fresh = srv.submit('a', cameras[i], 48, 40)
            fresh.result(timeout=60)
            assert torch.equal(rgba, fresh.rgba) and torch.equal(count, fresh.count)
        with pytest.raises(ValueError):
            srv.submit('a', cameras[0], 40, 48, out=(rgba, count))
