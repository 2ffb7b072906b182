"""Continuous-batching render server on one GPU.

Scenes (a SceneGrid plus a Decoder) stay resident in device memory.  Clients submit frame
requests (scene, camera, resolution) from any thread; a render loop thread packs the pending
requests of one scene (FIFO, up to ``max_requests`` requests and ``max_rays`` rays; frames above
``split_rays`` rays apart from smaller ones) into one ``diver_render_batch`` launch and keeps up to
``n_streams`` launches in flight.  Requests that
arrive while launches run are admitted into the next launch as soon as a stream slot and memory
are free, so their work overlaps the frames already on the GPU.  A second thread waits on the
oldest launch's completion event (blocking, without polling) and releases its frames and its
stream slot the moment the GPU finishes it.

Memory admission: every resident scene and every in-flight output buffer is charged against
``hbm_budget`` bytes.  A scene that does not fit is refused at ``add_scene``; a request whose
output could never fit next to the resident scenes is refused at ``submit``; a request that fits
only later waits in the queue until launches complete and release their outputs.

A ray's result depends only on its own intervals, so a frame rendered in a mixed launch is
bit-identical to the same frame rendered alone.

The reference has no serving layer (its real-time viewer renders one camera at a time,
eval_rt.py:106-135); this module is the serving path built around the same renderer.
"""
from __future__ import annotations

import collections
import dataclasses
import threading
import time
from typing import Deque, Dict, List, Optional

import numpy as np
import torch

from . import _lib, calls
from .decoder import Decoder
from .renderer import SceneGrid, focal_for

_OUT_BYTES_PER_PIXEL = 16 + 4  # rgba fp32 + int32 interval count


class AdmissionError(RuntimeError):
    """The request or scene can never fit the server's memory budget."""


@dataclasses.dataclass
class ServedScene:
    name: str
    grid: SceneGrid
    decoder: Decoder
    precision: str = 'bf16'
    traversal: str = 'auto'
    t_stop: float = 1e-4
    nbytes: int = 0


class FrameRequest:
    """One frame: filled in by the server; ``result()`` waits for it."""

    def __init__(self, scene: str, c2w, width: int, height: int, fov: float):
        self.scene = scene
        self.c2w = np.asarray(c2w, np.float32).reshape(3, 4)
        self.width = int(width)
        self.height = int(height)
        self.fov = float(fov)
        self.t_submit = time.perf_counter()
        self.t_launch: Optional[float] = None
        self.t_done: Optional[float] = None
        self.rgba: Optional[torch.Tensor] = None
        self.count: Optional[torch.Tensor] = None
        self.error: Optional[BaseException] = None
        self._done = threading.Event()

    @property
    def rays(self) -> int:
        return self.width * self.height

    @property
    def out_bytes(self) -> int:
        return self.rays * _OUT_BYTES_PER_PIXEL

    @property
    def latency(self) -> float:
        return (self.t_done - self.t_submit) if self.t_done is not None else float('nan')

    def done(self) -> bool:
        return self._done.is_set()

    def result(self, timeout: Optional[float] = None):
        """(rgb with white background [H, W, 3], transmittance [H, W], count [H, W])."""
        if not self._done.wait(timeout):
            raise TimeoutError('frame not rendered in time')
        if self.error is not None:
            raise self.error
        rgba = self.rgba
        return rgba[..., :3] + rgba[..., 3:], rgba[..., 3], self.count


@dataclasses.dataclass
class _Launch:
    event: torch.cuda.Event
    requests: List[FrameRequest]
    nbytes: int


class RenderServer:
    def __init__(self, device=None, hbm_budget: Optional[int] = None, n_streams: int = 2,
                 max_requests: int = _lib.MAX_REQUESTS, max_rays: int = 1 << 23,
                 split_rays: Optional[int] = 1 << 18):
        self.device = torch.device(device or 'cuda:0')
        if hbm_budget is None:
            free, _total = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(free * 0.9)
        self.hbm_budget = int(hbm_budget)
        self.max_requests = max(1, min(int(max_requests), _lib.MAX_REQUESTS))
        self.max_rays = int(max_rays)
        # frames above split_rays rays (512x512) and frames up to it go to separate launches, so
        # a small frame never waits for a large one's launch to finish; None mixes them
        self.split_rays = None if split_rays is None else int(split_rays)
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(max(1, n_streams))]
        self.scenes: Dict[str, ServedScene] = {}
        self._resident = 0
        self._inflight_bytes = 0
        self._pending: Deque[FrameRequest] = collections.deque()
        self._inflight: List[_Launch] = []
        self._cv = threading.Condition()
        self._stop = False
        self._next_stream = 0
        self.launches = 0
        self.completed: List[FrameRequest] = []
        self._thread = threading.Thread(target=self._loop, name='diver-render-loop', daemon=True)
        self._retirer = threading.Thread(target=self._retire_loop, name='diver-retire', daemon=True)
        self._thread.start()
        self._retirer.start()

    # ---- scenes ---------------------------------------------------------------------------
    def add_scene(self, name: str, grid: SceneGrid, decoder: Decoder, precision: str = 'bf16',
                  traversal: str = 'auto', t_stop: float = 1e-4) -> ServedScene:
        nbytes = grid.nbytes + sum(t.numel() * t.element_size()
                                   for t in (grid.blocks, grid.bricks) if t is not None)
        nbytes += decoder.nbytes
        with self._cv:
            if name in self.scenes:
                raise ValueError(f'scene {name!r} is already resident')
            if self._resident + nbytes > self.hbm_budget:
                raise AdmissionError(
                    f'scene {name!r} needs {nbytes} bytes; {self.hbm_budget - self._resident} '
                    f'of the {self.hbm_budget}-byte budget are free')
            sc = ServedScene(name, grid, decoder, precision, traversal, t_stop, nbytes)
            self.scenes[name] = sc
            self._resident += nbytes
        return sc

    # ---- requests -------------------------------------------------------------------------
    def submit(self, scene: str, c2w, width: int, height: int, fov: float = 0.69) -> FrameRequest:
        req = FrameRequest(scene, c2w, width, height, fov)
        with self._cv:
            if scene not in self.scenes:
                raise KeyError(f'unknown scene {scene!r}')
            if self._resident + req.out_bytes > self.hbm_budget:
                raise AdmissionError(f'a {width}x{height} frame needs {req.out_bytes} bytes of '
                                     f'output; {self.hbm_budget - self._resident} are free')
            if req.rays > self.max_rays:
                raise AdmissionError(f'{req.rays} rays exceed the per-launch limit {self.max_rays}')
            self._pending.append(req)
            self._cv.notify_all()
        return req

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        self._thread.join()
        self._retirer.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- render loop ----------------------------------------------------------------------
    def _size_class(self, req: FrameRequest) -> bool:
        return self.split_rays is not None and req.rays > self.split_rays

    def _take_batch(self) -> List[FrameRequest]:
        """Pending requests of the oldest request's scene and size class that fit the launch and
        memory limits (FIFO within the scene).  Called with the lock held."""
        if not self._pending:
            return []
        scene = self._pending[0].scene
        size_class = self._size_class(self._pending[0])
        free = self.hbm_budget - self._resident - self._inflight_bytes
        batch, rays, nbytes = [], 0, 0
        for req in list(self._pending):
            if req.scene != scene or self._size_class(req) != size_class:
                continue
            if (len(batch) == self.max_requests or rays + req.rays > self.max_rays or
                    nbytes + req.out_bytes > free):
                break
            batch.append(req)
            rays += req.rays
            nbytes += req.out_bytes
        for req in batch:
            self._pending.remove(req)
        return batch

    def _finished(self):
        """Nothing pending or in flight after close().  Called with the lock held."""
        return self._stop and not self._pending and not self._inflight

    def _launch(self, batch: List[FrameRequest]):
        sc = self.scenes[batch[0].scene]
        stream = self.streams[self._next_stream % len(self.streams)]
        self._next_stream += 1
        t = time.perf_counter()
        traversal = ('realtime' if sc.grid.octrees is not None else 'offline') \
            if sc.traversal == 'auto' else sc.traversal
        opts = calls.render_opts(sc.precision, traversal, sc.t_stop, 0.0)
        reqs = [dict(c2w=r.c2w, width=r.width, height=r.height, focal=focal_for(r.width, r.fov))
                for r in batch]
        with torch.cuda.stream(stream):
            try:
                outs = calls.render_batch(sc.grid, sc.decoder, reqs, opts)
            except Exception as e:  # a refused launch fails its requests, not the server
                for r in batch:
                    r.error = e
                    r.t_done = time.perf_counter()
                    r._done.set()
                return
            ev = torch.cuda.Event()
            ev.record(stream)
        for r, (rgba, count) in zip(batch, outs):
            r.rgba, r.count, r.t_launch = rgba, count, t
        nbytes = sum(r.out_bytes for r in batch)
        self._inflight.append(_Launch(ev, batch, nbytes))
        self._inflight_bytes += nbytes
        self.launches += 1
        self._cv.notify_all()  # the retire thread waits for a launch to watch

    def _loop(self):
        """Launch thread: packs admissible requests whenever a stream slot is free."""
        torch.cuda.set_device(self.device)
        with self._cv:
            while not self._finished():
                batch = self._take_batch() if len(self._inflight) < len(self.streams) else []
                if batch:
                    self._launch(batch)
                else:
                    # woken by an arrival, a completed launch (slot and memory freed) or close()
                    self._cv.wait()

    def _retire_loop(self):
        """Retire thread: blocks on the oldest launch's event (the GIL is released while it
        waits), then hands its frames to their clients and frees its slot and memory."""
        torch.cuda.set_device(self.device)
        while True:
            with self._cv:
                while not self._inflight:
                    if self._finished():
                        return
                    self._cv.wait()
                launch = self._inflight[0]
            launch.event.synchronize()
            now = time.perf_counter()
            with self._cv:
                self._inflight.remove(launch)
                self._inflight_bytes -= launch.nbytes
                for req in launch.requests:
                    req.t_done = now
                    self.completed.append(req)
                    req._done.set()
                self._cv.notify_all()

    # ---- statistics -----------------------------------------------------------------------
    def stats(self, since: float = 0.0) -> dict:
        reqs = [r for r in self.completed if r.t_submit >= since and r.error is None]
        if not reqs:
            return {'frames': 0}
        lat = np.array([r.latency for r in reqs]) * 1e3
        t0 = min(r.t_submit for r in reqs)
        t1 = max(r.t_done for r in reqs)
        rays = sum(r.rays for r in reqs)
        by_res = collections.defaultdict(list)
        for r in reqs:
            by_res[f'{r.width}x{r.height}'].append(r.latency * 1e3)
        return {
            'frames': len(reqs), 'launches': self.launches, 'rays': rays,
            'wall_s': t1 - t0, 'rays_per_s': rays / max(t1 - t0, 1e-9),
            'p50_ms': float(np.percentile(lat, 50)), 'p99_ms': float(np.percentile(lat, 99)),
            'by_resolution': {k: {'frames': len(v), 'p50_ms': float(np.percentile(v, 50)),
                                  'p99_ms': float(np.percentile(v, 99))}
                              for k, v in sorted(by_res.items())},
        }
