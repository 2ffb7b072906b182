"""Continuous-batching render server on one GPU: a Python face over the native serving loop
(``diver_server_*`` in include/diver.h).

Scenes (a SceneGrid plus a Decoder) stay resident in device memory.  Clients submit frame requests
(scene, camera, resolution) from any thread; the library's launch thread packs the pending frames
of one scene (FIFO, up to ``max_requests`` frames and ``max_rays`` ray ids; frames above
``split_rays`` pixels apart from smaller ones) into one ``diver_render_batch`` launch and keeps up
to ``n_streams`` launches in flight on streams of its own.  Frames that arrive while launches run
join the next launch as soon as a stream and memory are free, so their work overlaps the frames
already on the GPU.  The library's retire thread watches every launch's completion event and
completes its frames the moment the GPU finishes them.  Neither thread holds the GIL: a client's
``result()`` blocks inside the library, and the per-launch host work is native.

Memory admission: every resident scene and the output buffers of the frames on the GPU are
charged against ``hbm_budget`` bytes.  A scene that does not fit is refused at ``add_scene``; a
frame whose output could never fit next to the resident scenes is refused at ``submit``; a frame
that fits only later waits in the queue until launches complete.

A ray's result depends only on its own intervals, so a frame rendered in a mixed launch is
bit-identical to the same frame rendered alone.

The reference has no serving layer (its real-time viewer renders one camera at a time,
eval_rt.py:106-135); this module is the serving path built around the same renderer.
"""
from __future__ import annotations

import collections
import ctypes
import dataclasses
import threading
from typing import Dict, Optional

import numpy as np
import torch

from . import _lib, calls
from .decoder import Decoder
from .renderer import SceneGrid, focal_for

_OUT_BYTES_PER_PIXEL = 16 + 4  # rgba fp32 + int32 interval count


def clock() -> float:
    """Seconds on the clock frames are stamped with (CLOCK_MONOTONIC, as time.perf_counter)."""
    return _lib.lib().diver_server_clock()


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
    scene_id: int = -1


class FrameRequest:
    """One frame; ``result()`` waits for it.  ``rgba`` [H, W, 4] and ``count`` [H, W] are the
    device buffers the library renders into."""

    def __init__(self, server: 'RenderServer', scene: str, ticket: int, width: int, height: int,
                 rgba: torch.Tensor, count: torch.Tensor):
        self._server = server
        self.scene = scene
        self.ticket = ticket
        self.width = int(width)
        self.height = int(height)
        self.rgba = rgba
        self.count = count
        self.t_submit: Optional[float] = None
        self.t_launch: Optional[float] = None
        self.t_done: Optional[float] = None
        self.launch: Optional[int] = None
        self.error: Optional[BaseException] = None
        self._done = False
        self._lock = threading.Lock()

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
        return self._done

    def _wait(self, timeout: Optional[float]):
        with self._lock:
            if self._done:
                return
            info = _lib.FrameInfo()
            st, msg = _lib.status_call('diver_server_wait', self._server._handle, self.ticket,
                                       -1.0 if timeout is None else float(timeout),
                                       ctypes.byref(info))
            if st == _lib.DIVER_ETIMEDOUT:
                raise TimeoutError('frame not rendered in time')
            if st != 0 and info.ticket != self.ticket:
                # the wait itself failed (no record of the frame)
                self.error = _lib.DiverError('diver_server_wait', st, msg)
            else:
                self.t_submit, self.t_launch, self.t_done = info.t_submit, info.t_launch, info.t_done
                self.launch = info.launch
                if st != 0:
                    self.error = _lib.DiverError('diver_server_wait', st, msg)
            self._done = True
            self._server._release(self.ticket)

    def result(self, timeout: Optional[float] = None):
        """(rgb with white background [H, W, 3], transmittance [H, W], count [H, W])."""
        self._wait(timeout)
        if self.error is not None:
            raise self.error
        rgba = self.rgba
        return rgba[..., :3] + rgba[..., 3:], rgba[..., 3], self.count


class RenderServer:
    """``max_rays`` (ray ids per launch; a camera takes 64 per 8x8 tile) defaults to 2^21: one
    1920x1080 frame, three 800x800 or sixteen 200x200 frames per launch.  Larger launches add
    latency without adding throughput once two launches are in flight (measured at 800x800:
    2^21 gives 2.79 G rays/s saturated and a 1.7 ms p99 at 2400 frames/s, 2^23 gives 2.62 G
    rays/s and stalls the clients' output allocations).

    ``n_streams`` defaults to 3.  HIP maps streams onto a few hardware queues (4 per process by
    default) and shares a queue between streams once the process has more streams than queues;
    kernels on one queue run one after another.  Once torch's stream pool exists (any
    ``torch.cuda.Stream()``), two server streams were measured to share a queue: 800x800 frames
    saturated at 2.15 G rays/s with two streams and 2.67 G with three (2.56 G and 2.70 G in a
    process without the pool)."""

    def __init__(self, device=None, hbm_budget: Optional[int] = None, n_streams: int = 3,
                 max_requests: int = _lib.MAX_REQUESTS, max_rays: int = 1 << 21,
                 split_rays: Optional[int] = 1 << 18):
        self.device = torch.device(device or 'cuda:0')
        if hbm_budget is None:
            free, _total = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(free * 0.9)
        self.hbm_budget = int(hbm_budget)
        cfg = _lib.ServerConfig(self.hbm_budget, int(n_streams),
                                max(1, min(int(max_requests), _lib.MAX_REQUESTS)), int(max_rays),
                                0 if split_rays is None else int(split_rays))
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.call('diver_server_create', ctypes.byref(cfg), ctypes.byref(handle))
        self._handle = handle.value
        self.scenes: Dict[str, ServedScene] = {}
        self._lock = threading.Lock()
        # output buffers of frames the library may still write (dropped once complete)
        self._live: Dict[int, FrameRequest] = {}
        self._closed = False

    # ---- scenes ---------------------------------------------------------------------------
    def add_scene(self, name: str, grid: SceneGrid, decoder: Decoder, precision: str = 'bf16',
                  traversal: str = 'auto', t_stop: float = 1e-4) -> ServedScene:
        nbytes = grid.nbytes + sum(t.numel() * t.element_size()
                                   for t in (grid.blocks, grid.bricks) if t is not None)
        nbytes += decoder.nbytes
        walk = ('realtime' if grid.octrees is not None else 'offline') \
            if traversal == 'auto' else traversal
        opts = calls.render_opts(precision, walk, t_stop, 0.0)
        g = calls._grid_of(grid)
        sid = ctypes.c_int(-1)
        with self._lock:
            if name in self.scenes:
                raise ValueError(f'scene {name!r} is already resident')
            with torch.cuda.device(self.device):
                # the grid's structures were built on the current stream: launches wait for them
                st, msg = _lib.status_call('diver_server_add_scene', self._handle, ctypes.byref(g),
                                           decoder.handle, ctypes.byref(opts), nbytes,
                                           torch.cuda.current_stream(self.device).cuda_stream,
                                           ctypes.byref(sid))
            if st == _lib.DIVER_EBUDGET:
                raise AdmissionError(f'scene {name!r}: {msg}')
            if st != 0:
                raise _lib.DiverError('diver_server_add_scene', st, msg)
            sc = ServedScene(name, grid, decoder, precision, traversal, t_stop, nbytes, sid.value)
            self.scenes[name] = sc
        return sc

    # ---- requests -------------------------------------------------------------------------
    def submit(self, scene: str, c2w, width: int, height: int, fov: float = 0.69,
               out=None, intrinsics=None) -> FrameRequest:
        """Queue a frame.  ``out`` = (rgba [H, W, 4] float32, count [H, W] int32) device buffers to
        render into (a client that reuses its buffers); by default new ones are allocated.
        ``intrinsics`` = (fx, fy, cx, cy) selects a general pinhole camera (Tanks and Temples,
        BlendedMVS: diver_amd.cameras) instead of the Blender camera of ``fov``."""
        sc = self.scenes.get(scene)
        if sc is None:
            raise KeyError(f'unknown scene {scene!r}')
        c2w = np.ascontiguousarray(np.asarray(c2w, np.float32).reshape(3, 4))
        with torch.cuda.device(self.device):
            if out is None:
                rgba = torch.empty((height, width, 4), dtype=torch.float32, device=self.device)
                count = torch.empty((height, width), dtype=torch.int32, device=self.device)
            else:
                rgba, count = out
                if (tuple(rgba.shape) != (height, width, 4) or rgba.dtype != torch.float32
                        or tuple(count.shape) != (height, width) or count.dtype != torch.int32
                        or not rgba.is_contiguous() or not count.is_contiguous()
                        or rgba.device != self.device or count.device != self.device):
                    raise ValueError('out must be contiguous rgba [H, W, 4] float32 and count '
                                     '[H, W] int32 tensors on the server\'s device')
            ticket = ctypes.c_int64(0)
            m = c2w.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            # the buffers' memory may have been used by work still queued on the current stream
            stream = torch.cuda.current_stream(self.device).cuda_stream
            if intrinsics is None:
                st, msg = _lib.status_call(
                    'diver_server_submit', self._handle, sc.scene_id, m, int(width), int(height),
                    float(focal_for(width, fov)), rgba.data_ptr(), count.data_ptr(), stream,
                    ctypes.byref(ticket))
            else:
                fx, fy, cx, cy = (float(v) for v in intrinsics)
                st, msg = _lib.status_call(
                    'diver_server_submit_pinhole', self._handle, sc.scene_id, m, int(width),
                    int(height), fx, fy, cx, cy, rgba.data_ptr(), count.data_ptr(), stream,
                    ctypes.byref(ticket))
        if st == _lib.DIVER_EBUDGET:
            raise AdmissionError(msg)
        if st != 0:
            raise _lib.DiverError('diver_server_submit', st, msg)
        req = FrameRequest(self, scene, ticket.value, width, height, rgba, count)
        with self._lock:
            self._live[req.ticket] = req
        self._reap()
        return req

    def _release(self, ticket: int):
        with self._lock:
            self._live.pop(ticket, None)

    def _reap(self):
        """Drops the server's references to frames the library has completed (the clients keep
        theirs)."""
        buf = (ctypes.c_int64 * 256)()
        while True:
            n = _lib.lib().diver_server_take_completed(self._handle, buf, 256)
            with self._lock:
                for i in range(n):
                    self._live.pop(buf[i], None)
            if n < 256:
                break

    @property
    def launches(self) -> int:
        n = ctypes.c_int64(0)
        _lib.lib().diver_server_history(self._handle, None, 0, ctypes.byref(n))
        return n.value

    def close(self):
        if self._closed:
            return
        self._closed = True
        # every frame already submitted is rendered; its record is read before the library's
        # state goes away, so result() still answers after close()
        with self._lock:
            live = list(self._live.values())
        for req in live:
            try:
                req._wait(None)
            except Exception:  # noqa: BLE001 -- the error stays on the request
                pass
        st, msg = _lib.status_call('diver_server_destroy', self._handle)
        self._handle = None
        self._live.clear()
        if st != 0:
            raise _lib.DiverError('diver_server_destroy', st, msg)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- statistics -----------------------------------------------------------------------
    def history(self):
        """Records of the last completed frames (up to 65536), oldest first."""
        buf = (_lib.FrameInfo * 65536)()
        n = _lib.lib().diver_server_history(self._handle, buf, 65536, None)
        return [buf[i] for i in range(n)]

    def stats(self, since: float = 0.0) -> dict:
        recs = [r for r in self.history() if r.t_submit >= since and r.status == 0]
        if not recs:
            return {'frames': 0}
        lat = np.array([r.t_done - r.t_submit for r in recs]) * 1e3
        t0 = min(r.t_submit for r in recs)
        t1 = max(r.t_done for r in recs)
        rays = sum(r.width * r.height for r in recs)
        by_res = collections.defaultdict(list)
        for r, ms in zip(recs, lat):
            by_res[f'{r.width}x{r.height}'].append(ms)
        return {
            'frames': len(recs), 'launches': self.launches, 'rays': rays,
            'wall_s': t1 - t0, 'rays_per_s': rays / max(t1 - t0, 1e-9),
            'p50_ms': float(np.percentile(lat, 50)), 'p99_ms': float(np.percentile(lat, 99)),
            'by_resolution': {k: {'frames': len(v), 'p50_ms': float(np.percentile(v, 50)),
                                  'p99_ms': float(np.percentile(v, 99))}
                              for k, v in sorted(by_res.items())},
        }
