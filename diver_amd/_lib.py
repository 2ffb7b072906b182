"""ctypes binding of libdiver.so (the C ABI declared in include/diver.h).

The library is loaded from this package directory; if it is missing every call raises
``DiverUnavailable``: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get('DIVER_LIB') or os.path.join(_HERE, 'libdiver.so')

_lib = None

EXPORTS = [
    'diver_version', 'diver_last_error', 'diver_ray_voxel_intersect', 'diver_masked_intersect',
    'diver_aabb_intersect', 'diver_ray_march', 'diver_integrate', 'diver_integrate_backward',
    'diver_composite', 'diver_composite_backward', 'diver_decoder_create', 'diver_decoder_destroy',
    'diver_decoder_param_count', 'diver_mlp_eval', 'diver_render_rays', 'diver_render_camera',
    'diver_camera_rays', 'diver_camera_rays_pinhole', 'diver_block_bits', 'diver_tail_bits',
    'diver_walk_rays', 'diver_render_rows',
    'diver_decoder_update', 'diver_brick_bits', 'diver_render_batch', 'diver_alpha_map',
    'diver_corner_keys', 'diver_integrate_table', 'diver_integrate_table_backward',
    'diver_adam_step', 'diver_decoder_update_range', 'diver_ray_suffix',
    'diver_render_rays_sparsity', 'diver_server_clock', 'diver_server_create',
    'diver_server_destroy', 'diver_server_add_scene', 'diver_server_submit',
    'diver_server_submit_pinhole', 'diver_server_wait',
    'diver_server_take_completed', 'diver_server_history', 'diver_adam_step_rows',
    'diver_adjoint_decoder_forward', 'diver_adjoint_decoder_backward', 'diver_positional_encoding',
    'diver_adjoint_weight_grads', 'diver_adjoint_weight_grads_workspace',
    'diver_adjoint_weight_grads_noalloc', 'diver_render_rays_segments', 'diver_render_rows_segments',
]

# status codes of include/diver.h
DIVER_EBUDGET = 5
DIVER_ETIMEDOUT = 6
DIVER_ESTATE = 7

MAX_REQUESTS = 16


class DiverUnavailable(RuntimeError):
    pass


class DiverError(RuntimeError):
    def __init__(self, fn: str, status: int, msg: str):
        super().__init__(f'{fn} failed (status {status}): {msg}')
        self.status = status


class DecoderDesc(ctypes.Structure):
    _fields_ = [('channels', ctypes.c_int), ('hidden', ctypes.c_int), ('n_pre', ctypes.c_int),
                ('pre_identity', ctypes.c_int), ('density_act', ctypes.c_int),
                ('n_post', ctypes.c_int)]


class Grid(ctypes.Structure):
    _fields_ = [('voxels', ctypes.c_void_p), ('mask', ctypes.c_void_p),
                ('octrees', ctypes.c_void_p), ('blocks', ctypes.c_void_p),
                ('bricks', ctypes.c_void_p), ('n', ctypes.c_int),
                ('channels', ctypes.c_int), ('xyzmin', ctypes.c_float),
                ('voxel_size', ctypes.c_float), ('tail', ctypes.c_void_p)]


class RenderOpts(ctypes.Structure):
    _fields_ = [('precision', ctypes.c_int), ('traversal', ctypes.c_int),
                ('t_stop', ctypes.c_float), ('alpha_skip', ctypes.c_float)]


class ServerConfig(ctypes.Structure):
    _fields_ = [('hbm_budget', ctypes.c_int64), ('n_streams', ctypes.c_int),
                ('max_requests', ctypes.c_int), ('max_rays', ctypes.c_int64),
                ('split_rays', ctypes.c_int64)]


class FrameInfo(ctypes.Structure):
    _fields_ = [('ticket', ctypes.c_int64), ('scene', ctypes.c_int), ('width', ctypes.c_int),
                ('height', ctypes.c_int), ('status', ctypes.c_int), ('launch', ctypes.c_int64),
                ('t_submit', ctypes.c_double), ('t_launch', ctypes.c_double),
                ('t_done', ctypes.c_double)]


class Request(ctypes.Structure):
    _fields_ = [('c2w', ctypes.c_float * 12), ('width', ctypes.c_int), ('height', ctypes.c_int),
                ('focal', ctypes.c_float), ('n_rays', ctypes.c_int64), ('o', ctypes.c_void_p),
                ('d', ctypes.c_void_p), ('rgba', ctypes.c_void_p), ('count', ctypes.c_void_p),
                ('fy', ctypes.c_float), ('cx', ctypes.c_float), ('cy', ctypes.c_float)]


def _declare(L):
    vp, i64, i32, fl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_float
    L.diver_version.restype = ctypes.c_char_p
    L.diver_version.argtypes = []
    L.diver_last_error.restype = ctypes.c_char_p
    L.diver_last_error.argtypes = []
    sig = {
        'diver_ray_voxel_intersect': [i64, i32, fl, fl, fl, fl, vp, vp, vp, vp, vp, vp],
        'diver_masked_intersect': [i64, i32, fl, fl, fl, fl, fl, vp, vp, vp, vp, vp, vp, vp],
        'diver_aabb_intersect': [i64, vp, vp, vp, vp, vp, vp, i32, vp],
        'diver_ray_march': [i64, vp, vp, vp, vp, i32, i64, vp],
        'diver_integrate': [vp, i32, i32, i64, vp, vp, vp, vp],
        'diver_integrate_backward': [i32, i32, i64, vp, vp, vp, vp, vp],
        'diver_composite': [i64, i32, vp, vp, vp, i32, vp, vp, vp],
        'diver_composite_backward': [i64, i32, vp, vp, vp, i32, vp, vp, vp, vp, vp],
        'diver_decoder_create': [ctypes.POINTER(DecoderDesc), ctypes.POINTER(ctypes.c_float), i64,
                                 ctypes.POINTER(vp)],
        'diver_decoder_destroy': [vp],
        'diver_decoder_update': [vp, vp, i64, vp],
        'diver_decoder_update_range': [vp, vp, i64, i64, vp],
        'diver_ray_suffix': [i64, vp, vp, vp, vp, vp, vp, vp],
        'diver_mlp_eval': [vp, i64, vp, vp, vp, i32, vp, vp, vp, fl, fl, vp],
        'diver_render_rays': [ctypes.POINTER(Grid), vp, ctypes.POINTER(RenderOpts), i64, vp, vp,
                              vp, vp, vp],
        'diver_render_rays_sparsity': [ctypes.POINTER(Grid), vp, ctypes.POINTER(RenderOpts), i64,
                                       vp, vp, vp, vp, vp, vp],
        'diver_render_camera': [ctypes.POINTER(Grid), vp, ctypes.POINTER(RenderOpts),
                                ctypes.POINTER(ctypes.c_float), i32, i32, fl, vp, vp, vp],
        'diver_camera_rays': [ctypes.POINTER(ctypes.c_float), i32, i32, fl, vp, vp, vp],
        'diver_camera_rays_pinhole': [ctypes.POINTER(ctypes.c_float), i32, i32, fl, fl, fl, fl, vp, vp,
                                      vp],
        'diver_block_bits': [vp, i32, vp, vp],
        'diver_tail_bits': [vp, i32, vp, vp],
        'diver_brick_bits': [vp, i32, vp, vp],
        'diver_walk_rays': [ctypes.POINTER(Grid), i32, i64, vp, vp, i32, vp, vp, vp],
        'diver_render_rows': [ctypes.POINTER(Grid), vp, ctypes.POINTER(RenderOpts), i64, vp, vp,
                              vp, vp, vp, vp],
        'diver_render_batch': [ctypes.POINTER(Grid), vp, ctypes.POINTER(RenderOpts),
                               ctypes.POINTER(Request), i32, vp],
        'diver_alpha_map': [ctypes.POINTER(Grid), vp, i64, vp, vp, i32, vp, vp],
        'diver_corner_keys': [i32, i64, vp, vp, vp, vp],
        'diver_integrate_table': [vp, i32, i64, vp, vp, vp, vp, vp],
        'diver_integrate_table_backward': [i32, i64, vp, vp, vp, vp, vp, vp],
        'diver_adam_step': [vp, vp, vp, vp, i64, ctypes.c_double, ctypes.c_double, fl, fl, fl, fl,
                            i32, vp],
        'diver_adam_step_rows': [vp, vp, vp, vp, i64, vp, i64, i32, ctypes.c_double,
                                 ctypes.c_double, fl, fl, fl, fl, i32, vp],
        'diver_adjoint_decoder_forward': [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp],
        'diver_adjoint_decoder_backward': [vp, i64, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                           vp, vp, vp, vp],
        'diver_positional_encoding': [i64, vp, vp, vp],
        'diver_adjoint_weight_grads': [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i64,
                                       vp],
        'diver_adjoint_weight_grads_noalloc': [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp],
        'diver_render_rays_segments': [ctypes.POINTER(Grid), vp, ctypes.POINTER(RenderOpts), i64,
                                       vp, vp, vp, vp, vp, vp, i32, vp],
        'diver_render_rows_segments': [ctypes.POINTER(Grid), vp, ctypes.POINTER(RenderOpts), i64,
                                       vp, vp, vp, i32, vp, vp, vp],
    }
    sig.update({
        'diver_server_create': [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)],
        'diver_server_destroy': [vp],
        'diver_server_add_scene': [vp, ctypes.POINTER(Grid), vp, ctypes.POINTER(RenderOpts), i64,
                                   vp, ctypes.POINTER(ctypes.c_int)],
        'diver_server_submit': [vp, i32, ctypes.POINTER(ctypes.c_float), i32, i32, fl, vp, vp, vp,
                                ctypes.POINTER(ctypes.c_int64)],
        'diver_server_submit_pinhole': [vp, i32, ctypes.POINTER(ctypes.c_float), i32, i32, fl, fl, fl,
                                        fl, vp, vp, vp, ctypes.POINTER(ctypes.c_int64)],
        'diver_server_wait': [vp, i64, ctypes.c_double, ctypes.POINTER(FrameInfo)],
    })
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.diver_server_clock.argtypes = []
    L.diver_server_clock.restype = ctypes.c_double
    L.diver_server_take_completed.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), i64]
    L.diver_server_take_completed.restype = ctypes.c_int64
    L.diver_server_history.argtypes = [vp, ctypes.POINTER(FrameInfo), i64,
                                       ctypes.POINTER(ctypes.c_int64)]
    L.diver_server_history.restype = ctypes.c_int64
    L.diver_decoder_param_count.argtypes = [ctypes.POINTER(DecoderDesc)]
    L.diver_decoder_param_count.restype = ctypes.c_int64
    L.diver_adjoint_weight_grads_workspace.argtypes = [vp, i64]
    L.diver_adjoint_weight_grads_workspace.restype = ctypes.c_int64


def lib():
    """The loaded library (raises DiverUnavailable when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DiverUnavailable(
                f'{LIB_PATH} is missing: build it with `python -m diver_amd.build` '
                '(the render path has no CPU fallback)')
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def call(name: str, *args) -> None:
    L = lib()
    st = getattr(L, name)(*args)
    if st != 0:
        raise DiverError(name, st, L.diver_last_error().decode(errors='replace'))


def status_call(name: str, *args) -> tuple[int, str]:
    """Calls an entry point and returns (status, diver_last_error() text if it failed)."""
    L = lib()
    st = getattr(L, name)(*args)
    return st, (L.diver_last_error().decode(errors='replace') if st != 0 else '')


def stream_of(t) -> int:
    """Handle of torch's current stream on the tensor's device."""
    import torch
    return torch.cuda.current_stream(t.device).cuda_stream


def ptr(t) -> int:
    return t.data_ptr() if t is not None else None
