"""CPU ORACLE -- test infrastructure only.

Imported only by tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg, always as the
checker.  The product package (``diver_amd``) never imports it.

- ``traverse.c`` (built to ``liboracle.so`` by the Makefile): FP32 restatement of the reference's
  traversal kernels, wrapped here with numpy-facing helpers;
- ``fp64.py``: float64 restatements of the integral, decoders and compositing;
- ``cpu_path.py``: the reference's offline render path restated on the CPU with torch, FP32
  (timed as the CPU baseline).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    subprocess.run(['make', '-s', '-C', _HERE], check=True)
    return os.path.join(_HERE, 'liboracle.so')


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'liboracle.so')
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        f32p = ctypes.POINTER(ctypes.c_float)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        i32p = ctypes.POINTER(ctypes.c_int32)
        i64 = ctypes.c_int64
        fl = ctypes.c_float
        L.oracle_ray_voxel_intersect.argtypes = [i64, ctypes.c_int, fl, fl, fl, fl, f32p, f32p,
                                                 f32p, i32p, f32p]
        L.oracle_masked_intersect.argtypes = [i64, ctypes.c_int, fl, fl, fl, fl, fl, f32p, f32p,
                                              u8p, f32p, i32p, f32p]
        L.oracle_aabb_intersect.argtypes = [i64, f32p, f32p, f32p, u8p, f32p, u8p, ctypes.c_int]
        L.oracle_ray_march.argtypes = [i64, f32p, f32p, u8p, u8p, ctypes.c_int, i64]
        L.oracle_rt_walk.argtypes = [i64, ctypes.c_int, fl, fl, ctypes.c_int, f32p, f32p, u8p,
                                     u8p, f32p, i32p]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def ray_voxel_intersect(o, v, xyzmin, xyzmax, voxel_num, voxel_size, max_n=None):
    """Reference K1 on the CPU.  Returns untrimmed (isect [B,max_n,3], num [B], tns [B,max_n])
    initialised like the reference glue (zeros, zeros, -1)."""
    o, v = _f32(o), _f32(v)
    B = o.shape[0]
    max_n = int(voxel_num * 3) if max_n is None else max_n
    isect = np.zeros((B, max_n, 3), np.float32)
    num = np.zeros(B, np.int32)
    tns = np.full((B, max_n), -1.0, np.float32)
    lib().oracle_ray_voxel_intersect(B, max_n, xyzmin, xyzmax, float(voxel_num), voxel_size,
                                     _p(o, ctypes.c_float), _p(v, ctypes.c_float),
                                     _p(isect, ctypes.c_float), _p(num, ctypes.c_int32),
                                     _p(tns, ctypes.c_float))
    return isect, num, tns


def masked_intersect(o, v, xyzmin, xyzmax, voxel_num, voxel_size, mask, mask_scale, max_n=None):
    """Reference K2 on the CPU.  Returns untrimmed (isect [B,max_n,6], num [B], tns [B,max_n])."""
    o, v = _f32(o), _f32(v)
    mask = np.ascontiguousarray(mask, np.uint8)
    B = o.shape[0]
    max_n = int(voxel_num * 3) if max_n is None else max_n
    isect = np.zeros((B, max_n, 6), np.float32)
    num = np.zeros(B, np.int32)
    tns = np.full((B, max_n), -1.0, np.float32)
    lib().oracle_masked_intersect(B, max_n, xyzmin, xyzmax, float(voxel_num), voxel_size,
                                  float(mask_scale), _p(o, ctypes.c_float), _p(v, ctypes.c_float),
                                  _p(mask, ctypes.c_uint8), _p(isect, ctypes.c_float),
                                  _p(num, ctypes.c_int32), _p(tns, ctypes.c_float))
    return isect, num, tns


def aabb_intersect(coords, v, center, finish, rgba, octrees, voxel_num):
    """Reference K3 on the CPU, in place on numpy buffers."""
    B = v.shape[0] if v.ndim == 2 else v.shape[0] * v.shape[1]
    lib().oracle_aabb_intersect(B, _p(coords, ctypes.c_float), _p(_f32(v), ctypes.c_float),
                                _p(_f32(center), ctypes.c_float), _p(finish, ctypes.c_uint8),
                                _p(rgba, ctypes.c_float),
                                _p(np.ascontiguousarray(octrees, np.uint8), ctypes.c_uint8),
                                int(voxel_num))


REF_RAY_MARCH_STRIDE = 256 * 1024  # NUM_BLOCKS * NUM_THREADS (ray_march.cu:7-8,278)


def ray_march(coords, v, mask, finish, voxel_num, stride=REF_RAY_MARCH_STRIDE):
    """Reference K4 (one round) on the CPU, in place on ``coords``."""
    B = v.shape[0] if v.ndim == 2 else v.shape[0] * v.shape[1]
    lib().oracle_ray_march(B, _p(coords, ctypes.c_float), _p(_f32(v), ctypes.c_float),
                           _p(np.ascontiguousarray(mask, np.uint8), ctypes.c_uint8),
                           _p(np.ascontiguousarray(finish, np.uint8), ctypes.c_uint8),
                           int(voxel_num), int(stride))


def rt_walk(o, v, xyzmin, voxel_size, voxel_num, octrees, mask, max_n=None):
    """Whole real-time walk (aabb_intersect + ray_march rounds) per ray, own origin per ray.
    Returns (isect [B,max_n,6], num [B]) in masked_intersect's layout."""
    o, v = _f32(o), _f32(v)
    B = o.shape[0]
    max_n = 3 * voxel_num if max_n is None else max_n
    isect = np.zeros((B, max_n, 6), np.float32)
    num = np.zeros(B, np.int32)
    lib().oracle_rt_walk(B, max_n, np.float32(xyzmin), np.float32(voxel_size), int(voxel_num),
                         _p(o, ctypes.c_float), _p(v, ctypes.c_float),
                         _p(np.ascontiguousarray(octrees, np.uint8), ctypes.c_uint8),
                         _p(np.ascontiguousarray(mask, np.uint8), ctypes.c_uint8),
                         _p(isect, ctypes.c_float), _p(num, ctypes.c_int32))
    return isect, num


def octree_levels(mask: np.ndarray) -> np.ndarray:
    """eval_rt.py:236-239: max-pooled occupancy at block sizes 64, 16, 4, coarse first."""
    n = mask.shape[0]
    out = []
    for s in (64, 16, 4):
        m = n // s
        out.append(mask.reshape(m, s, m, s, m, s).any(axis=(1, 3, 5)).reshape(-1))
    return np.concatenate(out).astype(np.uint8)
