"""Camera conventions of the reference's datasets, turned into the device camera's terms.

The device camera (``diver_render_batch`` requests, ``diver_camera_rays[_pinhole]``) traces pixel
(x, y) along ((x + 0.5 - cx)/fx, -(y + 0.5 - cy)/fy, -1) in camera space, rotated by c2w[:, :3] and
normalised, from c2w[:, 3].  The datasets differ only in how they get there:

- Blender (utils/dataset/blender.py:25-50): c2w is the frame's ``transform_matrix[:3, :4]``;
  fx = fy = 0.5 * 800 / tan(0.5 * camera_angle_x) * W / 800, principal point at the centre.
- Tanks and Temples / BlendedMVS (utils/dataset/tanks.py:27-98, 172-230): intrinsics from
  ``intrinsics.txt`` scaled to the image size, OpenCV poses whose y and z axes are negated
  (``parse_extrinsics``), and ray origins moved into the unit box around the scene's bounding box,
  (o - mid) / scale, with the reference's ``scale = max(xmax - xmin, ymax - ymin, ymax - zmin) / 2``
  kept as written.  Directions are not rescaled, as there.

Everything here runs once per camera on the host; the per-pixel work is the device's.  Files are
read as text or JSON only.
"""
from __future__ import annotations

import json
import math
import os
from typing import Sequence, Union

import numpy as np

PathLike = Union[str, os.PathLike]


# ---- Blender ---------------------------------------------------------------------------------------

def blender_focal(camera_angle_x: float, width: int) -> float:
    """blender.py:28-31: the 800-pixel focal length, scaled to the image width."""
    return 0.5 * 800 / math.tan(0.5 * camera_angle_x) * width / 800


def blender_cameras(meta: Union[PathLike, dict], img_wh: Sequence[int]) -> list:
    """Requests for every frame of a ``transforms_*.json`` (blender.py:25-50, 85-97):
    [{'c2w', 'width', 'height', 'focal'}]."""
    if not isinstance(meta, dict):
        with open(meta, 'r') as f:
            meta = json.load(f)
    w, h = int(img_wh[0]), int(img_wh[1])
    focal = blender_focal(float(meta['camera_angle_x']), w)
    return [{'c2w': np.asarray(fr['transform_matrix'], np.float32)[:3, :4].copy(),
             'width': w, 'height': h, 'focal': float(np.float32(focal))}
            for fr in meta['frames']]


# ---- Tanks and Temples / BlendedMVS ----------------------------------------------------------------

def load_matrix(path: PathLike) -> np.ndarray:
    """tanks.py:183-184: whitespace-separated rows of floats."""
    with open(path, 'r') as f:
        return np.array([[float(w) for w in line.strip().split()] for line in f
                         if line.strip()]).astype(np.float32)


def load_intrinsics(path: PathLike) -> np.ndarray:
    """tanks.py:193-222: a 3x3 or 4x4 matrix, or a first line ``f cx cy _`` (fx = fy = f)."""
    try:
        k = load_matrix(path)
        if k.shape == (3, 3):
            k4 = np.zeros((4, 4), np.float32)
            k4[:3, :3] = k
            k4[3, 3] = 1
            k = k4
        return k
    except ValueError:
        pass
    with open(path, 'r') as f:
        fl, cx, cy, _ = map(float, f.readline().split())
    return np.array([[fl, 0., cx, 0.], [0., fl, cy, 0.], [0., 0., 1., 0.], [0., 0., 0., 1.]])


def parse_intrinsics(k: np.ndarray):
    """tanks.py:186-191: (fx, fy, cx, cy)."""
    return float(k[0, 0]), float(k[1, 1]), float(k[0, 2]), float(k[1, 2])


def scale_intrinsics(fx, fy, cx, cy, dataset: str, img_wh: Sequence[int]):
    """tanks.py:31-42: intrinsics given for 768x576 (BlendedMVS) or 1920x1080 (TanksAndTemple)
    images, scaled to img_wh = (w, h)."""
    w, h = float(img_wh[0]), float(img_wh[1])
    if dataset == 'BlendedMVS':
        return fx * w / 768., fy * h / 576., cx * w / 768., cy * h / 576.
    if dataset == 'TanksAndTemple':
        return fx * w / 1920., fy * h / 1080., cx * w / 1920., cy * h / 1080.
    raise ValueError(f'unknown dataset {dataset!r} (BlendedMVS or TanksAndTemple)')


def parse_extrinsics(extrinsics: np.ndarray, world2camera: bool = False) -> np.ndarray:
    """tanks.py:172-181: a 4x4 pose with the camera's y and z axes negated (OpenCV to the
    -z-forward, y-up camera); inverted first when the file holds world-to-camera."""
    e = np.asarray(extrinsics, np.float32)
    if e.shape == (3, 4):
        e = np.vstack([e, np.array([[0, 0, 0, 1.0]], np.float32)])
    if e.shape == (1, 16):
        e = e.reshape(4, 4)
    if world2camera:
        e = np.linalg.inv(e).astype(np.float32)
    e = e.copy()
    e[:3, 1:3] = -e[:3, 1:3]
    return e


def scene_normalisation(bounds: Sequence[float]):
    """tanks.py:46-52: (mid_point float32 [3], scale) of a bbox (xmin, ymin, zmin, xmax, ymax,
    zmax).  The scale's third term is ymax - zmin, as the reference has it."""
    xmin, ymin, zmin, xmax, ymax, zmax = [float(b) for b in bounds[:6]]
    mid = np.array([(xmin + xmax) / 2, (ymin + ymax) / 2, (zmin + zmax) / 2], np.float32)
    scale = max(xmax - xmin, ymax - ymin, ymax - zmin) * 0.5
    return mid, scale


def load_bounds(path: PathLike):
    """tanks.py:44-45: the first six numbers of ``bbox.txt``."""
    with open(path, 'r') as f:
        return [float(x) for x in f.readlines()[0].strip().split()[:6]]


def tanks_camera(pose: np.ndarray, intrinsics, bounds: Sequence[float], img_wh: Sequence[int]):
    """One request of a Tanks and Temples / BlendedMVS view: pose is the file's camera-to-world
    matrix (OpenCV axes), intrinsics the already scaled (fx, fy, cx, cy).  The returned c2w has the
    converted axes and the normalised origin (o - mid) / scale in FP32, as tanks.py:86-98 forms
    each ray's origin; directions are those of get_ray_directions / get_rays (tanks.py:120-169)."""
    c2w = parse_extrinsics(pose)[:3].astype(np.float32)
    mid, scale = scene_normalisation(bounds)
    c2w[:, 3] = (c2w[:, 3] - mid) / np.float32(scale)
    fx, fy, cx, cy = intrinsics
    return {'c2w': c2w, 'width': int(img_wh[0]), 'height': int(img_wh[1]),
            'focal': float(np.float32(fx)), 'fy': float(np.float32(fy)),
            'cx': float(np.float32(cx)), 'cy': float(np.float32(cy))}


def tanks_cameras(root: PathLike, img_wh: Sequence[int] = (1920, 1080), split: str = 'test'):
    """Requests for a dataset folder laid out as tanks.py:27-71 reads it (intrinsics.txt,
    bbox.txt, rgb/, pose/; train/extra views are the files starting with '0', test ones with
    '1'); the dataset name is the folder's parent, as there."""
    root = os.fspath(root)
    dataset = os.path.basename(os.path.dirname(os.path.normpath(root)))
    intr = scale_intrinsics(*parse_intrinsics(load_intrinsics(os.path.join(root, 'intrinsics.txt'))),
                            dataset, img_wh)
    bounds = load_bounds(os.path.join(root, 'bbox.txt'))
    prefix = '0' if split in ('train', 'extra') else '1'
    names = sorted(f for f in os.listdir(os.path.join(root, 'rgb')) if f[0] == prefix)
    return [tanks_camera(load_matrix(os.path.join(root, 'pose', f.replace('.png', '.txt'))), intr,
                         bounds, img_wh) for f in names]
