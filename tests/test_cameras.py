"""Dataset camera conventions on the host (diver_amd.cameras): Tanks and Temples / BlendedMVS
(utils/dataset/tanks.py:27-98, 120-230) and Blender (utils/dataset/blender.py:25-50).

Pinned by closed forms rather than by retyping the formulas: an identity OpenCV pose turns into a
camera looking down world +z with image rows going to world +y (OpenCV's x right, y down, z
forward); the pixel at the principal point looks straight down the camera axis; the Blender focal
length is 0.5 W / tan(fov / 2); the origin normalisation maps the bounding box's centre to 0 and its
widest axis onto [-1, 1] (with the reference's ymax - zmin term where it decides the scale).
"""
import math

import numpy as np
import pytest

from diver_amd import cameras as C


def _cam_dir(req, x, y):
    """Camera ray direction of pixel (x, y) in FP64 from a request's c2w and intrinsics."""
    fx = req['focal']
    fy = req.get('fy', fx)
    cx = req.get('cx', req['width'] / 2)
    cy = req.get('cy', req['height'] / 2)
    v = np.array([(x + 0.5 - cx) / fx, -(y + 0.5 - cy) / fy, -1.0])
    d = req['c2w'][:, :3].astype(np.float64) @ v
    return d / np.linalg.norm(d)


def test_parse_extrinsics_converts_opencv_axes():
    e = C.parse_extrinsics(np.eye(4, dtype=np.float32))
    np.testing.assert_array_equal(e[:3, :3], np.diag([1.0, -1.0, -1.0]).astype(np.float32))
    # 3x4 input, and a world-to-camera pose is inverted first
    t = np.array([[0, -1, 0, 2.0], [1, 0, 0, -1.0], [0, 0, 1, 0.5]], np.float32)
    e34 = C.parse_extrinsics(t)
    assert e34.shape == (4, 4) and np.array_equal(e34[3], [0, 0, 0, 1])
    w2c = np.linalg.inv(np.vstack([t, [0, 0, 0, 1]])).astype(np.float32)
    np.testing.assert_allclose(C.parse_extrinsics(w2c, world2camera=True), e34, atol=1e-6)


def test_tanks_camera_looks_along_opencv_forward():
    intr = (1000.0, 1000.0, 959.5, 539.5)  # principal point on a pixel centre
    req = C.tanks_camera(np.eye(4, dtype=np.float32), intr, (-1, -1, -1, 1, 1, 1), (1920, 1080))
    np.testing.assert_allclose(_cam_dir(req, 959, 539), [0, 0, 1], atol=1e-12)
    below = _cam_dir(req, 959, 1000)    # a lower image row: +y (OpenCV's y points down)
    right = _cam_dir(req, 1500, 539)    # a column to the right: +x
    assert below[1] > 0 and abs(below[0]) < 1e-12 and below[2] > 0
    assert right[0] > 0 and abs(right[1]) < 1e-12 and right[2] > 0


def test_origin_normalisation():
    # bbox centre (1, 2, 3); widest extent 4 (x) -> scale 2
    bounds = (-1, 1, 2, 3, 3, 4)
    mid, scale = C.scene_normalisation(bounds)
    np.testing.assert_array_equal(mid, np.float32([1, 2, 3]))
    assert scale == 2.0
    pose = np.eye(4, dtype=np.float32)
    pose[:3, 3] = [3, 2, 3]  # on the +x face of the box
    req = C.tanks_camera(pose, (10., 10., 5., 5.), bounds, (10, 10))
    np.testing.assert_array_equal(req['c2w'][:, 3], np.float32([1, 0, 0]))
    # the reference's third term is ymax - zmin: with zmin far below, it sets the scale
    _, s2 = C.scene_normalisation((0, 0, -10, 1, 1, 1))
    assert s2 == (1 - (-10)) * 0.5


def test_intrinsics_files_and_scaling(tmp_path):
    k3 = tmp_path / 'k3.txt'
    k3.write_text('1100 0 960\n0 1090 540\n0 0 1\n')
    k = C.load_intrinsics(k3)
    assert k.shape == (4, 4) and k[3, 3] == 1
    assert C.parse_intrinsics(k) == (1100.0, 1090.0, 960.0, 540.0)
    # the ragged DeepVoxels-style file (focal, centre, then other rows): only its first line is read
    kf = tmp_path / 'kf.txt'
    kf.write_text('581.5 384 288 0\n0. 0. 0.\n1.\n576 768\n')
    assert C.parse_intrinsics(C.load_intrinsics(kf)) == (581.5, 581.5, 384.0, 288.0)
    # intrinsics given for the native size scale with the image
    assert C.scale_intrinsics(1100, 1090, 960, 540, 'TanksAndTemple', (960, 540)) == \
        (550.0, 545.0, 480.0, 270.0)
    assert C.scale_intrinsics(581.5, 581.5, 384, 288, 'BlendedMVS', (768 * 2, 576 * 2)) == \
        (1163.0, 1163.0, 768.0, 576.0)
    with pytest.raises(ValueError):
        C.scale_intrinsics(1, 1, 1, 1, 'Blender', (1, 1))


def test_tanks_dataset_folder(tmp_path):
    root = tmp_path / 'TanksAndTemple' / 'Barn'
    (root / 'rgb').mkdir(parents=True)
    (root / 'pose').mkdir()
    (root / 'intrinsics.txt').write_text('1100 0 960\n0 1100 540\n0 0 1\n')
    (root / 'bbox.txt').write_text('-2 -2 -2 2 2 2 0.01\n')
    for name in ('0_000.png', '1_000.png', '1_001.png'):
        (root / 'rgb' / name).write_bytes(b'')
        pose = np.eye(4, dtype=np.float32)
        pose[0, 3] = float(name[2:5]) + 1.0
        np.savetxt(root / 'pose' / name.replace('.png', '.txt'), pose)
    reqs = C.tanks_cameras(root, (960, 540), 'test')
    assert len(reqs) == 2
    assert reqs[0]['focal'] == 550.0 and reqs[0]['fy'] == 550.0
    assert reqs[0]['cx'] == 480.0 and reqs[0]['cy'] == 270.0
    np.testing.assert_array_equal(reqs[1]['c2w'][:, 3], np.float32([1.0, 0, 0]))  # (2 - 0) / 2
    assert len(C.tanks_cameras(root, (960, 540), 'train')) == 1


def test_blender_cameras():
    fov = 0.6911112070083618
    meta = {'camera_angle_x': fov,
            'frames': [{'transform_matrix': np.eye(4).tolist()},
                       {'transform_matrix': np.eye(4)[[1, 0, 2, 3]].tolist()}]}
    reqs = C.blender_cameras(meta, (400, 400))
    assert len(reqs) == 2 and reqs[0]['c2w'].shape == (3, 4)
    # the half-angle of the horizontal field of view is fov / 2
    assert abs(math.atan(200 / reqs[0]['focal']) - fov / 2) < 1e-6
    assert abs(C.blender_focal(fov, 800) - 0.5 * 800 / math.tan(fov / 2)) < 1e-9
    np.testing.assert_allclose(_cam_dir(reqs[0], 199.5, 199.5), [0, 0, -1], atol=1e-12)
