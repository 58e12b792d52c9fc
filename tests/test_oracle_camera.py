"""Pins of the oracle's pinhole-camera ray generator (march_ref.c
ref_camera_rays, restating cam2world_ray, render_util.cuh:526-544) without
the reference: closed forms that fix the pixel centre, the c2w row/column
convention and the origin, and an FP64 evaluation of the camera model
(svox2.py:157-183) that bounds the float32 rounding.

The mistakes these catch: a pixel centre off by half a pixel (the +0.5),
fx/fy or cx/cy swapped (the cameras below have fx != fy, cx != cy and a
non-square image), a transposed rotation, a wrong origin column, and a
missing normalisation."""
import numpy as np
import pytest
import torch

import oracle.oracle as O
from svox2_amd import scenes
import svox2_amd

EPS24 = 2.0 ** -24  # half an ulp of 1.0: the rounding step of float32 values in [0.5, 1)


def _c2w(R, t):
    m = np.zeros((3, 4), np.float32)
    m[:, :3] = R
    m[:, 3] = t
    return m


def _fp64_dirs(c2w, fx, fy, cx, cy, width, height):
    """Unit world directions of every pixel centre in FP64, row-major pixels:
    R @ normalise(((ix + 1/2 - cx) / fx, (iy + 1/2 - cy) / fy, 1))."""
    iy, ix = np.meshgrid(np.arange(height, dtype=np.float64),
                         np.arange(width, dtype=np.float64), indexing="ij")
    v = np.stack([(ix + 0.5 - cx) / fx, (iy + 0.5 - cy) / fy, np.ones_like(ix)], -1)
    v = v.reshape(-1, 3)
    v /= np.sqrt((v * v).sum(1, keepdims=True))
    return v @ np.asarray(c2w, np.float64)[:, :3].T


def test_principal_point_looks_along_z(march):
    """Identity rotation: the pixel whose centre is the principal point looks
    exactly along +z, every origin is the translation column, and a pixel one
    step right / down turns exactly toward +x / +y."""
    t = np.array([0.25, -1.5, 3.0], np.float32)
    W, H, fx, fy = 9, 7, 8.0, 4.0
    cx, cy = 4.5, 3.5  # centre of pixel (4, 3)
    o, d = march.camera_rays(_c2w(np.eye(3), t), fx, fy, cx, cy, W, H)
    assert np.array_equal(o, np.broadcast_to(t, o.shape))
    centre = 3 * W + 4
    assert np.array_equal(d[centre], np.array([0, 0, 1], np.float32))
    right, down = d[centre + 1], d[centre + W]
    # (1/8, 0, 1) and (0, 1/4, 1), normalised
    assert right[1] == 0 and down[0] == 0
    assert abs(right[0] / right[2] - 1 / fx) < 4 * EPS24
    assert abs(down[1] / down[2] - 1 / fy) < 4 * EPS24
    assert np.abs(np.linalg.norm(d.astype(np.float64), axis=1) - 1).max() < 4 * EPS24


def test_signed_permutation_rotation_is_exact(march):
    """A signed permutation as c2w moves the camera-frame directions exactly
    (each world component is +-1 times one camera component): this fixes
    which index of c2w is the row and which the column."""
    W, H, fx, fy, cx, cy = 6, 5, 3.0, 5.0, 2.2, 2.9
    o0, d0 = march.camera_rays(_c2w(np.eye(3), np.zeros(3)), fx, fy, cx, cy, W, H)
    P = np.array([[0, 0, -1], [1, 0, 0], [0, -1, 0]], np.float32)  # not symmetric
    o1, d1 = march.camera_rays(_c2w(P, [1, 2, 3]), fx, fy, cx, cy, W, H)
    assert np.array_equal(d1, d0 @ P.T)
    assert np.array_equal(o1, np.broadcast_to(np.float32([1, 2, 3]), o1.shape))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_camera_rays_match_fp64_model(march, seed):
    """Random rotations, fx != fy, an off-centre principal point and a
    non-square image: every float32 direction is within 4 * 2^-24 (two
    float32 ulps of a unit-vector component) of the FP64 camera model, and of
    svox2_amd.Camera.gen_rays (the reference's FP64 Python generator)."""
    g = np.random.default_rng(seed)
    q = g.standard_normal(4)
    q /= np.linalg.norm(q)
    a, b, c, dq = q
    R = np.array([[a * a + b * b - c * c - dq * dq, 2 * (b * c - a * dq), 2 * (b * dq + a * c)],
                  [2 * (b * c + a * dq), a * a - b * b + c * c - dq * dq, 2 * (c * dq - a * b)],
                  [2 * (b * dq - a * c), 2 * (c * dq + a * b), a * a - b * b - c * c + dq * dq]])
    W, H = 37, 23
    fx, fy = 31.5 + seed, 44.25
    cx, cy = 15.3, 12.8
    c2w = _c2w(R, g.standard_normal(3))
    o, d = march.camera_rays(c2w, fx, fy, cx, cy, W, H)
    ref = _fp64_dirs(c2w, fx, fy, cx, cy, W, H)
    assert np.abs(d - ref).max() <= 4 * EPS24
    cam = svox2_amd.Camera(torch.from_numpy(np.vstack([c2w, [0, 0, 0, 1]])), fx=fx, fy=fy,
                           cx=cx, cy=cy, width=W, height=H)
    rays = cam.gen_rays()
    assert np.abs(d - rays.dirs.numpy()).max() <= 4 * EPS24
    assert np.array_equal(o, rays.origins.numpy())


def test_gen_rays_matches_fp64_model():
    """Camera.gen_rays itself (float64, rounded once to float32) is within
    half a float32 ulp of the FP64 camera model, so it can stand as the
    reference formula for the checks above and for the image kernel test."""
    cams = scenes.CameraSet(scenes.hemisphere_cameras(3, torch.Generator().manual_seed(5)),
                            width=40, height=30, focal=50.0)
    for k in range(3):
        cam = cams.camera(k)
        ref = _fp64_dirs(cam.c2w.numpy()[:3], cam.fx_val, cam.fy_val, cam.cx_val, cam.cy_val,
                         cam.width, cam.height)
        assert np.abs(cam.gen_rays().dirs.numpy() - ref).max() <= EPS24 * 1.0001
