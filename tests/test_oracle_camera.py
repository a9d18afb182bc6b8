"""The reference's camera ray tests (test/src/test_camera.cpp:29-97) restated
against the oracle's camera ray, which the device camera reproduces bit for
bit (the debug-view images are identical to the oracle's). The camera: 100 x
100 pixels of 0.1, focal length 1, principal point at 50.5 pixels so that
pixel (50, 50) sits on the optical axis; transform rows 0..2 of the 4x4."""
import numpy as np
import pytest

import oracle


def axis_camera(m=(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0)):
    cam = oracle.Camera()
    for i, x in enumerate(m):
        cam.transform[i] = x
    cam.focal_length = 1.0
    cam.center[0] = cam.center[1] = 50.5 * 0.1
    cam.pixel_size[0] = cam.pixel_size[1] = 0.1
    cam.width = cam.height = 100
    cam.pixel_sampling = 0
    return cam


def o(r):
    return np.float32([r["ox"], r["oy"], r["oz"]])


def d(r):
    return np.float32([r["dx"], r["dy"], r["dz"]])


def test_center_pixel_looks_down_optical_axis():
    r = oracle.camera_ray(axis_camera(), 50, 50)
    assert np.allclose(o(r), 0.0) and np.allclose(d(r), [0, 0, 1], atol=1e-5)


def test_origin_comes_from_transform_offset():
    r = oracle.camera_ray(axis_camera((1, 0, 0, 1, 0, 1, 0, 2, 0, 0, 1, 3)), 50, 50)
    assert np.allclose(o(r), [1, 2, 3]) and np.allclose(d(r), [0, 0, 1], atol=1e-5)


def test_off_axis_pixels_point_outward():
    right = oracle.camera_ray(axis_camera(), 80, 50)
    assert right["dx"] > 0 and abs(right["dy"]) < 1e-5
    up = oracle.camera_ray(axis_camera(), 50, 20)  # pixel y grows downward
    assert up["dy"] < 0 and abs(up["dx"]) < 1e-5


def test_direction_is_normalised():
    assert abs(np.linalg.norm(d(oracle.camera_ray(axis_camera(), 10, 90))) - 1.0) < 1e-5


def test_transform_rotates_view_direction():
    c, s = np.cos(np.pi / 2), np.sin(np.pi / 2)  # rotation about y by pi/2
    rot = np.float32([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    m = [rot[0, 0], rot[0, 1], rot[0, 2], 0, rot[1, 0], rot[1, 1], rot[1, 2], 0, rot[2, 0], rot[2, 1], rot[2, 2], 0]
    r = oracle.camera_ray(axis_camera(m), 50, 50)
    assert np.allclose(d(r), rot @ np.float32([0, 0, 1]), atol=1e-5)


@pytest.mark.parametrize("px", [(37, 64), (0, 0), (99, 13)])
def test_uniform_random_at_half_matches_center(px):
    center = oracle.camera_ray(axis_camera(), *px)
    jitter = oracle.camera_ray(axis_camera(), *px, 0.5, 0.5)
    assert np.array_equal(d(center), d(jitter))
    # and a jittered ray stays inside the pixel's cone: between its corners
    lo, hi = oracle.camera_ray(axis_camera(), *px, 0.0, 0.0), oracle.camera_ray(axis_camera(), *px, 1.0, 1.0)
    mid = oracle.camera_ray(axis_camera(), *px, 0.3, 0.7)
    assert lo["dx"] / lo["dz"] <= mid["dx"] / mid["dz"] <= hi["dx"] / hi["dz"]
