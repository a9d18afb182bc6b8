"""Pins the oracle's display step (Renderer::createPixels, src/Renderer.cpp:51-90)
and debug views (include/tracing/PathGeneration.hpp:274-299,377-429) before
they check the device.

The 8-bit values come from exact decision thresholds; they are checked against
a direct float evaluation of the reference's formula (equal except within
float rounding of a threshold), against known values, and for the packing,
clamping and row flip. The debug views are checked through properties fixed
by their definitions: barycentric weights sum to one, the winding colour
follows the vertex order, and UV-checker misses are not counted."""
import numpy as np
import pytest

import oracle


def reference_byte(c, srgb):
    """The reference's float arithmetic for one channel (toSRGB then round)."""
    c = np.float32(c)
    if srgb:
        r = np.float32(min(max(c, np.float32(0)), np.float32(1)))
        if float(r) < 0.0031308:
            v = np.float32(float(r) * 12.92)
        else:
            v = np.float32(float(np.power(r, np.float32(1) / np.float32(2.4), dtype=np.float32)) * 1.055 - float(np.float32(0.055)))
    else:
        v = np.float32(min(max(c, np.float32(0)), np.float32(1)))
    return int(np.floor(float(np.float32(v * np.float32(255))) + 0.5))


def one_pixel(c, srgb):
    a = np.zeros((1, 1, 4))
    a[0, 0] = [c, 0.0, 0.0, 1.0]
    return int(oracle.resolve_rgba8(a, srgb)[0, 0] & 0xFF)


def test_known_values():
    assert one_pixel(0.0, True) == 0 and one_pixel(1.0, True) == 255
    assert one_pixel(0.5, True) == 188        # 1.055 * 0.5^(1/2.4) - 0.055 = 0.7354 -> 187.5... -> 188
    assert one_pixel(0.5, False) == 128       # round(127.5) rounds half away from zero
    assert one_pixel(0.25, False) == 64
    assert one_pixel(0.001, True) == 3        # linear segment: 0.01292 * 255 = 3.29
    assert one_pixel(2.0, True) == 255 and one_pixel(-1.0, True) == 0 and one_pixel(float("nan"), True) == 0


@pytest.mark.parametrize("srgb", [True, False])
def test_matches_reference_formula_away_from_boundaries(srgb):
    rng = np.random.default_rng(5)
    vals = np.concatenate([rng.random(20000), rng.random(2000) * 0.01, rng.random(500) * 1.2 - 0.1])
    thr = oracle.display_thresholds(srgb)
    a = np.zeros((1, vals.size, 4))
    a[0, :, 0] = vals
    a[0, :, 3] = 1.0
    got = (oracle.resolve_rgba8(a, srgb)[0] & 0xFF).astype(int)
    ref = np.array([reference_byte(v, srgb) for v in vals])
    diff = np.nonzero(got != ref)[0]
    # any disagreement is one step, right at a rounding boundary
    assert np.all(np.abs(got[diff] - ref[diff]) == 1)
    if diff.size:
        near = np.min(np.abs(vals[diff, None] - thr[None, :]), axis=1)
        assert np.all(near < 1e-6 * np.maximum(vals[diff], 1e-3))
    assert diff.size <= 5


def test_thresholds_are_the_encoding_boundaries():
    for srgb in (True, False):
        t = oracle.display_thresholds(srgb)
        assert np.all(np.diff(t) > 0) and t[0] > 0 and t[-1] < 1
        for b in (1, 17, 128, 255):
            assert one_pixel(t[b - 1], srgb) == b
            assert one_pixel(np.nextafter(t[b - 1], 0), srgb) == b - 1


def test_packing_flip_and_empty_pixels():
    a = np.zeros((2, 3, 4))
    a[0, 0] = [1.0, 0.5, 0.0, 1.0]          # top-left of the sensor
    a[1, 2] = [0.4, 0.4, 0.4, 2.0]          # w = 2: the mean is 0.2
    a[0, 1] = [5.0, 5.0, 5.0, 0.0]          # no samples: black
    out = oracle.resolve_rgba8(a, srgb=False)
    assert out[1, 0] == (255 | (128 << 8) | (0 << 16) | (255 << 24))   # row flipped
    assert out[0, 2] == (51 | (51 << 8) | (51 << 16) | (255 << 24))
    assert out[1, 1] == 0xFF000000


def _two_triangle_scene():
    # two triangles facing +z at z = 0; the first wound counter-clockwise seen
    # from +z, the second clockwise; camera at z = 5 looking down -z
    v = np.array([[-2, -1, 0], [0, -1, 0], [-1, 1, 0],
                  [0.1, -1, 0], [1.1, 1, 0], [2.1, -1, 0]], np.float32)
    f = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    return v, f


def _camera(w=32, h=16):
    cam = oracle.Camera()
    # camera at (0, 0, 5), looking along -z: x right, y up, z forward = -z
    m = [1, 0, 0, 0,
         0, -1, 0, 0,
         0, 0, -1, 5]
    for i, x in enumerate(m):
        cam.transform[i] = x
    cam.focal_length = 1.0
    cam.pixel_size[0] = cam.pixel_size[1] = 2.0 / h
    cam.center[0], cam.center[1] = w / h, 1.0
    cam.width, cam.height, cam.pixel_sampling = w, h, 1
    return cam


def test_debug_views_properties():
    v, f = _two_triangle_scene()
    bvh = oracle.BVH(v, f)
    cam = _camera()
    tm = np.array([0, 1], np.int32)
    colours = np.array([[0.2, 0.4, 0.6, 1.0], [1.0, 0.0, 0.5, 1.0]], np.float32)
    spp = 3
    bary = oracle.debug_aov(bvh, tm, colours, cam, 2, spp, seed=4)
    hits = bary[..., :3].sum(-1)
    assert np.all(bary[..., 3] == spp)                       # every sample counted
    covered = hits > 0
    assert covered.any() and (~covered).any()
    assert np.allclose(hits[covered] % 1.0, 0.0, atol=1e-5) or np.allclose(hits[covered], np.round(hits[covered]), atol=1e-5)
    wind = oracle.debug_aov(bvh, tm, colours, cam, 3, spp, seed=4)
    ccw, cw = np.array([1.0, 0.73, 0.47]), np.array([0.53, 0.82, 1.0])
    # left triangle: (A - B) x (A - C) = (0, 0, 4) points at the camera, the
    # rays run along -z: not counter-clockwise; the right one is reversed
    left = wind[:, : cam.width // 2 - 2, :3].reshape(-1, 3)
    right = wind[:, cam.width // 2 + 2:, :3].reshape(-1, 3)
    for side, colour in ((left, cw), (right, ccw)):
        k = np.round(side[:, 2] / colour[2])          # hit samples in the pixel
        assert (k > 0).any()
        assert np.abs(side - np.outer(k, colour)).max() < 1e-5
    uv = oracle.debug_aov(bvh, tm, colours, cam, 1, spp, seed=4)
    assert np.all(uv[..., 3] <= spp)
    miss = uv[..., 3] == 0
    assert miss.any() and np.all(uv[miss] == 0)
    # a hit adds half the material's debug colour
    full = uv[..., 3] == spp
    got = uv[full][:, :3] / spp
    assert np.all(np.isclose(got, 0.5 * colours[0, :3], atol=1e-6).all(1) | np.isclose(got, 0.5 * colours[1, :3], atol=1e-6).all(1))
