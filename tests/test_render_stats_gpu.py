"""Render statistics and device fills through the C ABI: dr_scene_ray_casts
(the reference's per-pixel casts behind Renderer::getTotalCasts,
src/Renderer.cpp:26-33, counting one cast per path vertex as
PathGeneration.hpp:100-111, shadow rays excluded) and dr_fill_zero.

The cast counts are pinned by closed forms: inside a closed sphere with no
emitters and no Russian roulette every path traces exactly max_depth rays; a
debug view traces one per sample; and the count is the same however the
samples are split over calls."""
import numpy as np
import pytest

import devilray_amd as dr

pytestmark = pytest.mark.gpu


def _inside_sphere(ctx, albedo=0.5):
    v, f = dr.gen_icosphere(3)
    m = dr.DrMaterial()
    m.kind = dr.DR_MAT_LAMBERT
    m.albedo[:] = [albedo] * 3
    m.emission[:] = [0.0] * 3
    m.ior = 1.0
    s = dr.Scene(ctx, v * 5.0, f, np.zeros(len(f), np.uint32), [m])
    s.build_bvh()
    cam = dr.camera_lookat((0.1, 0.2, 0.3), (0.1, 0.2, 1.3), (0, 1, 0), 60.0, 40, 30)
    return s, cam


def test_fill_zero(gpu_ctx):
    print("[case] fill_zero", flush=True)
    a = np.arange(4096, dtype=np.float64) + 1.0
    d = gpu_ctx.upload(a)
    gpu_ctx.fill_zero(d, 16 * 100)
    gpu_ctx.synchronize()
    got = d.to_numpy(np.float64, 4096)
    assert np.all(got[:200] == 0.0) and np.array_equal(got[200:], a[200:])
    gpu_ctx.fill_zero(d)
    gpu_ctx.synchronize()
    assert np.all(d.to_numpy(np.float64, 4096) == 0.0)
    with pytest.raises(dr.DevilRayError):
        gpu_ctx.fill_zero(d, 8)  # not a multiple of 16
    d.free()


def test_ray_casts_closed_sphere(gpu_ctx):
    print("[case] ray casts in a closed sphere", flush=True)
    s, cam = _inside_sphere(gpu_ctx)
    npix = cam.width * cam.height
    assert s.ray_casts() == 0
    img = s.render(cam, spp=3, max_depth=7, rr_depth=7)
    assert np.all(img[..., 3] == 3)
    assert s.ray_casts() == 7 * 3 * npix               # every path: max_depth rays
    s.render(cam, spp=2, max_depth=1, rr_depth=1)      # camera rays only
    assert s.ray_casts(reset=True) == 7 * 3 * npix + 2 * npix
    assert s.ray_casts() == 0
    s.render_debug(cam, dr.DR_DEBUG_BARYCENTRIC, spp=5)
    assert s.ray_casts(reset=True) == 5 * npix
    s.close()


def test_ray_casts_independent_of_split(gpu_ctx):
    print("[case] ray casts with roulette, one call vs four", flush=True)
    s, cam = _inside_sphere(gpu_ctx)
    npix = cam.width * cam.height
    one = s.render(cam, spp=4, max_depth=9, rr_depth=2, seed=3)
    c1 = s.ray_casts(reset=True)
    acc = None
    for k in range(4):
        acc = s.render(cam, spp=1, max_depth=9, rr_depth=2, seed=3, sample_offset=k, accum=acc)
    c4 = s.ray_casts()
    assert np.array_equal(one, acc)
    assert c4 == c1 and 2 * 4 * npix < c1 < 9 * 4 * npix   # roulette ends some paths early
    s.close()
