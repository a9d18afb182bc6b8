"""dr_trace_closest_noalloc (the zero-allocation closest hit, its scratch in
a static device array) against dr_trace_closest through the C ABI: two calls
in a row on one scene, the second with a smaller batch, bit-identical hits;
a second scene on the same device; a tree deeper than the static scratch
covers is refused with DR_E_OVERFLOW and leaves the output untouched."""
import ctypes as C

import numpy as np
import pytest

import devilray_amd as dr
from devilray_amd import _lib

pytestmark = pytest.mark.gpu


def trace_both(ctx, scene, rays):
    n = rays.shape[0]
    d_rays = ctx.upload(rays)
    d_ref = ctx.alloc(n * 16)
    d_got = ctx.alloc(n * 16)
    d_got.copy_from(np.full(4 * n, -7, np.int32))
    dr.check(dr.lib().dr_trace_closest(scene.handle, C.c_void_p(d_rays.ptr), n, C.c_void_p(d_ref.ptr), None))
    dr.check(dr.lib().dr_trace_closest_noalloc(scene.handle, C.c_void_p(d_rays.ptr), n, C.c_void_p(d_got.ptr),
                                               None))
    ctx.synchronize()
    return d_got.to_numpy(dr.HIT_DTYPE, n), d_ref.to_numpy(dr.HIT_DTYPE, n)


def test_noalloc_equals_trace_closest(gpu_ctx):
    print("[case] zero-allocation closest hit vs dr_trace_closest", flush=True)
    v, f = dr.gen_icosphere(6, 0.05, 2)
    scene = dr.Scene(gpu_ctx, v, f)
    scene.build_bvh()
    assert scene.info().depth <= 21
    big = np.concatenate([dr.gen_rays_surface(v, f, 1 << 18, seed=1, inward=True, offset=1e-5),
                          dr.gen_rays_surface(v, f, 1 << 16, seed=2, inward=False, offset=1e-5)])
    small = dr.gen_rays_surface(v, f, 3000, seed=3, inward=True, offset=1e-5)
    for rays in (big, small):
        got, ref = trace_both(gpu_ctx, scene, rays)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
        assert (ref["id"] >= 0).mean() > 0.5
    scene.close()
    # another scene (per-ray schedule requested: the variant runs its
    # resident form) on the same device's array
    v2, f2 = dr.gen_icosphere(3, 0.0, 0)
    s2 = dr.Scene(gpu_ctx, v2, f2)
    s2.build_bvh(leaf_max=2)
    s2.set_schedule(dr.SCHED_PER_RAY)
    got, ref = trace_both(gpu_ctx, s2, dr.gen_rays_surface(v2, f2, 5000, seed=4, inward=True))
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
    assert s2.info().overflow_flag == 0
    s2.close()


def test_noalloc_refuses_deep_tree(gpu_ctx):
    print("[case] zero-allocation closest hit: deep tree refused", flush=True)
    v, f = [], []
    for k in range(150):
        s = 0.9 ** k
        v += [(-s, -s, s), (2 * s, -s, s), (-s, 2 * s, s)]
        f.append((3 * k, 3 * k + 1, 3 * k + 2))
    scene = dr.Scene(gpu_ctx, np.float32(v), np.uint32(f))
    scene.build_bvh(leaf_max=1, split=0.0)
    assert scene.info().depth > 21
    rays = np.zeros(64, dr.RAY_DTYPE)
    rays["dz"] = 1.0
    rays["tmax"] = np.inf
    d_rays = gpu_ctx.upload(rays)
    d_hits = gpu_ctx.alloc(64 * 16)
    d_hits.copy_from(np.full(256, -7, np.int32))
    st = dr.lib().dr_trace_closest_noalloc(scene.handle, C.c_void_p(d_rays.ptr), 64, C.c_void_p(d_hits.ptr), None)
    assert st == _lib.DR_E_OVERFLOW
    assert b"wide levels" in dr.lib().dr_last_error()
    gpu_ctx.synchronize()
    assert np.all(d_hits.to_numpy(np.int32, 256) == -7)
    scene.close()
