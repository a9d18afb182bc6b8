"""DevilRay render core for MI355X (gfx950).

Python-side host helpers over the C ABI (include/dr_api.h): a device context,
scene upload + on-device BVH build, and batched closest/any-hit traversal.
Every call goes to libdevilray.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (  # noqa: F401
    ALIAS_DTYPE, DR_BATCH_LAUNCH_PERSISTENT, DR_BATCH_LAUNCH_SEQUENCE, DrBatchOptions, DrBatchView, DR_DEBUG_BARYCENTRIC, DR_DEBUG_UV_CHECKER, DR_DEBUG_WINDING,
    DR_GUIDE_COLLECT, DR_GUIDE_SELECT_KL, DR_GUIDE_SELECT_VARIANCE, DR_GUIDE_USE, DR_MAT_DIELECTRIC, DR_MAT_LAMBERT, DR_OUTPUT_LINEAR, DR_OUTPUT_SRGB, GUIDE_HEAD_FLOATS,
    GUIDE_N_PARAMS, GUIDE_SAMPLE_DTYPE, HIT_DTYPE, INSTANCE_DTYPE, RAY_DTYPE, DevilRayError, DrBvhParams, DrCamera, DrGuideParams,
    DrMaterial, DrRenderParams, DrSceneDesc, DrSceneInfo, SCHED_PER_RAY, SCHED_PER_RAY_SPECULATIVE, SCHED_PERSISTENT,
    SCHED_PERSISTENT_SPECULATIVE, SCHED_REORDER, SCHED_XCD_SLICES, check, lib, ptr,
)


def camera_lookat(eye, target, up, fov_y_deg: float, width: int, height: int,
                  pixel_sampling: int = 1) -> DrCamera:
    """Reference-model pinhole camera (Camera / cameraRay) looking from eye
    to target: rows of the image go down the screen, so camera +y maps to
    world 'down'."""
    e = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - e
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, np.float64))
    right /= np.linalg.norm(right)
    down = np.cross(fwd, right)
    m = np.zeros(12, np.float32)
    for r in range(3):
        m[4 * r + 0] = right[r]
        m[4 * r + 1] = down[r]
        m[4 * r + 2] = fwd[r]
        m[4 * r + 3] = e[r]
    cam = DrCamera()
    cam.transform[:] = list(m)
    cam.focal_length = float((height / 2.0) / np.tan(np.radians(fov_y_deg) / 2.0))
    cam.center[:] = [width / 2.0, height / 2.0]
    cam.pixel_size[:] = [1.0, 1.0]
    cam.width, cam.height, cam.pixel_sampling = width, height, pixel_sampling
    return cam


def camera_reference(width: int, height: int, focal_length: float, pixel_size: float,
                     pixel_sampling: int = 1) -> DrCamera:
    """The reference tests' camera (test/src/test_render.cpp:21-36): identity
    transform, principal point at the image centre."""
    cam = DrCamera()
    cam.transform[:] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
    cam.focal_length = focal_length
    cam.center[:] = [width / 2.0 * pixel_size, height / 2.0 * pixel_size]
    cam.pixel_size[:] = [pixel_size, pixel_size]
    cam.width, cam.height, cam.pixel_sampling = width, height, pixel_sampling
    return cam

__all__ = [
    "Context", "Scene", "Instances", "DeviceArray", "gen_icosphere", "gen_rays_surface", "gen_rays_camera",
    "RAY_DTYPE", "HIT_DTYPE", "INSTANCE_DTYPE", "DevilRayError",
]


# ---------------------------------------------------------------- generators
def gen_icosphere(subdiv: int, sigma: float = 0.0, seed: int = 0):
    """Displaced icosphere: (vertices float32[V,3], indices uint32[T,3])."""
    nv, nt = C.c_uint32(), C.c_uint32()
    check(lib().dr_gen_icosphere(subdiv, sigma, seed, None, C.byref(nv), None, C.byref(nt)))
    v = np.empty((nv.value, 3), np.float32)
    f = np.empty((nt.value, 3), np.uint32)
    check(lib().dr_gen_icosphere(subdiv, sigma, seed, ptr(v), C.byref(nv), ptr(f), C.byref(nt)))
    return v, f


def gen_rays_surface(vertices, indices, n_rays: int, seed: int = 0, inward: bool = True,
                     offset: float = 1e-5) -> np.ndarray:
    v = np.ascontiguousarray(vertices, np.float32)
    f = np.ascontiguousarray(indices, np.uint32)
    rays = np.empty(n_rays, RAY_DTYPE)
    check(lib().dr_gen_rays_surface(ptr(v), ptr(f), f.shape[0], n_rays, seed, int(inward), offset, ptr(rays)))
    return rays


def gen_rays_camera(eye, target, up, fov_y_deg: float, width: int, height: int) -> np.ndarray:
    e = np.asarray(eye, np.float32)
    t = np.asarray(target, np.float32)
    u = np.asarray(up, np.float32)
    rays = np.empty(width * height, RAY_DTYPE)
    check(lib().dr_gen_rays_camera(ptr(e), ptr(t), ptr(u), fov_y_deg, width, height, ptr(rays)))
    return rays


# ---------------------------------------------------------------- device side
class Context:
    """One HIP device (dr_context)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().dr_context_create(device, C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            check(lib().dr_context_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self):
        check(lib().dr_synchronize(self.handle))

    def resolve_rgba8(self, accum: np.ndarray, srgb: bool = True) -> np.ndarray:
        """Display pixels of an FP64 RGBW image (height, width, 4) through
        dr_resolve_rgba8: packed RGBA8 u32 (height, width), rows flipped."""
        a = np.ascontiguousarray(accum, np.float64)
        h, w = a.shape[0], a.shape[1]
        d = self.upload(a)
        o = self.alloc(h * w * 4)
        check(lib().dr_resolve_rgba8(self.handle, C.c_void_p(d.ptr), w, h, DR_OUTPUT_SRGB if srgb else DR_OUTPUT_LINEAR,
                                     C.c_void_p(o.ptr), None))
        self.synchronize()
        out = o.to_numpy(np.uint32, h * w).reshape(h, w)
        d.free()
        o.free()
        return out

    def fill_zero(self, d: "DeviceArray", nbytes: int | None = None, stream=None):
        """Zero the first nbytes (a multiple of 16; default all) of a device array with a kernel."""
        n = d.nbytes if nbytes is None else int(nbytes)
        check(lib().dr_fill_zero(self.handle, C.c_void_p(d.ptr), n, stream))

    def alloc(self, nbytes: int) -> "DeviceArray":
        return DeviceArray(self, nbytes)

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.nbytes)
        d.copy_from(a)
        return d


class DeviceArray:
    """A raw device allocation owned by a Context."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        h = C.c_void_p()
        check(lib().dr_device_alloc(ctx.handle, self.nbytes, C.byref(h)))
        self.ptr = h.value

    def free(self):
        if self.ptr:
            check(lib().dr_device_free(self.ctx.handle, C.c_void_p(self.ptr)))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def copy_from(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().dr_memcpy_htod(self.ctx.handle, C.c_void_p(self.ptr), ptr(a), a.nbytes))

    def to_numpy(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        check(lib().dr_memcpy_dtoh(self.ctx.handle, ptr(out), C.c_void_p(self.ptr), out.nbytes))
        return out


class Scene:
    """Triangle scene on the device (dr_scene) with its BVH."""

    def __init__(self, ctx: Context, vertices, indices, tri_material=None, materials=None):
        self.ctx = ctx
        self._v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        self._f = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        d = DrSceneDesc()
        d.vertices = self._v.ctypes.data_as(C.POINTER(C.c_float))
        d.n_vertices = self._v.shape[0]
        d.indices = self._f.ctypes.data_as(C.POINTER(C.c_uint32))
        d.n_triangles = self._f.shape[0]
        if tri_material is not None:
            self._m = np.ascontiguousarray(tri_material, np.uint32)
            d.tri_material = self._m.ctypes.data_as(C.POINTER(C.c_uint32))
        if materials:
            arr = (DrMaterial * len(materials))(*materials)
            self._mats = arr
            d.materials = C.cast(arr, C.POINTER(DrMaterial))
            d.n_materials = len(materials)
        h = C.c_void_p()
        check(lib().dr_scene_create(ctx.handle, C.byref(d), C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            check(lib().dr_scene_destroy(self.handle))
            self.handle = None

    def build_bvh(self, leaf_max: int = 4, builder: str = "ploc", radius: int = 0, split: float | None = None,
                  split_ratio: float | None = None, stream=None):
        """builder: "ploc" (clustering, default) or "lbvh" (radix tree);
        radius: PLOC search radius (0 = the library default); split: cap on
        the extra triangle references as a multiple of the triangle count, in
        steps of 1/4 up to 63.75 (None = the library default, 0 = no
        splitting); split_ratio: r0, a triangle whose box is `ratio` times its
        ideal half area gets floor(ratio / r0) pieces (steps of 1/4, None =
        the library default)."""
        if builder not in ("ploc", "lbvh"):
            raise ValueError("builder must be 'ploc' or 'lbvh'")
        flags = (1 if builder == "lbvh" else 0) | ((radius & 0xFF) << 8)
        if split is not None:
            q = int(round(split * 4))
            if not 0 <= q <= 255:
                raise ValueError("split must be in [0, 63.75]")
            flags |= (q << 16) if q else 2
        if split_ratio is not None:
            q = int(round(split_ratio * 4))
            if not 1 <= q <= 255:
                raise ValueError("split_ratio must be in [0.25, 63.75]")
            flags |= q << 24
        p = DrBvhParams(leaf_max, flags)
        check(lib().dr_scene_build_bvh(self.handle, C.byref(p), stream))

    def refit(self, vertices, stream=None):
        """Move the vertices (same count and triangles) and refit the built BVH to them."""
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        if v.shape != self._v.shape:
            raise ValueError(f"refit needs {self._v.shape[0]} vertices, got {v.shape[0]}")
        check(lib().dr_scene_refit(self.handle, ptr(v), stream))
        self._v = v

    def set_schedule(self, closest: int, any_hit: int | None = None):
        check(lib().dr_scene_set_trace_schedule(self.handle, closest, closest if any_hit is None else any_hit))

    def set_tuning(self, refill_min: int = 16, blocks_per_cu: int = 6, sort_key: int = 0):
        check(lib().dr_scene_set_trace_tuning(self.handle, refill_min, blocks_per_cu, sort_key))

    def build_lights(self, stream=None):
        check(lib().dr_scene_build_lights(self.handle, stream))

    def set_render_lanes(self, lanes: int):
        """Lanes (streams) a render deals its paths to (1..4, default 4)."""
        check(lib().dr_scene_set_render_lanes(self.handle, lanes))

    def lights(self):
        """(alias entries, per-triangle area pdf) copied to the host."""
        n = C.c_uint32()
        check(lib().dr_scene_light_count(self.handle, C.byref(n)))
        ent = np.zeros(n.value, ALIAS_DTYPE)
        pdf = np.zeros(self._f.shape[0], np.float32)
        check(lib().dr_scene_copy_lights(self.handle, ptr(ent) if n.value else None, ptr(pdf)))
        return ent, pdf

    def render_dev(self, camera: DrCamera, params: DrRenderParams, accum_ptr: int, stream=None):
        check(lib().dr_render(self.handle, C.byref(camera), C.byref(params), C.c_void_p(accum_ptr), stream))

    def ray_casts(self, reset: bool = False) -> int:
        """Camera and bounce rays traced by this scene's renders since creation or the last reset
        (shadow rays excluded); synchronises the null stream."""
        n = C.c_uint64()
        check(lib().dr_scene_ray_casts(self.handle, C.byref(n), 1 if reset else 0, None))
        return n.value

    def set_debug_colors(self, rgba):
        c = np.ascontiguousarray(rgba, np.float32).reshape(-1, 4)
        check(lib().dr_scene_set_debug_colors(self.handle, ptr(c), c.shape[0]))

    def render_debug(self, camera: DrCamera, mode: int, spp: int, seed: int = 0, sample_offset: int = 0,
                     accum: np.ndarray | None = None) -> np.ndarray:
        """Debug view (DR_DEBUG_*) into an FP64 RGBW accumulation image (height, width, 4)."""
        npix = camera.width * camera.height
        d = self.ctx.alloc(npix * 32)
        init = np.zeros((npix, 4), np.float64) if accum is None else np.ascontiguousarray(accum, np.float64)
        d.copy_from(init)
        p = DrRenderParams(spp, sample_offset, 1, 1, seed, 0, 0)
        check(lib().dr_render_debug(self.handle, C.byref(camera), C.byref(p), mode, C.c_void_p(d.ptr), None))
        self.ctx.synchronize()
        out = d.to_numpy(np.float64, npix * 4).reshape(camera.height, camera.width, 4)
        d.free()
        return out

    def render(self, camera: DrCamera, spp: int, max_depth: int = 8, rr_depth: int = 3, seed: int = 0,
               fresnel: int = 0, sample_offset: int = 0, accum: np.ndarray | None = None) -> np.ndarray:
        """Render into an FP64 RGBW accumulation image (height, width, 4)."""
        npix = camera.width * camera.height
        d = self.ctx.alloc(npix * 32)
        init = np.zeros((npix, 4), np.float64) if accum is None else np.ascontiguousarray(accum, np.float64)
        d.copy_from(init)
        p = DrRenderParams(spp, sample_offset, max_depth, rr_depth, seed, fresnel, 0)
        self.render_dev(camera, p, d.ptr)
        self.ctx.synchronize()
        return d.to_numpy(np.float64, npix * 4).reshape(camera.height, camera.width, 4)

    def render_guided_dev(self, camera: DrCamera, params: DrRenderParams, guide: "Guide", flags: int,
                          accum_ptr: int, stream=None):
        check(lib().dr_render_guided(self.handle, C.byref(camera), C.byref(params),
                                     guide.handle if guide is not None else None, flags, C.c_void_p(accum_ptr),
                                     stream))

    def render_guided(self, camera: DrCamera, guide: "Guide", flags: int, spp: int, max_depth: int = 8,
                      rr_depth: int = 3, seed: int = 0, fresnel: int = 0, sample_offset: int = 0,
                      accum: np.ndarray | None = None) -> np.ndarray:
        """render() with the learned guide (flags: DR_GUIDE_USE | DR_GUIDE_COLLECT)."""
        npix = camera.width * camera.height
        d = self.ctx.alloc(npix * 32)
        init = np.zeros((npix, 4), np.float64) if accum is None else np.ascontiguousarray(accum, np.float64)
        d.copy_from(init)
        p = DrRenderParams(spp, sample_offset, max_depth, rr_depth, seed, fresnel, 0)
        self.render_guided_dev(camera, p, guide, flags, d.ptr)
        self.ctx.synchronize()
        return d.to_numpy(np.float64, npix * 4).reshape(camera.height, camera.width, 4)

    def bounds(self):
        """Axis-aligned box of the vertices (lo xyz, hi xyz) as float32."""
        return np.concatenate([self._v.min(0), self._v.max(0)]).astype(np.float32)

    def info(self) -> DrSceneInfo:
        i = DrSceneInfo()
        check(lib().dr_scene_get_info(self.handle, C.byref(i)))
        return i

    # device-pointer entry points
    def trace_closest_dev(self, rays_ptr: int, n: int, hits_ptr: int, stream=None):
        check(lib().dr_trace_closest(self.handle, C.c_void_p(rays_ptr), n, C.c_void_p(hits_ptr), stream))

    def trace_any_dev(self, rays_ptr: int, n: int, occ_ptr: int, stream=None):
        check(lib().dr_trace_any(self.handle, C.c_void_p(rays_ptr), n, C.c_void_p(occ_ptr), stream))

    # host convenience wrappers (copy in, trace, copy out)
    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        n = rays.shape[0]
        dr = self.ctx.upload(rays)
        dh = self.ctx.alloc(max(n, 1) * HIT_DTYPE.itemsize)
        self.trace_closest_dev(dr.ptr, n, dh.ptr)
        self.ctx.synchronize()
        return dh.to_numpy(HIT_DTYPE, n)

    def trace_any(self, rays: np.ndarray) -> np.ndarray:
        n = rays.shape[0]
        dr = self.ctx.upload(rays)
        do = self.ctx.alloc(max(n, 1) * 4)
        self.trace_any_dev(dr.ptr, n, do.ptr)
        self.ctx.synchronize()
        return do.to_numpy(np.uint32, n)

    def trace_closest_counted(self, rays: np.ndarray):
        n = rays.shape[0]
        dr = self.ctx.upload(rays)
        dh = self.ctx.alloc(max(n, 1) * HIT_DTYPE.itemsize)
        db = self.ctx.alloc(max(n, 1) * 4)
        dt = self.ctx.alloc(max(n, 1) * 4)
        check(lib().dr_trace_closest_counted(self.handle, C.c_void_p(dr.ptr), n, C.c_void_p(dh.ptr),
                                             C.c_void_p(db.ptr), C.c_void_p(dt.ptr), None))
        self.ctx.synchronize()
        return dh.to_numpy(HIT_DTYPE, n), db.to_numpy(np.uint32, n), dt.to_numpy(np.uint32, n)


class Guide:
    """Learned path guide (dr_guide): an MLP proposing an 8-lobe vMF mixture
    per Lambert vertex, trained on the device with BF16 MFMA kernels."""

    def __init__(self, ctx: Context, bounds, capacity: int = 1 << 18, lr: float = 1e-2, beta1: float = 0.9,
                 beta2: float = 0.99, eps: float = 1e-8, seed: int = 0):
        self.ctx = ctx
        self.capacity = capacity
        p = DrGuideParams((C.c_float * 6)(*[float(b) for b in bounds]), capacity, lr, beta1, beta2, eps, seed)
        h = C.c_void_p()
        check(lib().dr_guide_create(ctx.handle, C.byref(p), C.byref(h)))
        self.handle = h
        self.collect_fraction, self.guided_vertices, self.train_batch = 1.0, 64, 0

    def close(self):
        if self.handle:
            check(lib().dr_guide_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        out = np.zeros(GUIDE_N_PARAMS, np.float32)
        check(lib().dr_guide_get_params(self.handle, ptr(out)))
        return out

    def set_params(self, params):
        a = np.ascontiguousarray(params, np.float32)
        assert a.size == GUIDE_N_PARAMS
        check(lib().dr_guide_set_params(self.handle, ptr(a)))

    def eval(self, samples: np.ndarray) -> np.ndarray:
        """Decoded heads (n x 44) for host samples (GUIDE_SAMPLE_DTYPE)."""
        s = np.ascontiguousarray(samples, GUIDE_SAMPLE_DTYPE)
        n = len(s)
        ds = self.ctx.upload(s)
        dh = self.ctx.alloc(max(n, 1) * GUIDE_HEAD_FLOATS * 4)
        check(lib().dr_guide_eval(self.handle, C.c_void_p(ds.ptr), n, C.c_void_p(dh.ptr), None))
        self.ctx.synchronize()
        return dh.to_numpy(np.float32, n * GUIDE_HEAD_FLOATS).reshape(n, GUIDE_HEAD_FLOATS)

    def gradient(self, samples: np.ndarray):
        """(gradient of sum w * -log p, sum of weights) from the device kernels."""
        s = np.ascontiguousarray(samples, GUIDE_SAMPLE_DTYPE)
        n = len(s)
        ds = self.ctx.upload(s)
        dg = self.ctx.alloc((GUIDE_N_PARAMS + 1) * 4)
        check(lib().dr_guide_gradient(self.handle, C.c_void_p(ds.ptr), n, C.c_void_p(dg.ptr), None))
        self.ctx.synchronize()
        g = dg.to_numpy(np.float32, GUIDE_N_PARAMS + 1)
        return g[:GUIDE_N_PARAMS], float(g[GUIDE_N_PARAMS])

    def train_step_dev(self, samples_ptr: int | None = None, n: int = 0, stream=None):
        check(lib().dr_guide_train_step(self.handle, C.c_void_p(samples_ptr) if samples_ptr else None, n, stream))

    def train_step(self, samples: np.ndarray | None = None):
        """One Adam step on host samples, or (None) on the last collection."""
        if samples is None:
            self.train_step_dev()
        else:
            s = np.ascontiguousarray(samples, GUIDE_SAMPLE_DTYPE)
            ds = self.ctx.upload(s)
            self.train_step_dev(ds.ptr, len(s))
            self.ctx.synchronize()

    def copy_samples(self, max_n: int | None = None):
        """(samples kept by the last collecting render, number offered)."""
        kept, total = C.c_uint32(), C.c_uint32()
        check(lib().dr_guide_copy_samples(self.handle, None, 0, C.byref(kept), C.byref(total)))
        n = kept.value if max_n is None else min(kept.value, max_n)
        out = np.zeros(n, GUIDE_SAMPLE_DTYPE)
        if n:
            check(lib().dr_guide_copy_samples(self.handle, ptr(out), n, C.byref(kept), C.byref(total)))
        return out, total.value

    def set_learning_rate(self, lr: float):
        check(lib().dr_guide_set_learning_rate(self.handle, lr))

    def set_selection_objective(self, objective: int):
        """DR_GUIDE_SELECT_KL (default) or DR_GUIDE_SELECT_VARIANCE for the
        selection logit's training loss."""
        check(lib().dr_guide_set_selection_objective(self.handle, int(objective)))

    def set_options(self, collect_fraction: float = 1.0, guided_vertices: int = 64, train_batch: int = 0):
        check(lib().dr_guide_set_options(self.handle, collect_fraction, guided_vertices, train_batch))
        self.collect_fraction, self.guided_vertices, self.train_batch = collect_fraction, guided_vertices, train_batch


def make_instances(scene, scale, translate) -> np.ndarray:
    """INSTANCE_DTYPE array from per-instance scene indices, scales (3) and
    translations (3); scalars broadcast."""
    sc = np.atleast_1d(np.asarray(scene, np.uint32))
    n = max(sc.shape[0], np.atleast_2d(scale).shape[0], np.atleast_2d(translate).shape[0])
    out = np.zeros(n, INSTANCE_DTYPE)
    out["scene"] = np.broadcast_to(sc, (n,))
    out["scale"] = np.broadcast_to(np.asarray(scale, np.float32), (n, 3))
    out["translate"] = np.broadcast_to(np.asarray(translate, np.float32), (n, 3))
    return out


class Instances:
    """Built scenes placed with a scale and a translation each (dr_instances):
    a top-level tree over the instances above each scene's own tree. Hit ids
    number the instances' triangles in instance order."""

    def __init__(self, ctx: Context, scenes, instances):
        self.ctx = ctx
        self._inst = np.ascontiguousarray(instances, INSTANCE_DTYPE)
        arr = (C.c_void_p * len(scenes))(*[s.handle.value if hasattr(s.handle, "value") else s.handle
                                           for s in scenes])
        h = C.c_void_p()
        check(lib().dr_instances_create(ctx.handle, arr, len(scenes), self._inst.ctypes.data_as(C.c_void_p),
                                        self._inst.shape[0], C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            check(lib().dr_instances_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> DrSceneInfo:
        i = DrSceneInfo()
        check(lib().dr_instances_get_info(self.handle, C.byref(i)))
        return i

    def set_trace_schedule(self, closest: int, any_hit: int):
        check(lib().dr_instances_set_trace_schedule(self.handle, closest, any_hit))

    def trace_closest_dev(self, rays_ptr: int, n: int, hits_ptr: int, stream=None):
        check(lib().dr_instances_trace_closest(self.handle, C.c_void_p(rays_ptr), n, C.c_void_p(hits_ptr), stream))

    def trace_any_dev(self, rays_ptr: int, n: int, occ_ptr: int, stream=None):
        check(lib().dr_instances_trace_any(self.handle, C.c_void_p(rays_ptr), n, C.c_void_p(occ_ptr), stream))

    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        n = rays.shape[0]
        dr = self.ctx.upload(rays)
        dh = self.ctx.alloc(max(n, 1) * HIT_DTYPE.itemsize)
        self.trace_closest_dev(dr.ptr, n, dh.ptr)
        self.ctx.synchronize()
        return dh.to_numpy(HIT_DTYPE, n)

    def trace_any(self, rays: np.ndarray) -> np.ndarray:
        n = rays.shape[0]
        dr = self.ctx.upload(rays)
        do = self.ctx.alloc(max(n, 1) * 4)
        self.trace_any_dev(dr.ptr, n, do.ptr)
        self.ctx.synchronize()
        return do.to_numpy(np.uint32, n)

    def lights(self):
        """(alias entries naming emitters by index, the emitters' instanced
        triangle ids, their area pdfs) copied to the host."""
        n = C.c_uint32()
        check(lib().dr_instances_light_count(self.handle, C.byref(n)))
        ent = np.zeros(n.value, ALIAS_DTYPE)
        tri = np.zeros(n.value, np.uint32)
        pdf = np.zeros(n.value, np.float32)
        if n.value:
            check(lib().dr_instances_copy_lights(self.handle, ptr(ent), ptr(tri), ptr(pdf)))
        return ent, tri, pdf

    def render_dev(self, camera: DrCamera, params: DrRenderParams, accum_ptr: int, stream=None):
        check(lib().dr_instances_render(self.handle, C.byref(camera), C.byref(params), C.c_void_p(accum_ptr),
                                        stream))

    def render(self, camera: DrCamera, spp: int, max_depth: int = 8, rr_depth: int = 3, seed: int = 0,
               fresnel: int = 0, sample_offset: int = 0) -> np.ndarray:
        """Render into an FP64 RGBW accumulation image (height, width, 4), as Scene.render."""
        npix = camera.width * camera.height
        d = self.ctx.alloc(npix * 32)
        d.copy_from(np.zeros((npix, 4), np.float64))
        self.render_dev(camera, DrRenderParams(spp, sample_offset, max_depth, rr_depth, seed, fresnel, 0), d.ptr)
        self.ctx.synchronize()
        return d.to_numpy(np.float64, npix * 4).reshape(camera.height, camera.width, 4)


class Batch:
    """Scenes merged for batched rendering (dr_batch): views of all scenes
    render in one sequence of wavefront launches."""

    def __init__(self, ctx: Context, scenes):
        self.ctx = ctx
        arr = (C.c_void_p * len(scenes))(*[s.handle.value if hasattr(s.handle, "value") else s.handle
                                           for s in scenes])
        h = C.c_void_p()
        check(lib().dr_batch_create(ctx.handle, arr, len(scenes), C.byref(h)))
        self.handle = h
        self.n_scenes = len(scenes)

    def close(self):
        if self.handle:
            check(lib().dr_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render_dev(self, views, max_depth: int = 12, rr_depth: int = 3, seed: int = 0, fresnel: int = 0,
                   chunk_paths: int = 1 << 24, trace_sched: int = 1, refill_min: int = 16, use_graph: bool = False,
                   launch: int = 0, blocks_per_cu: int = 0, stream=None):
        """views: [(scene index, DrCamera, spp, sample_offset, accum device pointer)]. launch:
        DR_BATCH_LAUNCH_SEQUENCE (a launch per stage, per chunk) or DR_BATCH_LAUNCH_PERSISTENT (one
        launch for the whole render, blocks_per_cu resident blocks per CU, 0 = occupancy limit)."""
        vs = (DrBatchView * len(views))()
        for i, (sc, cam, spp, off, acc) in enumerate(views):
            vs[i].scene, vs[i].spp, vs[i].sample_offset = sc, spp, off
            vs[i].camera = cam
            vs[i].accum = acc
        p = DrRenderParams(0, 0, max_depth, rr_depth, seed, fresnel, 0)
        o = DrBatchOptions(chunk_paths, trace_sched, refill_min, int(use_graph), launch, blocks_per_cu)
        check(lib().dr_batch_render(self.handle, vs, len(views), C.byref(p), C.byref(o), stream))

    def autotune(self, views, max_depth: int = 12, rr_depth: int = 3, seed: int = 0, fresnel: int = 0,
                 budget_spp: int = 4, reps: int = 1):
        """dr_batch_autotune over views as for render_dev (their accumulation
        pointers are not written): (best options as render_dev keyword
        arguments, [(milliseconds, options)] for every candidate timed)."""
        vs = (DrBatchView * len(views))()
        for i, (sc, cam, spp, off, acc) in enumerate(views):
            vs[i].scene, vs[i].spp, vs[i].sample_offset = sc, spp, off
            vs[i].camera = cam
            vs[i].accum = acc
        p = DrRenderParams(0, 0, max_depth, rr_depth, seed, fresnel, 0)
        best = DrBatchOptions()
        cap = 256
        tried = (DrBatchOptions * cap)()
        ms = (C.c_float * cap)()
        n = C.c_uint32()
        check(lib().dr_batch_autotune(self.handle, vs, len(views), C.byref(p), budget_spp, reps, C.byref(best),
                                      tried, ms, cap, C.byref(n)))

        def kw(o):
            return dict(chunk_paths=o.chunk_paths, trace_sched=o.trace_sched, refill_min=o.refill_min,
                        use_graph=bool(o.use_graph), launch=o.launch, blocks_per_cu=o.blocks_per_cu)
        return kw(best), [(float(ms[k]), kw(tried[k])) for k in range(min(n.value, cap))]
