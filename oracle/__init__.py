"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker or the baseline, never as
the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

RAY_DTYPE = np.dtype([(n, np.float32) for n in ("ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax")])
HIT_DTYPE = np.dtype([("id", np.int32), ("t", np.float32), ("u", np.float32), ("v", np.float32)])

_vp = C.c_void_p
_i32 = C.c_int32
PROTOTYPES = {
    "or_refscene_build": (_vp, [_vp, C.c_int32, _vp, C.c_int32, _vp, C.c_int32]),
    "or_refscene_free": (None, [_vp]),
    "or_render_ref": (None, [_vp, _vp, _vp, _vp, C.c_int32, _vp, _vp, C.c_int32, C.c_int32, _vp, _vp, C.c_int32]),
    "or_power_heuristic": (C.c_float, [C.c_float, C.c_float]),
    "or_fresnel": (C.c_float, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32]),
    "or_dielectric_scatter": (C.c_int32, [_vp, _vp, C.c_int32, C.c_float, C.c_float, C.c_int32, _vp]),
    "or_roulette": (C.c_int32, [_vp, C.c_float, C.c_uint32, C.c_uint32, C.c_uint32]),
    "or_emission_weight": (C.c_float, [C.c_int, C.c_float, C.c_float]),
    "or_direct_light": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_float, _vp, _vp, _vp]),
    "or_camera_ray": (None, [_vp, C.c_uint32, C.c_uint32, C.c_float, C.c_float, _vp]),
    "or_philox4x32_10": (None, [_vp, _vp, _vp]),
    "or_u01": (C.c_float, [C.c_uint32]),
    "or_sample_sphere": (None, [C.c_float, C.c_float, _vp]),
    "or_sample_cosine": (None, [_vp, C.c_float, C.c_float, _vp]),
    "or_sample_triangle": (None, [_vp, _vp, _vp, C.c_float, C.c_float, _vp]),
    "or_tri_watertight": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "or_tri_reference": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "or_box_interval_ref": (None, [C.c_float, C.c_float, C.c_float, C.c_float, _vp, _vp]),
    "or_box_ref": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "or_bvh_build": (_vp, [_vp, _i32, _vp, _i32, _i32, _i32]),
    "or_bvh_free": (None, [_vp]),
    "or_bvh_node_count": (_i32, [_vp]),
    "or_bvh_sah_cost": (C.c_double, [_vp]),
    "or_bvh_validate": (_i32, [_vp, _vp, _vp]),
    "or_trace_bvh": (None, [_vp, _vp, _vp, _vp, _i32, _i32, _vp, _vp, _vp, _vp, _i32]),
    "or_trace_brute": (None, [_vp, _vp, _i32, _vp, _i32, _i32, _vp, _vp, _i32]),
    "or_trace_instances_brute": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _i32, _i32, _vp, _vp, _i32]),
    "or_refbbh_build": (_vp, [_vp, _i32, _vp, _i32]),
    "or_refbbh_free": (None, [_vp]),
    "or_refbbh_depth": (_i32, [_vp]),
    "or_refbbh_node_count": (_i32, [_vp]),
    "or_refbbh_nodes": (None, [_vp, _vp, _vp]),
    "or_refbbh_order": (None, [_vp, _i32, _vp]),
    "or_refbbh_trace": (None, [_vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _i32]),
    "or_lights_build": (_i32, [_vp, _vp, _i32, _vp, _vp, _vp, _vp]),
    "or_sincos_2pi": (None, [C.c_float, _vp, _vp]),
    "or_render": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _i32, _i32, _vp, _vp, _vp, _i32]),
    "or_guide_encode": (None, [_vp, _vp, _vp]),
    "or_guide_forward": (None, [_vp, _vp, _vp, _i32, _i32, _vp, _vp]),
    "or_guide_loss": (C.c_double, [_vp, _vp, _vp, _vp, _i32, _i32]),
    "or_guide_grad": (C.c_double, [_vp, _vp, _vp, _i32, _i32, _vp]),
    "or_guide_loss_sel": (C.c_double, [_vp, _vp, _vp, _vp, _i32, _i32, _i32]),
    "or_guide_grad_sel": (C.c_double, [_vp, _vp, _vp, _i32, _i32, _i32, _vp]),
    "or_vmf_pdf": (C.c_double, [_vp, C.c_double, _vp]),
    "or_vmf_sample": (None, [_vp, C.c_double, C.c_double, C.c_double, _vp]),
    "or_display_thresholds": (None, [_i32, _vp]),
    "or_resolve_rgba8": (None, [_vp, C.c_uint32, C.c_uint32, _i32, _vp]),
    "or_debug_aov": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp]),
}

MATERIAL_DTYPE = np.dtype([("kind", np.int32), ("albedo", np.float32, 3), ("emission", np.float32, 3),
                           ("ior", np.float32)])
ALIAS_DTYPE = np.dtype([("p_a", np.float32), ("pdf_a", np.float32), ("pdf_b", np.float32),
                        ("a", np.int32), ("b", np.int32)])


class Camera(C.Structure):
    _fields_ = [("transform", C.c_float * 12), ("focal_length", C.c_float), ("center", C.c_float * 2),
                ("pixel_size", C.c_float * 2), ("width", C.c_uint32), ("height", C.c_uint32),
                ("pixel_sampling", C.c_uint32)]


class RenderParams(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("sample_offset", C.c_uint32), ("max_depth", C.c_uint32),
                ("rr_depth", C.c_uint32), ("seed", C.c_uint64), ("fresnel", C.c_uint32), ("flags", C.c_uint32)]


def sincos_2pi(u: float):
    s, c = np.zeros(1, np.float32), np.zeros(1, np.float32)
    lib().or_sincos_2pi(u, _p(s), _p(c))
    return float(s[0]), float(c[0])


def camera_ray(cam: "Camera", x: int, y: int, sx: float = 0.5, sy: float = 0.5) -> np.ndarray:
    """One pinhole ray (RAY_DTYPE record) through sensor point (x + sx, y + sy)."""
    out = np.zeros(1, RAY_DTYPE)
    lib().or_camera_ray(C.byref(cam), x, y, sx, sy, _p(out))
    return out[0]


def power_heuristic(a: float, b: float) -> float:
    return lib().or_power_heuristic(a, b)


def emission_weight(spec_prev: bool, pdf_bsdf: float, pdf_light: float) -> float:
    return lib().or_emission_weight(1 if spec_prev else 0, pdf_bsdf, pdf_light)


def direct_light(p, n, albedo, xl, nl, emission, pdf_area: float):
    """(visible side?, unweighted Ld rgb, light solid-angle pdf, cosine bounce pdf)"""
    args = [np.ascontiguousarray(x, np.float32) for x in (p, n, albedo, xl, nl, emission)]
    ld, pl, pb = np.zeros(3, np.float32), np.zeros(1, np.float32), np.zeros(1, np.float32)
    ok = lib().or_direct_light(*[_p(a) for a in args], pdf_area, _p(ld), _p(pl), _p(pb))
    return bool(ok), ld, float(pl[0]), float(pb[0])


PLANT_RR_NO_REWEIGHT, PLANT_TRANSMIT_WEIGHT = 0x10000, 0x20000


def fresnel(cos_i: float, n1: float, n2: float, cos_t: float, schlick: bool) -> float:
    return lib().or_fresnel(cos_i, n1, n2, cos_t, 1 if schlick else 0)


def dielectric_scatter(d, n, entering: bool, ior: float, u: float, schlick: bool):
    """(reflected?, new direction) of one dielectric event; n faces against d."""
    dd, nn = np.ascontiguousarray(d, np.float32), np.ascontiguousarray(n, np.float32)
    out = np.zeros(3, np.float32)
    r = lib().or_dielectric_scatter(_p(dd), _p(nn), 1 if entering else 0, ior, u, 1 if schlick else 0, _p(out))
    return bool(r), out


def roulette(beta, u: float, vertex: int, rr_depth: int, max_depth: int):
    """(alive?, throughput after the step) of Russian roulette after a vertex."""
    b = np.array(beta, np.float32)
    alive = lib().or_roulette(_p(b), u, vertex, rr_depth, max_depth)
    return bool(alive), b


def sample_sphere(r1: float, r2: float) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().or_sample_sphere(r1, r2, _p(out))
    return out


def sample_cosine(n, r1: float, r2: float) -> np.ndarray:
    nn, out = np.ascontiguousarray(n, np.float32), np.zeros(3, np.float32)
    lib().or_sample_cosine(_p(nn), r1, r2, _p(out))
    return out


def sample_triangle(a, b, c, r1: float, r2: float) -> np.ndarray:
    a, b, c = (np.ascontiguousarray(x, np.float32) for x in (a, b, c))
    out = np.zeros(3, np.float32)
    lib().or_sample_triangle(_p(a), _p(b), _p(c), r1, r2, _p(out))
    return out


def materials_array(mats) -> np.ndarray:
    """[(kind, albedo3, emission3, ior), ...] -> structured array (dr_material layout)."""
    a = np.zeros(len(mats), MATERIAL_DTYPE)
    for i, (k, alb, em, ior) in enumerate(mats):
        a[i] = (k, alb, em, ior)
    return a


def lights_build(vertices, indices, tri_mat, mats):
    v = np.ascontiguousarray(vertices, np.float32)
    f = np.ascontiguousarray(indices, np.int32)
    tm = np.ascontiguousarray(tri_mat, np.int32)
    m = np.ascontiguousarray(mats, MATERIAL_DTYPE)
    ent = np.zeros(max(1, f.shape[0]), ALIAS_DTYPE)
    pdf = np.zeros(f.shape[0], np.float32)
    ne = lib().or_lights_build(_p(v), _p(f), f.shape[0], _p(tm), _p(m), _p(ent), _p(pdf))
    return ent[:ne].copy(), pdf


def render(bvh: "BVH", tri_mat, mats, camera: Camera, spp: int, max_depth: int = 8, rr_depth: int = 3,
           seed: int = 0, fresnel: int = 0, sample_offset: int = 0, pixels=None, accum=None,
           vertex_hist: bool = False, nthreads: int | None = None, plant: int = 0, counts: bool = False):
    """Build-semantics path tracer; returns (accum[h, w, 4] float64, hist),
    plus the traversal work (extension rays, their interior-node visits and
    triangle tests in the canonical BVH, then the same for shadow rays) when
    counts is set. plant: PLANT_* bits that insert a known mistake (for the
    tests' power checks)."""
    tm = np.ascontiguousarray(tri_mat, np.int32)
    m = np.ascontiguousarray(mats, MATERIAL_DTYPE)
    ent, pdf = lights_build(bvh.v, bvh.f, tm, m)
    npix = camera.width * camera.height
    acc = np.zeros((npix, 4), np.float64) if accum is None else np.ascontiguousarray(accum, np.float64).reshape(npix, 4)
    prm = RenderParams(spp, sample_offset, max_depth, rr_depth, seed, fresnel, plant)
    lo, hi = (0, npix) if pixels is None else pixels
    hist = np.zeros(max_depth, np.int32) if vertex_hist else None
    cnt = np.zeros(6, np.int64)
    lib().or_render(bvh.h, _p(bvh.v), _p(bvh.f), _p(tm), _p(m), _p(ent) if len(ent) else None, _p(pdf), len(ent),
                    C.byref(camera), C.byref(prm), lo, hi, _p(acc), _p(hist), _p(cnt) if counts else None,
                    nthreads or threads())
    if counts:
        return acc.reshape(camera.height, camera.width, 4), hist, cnt
    return acc.reshape(camera.height, camera.width, 4), hist

class RefScene:
    """The reference's scene model for its integrator: objects (contiguous
    triangle ranges, scenes.SceneData.object_ranges()) with one median-split
    BBH each."""

    def __init__(self, vertices, indices, objects):
        self.v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        self.f = np.ascontiguousarray(indices, np.int32).reshape(-1, 3)
        self.objects = np.ascontiguousarray(objects, np.int32).reshape(-1, 2)
        self.h = lib().or_refscene_build(_p(self.v), self.v.shape[0], _p(self.f), self.f.shape[0],
                                         _p(self.objects), self.objects.shape[0])

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_refscene_free(self.h)
            self.h = None


def render_ref(scene: RefScene, tri_mat, mats, camera: Camera, spp: int, max_depth: int = 10, seed: int = 0,
               sample_offset: int = 0, pixels=None, accum=None, nthreads: int | None = None, counts: bool = False):
    """The reference's integrator (sampleColor, reference semantics); returns
    accum[h, w, 4] float64, plus (extension casts, shadow casts, box tests,
    triangle tests) when counts is set."""
    tm = np.ascontiguousarray(tri_mat, np.int32)
    m = np.ascontiguousarray(mats, MATERIAL_DTYPE)
    ent, _ = lights_build(scene.v, scene.f, tm, m)
    npix = camera.width * camera.height
    acc = np.zeros((npix, 4), np.float64) if accum is None else np.ascontiguousarray(accum, np.float64).reshape(npix, 4)
    prm = RenderParams(spp, sample_offset, max_depth, max_depth, seed, 1, 0)
    lo, hi = (0, npix) if pixels is None else pixels
    cnt = np.zeros(4, np.int64)
    lib().or_render_ref(scene.h, _p(tm), _p(m), _p(ent) if len(ent) else None, len(ent), C.byref(camera),
                        C.byref(prm), lo, hi, _p(acc), _p(cnt), nthreads or threads())
    out = acc.reshape(camera.height, camera.width, 4)
    return (out, cnt) if counts else out


def display_thresholds(srgb: bool) -> np.ndarray:
    """The 255 decision thresholds of the 8-bit linear or sRGB encoding."""
    t = np.zeros(255, np.float64)
    lib().or_display_thresholds(1 if srgb else 0, _p(t))
    return t


def resolve_rgba8(accum, srgb: bool = True) -> np.ndarray:
    """Renderer::createPixels restated (src/Renderer.cpp:51-90): FP64 RGBW
    [h, w, 4] -> packed RGBA8 u32 [h, w] with the rows flipped."""
    a = np.ascontiguousarray(accum, np.float64)
    h, w = a.shape[0], a.shape[1]
    out = np.zeros((h, w), np.uint32)
    lib().or_resolve_rgba8(_p(a), w, h, 1 if srgb else 0, _p(out))
    return out


def debug_aov(bvh: "BVH", tri_mat, debug_rgba, camera: Camera, mode: int, spp: int, seed: int = 0,
              sample_offset: int = 0, accum=None):
    """Debug views of the first hit (mode 1 UV checker, 2 barycentric, 3
    winding), FP64 RGBW accumulation [h, w, 4] in sample order."""
    tm = np.ascontiguousarray(tri_mat, np.int32)
    dc = np.ascontiguousarray(debug_rgba, np.float32)
    npix = camera.width * camera.height
    acc = np.zeros((npix, 4), np.float64) if accum is None else np.ascontiguousarray(accum, np.float64).reshape(npix, 4)
    prm = RenderParams(spp, sample_offset, 1, 1, seed, 0, 0)
    lib().or_debug_aov(bvh.h, _p(bvh.v), _p(bvh.f), _p(tm), _p(dc), C.byref(camera), C.byref(prm), mode, _p(acc))
    return acc.reshape(camera.height, camera.width, 4)


_lib = None


def build() -> None:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _p(a: np.ndarray):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def threads() -> int:
    return max(1, min(os.cpu_count() or 1, 64))


def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def tri_watertight(ray, v0, v1, v2):
    r = np.asarray(ray, np.float32)
    a, b, c = (np.asarray(x, np.float32) for x in (v0, v1, v2))
    t, u, v = (np.zeros(1, np.float32) for _ in range(3))
    hit = lib().or_tri_watertight(_p(r), _p(a), _p(b), _p(c), _p(t), _p(u), _p(v))
    return (float(t[0]), float(u[0]), float(v[0])) if hit else None


def tri_reference(p, d, a, b, c):
    arrs = [np.asarray(x, np.float32) for x in (p, d, a, b, c)]
    t = np.zeros(1, np.float32)
    bari = np.zeros(3, np.float32)
    hit = lib().or_tri_reference(*[_p(x) for x in arrs], _p(t), _p(bari))
    return (float(t[0]), bari.copy()) if hit else None


def box_interval_ref(lo, hi, p, v):
    e, x = np.zeros(1, np.float32), np.zeros(1, np.float32)
    lib().or_box_interval_ref(lo, hi, p, v, _p(e), _p(x))
    return float(e[0]), float(x[0])


def box_ref(lo, hi, p, d):
    arrs = [np.asarray(x, np.float32) for x in (lo, hi, p, d)]
    t = np.zeros(1, np.float32)
    hit = lib().or_box_ref(*[_p(x) for x in arrs], _p(t))
    return float(t[0]) if hit else None


class BVH:
    """Canonical binary binned-SAH BVH (32 bins, <= max_leaf tris per leaf)."""

    def __init__(self, vertices, indices, max_leaf: int = 4, bins: int = 32):
        self.v = np.ascontiguousarray(vertices, np.float32)
        self.f = np.ascontiguousarray(indices, np.int32)
        self.h = lib().or_bvh_build(_p(self.v), self.v.shape[0], _p(self.f), self.f.shape[0], max_leaf, bins)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_bvh_free(self.h)
            self.h = None

    def node_count(self) -> int:
        return lib().or_bvh_node_count(self.h)

    def sah_cost(self) -> float:
        return lib().or_bvh_sah_cost(self.h)

    def validate(self) -> int:
        return lib().or_bvh_validate(self.h, _p(self.v), _p(self.f))

    def closest(self, rays: np.ndarray, counts: bool = False, nthreads: int | None = None):
        n = rays.shape[0]
        hits = np.empty(n, HIT_DTYPE)
        nt = np.empty(n, np.int32) if counts else None
        tt = np.empty(n, np.int32) if counts else None
        lib().or_trace_bvh(self.h, _p(self.v), _p(self.f), _p(rays), n, 0, _p(hits), None, _p(nt), _p(tt),
                           nthreads or threads())
        return (hits, nt, tt) if counts else hits

    def any(self, rays: np.ndarray, counts: bool = False, nthreads: int | None = None):
        n = rays.shape[0]
        occ = np.empty(n, np.uint32)
        nt = np.empty(n, np.int32) if counts else None
        tt = np.empty(n, np.int32) if counts else None
        lib().or_trace_bvh(self.h, _p(self.v), _p(self.f), _p(rays), n, 1, None, _p(occ), _p(nt), _p(tt),
                           nthreads or threads())
        return (occ, nt, tt) if counts else occ


def brute_closest(vertices, indices, rays, nthreads=None):
    v = np.ascontiguousarray(vertices, np.float32)
    f = np.ascontiguousarray(indices, np.int32)
    n = rays.shape[0]
    hits = np.empty(n, HIT_DTYPE)
    lib().or_trace_brute(_p(v), _p(f), f.shape[0], _p(rays), n, 0, _p(hits), None, nthreads or threads())
    return hits


def brute_any(vertices, indices, rays, nthreads=None):
    v = np.ascontiguousarray(vertices, np.float32)
    f = np.ascontiguousarray(indices, np.int32)
    n = rays.shape[0]
    occ = np.empty(n, np.uint32)
    lib().or_trace_brute(_p(v), _p(f), f.shape[0], _p(rays), n, 1, None, _p(occ), nthreads or threads())
    return occ


def _instances_call(scenes, instances, rays, any_hit, nthreads):
    """scenes: [(vertices, indices)]; instances: a structured array with
    fields scene (u32), scale (3 f32), translate (3 f32), 28 B each."""
    vs = [np.ascontiguousarray(v, np.float32).reshape(-1, 3) for v, _ in scenes]
    fs = [np.ascontiguousarray(f, np.int32).reshape(-1, 3) for _, f in scenes]
    vofs = np.cumsum([0] + [v.shape[0] for v in vs[:-1]]).astype(np.int32)
    tofs = np.cumsum([0] + [f.shape[0] for f in fs[:-1]]).astype(np.int32)
    ntri = np.array([f.shape[0] for f in fs], np.int32)
    v = np.ascontiguousarray(np.concatenate(vs))
    f = np.ascontiguousarray(np.concatenate(fs))
    inst = np.ascontiguousarray(instances)
    if inst.dtype.itemsize != 28 or int(np.max(inst["scene"], initial=0)) >= len(scenes):
        raise ValueError("instances: 28-byte records naming scenes in range")
    n = rays.shape[0]
    hits = None if any_hit else np.empty(n, HIT_DTYPE)
    occ = np.empty(n, np.uint32) if any_hit else None
    lib().or_trace_instances_brute(_p(v), _p(f), _p(vofs), _p(tofs), _p(ntri), _p(inst), inst.shape[0],
                                   _p(rays), n, int(any_hit), _p(hits) if hits is not None else None,
                                   _p(occ) if occ is not None else None, nthreads or threads())
    return occ if any_hit else hits


def instances_closest(scenes, instances, rays, nthreads=None):
    """Closest hit over instanced scenes by brute force in model space."""
    return _instances_call(scenes, instances, rays, False, nthreads)


def instances_any(scenes, instances, rays, nthreads=None):
    """Occlusion over instanced scenes by brute force in model space."""
    return _instances_call(scenes, instances, rays, True, nthreads)


class RefBBH:
    """Reference-semantics median-split BBH with the stackless bit-trail walk."""

    def __init__(self, vertices, indices):
        self.v = np.ascontiguousarray(vertices, np.float32)
        self.f = np.ascontiguousarray(indices, np.int32)
        self.h = lib().or_refbbh_build(_p(self.v), self.v.shape[0], _p(self.f), self.f.shape[0])

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_refbbh_free(self.h)
            self.h = None

    def depth(self) -> int:
        return lib().or_refbbh_depth(self.h)

    def nodes(self):
        """(boxes float32 [n, 6] lo/hi, links int32 [n, 5] left/right/parent/tb/te, triangle order)"""
        n = lib().or_refbbh_node_count(self.h)
        boxes, links = np.zeros((n, 6), np.float32), np.zeros((n, 5), np.int32)
        lib().or_refbbh_nodes(self.h, _p(boxes), _p(links))
        order = np.zeros(self.f.shape[0], np.int32)
        lib().or_refbbh_order(self.h, self.f.shape[0], _p(order))
        return boxes, links, order

    def closest(self, rays, counts: bool = False, nthreads: int | None = None):
        n = rays.shape[0]
        hits = np.empty(n, HIT_DTYPE)
        bt = np.empty(n, np.int64) if counts else None
        tt = np.empty(n, np.int64) if counts else None
        lib().or_refbbh_trace(self.h, _p(self.v), _p(self.f), _p(rays), n, _p(hits), _p(bt), _p(tt),
                              nthreads or threads())
        return (hits, bt, tt) if counts else hits


# ---- learned path guide (FP64) ---------------------------------------------
GUIDE_N_PARAMS = 17712
GUIDE_SAMPLE_DTYPE = np.dtype([("pos", np.float32, 3), ("weight", np.float32), ("normal", np.float32, 3),
                               ("pdf", np.float32), ("wo", np.float32, 3), ("pad1", np.float32),
                               ("wi", np.float32, 3), ("pad2", np.float32)])


def _guide_args(params, bounds, samples):
    P = np.ascontiguousarray(params, np.float64)
    assert P.size == GUIDE_N_PARAMS
    b = np.ascontiguousarray(bounds, np.float32)
    s = np.ascontiguousarray(samples, GUIDE_SAMPLE_DTYPE)
    return P, b, s


def guide_encode(bounds, sample) -> np.ndarray:
    b = np.ascontiguousarray(bounds, np.float32)
    s = np.ascontiguousarray(np.atleast_1d(sample), GUIDE_SAMPLE_DTYPE)
    out = np.zeros(32, np.float64)
    lib().or_guide_encode(_p(b), _p(s), _p(out))
    return out


def guide_forward(params, bounds, samples, emulate_bf16: bool = False):
    """(raw outputs n x 48, decoded head n x 44)."""
    P, b, s = _guide_args(params, bounds, samples)
    raw = np.zeros((len(s), 48), np.float64)
    head = np.zeros((len(s), 44), np.float64)
    lib().or_guide_forward(_p(P), _p(b), _p(s), len(s), int(emulate_bf16), _p(raw), _p(head))
    return raw, head


def guide_loss(params, bounds, samples, emulate_bf16: bool = False, frozen=None, objective: int = 0) -> float:
    """Training loss; `frozen` parameters give the mixture inside the
    selection term (default: params themselves). objective: 0 KL selection
    term, 1 second-moment selection term, 2 no selection term."""
    P, b, s = _guide_args(params, bounds, samples)
    F = None if frozen is None else np.ascontiguousarray(frozen, np.float64)
    return float(lib().or_guide_loss_sel(_p(P), _p(F) if F is not None else None, _p(b), _p(s), len(s),
                                         int(emulate_bf16), int(objective)))


def guide_grad(params, bounds, samples, emulate_bf16: bool = False, objective: int = 0):
    """(gradient of the loss, sum of sample weights); objective 0 or 1 as in
    guide_loss."""
    P, b, s = _guide_args(params, bounds, samples)
    g = np.zeros(GUIDE_N_PARAMS, np.float64)
    sw = lib().or_guide_grad_sel(_p(P), _p(b), _p(s), len(s), int(emulate_bf16), int(objective), _p(g))
    return g, float(sw)


def vmf_pdf(mu, kappa: float, w) -> float:
    m = np.ascontiguousarray(mu, np.float64)
    x = np.ascontiguousarray(w, np.float64)
    return float(lib().or_vmf_pdf(_p(m), float(kappa), _p(x)))


def vmf_sample(mu, kappa: float, x1: float, x2: float) -> np.ndarray:
    m = np.ascontiguousarray(mu, np.float64)
    out = np.zeros(3, np.float64)
    lib().or_vmf_sample(_p(m), float(kappa), float(x1), float(x2), _p(out))
    return out
