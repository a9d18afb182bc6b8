"""ctypes binding of libdevilray.so (the C ABI declared in include/dr_api.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libdevilray.so")

DR_OK, DR_E_INVALID, DR_E_OOM, DR_E_DEVICE, DR_E_UNSUPPORTED, DR_E_OVERFLOW = range(6)
DR_MAT_LAMBERT, DR_MAT_DIELECTRIC = 0, 1


class DevilRayError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"devilray status {status}: {msg}")
        self.status = status


class DrRay(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax")]


class DrHit(C.Structure):
    _fields_ = [("id", C.c_int32), ("t", C.c_float), ("u", C.c_float), ("v", C.c_float)]


class DrMaterial(C.Structure):
    _fields_ = [("kind", C.c_int32), ("albedo", C.c_float * 3), ("emission", C.c_float * 3), ("ior", C.c_float)]


class DrSceneDesc(C.Structure):
    _fields_ = [
        ("vertices", C.POINTER(C.c_float)), ("n_vertices", C.c_uint32),
        ("indices", C.POINTER(C.c_uint32)), ("n_triangles", C.c_uint32),
        ("tri_material", C.POINTER(C.c_uint32)),
        ("materials", C.POINTER(DrMaterial)), ("n_materials", C.c_uint32),
    ]


class DrBvhParams(C.Structure):
    _fields_ = [("leaf_max", C.c_uint32), ("flags", C.c_uint32)]


class DrSceneInfo(C.Structure):
    _fields_ = [
        ("n_triangles", C.c_uint32), ("n_nodes", C.c_uint32), ("n_leaves", C.c_uint32),
        ("depth", C.c_uint32), ("overflow_flag", C.c_uint32), ("simd_stats", C.c_uint32 * 4),
        ("build_ms", C.c_double), ("sah_cost", C.c_double), ("n_references", C.c_uint32),
        ("stack_high_water", C.c_uint32),
    ]


class DrCamera(C.Structure):
    _fields_ = [("transform", C.c_float * 12), ("focal_length", C.c_float), ("center", C.c_float * 2),
                ("pixel_size", C.c_float * 2), ("width", C.c_uint32), ("height", C.c_uint32),
                ("pixel_sampling", C.c_uint32)]


class DrRenderParams(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("sample_offset", C.c_uint32), ("max_depth", C.c_uint32),
                ("rr_depth", C.c_uint32), ("seed", C.c_uint64), ("fresnel", C.c_uint32), ("flags", C.c_uint32)]


class DrBatchView(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("spp", C.c_uint32), ("sample_offset", C.c_uint32), ("pad", C.c_uint32),
                ("camera", DrCamera), ("accum", C.c_void_p)]


class DrBatchOptions(C.Structure):
    _fields_ = [("chunk_paths", C.c_uint32), ("trace_sched", C.c_uint32), ("refill_min", C.c_uint32),
                ("use_graph", C.c_uint32), ("launch", C.c_uint32), ("blocks_per_cu", C.c_uint32)]


DR_BATCH_LAUNCH_SEQUENCE = 0
DR_BATCH_LAUNCH_PERSISTENT = 1


class DrGuideParams(C.Structure):
    _fields_ = [("bounds", C.c_float * 6), ("capacity", C.c_uint32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint64)]


GUIDE_N_PARAMS = 17712
GUIDE_HEAD_FLOATS = 44
DR_GUIDE_USE, DR_GUIDE_COLLECT = 1, 2
DR_GUIDE_SELECT_KL, DR_GUIDE_SELECT_VARIANCE = 0, 1
DR_OUTPUT_LINEAR, DR_OUTPUT_SRGB = 0, 1
DR_DEBUG_UV_CHECKER, DR_DEBUG_BARYCENTRIC, DR_DEBUG_WINDING = 1, 2, 3
SCHED_PER_RAY, SCHED_PER_RAY_SPECULATIVE, SCHED_PERSISTENT, SCHED_PERSISTENT_SPECULATIVE = 0, 1, 2, 3
SCHED_REORDER, SCHED_XCD_SLICES = 4, 8
GUIDE_SAMPLE_DTYPE = np.dtype([("pos", np.float32, 3), ("weight", np.float32), ("normal", np.float32, 3),
                               ("pdf", np.float32), ("wo", np.float32, 3), ("pad1", np.float32),
                               ("wi", np.float32, 3), ("pad2", np.float32)])

ALIAS_DTYPE = np.dtype([("p_a", np.float32), ("pdf_a", np.float32), ("pdf_b", np.float32),
                        ("a", np.int32), ("b", np.int32)])

RAY_DTYPE = np.dtype([(n, np.float32) for n in ("ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax")])
HIT_DTYPE = np.dtype([("id", np.int32), ("t", np.float32), ("u", np.float32), ("v", np.float32)])
# dr_instance: world = model * scale + translate
INSTANCE_DTYPE = np.dtype([("scene", np.uint32), ("scale", np.float32, 3), ("translate", np.float32, 3)])

_vp = C.c_void_p
_u32 = C.c_uint32
_i32 = C.c_int32
_st = C.c_int

# name -> (restype, argtypes); every symbol include/dr_api.h declares
PROTOTYPES = {
    "dr_status_string": (C.c_char_p, [_st]),
    "dr_last_error": (C.c_char_p, []),
    "dr_device_count": (_st, [C.POINTER(_i32)]),
    "dr_context_create": (_st, [_i32, C.POINTER(_vp)]),
    "dr_context_destroy": (_st, [_vp]),
    "dr_device_alloc": (_st, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "dr_device_free": (_st, [_vp, _vp]),
    "dr_memcpy_htod": (_st, [_vp, _vp, _vp, C.c_size_t]),
    "dr_memcpy_dtoh": (_st, [_vp, _vp, _vp, C.c_size_t]),
    "dr_memcpy_dtod": (_st, [_vp, _vp, _vp, C.c_size_t, _vp]),
    "dr_synchronize": (_st, [_vp]),
    "dr_fill_zero": (_st, [_vp, _vp, C.c_size_t, _vp]),
    "dr_scene_create": (_st, [_vp, C.POINTER(DrSceneDesc), C.POINTER(_vp)]),
    "dr_scene_destroy": (_st, [_vp]),
    "dr_scene_get_info": (_st, [_vp, C.POINTER(DrSceneInfo)]),
    "dr_scene_build_bvh": (_st, [_vp, C.POINTER(DrBvhParams), _vp]),
    "dr_scene_refit": (_st, [_vp, _vp, _vp]),
    "dr_trace_closest": (_st, [_vp, _vp, _u32, _vp, _vp]),
    "dr_trace_closest_noalloc": (_st, [_vp, _vp, _u32, _vp, _vp]),
    "dr_trace_any": (_st, [_vp, _vp, _u32, _vp, _vp]),
    "dr_trace_closest_counted": (_st, [_vp, _vp, _u32, _vp, _vp, _vp, _vp]),
    "dr_scene_set_trace_schedule": (_st, [_vp, _u32, _u32]),
    "dr_scene_set_trace_tuning": (_st, [_vp, _u32, _u32, _u32]),
    "dr_scene_set_render_lanes": (_st, [_vp, _u32]),
    "dr_scene_build_lights": (_st, [_vp, _vp]),
    "dr_scene_light_count": (_st, [_vp, C.POINTER(_u32)]),
    "dr_scene_copy_lights": (_st, [_vp, _vp, _vp]),
    "dr_render": (_st, [_vp, C.POINTER(DrCamera), C.POINTER(DrRenderParams), _vp, _vp]),
    "dr_resolve_rgba8": (_st, [_vp, _vp, _u32, _u32, _u32, _vp, _vp]),
    "dr_render_debug": (_st, [_vp, C.POINTER(DrCamera), C.POINTER(DrRenderParams), _u32, _vp, _vp]),
    "dr_scene_ray_casts": (_st, [_vp, C.POINTER(C.c_uint64), C.c_int32, _vp]),
    "dr_scene_set_debug_colors": (_st, [_vp, _vp, _u32]),
    "dr_batch_create": (_st, [_vp, _vp, _u32, C.POINTER(_vp)]),
    "dr_batch_destroy": (_st, [_vp]),
    "dr_batch_render": (_st, [_vp, _vp, _u32, C.POINTER(DrRenderParams), C.POINTER(DrBatchOptions), _vp]),
    "dr_batch_autotune": (_st, [_vp, _vp, _u32, C.POINTER(DrRenderParams), _u32, _u32, C.POINTER(DrBatchOptions),
                                _vp, _vp, _u32, C.POINTER(_u32)]),
    "dr_instances_create": (_st, [_vp, _vp, _u32, _vp, _u32, C.POINTER(_vp)]),
    "dr_instances_destroy": (_st, [_vp]),
    "dr_instances_trace_closest": (_st, [_vp, _vp, _u32, _vp, _vp]),
    "dr_instances_trace_any": (_st, [_vp, _vp, _u32, _vp, _vp]),
    "dr_instances_set_trace_schedule": (_st, [_vp, _u32, _u32]),
    "dr_instances_get_info": (_st, [_vp, C.POINTER(DrSceneInfo)]),
    "dr_instances_render": (_st, [_vp, C.POINTER(DrCamera), C.POINTER(DrRenderParams), _vp, _vp]),
    "dr_instances_light_count": (_st, [_vp, C.POINTER(_u32)]),
    "dr_instances_copy_lights": (_st, [_vp, _vp, _vp, _vp]),
    "dr_render_guided": (_st, [_vp, C.POINTER(DrCamera), C.POINTER(DrRenderParams), _vp, _u32, _vp, _vp]),
    "dr_guide_create": (_st, [_vp, C.POINTER(DrGuideParams), C.POINTER(_vp)]),
    "dr_guide_destroy": (_st, [_vp]),
    "dr_guide_get_params": (_st, [_vp, _vp]),
    "dr_guide_set_params": (_st, [_vp, _vp]),
    "dr_guide_eval": (_st, [_vp, _vp, _u32, _vp, _vp]),
    "dr_guide_gradient": (_st, [_vp, _vp, _u32, _vp, _vp]),
    "dr_guide_train_step": (_st, [_vp, _vp, _u32, _vp]),
    "dr_guide_copy_samples": (_st, [_vp, _vp, _u32, C.POINTER(_u32), C.POINTER(_u32)]),
    "dr_guide_set_options": (_st, [_vp, C.c_float, _u32, _u32]),
    "dr_guide_set_learning_rate": (_st, [_vp, C.c_float]),
    "dr_guide_set_selection_objective": (_st, [_vp, C.c_int32]),
    "dr_gen_icosphere": (_st, [_u32, C.c_float, C.c_uint64, _vp, C.POINTER(_u32), _vp, C.POINTER(_u32)]),
    "dr_gen_rays_surface": (_st, [_vp, _vp, _u32, _u32, C.c_uint64, _i32, C.c_float, _vp]),
    "dr_gen_rays_camera": (_st, [_vp, _vp, _vp, C.c_float, _u32, _u32, _vp]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libdevilray.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `python -m devilray_amd.build` (no CPU fallback exists)")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def check(status: int) -> None:
    if status != DR_OK:
        msg = lib().dr_last_error()
        raise DevilRayError(status, msg.decode() if msg else "")


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data
