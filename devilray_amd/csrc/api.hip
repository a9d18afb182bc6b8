// C-ABI entry points: status strings, device memory, contexts, scenes and
// dispatch to the traversal and build kernels.
#include "dr_internal.hpp"

#include <algorithm>
#include <cstring>
#include <new>

namespace dr {

namespace {
thread_local std::string g_last_error;

__global__ __launch_bounds__(256) void k_zero16(uint4 *p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}
}

void set_error(const std::string &msg) { g_last_error = msg; }

dr_status fail(dr_status s, const std::string &msg)
{
    g_last_error = msg;
    return s;
}

dr_status DevBuf::ensure(size_t n)
{
    if (n <= bytes && ptr) return DR_OK;
    release();
    if (n == 0) n = 16;
    hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        return fail(DR_E_OOM, std::string("hipMalloc(") + std::to_string(n) + "): " + hipGetErrorString(e));
    }
    bytes = n;
    return DR_OK;
}

void DevBuf::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

}  // namespace dr

using namespace dr;

extern "C" {

const char *dr_status_string(dr_status s)
{
    switch (s) {
    case DR_OK: return "ok";
    case DR_E_INVALID: return "invalid argument";
    case DR_E_OOM: return "out of device memory";
    case DR_E_DEVICE: return "device error";
    case DR_E_UNSUPPORTED: return "unsupported";
    case DR_E_OVERFLOW: return "overflow";
    }
    return "unknown status";
}

const char *dr_last_error(void) { return g_last_error.c_str(); }

dr_status dr_device_count(int32_t *count)
{
    if (!count) return fail(DR_E_INVALID, "count is null");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        *count = 0;
        return fail(DR_E_DEVICE, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    }
    *count = c;
    return DR_OK;
}

dr_status dr_context_create(int32_t device, dr_context **out)
{
    if (!out) return fail(DR_E_INVALID, "out is null");
    *out = nullptr;
    int count = 0;
    DR_HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(DR_E_INVALID, "device index out of range");
    DR_HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    DR_HIP_TRY(hipGetDeviceProperties(&prop, device));
    auto *ctx = new (std::nothrow) Context();
    if (!ctx) return fail(DR_E_OOM, "host allocation failed");
    ctx->device = device;
    ctx->cu_count = prop.multiProcessorCount;
    if (dr_status st = noalloc_init(ctx)) {
        delete ctx;
        return st;
    }
    *out = reinterpret_cast<dr_context *>(ctx);
    return DR_OK;
}

dr_status dr_context_destroy(dr_context *c)
{
    if (!c) return fail(DR_E_INVALID, "context is null");
    auto *ctx = reinterpret_cast<Context *>(c);
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    delete ctx;
    return DR_OK;
}

dr_status dr_device_alloc(dr_context *c, size_t bytes, void **out)
{
    if (!c || !out) return fail(DR_E_INVALID, "null argument");
    auto *ctx = reinterpret_cast<Context *>(c);
    DR_HIP_TRY(hipSetDevice(ctx->device));
    *out = nullptr;
    DR_HIP_TRY(hipMalloc(out, bytes ? bytes : 16));
    return DR_OK;
}

dr_status dr_device_free(dr_context *c, void *ptr)
{
    if (!c) return fail(DR_E_INVALID, "context is null");
    if (ptr) DR_HIP_TRY(hipFree(ptr));
    return DR_OK;
}

dr_status dr_memcpy_htod(dr_context *c, void *dst, const void *src, size_t bytes)
{
    if (!c || (!dst && bytes) || (!src && bytes)) return fail(DR_E_INVALID, "null argument");
    if (bytes) DR_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return DR_OK;
}

dr_status dr_memcpy_dtoh(dr_context *c, void *dst, const void *src, size_t bytes)
{
    if (!c || (!dst && bytes) || (!src && bytes)) return fail(DR_E_INVALID, "null argument");
    if (bytes) DR_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return DR_OK;
}

dr_status dr_memcpy_dtod(dr_context *c, void *dst, const void *src, size_t bytes, dr_stream stream)
{
    if (!c || (!dst && bytes) || (!src && bytes)) return fail(DR_E_INVALID, "null argument");
    if (bytes)
        DR_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, reinterpret_cast<hipStream_t>(stream)));
    return DR_OK;
}

dr_status dr_fill_zero(dr_context *c, void *dst, size_t bytes, dr_stream stream)
{
    if (!c || (!dst && bytes)) return fail(DR_E_INVALID, "null argument");
    if ((reinterpret_cast<uintptr_t>(dst) & 15u) || (bytes & 15u))
        return fail(DR_E_INVALID, "dst and bytes must be multiples of 16");
    if (!bytes) return DR_OK;
    DR_HIP_TRY(hipSetDevice(reinterpret_cast<Context *>(c)->device));
    const size_t n = bytes / 16;
    const uint32_t grid = (uint32_t)std::min<size_t>((n + 255) / 256, 8192);
    k_zero16<<<grid, 256, 0, reinterpret_cast<hipStream_t>(stream)>>>(static_cast<uint4 *>(dst), n);
    DR_HIP_TRY(hipGetLastError());
    return DR_OK;
}

dr_status dr_synchronize(dr_context *c)
{
    if (!c) return fail(DR_E_INVALID, "context is null");
    DR_HIP_TRY(hipDeviceSynchronize());
    return DR_OK;
}

dr_status dr_scene_create(dr_context *c, const dr_scene_desc *d, dr_scene **out)
{
    if (!c || !d || !out) return fail(DR_E_INVALID, "null argument");
    *out = nullptr;
    if (d->n_triangles && (!d->vertices || !d->indices)) return fail(DR_E_INVALID, "null geometry");
    if (d->n_materials && !d->materials) return fail(DR_E_INVALID, "null materials");
    for (uint64_t k = 0; k < 3ull * d->n_triangles; ++k)
        if (d->indices[k] >= d->n_vertices) return fail(DR_E_INVALID, "vertex index out of range");
    if (d->tri_material) {
        for (uint32_t k = 0; k < d->n_triangles; ++k)
            if (d->tri_material[k] >= (d->n_materials ? d->n_materials : 1))
                return fail(DR_E_INVALID, "material index out of range");
    }
    auto *ctx = reinterpret_cast<Context *>(c);
    DR_HIP_TRY(hipSetDevice(ctx->device));
    auto *s = new (std::nothrow) Scene();
    if (!s) return fail(DR_E_OOM, "host allocation failed");
    s->ctx = ctx;
    s->n_tris = d->n_triangles;
    s->n_verts = d->n_vertices;
    s->n_mats = d->n_materials ? d->n_materials : 1;
    dr_status st = DR_OK;
    const size_t vb = sizeof(float) * 3 * (size_t)d->n_vertices;
    const size_t ib = sizeof(uint32_t) * 3 * (size_t)d->n_triangles;
    const size_t mb = sizeof(uint32_t) * (size_t)d->n_triangles;
    if (!st) st = s->verts.ensure(vb);
    if (!st) st = s->indices.ensure(ib);
    if (!st) st = s->tri_mat.ensure(mb);
    if (!st) st = s->materials.ensure(sizeof(dr_material) * s->n_mats);
    if (!st) st = s->flags.ensure(2048);
    if (st) { delete s; return st; }
    hipError_t e = hipSuccess;
    if (vb) e = hipMemcpy(s->verts.ptr, d->vertices, vb, hipMemcpyHostToDevice);
    if (e == hipSuccess && ib) e = hipMemcpy(s->indices.ptr, d->indices, ib, hipMemcpyHostToDevice);
    if (e == hipSuccess && mb) {
        if (d->tri_material) e = hipMemcpy(s->tri_mat.ptr, d->tri_material, mb, hipMemcpyHostToDevice);
        else e = hipMemset(s->tri_mat.ptr, 0, mb);
    }
    s->host_materials.resize(s->n_mats);
    if (d->n_materials) std::memcpy(s->host_materials.data(), d->materials, sizeof(dr_material) * d->n_materials);
    else {
        dr_material m{};
        m.kind = DR_MAT_LAMBERT;
        m.albedo[0] = m.albedo[1] = m.albedo[2] = 0.5f;
        m.ior = 1.0f;
        s->host_materials[0] = m;
    }
    if (e == hipSuccess)
        e = hipMemcpy(s->materials.ptr, s->host_materials.data(), sizeof(dr_material) * s->n_mats, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->flags.ptr, 0, 2048);
    if (e != hipSuccess) {
        delete s;
        return fail(DR_E_DEVICE, std::string("scene upload: ") + hipGetErrorString(e));
    }
    *out = reinterpret_cast<dr_scene *>(s);
    return DR_OK;
}

dr_status dr_scene_destroy(dr_scene *sc)
{
    if (!sc) return fail(DR_E_INVALID, "scene is null");
    auto *s = reinterpret_cast<Scene *>(sc);
    (void)hipSetDevice(s->ctx->device);
    (void)hipDeviceSynchronize();
    delete s;
    return DR_OK;
}

dr_status dr_scene_get_info(dr_scene *sc, dr_scene_info *info)
{
    if (!sc || !info) return fail(DR_E_INVALID, "null argument");
    auto *s = reinterpret_cast<Scene *>(sc);
    std::memset(info, 0, sizeof(*info));
    info->n_triangles = s->n_tris;
    info->n_nodes = s->n_nodes;
    info->n_leaves = s->n_leaves;
    info->depth = s->depth;
    info->build_ms = s->build_ms;
    info->sah_cost = s->sah_cost;
    info->n_references = s->n_refs;
    info->stack_high_water = 0;
    if (s->flags.ptr) {
        uint32_t f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        DR_HIP_TRY(hipMemcpy(f, s->flags.ptr, 32, hipMemcpyDeviceToHost));
        info->overflow_flag = f[0];
        for (int k = 0; k < 4; ++k) info->simd_stats[k] = f[2 + k];
        info->stack_high_water = f[6];
    }
    return DR_OK;
}

dr_status dr_scene_build_bvh(dr_scene *sc, const dr_bvh_params *p, dr_stream stream)
{
    if (!sc) return fail(DR_E_INVALID, "scene is null");
    auto *s = reinterpret_cast<Scene *>(sc);
    DR_HIP_TRY(hipSetDevice(s->ctx->device));
    return build_bvh(s, p, reinterpret_cast<hipStream_t>(stream));
}

dr_status dr_scene_refit(dr_scene *sc, const float *vertices, dr_stream stream)
{
    if (!sc || !vertices) return fail(DR_E_INVALID, "null argument");
    auto *s = reinterpret_cast<Scene *>(sc);
    if (!s->built) return fail(DR_E_INVALID, "scene BVH not built (call dr_scene_build_bvh)");
    DR_HIP_TRY(hipSetDevice(s->ctx->device));
    const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    DR_HIP_TRY(hipStreamSynchronize(st));
    DR_HIP_TRY(hipMemcpy(s->verts.ptr, vertices, sizeof(float) * 3 * (size_t)s->n_verts, hipMemcpyHostToDevice));
    s->lights_built = false;  // triangle areas changed
    return refit_bvh(s, st);
}

dr_status dr_scene_set_trace_schedule(dr_scene *sc, uint32_t closest_sched, uint32_t any_sched)
{
    if (!sc) return fail(DR_E_INVALID, "scene is null");
    if (closest_sched > 15 || any_sched > 15) return fail(DR_E_INVALID, "schedule must be 0..15");
    auto *s = reinterpret_cast<Scene *>(sc);
    s->sched_closest = closest_sched;
    s->sched_any = any_sched;
    return DR_OK;
}

dr_status dr_scene_set_trace_tuning(dr_scene *sc, uint32_t refill_min, uint32_t blocks_per_cu, uint32_t sort_key)
{
    if (!sc) return fail(DR_E_INVALID, "scene is null");
    if (refill_min < 1 || refill_min > 64) return fail(DR_E_INVALID, "refill_min must be 1..64");
    if (sort_key > 1) return fail(DR_E_INVALID, "sort_key must be 0 or 1");
    auto *s = reinterpret_cast<Scene *>(sc);
    s->refill_min = refill_min;
    s->blocks_per_cu = blocks_per_cu;
    s->sort_key = sort_key;
    return DR_OK;
}

dr_status dr_scene_set_render_lanes(dr_scene *sc, uint32_t lanes)
{
    if (!sc) return fail(DR_E_INVALID, "scene is null");
    if (lanes < 1 || lanes > kMaxRenderLanes) return fail(DR_E_INVALID, "render lanes must be 1..4");
    reinterpret_cast<Scene *>(sc)->render_lanes = lanes;
    return DR_OK;
}

dr_status dr_trace_closest(dr_scene *sc, const dr_ray *rays, uint32_t n, dr_hit *hits, dr_stream stream)
{
    if (!sc) return fail(DR_E_INVALID, "scene is null");
    return launch_trace(reinterpret_cast<Scene *>(sc), rays, n, hits, nullptr, nullptr, nullptr, false,
                        reinterpret_cast<hipStream_t>(stream));
}

dr_status dr_trace_closest_noalloc(dr_scene *sc, const dr_ray *rays, uint32_t n, dr_hit *hits, dr_stream stream)
{
    if (!sc) return fail(DR_E_INVALID, "scene is null");
    return launch_trace_noalloc(reinterpret_cast<Scene *>(sc), rays, n, hits, reinterpret_cast<hipStream_t>(stream));
}

dr_status dr_trace_any(dr_scene *sc, const dr_ray *rays, uint32_t n, uint32_t *occluded, dr_stream stream)
{
    if (!sc) return fail(DR_E_INVALID, "scene is null");
    return launch_trace(reinterpret_cast<Scene *>(sc), rays, n, nullptr, occluded, nullptr, nullptr, true,
                        reinterpret_cast<hipStream_t>(stream));
}

dr_status dr_trace_closest_counted(dr_scene *sc, const dr_ray *rays, uint32_t n, dr_hit *hits,
                                   uint32_t *box_tests, uint32_t *tri_tests, dr_stream stream)
{
    if (!sc) return fail(DR_E_INVALID, "scene is null");
    if (!box_tests || !tri_tests) return fail(DR_E_INVALID, "null counter buffer");
    return launch_trace(reinterpret_cast<Scene *>(sc), rays, n, hits, nullptr, box_tests, tri_tests, false,
                        reinterpret_cast<hipStream_t>(stream));
}

dr_status dr_scene_build_lights(dr_scene *sc, dr_stream stream)
{
    if (!sc) return fail(DR_E_INVALID, "scene is null");
    auto *s = reinterpret_cast<Scene *>(sc);
    DR_HIP_TRY(hipSetDevice(s->ctx->device));
    return build_lights(s, reinterpret_cast<hipStream_t>(stream));
}

dr_status dr_scene_light_count(dr_scene *sc, uint32_t *n_emit)
{
    if (!sc || !n_emit) return fail(DR_E_INVALID, "null argument");
    auto *s = reinterpret_cast<Scene *>(sc);
    if (!s->lights_built)
        if (dr_status st = build_lights(s, nullptr)) return st;
    *n_emit = s->n_emit;
    return DR_OK;
}

dr_status dr_scene_copy_lights(dr_scene *sc, void *entries, float *pdf_area)
{
    if (!sc) return fail(DR_E_INVALID, "scene is null");
    auto *s = reinterpret_cast<Scene *>(sc);
    if (!s->lights_built)
        if (dr_status st = build_lights(s, nullptr)) return st;
    DR_HIP_TRY(hipDeviceSynchronize());
    if (entries && s->n_emit) DR_HIP_TRY(hipMemcpy(entries, s->alias.ptr, 20ull * s->n_emit, hipMemcpyDeviceToHost));
    if (pdf_area) DR_HIP_TRY(hipMemcpy(pdf_area, s->light_pdf_area.ptr, 4ull * s->n_tris, hipMemcpyDeviceToHost));
    return DR_OK;
}

dr_status dr_render(dr_scene *sc, const dr_camera *camera, const dr_render_params *params, double *accum,
                    dr_stream stream)
{
    if (!sc) return fail(DR_E_INVALID, "scene is null");
    auto *s = reinterpret_cast<Scene *>(sc);
    DR_HIP_TRY(hipSetDevice(s->ctx->device));
    return render(s, camera, params, nullptr, 0u, accum, reinterpret_cast<hipStream_t>(stream));
}

dr_status dr_scene_ray_casts(dr_scene *sc, uint64_t *casts, int32_t reset, dr_stream stream)
{
    if (!sc || !casts) return fail(DR_E_INVALID, "null argument");
    auto *s = reinterpret_cast<Scene *>(sc);
    DR_HIP_TRY(hipSetDevice(s->ctx->device));
    const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    DR_HIP_TRY(hipStreamSynchronize(st));
    *casts = 0;
    if (!s->ray_casts.ptr) return DR_OK;
    DR_HIP_TRY(hipMemcpy(casts, s->ray_casts.ptr, sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (reset)
        if (dr_status e = count_ray_casts(s, 0, st, true)) return e;
    return DR_OK;
}

dr_status dr_render_guided(dr_scene *sc, const dr_camera *camera, const dr_render_params *params, dr_guide *guide,
                           uint32_t guide_flags, double *accum, dr_stream stream)
{
    if (!sc) return fail(DR_E_INVALID, "scene is null");
    auto *s = reinterpret_cast<Scene *>(sc);
    auto *g = reinterpret_cast<Guide *>(guide);
    if (g && g->ctx->device != s->ctx->device) return fail(DR_E_INVALID, "guide and scene on different devices");
    DR_HIP_TRY(hipSetDevice(s->ctx->device));
    return render(s, camera, params, g, guide_flags, accum, reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
