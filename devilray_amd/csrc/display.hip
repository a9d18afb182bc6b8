// Display step after accumulation, and the debug views of the first hit.
//
// dr_resolve_rgba8 replaces Renderer::createPixels (src/Renderer.cpp:51-90):
// per pixel rgb x (1 / w) (black where w = 0), clamped to [0, 1], sRGB-encoded or
// kept linear, rounded to 8 bits, packed R | G << 8 | B << 16 | 255 << 24,
// rows flipped. The 8-bit value is round(255 f(c)) for the encoding f: the
// number of decision thresholds f^-1((b - 1/2) / 255), b = 1..255, that c
// reaches. The thresholds are computed once on the host in double precision,
// so the result is the exact rounding of the encoding and the same on any
// device (a pow alone, on the device or the host, rounds differently next to
// a rounding boundary); a float pow only gives the first guess.
//
// dr_render_debug replaces the debug branch of sampleColor and
// sampleColorDebug (include/tracing/PathGeneration.hpp:274-299,377-429) for
// the modes of DebugOptions (include/DebugOptions.hpp:3-8). Per sample, the
// camera ray's closest hit gives
//   UV checker:  the checker pattern of the hit's uv (the reference's uv is
//                always (0, 0), so the dark square, 0.5) times the material's
//                debug colour; a miss adds nothing (not even to w);
//   barycentric: (1 - u - v, u, v), the weights of the triangle's A, B, C;
//   winding:     (1, 0.73, 0.47) when ((A - B) x (A - C)) . d > 0, else
//                (0.53, 0.82, 1); a miss adds black;
// and w grows by one per sample (apart from UV-checker misses). Samples are
// added in sample order in FP64.
#include "dr_internal.hpp"
#include "shade.hpp"

#include <cmath>
#include <vector>

namespace dr {

namespace {

constexpr int kT = 256;
constexpr int kThr = 255;

inline uint32_t blocks(uint64_t n) { return (uint32_t)((n + kT - 1) / kT); }

// the reference's sRGB encoding (src/Renderer.cpp:55-60) and its inverse
double srgb_inverse(double y)
{
    const double knee = 0.0031308 * 12.92;
    return y < knee ? y / 12.92 : std::pow((y + 0.055) / 1.055, 2.4);
}

dr_status ensure_thresholds(Context *ctx)
{
    if (ctx->display_thr.bytes >= sizeof(double) * 2 * kThr) return DR_OK;
    std::vector<double> t(2 * kThr);
    for (int b = 1; b <= kThr; ++b) {
        const double y = (b - 0.5) / 255.0;
        t[b - 1] = y;                          // linear
        t[kThr + b - 1] = srgb_inverse(y);     // sRGB
    }
    if (dr_status st = ctx->display_thr.ensure(sizeof(double) * t.size())) return st;
    DR_HIP_TRY(hipMemcpy(ctx->display_thr.ptr, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
    return DR_OK;
}

// The 8-bit value: the number of thresholds <= c (ascending). A float
// evaluation of the encoding gives a first guess within a step or two of it;
// the exact thresholds then settle it.
__device__ __forceinline__ uint32_t quantize(double c, const double *thr, bool srgb)
{
    if (!(c > 0.0)) return 0u;  // also NaN
    if (c >= 1.0) return 255u;
    const float x = (float)c;
    // fast exp/log suffice for the guess; the thresholds make the result exact
    const float y = !srgb ? x : (x < 0.0031308f ? x * 12.92f : 1.055f * __expf(__logf(x) * (1.0f / 2.4f)) - 0.055f);
    int b = (int)(y * 255.0f + 0.5f);
    b = b < 0 ? 0 : (b > kThr ? kThr : b);
    while (b > 0 && c < thr[b - 1]) --b;
    while (b < kThr && c >= thr[b]) ++b;
    return (uint32_t)b;
}

__global__ __launch_bounds__(kT) void k_resolve_rgba8(const double4 *__restrict__ acc, uint32_t w, uint32_t h,
                                                      const double *__restrict__ thr_g, uint32_t srgb,
                                                      uint32_t *__restrict__ out)
{
    __shared__ double thr[kThr];
    for (int k = threadIdx.x; k < kThr; k += kT) thr[k] = thr_g[k];
    __syncthreads();
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i >= w * h) return;
    const uint32_t x = i % w, y = i / w;
    const double4 a = acc[i];
    uint32_t r = 0, g = 0, b = 0;
    if (a.w != 0.0) {
        const double inv = 1.0 / a.w;
        r = quantize(a.x * inv, thr, srgb != 0);
        g = quantize(a.y * inv, thr, srgb != 0);
        b = quantize(a.z * inv, thr, srgb != 0);
    }
    out[(size_t)(h - 1 - y) * w + x] = r | (g << 8) | (b << 16) | (255u << 24);
}

__global__ __launch_bounds__(kT) void k_debug_rays(CamParams cam, uint32_t sample, uint64_t seed,
                                                   dr_ray *__restrict__ rays)
{
    const uint32_t p = blockIdx.x * kT + threadIdx.x;
    if (p >= cam.w * cam.h) return;
    rays[p] = camera_ray(cam, p, sample, seed);
}

__global__ __launch_bounds__(kT) void k_debug_accum(const dr_ray *__restrict__ rays, const dr_hit *__restrict__ hits,
                                                    uint32_t npix, const float *__restrict__ verts,
                                                    const uint32_t *__restrict__ idx,
                                                    const uint32_t *__restrict__ tri_mat,
                                                    const float4 *__restrict__ dcol, uint32_t mode,
                                                    double4 *__restrict__ acc)
{
    const uint32_t p = blockIdx.x * kT + threadIdx.x;
    if (p >= npix) return;
    const dr_hit hh = hits[p];
    float cr = 0.0f, cg = 0.0f, cb = 0.0f;
    if (hh.id >= 0) {
        const uint32_t t = (uint32_t)hh.id;
        if (mode == DR_DEBUG_UV_CHECKER) {
            const float4 c = dcol[tri_mat[t]];
            cr = 0.5f * c.x; cg = 0.5f * c.y; cb = 0.5f * c.z;
        } else if (mode == DR_DEBUG_BARYCENTRIC) {
            cr = 1.0f - hh.u - hh.v; cg = hh.u; cb = hh.v;
        } else {
            const uint32_t ia = idx[3 * (size_t)t], ib = idx[3 * (size_t)t + 1], ic = idx[3 * (size_t)t + 2];
            const V3 A{verts[3 * (size_t)ia], verts[3 * (size_t)ia + 1], verts[3 * (size_t)ia + 2]};
            const V3 B{verts[3 * (size_t)ib], verts[3 * (size_t)ib + 1], verts[3 * (size_t)ib + 2]};
            const V3 C{verts[3 * (size_t)ic], verts[3 * (size_t)ic + 1], verts[3 * (size_t)ic + 2]};
            const V3 n = cross(sub(A, B), sub(A, C));
            const dr_ray r = rays[p];
            const bool ccw = n.x * r.dx + n.y * r.dy + n.z * r.dz > 0.0f;
            cr = ccw ? 1.0f : 0.53f; cg = ccw ? 0.73f : 0.82f; cb = ccw ? 0.47f : 1.0f;
        }
    } else if (mode == DR_DEBUG_UV_CHECKER) {
        return;  // sampleColorDebug returns before counting the sample
    }
    double4 a = acc[p];
    a.x += (double)cr; a.y += (double)cg; a.z += (double)cb; a.w += 1.0;
    acc[p] = a;
}

}  // namespace

}  // namespace dr

using namespace dr;

extern "C" {

dr_status dr_resolve_rgba8(dr_context *cx, const double *accum, uint32_t width, uint32_t height, uint32_t output,
                           uint32_t *rgba, dr_stream stream)
{
    if (!cx || !accum || !rgba) return fail(DR_E_INVALID, "null argument");
    if (width == 0 || height == 0 || (uint64_t)width * height > 0xFFFFFFFFull)
        return fail(DR_E_INVALID, "image size out of range");
    if (output != DR_OUTPUT_LINEAR && output != DR_OUTPUT_SRGB) return fail(DR_E_INVALID, "unknown output encoding");
    auto *ctx = reinterpret_cast<Context *>(cx);
    DR_HIP_TRY(hipSetDevice(ctx->device));
    if (dr_status st = ensure_thresholds(ctx)) return st;
    const uint32_t n = width * height;
    k_resolve_rgba8<<<blocks(n), kT, 0, (hipStream_t)stream>>>(
        reinterpret_cast<const double4 *>(accum), width, height,
        ctx->display_thr.as<double>() + (output == DR_OUTPUT_SRGB ? kThr : 0), output == DR_OUTPUT_SRGB ? 1u : 0u,
        rgba);
    DR_HIP_TRY(hipGetLastError());
    return DR_OK;
}

dr_status dr_scene_set_debug_colors(dr_scene *sc, const float *rgba, uint32_t n_materials)
{
    if (!sc || !rgba) return fail(DR_E_INVALID, "null argument");
    auto *s = reinterpret_cast<Scene *>(sc);
    if (n_materials != s->n_mats) return fail(DR_E_INVALID, "one colour per material is needed");
    DR_HIP_TRY(hipSetDevice(s->ctx->device));
    if (dr_status st = s->debug_colors.ensure(16ull * (n_materials ? n_materials : 1))) return st;
    if (n_materials) DR_HIP_TRY(hipMemcpy(s->debug_colors.ptr, rgba, 16ull * n_materials, hipMemcpyHostToDevice));
    s->debug_colors_set = true;
    return DR_OK;
}

dr_status dr_render_debug(dr_scene *sc, const dr_camera *camera, const dr_render_params *params, uint32_t mode,
                          double *accum, dr_stream stream)
{
    if (!sc || !camera || !params || !accum) return fail(DR_E_INVALID, "null argument");
    if (mode != DR_DEBUG_UV_CHECKER && mode != DR_DEBUG_BARYCENTRIC && mode != DR_DEBUG_WINDING)
        return fail(DR_E_INVALID, "unknown debug mode");
    auto *s = reinterpret_cast<Scene *>(sc);
    if (!s->built) return fail(DR_E_INVALID, "BVH not built");
    if (camera->width == 0 || camera->height == 0 || (uint64_t)camera->width * camera->height > (1ull << 30))
        return fail(DR_E_INVALID, "image size out of range");
    DR_HIP_TRY(hipSetDevice(s->ctx->device));
    const hipStream_t st = (hipStream_t)stream;
    const uint32_t npix = camera->width * camera->height;
    if (!s->debug_colors_set) {  // default: each material's albedo
        std::vector<float> c(4 * std::max<size_t>(s->host_materials.size(), 1), 0.0f);
        for (size_t m = 0; m < s->host_materials.size(); ++m)
            for (int k = 0; k < 3; ++k) c[4 * m + k] = s->host_materials[m].albedo[k];
        if (dr_status e = s->debug_colors.ensure(sizeof(float) * c.size())) return e;
        DR_HIP_TRY(hipMemcpy(s->debug_colors.ptr, c.data(), sizeof(float) * c.size(), hipMemcpyHostToDevice));
        s->debug_colors_set = true;
    }
    if (dr_status e = s->dbg_rays.ensure(sizeof(dr_ray) * (size_t)npix)) return e;
    if (dr_status e = s->dbg_hits.ensure(sizeof(dr_hit) * (size_t)npix)) return e;
    const CamParams cam = cam_params(*camera);
    if (dr_status e = count_ray_casts(s, (uint64_t)npix * params->spp, st)) return e;
    for (uint32_t k = 0; k < params->spp; ++k) {
        k_debug_rays<<<blocks(npix), kT, 0, st>>>(cam, params->sample_offset + k, params->seed,
                                                  s->dbg_rays.as<dr_ray>());
        DR_HIP_TRY(hipGetLastError());
        if (dr_status e = launch_trace(s, s->dbg_rays.as<dr_ray>(), npix, s->dbg_hits.as<dr_hit>(), nullptr, nullptr,
                                       nullptr, false, st))
            return e;
        k_debug_accum<<<blocks(npix), kT, 0, st>>>(s->dbg_rays.as<dr_ray>(), s->dbg_hits.as<dr_hit>(), npix,
                                                   s->verts.as<float>(), s->indices.as<uint32_t>(),
                                                   s->tri_mat.as<uint32_t>(), s->debug_colors.as<float4>(), mode,
                                                   reinterpret_cast<double4 *>(accum));
        DR_HIP_TRY(hipGetLastError());
    }
    return DR_OK;
}

}  // extern "C"
