// Instancing: built scenes placed several times with a scale-and-translate
// transform each, traced through a two-level forest.
//
// The reference lets several TriangleMesh views, each with its own
// model_to_world Transform, share one GpuTris storage (src/GpuTris.cpp:61-85)
// and tests a ray against every object in model space
// (include/tracing/IntersectionTestsImpl.hpp:143-151, 234-251; the transform
// of include/Transform.hpp:5-26). Here each scene's tree is stored once and
// a top-level 4-wide tree over the instances' world boxes selects which
// instances a ray enters (trace.hip, INST).
//
// Forest layout: nodes [0, n_top) are the top tree, its leaf references name
// an instance; then each scene's nodes with their references moved by the
// scene's node and triangle-reference offsets. Triangle references keep their
// scene-local ids; an instance adds its id_delta, the triangle count of the
// instances before it, so hit ids number the instances' triangles in
// instance order.
//
// The top tree is built by the scene builder (bvh_build.hip) over one
// degenerate triangle per instance whose box is the instance's world box
// (vertices lo, hi, lo), one reference per leaf and no splits; its leaf
// references are then renamed to the instances they hold.
#include "dr_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

namespace dr {

struct Instances {
    Context *ctx = nullptr;
    Scene forest;
    uint32_t n_scenes = 0;
};

namespace {

__global__ void k_offset_u32(uint32_t *__restrict__ dst, const uint32_t *__restrict__ src, size_t n, uint32_t off)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] + off;
}

// The shading data of a forest: every scene's vertices, triangles (indices
// moved by the scene's vertex offset) and materials (ids moved by its
// material offset) once, and one InstShade per instance.
dr_status shading_data(Scene &f, Scene *const *sc, uint32_t n_sc, const dr_instance *in, uint32_t n_in)
{
    std::vector<uint32_t> voff(n_sc), moff(n_sc);
    f.cat_tri_off.assign(n_sc + 1, 0u);
    uint64_t nv = 0, nt = 0, nm = 0;
    for (uint32_t k = 0; k < n_sc; ++k) {
        voff[k] = (uint32_t)nv;
        moff[k] = (uint32_t)nm;
        f.cat_tri_off[k] = (uint32_t)nt;
        nv += sc[k]->n_verts;
        nt += sc[k]->n_tris;
        nm += sc[k]->n_mats;
    }
    if (nv > 0xFFFFFFFFull || nt > 0x7FFFFFFFull || nm > 0xFFFFFFFFull)
        return fail(DR_E_INVALID, "instanced scenes hold too many vertices, triangles or materials");
    f.cat_tri_off[n_sc] = (uint32_t)nt;
    f.n_verts = (uint32_t)nv;
    f.n_cat_tris = (uint32_t)nt;
    f.n_mats = (uint32_t)nm;
    dr_status st = f.verts.ensure(12ull * std::max<uint64_t>(nv, 1));
    if (!st) st = f.indices.ensure(12ull * std::max<uint64_t>(nt, 1));
    if (!st) st = f.tri_mat.ensure(4ull * std::max<uint64_t>(nt, 1));
    if (!st) st = f.materials.ensure(sizeof(dr_material) * std::max<uint64_t>(nm, 1));
    if (!st) st = f.inst_shade.ensure(sizeof(InstShade) * (size_t)n_in);
    if (st) return st;
    f.host_materials.clear();
    for (uint32_t k = 0; k < n_sc; ++k) {
        const Scene *s = sc[k];
        if (s->n_verts)
            DR_HIP_TRY(hipMemcpy(f.verts.as<float>() + 3ull * voff[k], s->verts.ptr, 12ull * s->n_verts,
                                 hipMemcpyDeviceToDevice));
        if (s->n_tris) {
            const size_t ni = 3ull * s->n_tris;
            k_offset_u32<<<(unsigned)((ni + 255) / 256), 256>>>(f.indices.as<uint32_t>() + 3ull * f.cat_tri_off[k],
                                                               s->indices.as<uint32_t>(), ni, voff[k]);
            k_offset_u32<<<(unsigned)((s->n_tris + 255) / 256), 256>>>(f.tri_mat.as<uint32_t>() + f.cat_tri_off[k],
                                                                      s->tri_mat.as<uint32_t>(), s->n_tris, moff[k]);
            DR_HIP_TRY(hipGetLastError());
        }
        f.host_materials.insert(f.host_materials.end(), s->host_materials.begin(), s->host_materials.end());
    }
    if (nm)
        DR_HIP_TRY(hipMemcpy(f.materials.ptr, f.host_materials.data(), sizeof(dr_material) * nm,
                             hipMemcpyHostToDevice));
    std::vector<InstShade> rec(n_in);
    f.inst_scene.resize(n_in);
    uint64_t id_off = 0;
    for (uint32_t i = 0; i < n_in; ++i) {
        const uint32_t k = in[i].scene;
        InstShade &r = rec[i];
        r.sx = in[i].scale[0];
        r.sy = in[i].scale[1];
        r.sz = in[i].scale[2];
        r.tx = in[i].translate[0];
        r.ty = in[i].translate[1];
        r.tz = in[i].translate[2];
        r.tri_off = f.cat_tri_off[k];
        r.id_delta = (uint32_t)id_off;
        f.inst_scene[i] = k;
        id_off += sc[k]->n_tris;
    }
    DR_HIP_TRY(hipMemcpy(f.inst_shade.ptr, rec.data(), sizeof(InstShade) * n_in, hipMemcpyHostToDevice));
    DR_HIP_TRY(hipDeviceSynchronize());
    return DR_OK;
}

// The world box of an instance: the scene's root box through p*s + t, padded
// by `pad` on every side.
void world_box(const Scene *s, const dr_instance &in, float pad, float lo[3], float hi[3])
{
    for (int a = 0; a < 3; ++a) {
        const float e0 = s->bounds[a] * in.scale[a] + in.translate[a];
        const float e1 = s->bounds[3 + a] * in.scale[a] + in.translate[a];
        lo[a] = std::min(e0, e1) - pad;
        hi[a] = std::max(e0, e1) + pad;
    }
}

dr_status build_forest(Instances *I, Scene *const *sc, uint32_t n_sc, const dr_instance *in, uint32_t n_in)
{
    Scene &f = I->forest;
    f.ctx = I->ctx;
    // per-scene offsets in the forest (after the top tree) and triangle counts
    std::vector<uint64_t> node_off(n_sc), ref_off(n_sc);
    uint64_t nn = 0, nr = 0;
    uint32_t blas_depth = 0;
    for (uint32_t k = 0; k < n_sc; ++k) {
        node_off[k] = nn;
        ref_off[k] = nr;
        nn += sc[k]->n_nodes;
        nr += sc[k]->n_refs;
        blas_depth = std::max(blas_depth, sc[k]->depth);
    }
    if (nr > kMaxLeafRefs) return fail(DR_E_INVALID, "instanced scenes hold more than 2^28 - 2 triangle references");
    uint64_t total_tris = 0;
    for (uint32_t i = 0; i < n_in; ++i) total_tris += sc[in[i].scene]->n_tris;
    if (total_tris > 0x7FFFFFFFull) return fail(DR_E_INVALID, "instances hold more than 2^31 - 1 triangles in all");

    // world boxes; the pad covers the rounding of the model-space ray (a few
    // ulps of the coordinates' magnitude) so the top tree never drops an
    // instance that the exact ray touches
    float mag = 0.0f;
    for (uint32_t i = 0; i < n_in; ++i) {
        float lo[3], hi[3];
        world_box(sc[in[i].scene], in[i], 0.0f, lo, hi);
        for (int a = 0; a < 3; ++a) mag = std::max({mag, std::fabs(lo[a]), std::fabs(hi[a])});
    }
    if (!std::isfinite(mag)) return fail(DR_E_INVALID, "instance world box is not finite");
    const float pad = mag * 0x1p-16f;
    std::vector<float> tv(9ull * n_in);
    std::vector<uint32_t> ti(3ull * n_in);
    for (uint32_t i = 0; i < n_in; ++i) {
        float lo[3], hi[3];
        world_box(sc[in[i].scene], in[i], pad, lo, hi);
        for (int a = 0; a < 3; ++a) {
            tv[9ull * i + a] = lo[a];
            tv[9ull * i + 3 + a] = hi[a];
            tv[9ull * i + 6 + a] = lo[a];
        }
        for (int c = 0; c < 3; ++c) ti[3ull * i + c] = 3u * i + (uint32_t)c;
    }

    // the top tree, built as a scene of instance boxes
    Scene top;
    top.ctx = I->ctx;
    top.n_tris = n_in;
    top.n_verts = 3u * n_in;
    dr_status st = top.verts.ensure(sizeof(float) * tv.size());
    if (!st) st = top.indices.ensure(sizeof(uint32_t) * ti.size());
    if (st) return st;
    DR_HIP_TRY(hipMemcpy(top.verts.ptr, tv.data(), sizeof(float) * tv.size(), hipMemcpyHostToDevice));
    DR_HIP_TRY(hipMemcpy(top.indices.ptr, ti.data(), sizeof(uint32_t) * ti.size(), hipMemcpyHostToDevice));
    dr_bvh_params bp{};
    bp.leaf_max = 1;
    bp.flags = DR_BVH_NO_SPLIT;
    if ((st = build_bvh(&top, &bp, nullptr))) return st;
    const uint32_t n_top = top.n_nodes;
    std::vector<QNode> tn(n_top);
    // leaf-ordered references, 12 words each; word 3 holds the id
    std::vector<uint32_t> tref(12ull * top.n_refs);
    DR_HIP_TRY(hipMemcpy(tn.data(), top.nodes.ptr, sizeof(QNode) * n_top, hipMemcpyDeviceToHost));
    DR_HIP_TRY(hipMemcpy(tref.data(), top.tri_v.ptr, sizeof(uint32_t) * tref.size(), hipMemcpyDeviceToHost));
    for (QNode &w : tn)
        for (int c = 0; c < 4; ++c) {
            const int32_t r = w.child[c];
            if (r >= 0 || r == DR_EMPTY_REF) continue;
            const uint32_t first = leaf_first(r);
            if (leaf_count(r) != 1u || first >= top.n_refs)
                return fail(DR_E_DEVICE, "top-level tree leaf does not hold exactly one instance");
            const uint32_t id = tref[12ull * first + 3];
            if (id >= n_in) return fail(DR_E_DEVICE, "top-level tree names an unknown instance");
            w.child[c] = make_leaf_ref(id, 1u);
        }
    if ((uint64_t)n_top + nn >= (1ull << 31)) return fail(DR_E_INVALID, "forest too large");

    // the forest
    f.n_inst = n_in;
    f.n_top = n_top;
    f.n_nodes = (uint32_t)(n_top + nn);
    f.n_refs = (uint32_t)nr;
    f.n_tris = (uint32_t)total_tris;
    f.n_leaves = 0;
    for (uint32_t k = 0; k < n_sc; ++k) f.n_leaves += sc[k]->n_leaves;
    if (!st) st = f.nodes.ensure(sizeof(QNode) * (size_t)f.n_nodes);
    if (!st) st = f.tri_v.ensure(48ull * std::max<uint64_t>(nr, 1));
    if (!st) st = f.inst_recs.ensure(sizeof(InstRec) * (size_t)n_in);
    if (!st) st = f.flags.ensure(2048);
    if (st) return st;
    DR_HIP_TRY(hipMemset(f.flags.ptr, 0, 2048));
    DR_HIP_TRY(hipMemcpy(f.nodes.ptr, tn.data(), sizeof(QNode) * n_top, hipMemcpyHostToDevice));
    for (uint32_t k = 0; k < n_sc; ++k) {
        const Scene *s = sc[k];
        if ((st = rebase_nodes(s->nodes.as<QNode>(), s->n_nodes, (uint32_t)(n_top + node_off[k]),
                               (uint32_t)ref_off[k], f.nodes.as<QNode>() + n_top + node_off[k], nullptr)))
            return st;
        if (s->n_refs)
            DR_HIP_TRY(hipMemcpy(f.tri_v.as<float4>() + 3 * ref_off[k], s->tri_v.ptr, 48ull * s->n_refs,
                                 hipMemcpyDeviceToDevice));
    }
    std::vector<InstRec> rec(n_in);
    uint64_t id_off = 0;
    for (uint32_t i = 0; i < n_in; ++i) {
        const uint32_t k = in[i].scene;
        InstRec &r = rec[i];
        r.root = (int32_t)(n_top + node_off[k]);  // a scene's root is its node 0
        // inverse scale as Transform::applyInverse takes it (s.inv(), then products)
        r.isx = 1.0f / in[i].scale[0];
        r.isy = 1.0f / in[i].scale[1];
        r.isz = 1.0f / in[i].scale[2];
        r.tx = in[i].translate[0];
        r.ty = in[i].translate[1];
        r.tz = in[i].translate[2];
        r.id_delta = (uint32_t)id_off;
        id_off += sc[k]->n_tris;
    }
    DR_HIP_TRY(hipMemcpy(f.inst_recs.ptr, rec.data(), sizeof(InstRec) * n_in, hipMemcpyHostToDevice));
    DR_HIP_TRY(hipDeviceSynchronize());
    for (int a = 0; a < 6; ++a) f.bounds[a] = top.bounds[a];
    // both levels on one stack: the top tree's levels, the marker, the
    // deepest instanced tree's levels
    f.depth = top.depth + blas_depth;
    f.build_ms = top.build_ms;
    if ((st = shading_data(f, sc, n_sc, in, n_in))) return st;
    f.built = true;
    I->n_scenes = n_sc;
    return DR_OK;
}

}  // namespace

}  // namespace dr

using namespace dr;

extern "C" {

dr_status dr_instances_create(dr_context *cx, dr_scene *const *scenes, uint32_t n_scenes,
                              const dr_instance *instances, uint32_t n_instances, dr_instances **out)
{
    if (!cx || !scenes || !instances || !out || n_scenes == 0 || n_instances == 0)
        return fail(DR_E_INVALID, "null argument, no scenes or no instances");
    *out = nullptr;
    if (n_instances > kMaxLeafRefs) return fail(DR_E_INVALID, "more than 2^28 - 2 instances");
    auto *ctx = reinterpret_cast<Context *>(cx);
    DR_HIP_TRY(hipSetDevice(ctx->device));
    std::vector<Scene *> sc(n_scenes);
    for (uint32_t k = 0; k < n_scenes; ++k) {
        sc[k] = reinterpret_cast<Scene *>(scenes[k]);
        if (!sc[k]) return fail(DR_E_INVALID, "null scene");
        if (sc[k]->ctx != ctx) return fail(DR_E_INVALID, "scene of another context");
        if (!sc[k]->built) return fail(DR_E_INVALID, "scene BVH not built (call dr_scene_build_bvh)");
        if (sc[k]->n_inst) return fail(DR_E_INVALID, "a scene of instances cannot be instanced");
    }
    for (uint32_t i = 0; i < n_instances; ++i) {
        const dr_instance &in = instances[i];
        if (in.scene >= n_scenes) return fail(DR_E_INVALID, "instance scene index out of range");
        for (int a = 0; a < 3; ++a) {
            if (!std::isfinite(in.scale[a]) || in.scale[a] == 0.0f || !std::isfinite(1.0f / in.scale[a]))
                return fail(DR_E_INVALID, "instance scale must be finite, non-zero and invertible");
            if (!std::isfinite(in.translate[a])) return fail(DR_E_INVALID, "instance translation must be finite");
        }
    }
    auto *I = new (std::nothrow) Instances();
    if (!I) return fail(DR_E_OOM, "host allocation failed");
    I->ctx = ctx;
    if (dr_status st = build_forest(I, sc.data(), n_scenes, instances, n_instances)) {
        delete I;
        return st;
    }
    *out = reinterpret_cast<dr_instances *>(I);
    return DR_OK;
}

dr_status dr_instances_destroy(dr_instances *ih)
{
    if (!ih) return DR_OK;
    auto *I = reinterpret_cast<Instances *>(ih);
    (void)hipSetDevice(I->ctx->device);
    (void)hipDeviceSynchronize();
    delete I;
    return DR_OK;
}

dr_status dr_instances_set_trace_schedule(dr_instances *ih, uint32_t closest_sched, uint32_t any_sched)
{
    if (!ih) return fail(DR_E_INVALID, "instances handle is null");
    if (closest_sched > 15u || any_sched > 15u) return fail(DR_E_INVALID, "unknown trace schedule");
    auto *I = reinterpret_cast<Instances *>(ih);
    I->forest.sched_closest = closest_sched;
    I->forest.sched_any = any_sched;
    return DR_OK;
}

dr_status dr_instances_trace_closest(dr_instances *ih, const dr_ray *rays, uint32_t n, dr_hit *hits,
                                     dr_stream stream)
{
    if (!ih) return fail(DR_E_INVALID, "instances handle is null");
    auto *I = reinterpret_cast<Instances *>(ih);
    return launch_trace(&I->forest, rays, n, hits, nullptr, nullptr, nullptr, false,
                        reinterpret_cast<hipStream_t>(stream));
}

dr_status dr_instances_trace_any(dr_instances *ih, const dr_ray *rays, uint32_t n, uint32_t *occluded,
                                 dr_stream stream)
{
    if (!ih) return fail(DR_E_INVALID, "instances handle is null");
    auto *I = reinterpret_cast<Instances *>(ih);
    return launch_trace(&I->forest, rays, n, nullptr, occluded, nullptr, nullptr, true,
                        reinterpret_cast<hipStream_t>(stream));
}

dr_status dr_instances_render(dr_instances *ih, const dr_camera *camera, const dr_render_params *params,
                              double *accum, dr_stream stream)
{
    if (!ih) return fail(DR_E_INVALID, "instances handle is null");
    auto *I = reinterpret_cast<Instances *>(ih);
    DR_HIP_TRY(hipSetDevice(I->ctx->device));
    return render_instances(&I->forest, camera, params, accum, reinterpret_cast<hipStream_t>(stream));
}

dr_status dr_instances_light_count(dr_instances *ih, uint32_t *n_emit)
{
    if (!ih || !n_emit) return fail(DR_E_INVALID, "null argument");
    auto *I = reinterpret_cast<Instances *>(ih);
    DR_HIP_TRY(hipSetDevice(I->ctx->device));
    if (!I->forest.lights_built)
        if (dr_status st = build_instance_lights(&I->forest, nullptr)) return st;
    *n_emit = I->forest.n_emit;
    return DR_OK;
}

dr_status dr_instances_copy_lights(dr_instances *ih, void *entries, uint32_t *emit_tri, float *emit_pdf)
{
    if (!ih) return fail(DR_E_INVALID, "instances handle is null");
    auto *I = reinterpret_cast<Instances *>(ih);
    Scene &f = I->forest;
    DR_HIP_TRY(hipSetDevice(I->ctx->device));
    if (!f.lights_built)
        if (dr_status st = build_instance_lights(&f, nullptr)) return st;
    DR_HIP_TRY(hipDeviceSynchronize());
    if (f.n_emit) {
        if (entries) DR_HIP_TRY(hipMemcpy(entries, f.alias.ptr, 20ull * f.n_emit, hipMemcpyDeviceToHost));
        if (emit_tri) DR_HIP_TRY(hipMemcpy(emit_tri, f.emit_tri.ptr, 4ull * f.n_emit, hipMemcpyDeviceToHost));
        if (emit_pdf) DR_HIP_TRY(hipMemcpy(emit_pdf, f.emit_pdf.ptr, 4ull * f.n_emit, hipMemcpyDeviceToHost));
    }
    return DR_OK;
}

dr_status dr_instances_get_info(dr_instances *ih, dr_scene_info *info)
{
    if (!ih || !info) return fail(DR_E_INVALID, "null argument");
    auto *I = reinterpret_cast<Instances *>(ih);
    const Scene &f = I->forest;
    std::memset(info, 0, sizeof(*info));
    info->n_triangles = f.n_tris;
    info->n_nodes = f.n_nodes;
    info->n_leaves = f.n_leaves;
    info->depth = f.depth;
    info->build_ms = f.build_ms;
    info->n_references = f.n_refs;
    uint32_t flag = 0;
    DR_HIP_TRY(hipMemcpy(&flag, f.flags.ptr, 4, hipMemcpyDeviceToHost));
    info->overflow_flag = flag;
    return DR_OK;
}

}  // extern "C"
