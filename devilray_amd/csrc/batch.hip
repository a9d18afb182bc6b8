// Batched rendering: built scenes merged into one device forest (vertex,
// triangle, material, node and light arrays concatenated, references rebased)
// plus a per-scene table of BVH roots and light-table ranges; views of all
// scenes then render through the same wavefront kernels in chunks of paths,
// the bounce loop optionally recorded once as a HIP graph.
#include "dr_internal.hpp"

#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include <vector>

namespace dr {

namespace {

constexpr int kT = 256;

struct AliasEntryB {
    float p_a, pdf_a, pdf_b;
    int32_t a, b;
};
static_assert(sizeof(AliasEntryB) == 20, "alias entry is 20 B");

inline uint32_t blocks(uint64_t n) { return (uint32_t)((n + kT - 1) / kT); }

__global__ __launch_bounds__(kT) void k_add_u32(const uint32_t *__restrict__ src, uint64_t n, uint32_t off,
                                                uint32_t *__restrict__ dst)
{
    const uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
    if (i < n) dst[i] = src[i] + off;
}

// leaf-ordered triangles: the id bits in the first vertex's w move by off
__global__ __launch_bounds__(kT) void k_rebase_triv(const float4 *__restrict__ src, uint32_t n, uint32_t off,
                                                    float4 *__restrict__ dst)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    float4 a = src[3 * (size_t)i];
    a.w = __int_as_float((int32_t)((uint32_t)__float_as_int(a.w) + off));
    dst[3 * (size_t)i] = a;
    dst[3 * (size_t)i + 1] = src[3 * (size_t)i + 1];
    dst[3 * (size_t)i + 2] = src[3 * (size_t)i + 2];
}

// interior references move by node_off, leaf references by ref_off
__global__ __launch_bounds__(kT) void k_rebase_nodes(const QNode *__restrict__ src, uint32_t n, uint32_t node_off,
                                                     uint32_t ref_off, QNode *__restrict__ dst)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    QNode w = src[i];
    for (int c = 0; c < 4; ++c) {
        const int32_t r = w.child[c];
        if (r >= 0) w.child[c] = r + (int32_t)node_off;
        else if (r != DR_EMPTY_REF) w.child[c] = make_leaf_ref(leaf_first(r) + ref_off, leaf_count(r));
    }
    dst[i] = w;
}

__global__ __launch_bounds__(kT) void k_rebase_alias(const AliasEntryB *__restrict__ src, uint32_t n, uint32_t tri_off,
                                                     AliasEntryB *__restrict__ dst)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    AliasEntryB e = src[i];
    if (e.a >= 0) e.a += (int32_t)tri_off;
    if (e.b >= 0) e.b += (int32_t)tri_off;
    dst[i] = e;
}

}  // namespace

dr_status rebase_nodes(const QNode *src, uint32_t n, uint32_t node_off, uint32_t ref_off, QNode *dst,
                       hipStream_t stream)
{
    if (n == 0) return DR_OK;
    k_rebase_nodes<<<blocks(n), kT, 0, stream>>>(src, n, node_off, ref_off, dst);
    DR_HIP_TRY(hipGetLastError());
    return DR_OK;
}

struct Batch {
    Context *ctx = nullptr;
    Scene merged;
    uint32_t n_scenes = 0;
    DevBuf ranges;  // SceneRange per scene
    DevBuf views;   // ViewInfo table of the last render call
    hipStream_t stream = nullptr;  // capture stream when the caller passes the null stream
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    ~Batch()
    {
        merged.release_graph();  // before the stream its replays ran on goes away
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

dr_status merge(Batch *b, Scene *const *sc, uint32_t n)
{
    Scene &m = b->merged;
    m.ctx = b->ctx;
    uint64_t nv = 0, nt = 0, nm = 0, nn = 0, ne = 0, nr = 0;
    for (uint32_t k = 0; k < n; ++k) {
        nv += sc[k]->n_verts;
        nt += sc[k]->n_tris;
        nr += sc[k]->n_refs;
        nm += sc[k]->n_mats;
        nn += sc[k]->n_nodes;
        ne += sc[k]->n_emit;
    }
    if (nr > kMaxLeafRefs || nt >= (1ull << 31) || nn >= (1ull << 31) || nv >= (1ull << 32))
        return fail(DR_E_INVALID, "batch too large: more than 2^28 - 2 triangle references in all");
    m.n_verts = (uint32_t)nv;
    m.n_tris = (uint32_t)nt;
    m.n_refs = (uint32_t)nr;
    m.n_mats = (uint32_t)nm;
    m.n_nodes = (uint32_t)nn;
    m.n_emit = (uint32_t)ne;
    dr_status st = DR_OK;
    auto need = [&](DevBuf &buf, size_t bytes) { if (st == DR_OK) st = buf.ensure(bytes); };
    need(m.verts, 12ull * nv);
    need(m.indices, 12ull * nt);
    need(m.tri_mat, 4ull * nt);
    need(m.materials, sizeof(dr_material) * nm);
    need(m.tri_v, 48ull * nr);
    need(m.nodes, sizeof(QNode) * nn);
    need(m.light_pdf_area, 4ull * nt);
    need(m.alias, sizeof(AliasEntryB) * std::max<uint64_t>(ne, 1));
    need(m.flags, 2048);
    need(b->ranges, sizeof(SceneRange) * n);
    if (st) return st;
    DR_HIP_TRY(hipMemset(m.flags.ptr, 0, 2048));
    std::vector<SceneRange> ranges(n);
    m.host_materials.clear();
    uint64_t vo = 0, to = 0, mo = 0, no = 0, eo = 0, ro = 0;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint32_t depth = 0, leaves = 0;
    const hipStream_t s0 = nullptr;
    for (uint32_t k = 0; k < n; ++k) {
        const Scene *s = sc[k];
        if (s->n_verts)
            DR_HIP_TRY(hipMemcpyAsync(m.verts.as<float>() + 3 * vo, s->verts.ptr, 12ull * s->n_verts,
                                      hipMemcpyDeviceToDevice, s0));
        if (s->n_tris) {
            k_add_u32<<<blocks(3ull * s->n_tris), kT, 0, s0>>>(s->indices.as<uint32_t>(), 3ull * s->n_tris,
                                                               (uint32_t)vo, m.indices.as<uint32_t>() + 3 * to);
            k_add_u32<<<blocks(s->n_tris), kT, 0, s0>>>(s->tri_mat.as<uint32_t>(), s->n_tris, (uint32_t)mo,
                                                        m.tri_mat.as<uint32_t>() + to);
            k_rebase_triv<<<blocks(s->n_refs), kT, 0, s0>>>(s->tri_v.as<float4>(), s->n_refs, (uint32_t)to,
                                                            m.tri_v.as<float4>() + 3 * ro);
            DR_HIP_TRY(hipMemcpyAsync(m.light_pdf_area.as<float>() + to, s->light_pdf_area.ptr, 4ull * s->n_tris,
                                      hipMemcpyDeviceToDevice, s0));
        }
        if (dr_status st2 = rebase_nodes(s->nodes.as<QNode>(), s->n_nodes, (uint32_t)no, (uint32_t)ro,
                                         m.nodes.as<QNode>() + no, s0))
            return st2;
        if (s->n_emit)
            k_rebase_alias<<<blocks(s->n_emit), kT, 0, s0>>>(s->alias.as<AliasEntryB>(), s->n_emit, (uint32_t)to,
                                                             m.alias.as<AliasEntryB>() + eo);
        DR_HIP_TRY(hipGetLastError());
        m.host_materials.insert(m.host_materials.end(), s->host_materials.begin(), s->host_materials.end());
        ranges[k] = SceneRange{(uint32_t)no, (uint32_t)eo, s->n_emit, 0u};
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], s->bounds[a]);
            hi[a] = std::max(hi[a], s->bounds[3 + a]);
        }
        depth = std::max(depth, s->depth);
        leaves += s->n_leaves;
        vo += s->n_verts;
        to += s->n_tris;
        ro += s->n_refs;
        mo += s->n_mats;
        no += s->n_nodes;
        eo += s->n_emit;
    }
    DR_HIP_TRY(hipMemcpy(m.materials.ptr, m.host_materials.data(), sizeof(dr_material) * nm, hipMemcpyHostToDevice));
    DR_HIP_TRY(hipMemcpy(b->ranges.ptr, ranges.data(), sizeof(SceneRange) * n, hipMemcpyHostToDevice));
    DR_HIP_TRY(hipDeviceSynchronize());
    for (int a = 0; a < 3; ++a) {
        m.bounds[a] = lo[a];
        m.bounds[3 + a] = hi[a];
    }
    m.depth = depth;
    m.n_leaves = leaves;
    m.built = true;
    m.lights_built = true;
    b->n_scenes = n;
    return DR_OK;
}

}  // namespace

}  // namespace dr

using namespace dr;

extern "C" {

dr_status dr_batch_create(dr_context *cx, dr_scene *const *scenes, uint32_t n_scenes, dr_batch **out)
{
    if (!cx || !scenes || !out || n_scenes == 0) return fail(DR_E_INVALID, "null argument or no scenes");
    *out = nullptr;
    auto *ctx = reinterpret_cast<Context *>(cx);
    DR_HIP_TRY(hipSetDevice(ctx->device));
    std::vector<Scene *> sc(n_scenes);
    for (uint32_t k = 0; k < n_scenes; ++k) {
        sc[k] = reinterpret_cast<Scene *>(scenes[k]);
        if (!sc[k]) return fail(DR_E_INVALID, "null scene");
        if (sc[k]->ctx != ctx) return fail(DR_E_INVALID, "scene of another context");
        if (!sc[k]->built) return fail(DR_E_INVALID, "scene BVH not built (call dr_scene_build_bvh)");
        if (!sc[k]->lights_built)
            if (dr_status st = build_lights(sc[k], nullptr)) return st;
    }
    auto *b = new (std::nothrow) Batch();
    if (!b) return fail(DR_E_OOM, "host allocation failed");
    b->ctx = ctx;
    if (dr_status st = merge(b, sc.data(), n_scenes)) {
        delete b;
        return st;
    }
    *out = reinterpret_cast<dr_batch *>(b);
    return DR_OK;
}

dr_status dr_batch_destroy(dr_batch *bt)
{
    if (!bt) return DR_OK;
    auto *b = reinterpret_cast<Batch *>(bt);
    (void)hipSetDevice(b->ctx->device);
    (void)hipDeviceSynchronize();
    delete b;
    return DR_OK;
}

dr_status dr_batch_render(dr_batch *bt, const dr_batch_view *views, uint32_t n_views, const dr_render_params *params,
                          const dr_batch_options *options, dr_stream stream)
{
    if (!bt || !params || (n_views && !views)) return fail(DR_E_INVALID, "null argument");
    auto *b = reinterpret_cast<Batch *>(bt);
    DR_HIP_TRY(hipSetDevice(b->ctx->device));
    dr_batch_options opt{kPathCap, DR_SCHED_PER_RAY_SPECULATIVE, 16u, 0u, DR_BATCH_LAUNCH_SEQUENCE, 0u};
    if (options) opt = *options;
    if (opt.chunk_paths < 4096 || opt.chunk_paths > kMaxChunkPaths) return fail(DR_E_INVALID, "chunk_paths out of range");
    if (opt.trace_sched > 15u) return fail(DR_E_INVALID, "unknown trace schedule");
    if (opt.refill_min < 1u || opt.refill_min > 64u) return fail(DR_E_INVALID, "refill_min must be 1..64");
    if (opt.launch > DR_BATCH_LAUNCH_PERSISTENT) return fail(DR_E_INVALID, "unknown launch mode");
    if (opt.blocks_per_cu > 16u) return fail(DR_E_INVALID, "blocks_per_cu must be 0..16");
    std::vector<ViewInfo> vt(n_views);
    uint64_t total = 0;
    for (uint32_t v = 0; v < n_views; ++v) {
        const dr_batch_view &w = views[v];
        if (w.scene >= b->n_scenes) return fail(DR_E_INVALID, "view scene index out of range");
        if (dr_status st = check_render_args(&w.camera, params, w.accum)) return st;
        ViewInfo &i = vt[v];
        i.cam = cam_params(w.camera);
        i.item_off = total;
        i.seed = params->seed ^ ((uint64_t)v << 32);
        i.accum = w.accum;
        i.npix = w.camera.width * w.camera.height;
        i.spp = w.spp;
        i.sample_offset = w.sample_offset;
        i.scene = w.scene;
        total += (uint64_t)i.npix * w.spp;
    }
    if (total == 0) return DR_OK;
    hipStream_t user = reinterpret_cast<hipStream_t>(stream);
    // the view table is rewritten: earlier work on the stream must be done
    DR_HIP_TRY(hipStreamSynchronize(user));
    if (dr_status st = b->views.ensure(sizeof(ViewInfo) * n_views)) return st;
    DR_HIP_TRY(hipMemcpy(b->views.ptr, vt.data(), sizeof(ViewInfo) * n_views, hipMemcpyHostToDevice));
    b->merged.refill_min = opt.refill_min;
    RenderJob j;
    j.s = &b->merged;
    j.scenes = b->ranges.as<SceneRange>();
    j.views = b->views.as<ViewInfo>();
    j.n_views = n_views;
    j.total = total;
    j.prm = params;
    j.cap = (uint32_t)std::min<uint64_t>(total, opt.chunk_paths);
    j.sched = opt.trace_sched;
    j.graph = opt.use_graph != 0;
    j.persistent = opt.launch == DR_BATCH_LAUNCH_PERSISTENT;
    j.blocks_per_cu = opt.blocks_per_cu;
    hipStream_t run = user;
    if (j.graph && !j.persistent && user == nullptr) {
        // the null stream cannot be recorded: run on the batch's own stream
        if (!b->stream) {
            DR_HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
            DR_HIP_TRY(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming));
            DR_HIP_TRY(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming));
        }
        DR_HIP_TRY(hipEventRecord(b->ev_in, user));
        DR_HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_in, 0));
        run = b->stream;
    }
    dr_status st = render_job(j, run);
    if (run != user) {
        DR_HIP_TRY(hipEventRecord(b->ev_out, run));
        DR_HIP_TRY(hipStreamWaitEvent(user, b->ev_out, 0));
    }
    return st;
}

dr_status dr_batch_autotune(dr_batch *bt, const dr_batch_view *views, uint32_t n_views,
                            const dr_render_params *params, uint32_t budget_spp, uint32_t reps,
                            dr_batch_options *best, dr_batch_options *tried, float *ms, uint32_t max_tried,
                            uint32_t *n_tried)
{
    if (!bt || !views || !n_views || !params || !best) return fail(DR_E_INVALID, "null argument or no views");
    if (budget_spp < 1 || reps < 1 || reps > 16) return fail(DR_E_INVALID, "budget_spp >= 1 and reps 1..16");
    if (max_tried && (!tried || !ms)) return fail(DR_E_INVALID, "null candidate output arrays");
    auto *b = reinterpret_cast<Batch *>(bt);
    DR_HIP_TRY(hipSetDevice(b->ctx->device));
    // the trial renders: each view at min(spp, budget_spp) into one scratch
    // accumulation buffer (the caller's images are not touched)
    std::vector<dr_batch_view> tv(views, views + n_views);
    uint64_t npix = 0;
    for (const dr_batch_view &v : tv) npix += (uint64_t)v.camera.width * v.camera.height;
    DevBuf scratch;
    if (dr_status st = scratch.ensure(32ull * std::max<uint64_t>(npix, 1))) return st;
    DR_HIP_TRY(hipMemset(scratch.ptr, 0, 32ull * std::max<uint64_t>(npix, 1)));
    uint64_t off = 0;
    for (dr_batch_view &v : tv) {
        v.spp = std::min(v.spp, budget_spp);
        v.accum = scratch.as<double>() + 4 * off;
        off += (uint64_t)v.camera.width * v.camera.height;
    }
    // candidates: the launch sequence over queue sizes, traversal schedules
    // (per-ray or resident-grid speculative; the refill threshold for the
    // latter) and graph recording; the persistent launch over queue sizes,
    // the resident-grid schedule with one ray counter or one per XCD, refill
    // thresholds and resident blocks per CU (more cost more at its barriers)
    std::vector<dr_batch_options> cand;
    for (uint32_t chunk : {1u << 21, 1u << 22, 1u << 23, 1u << 24})
        for (uint32_t graph : {0u, 1u}) {
            cand.push_back(dr_batch_options{chunk, DR_SCHED_PER_RAY_SPECULATIVE, 16u, graph,
                                            DR_BATCH_LAUNCH_SEQUENCE, 0u});
            for (uint32_t refill : {8u, 16u, 32u})
                cand.push_back(dr_batch_options{chunk, DR_SCHED_PERSISTENT_SPECULATIVE, refill, graph,
                                                DR_BATCH_LAUNCH_SEQUENCE, 0u});
        }
    for (uint32_t chunk : {1u << 22, 1u << 23})
        for (uint32_t sched : {(uint32_t)DR_SCHED_PERSISTENT_SPECULATIVE,
                               (uint32_t)DR_SCHED_PERSISTENT_SPECULATIVE | DR_SCHED_XCD_SLICES})
            for (uint32_t refill : {32u, 64u})
                for (uint32_t bpc : {1u, 2u})
                    cand.push_back(dr_batch_options{chunk, sched, refill, 0u, DR_BATCH_LAUNCH_PERSISTENT, bpc});
    hipEvent_t e0 = nullptr, e1 = nullptr;
    DR_HIP_TRY(hipEventCreate(&e0));
    DR_HIP_TRY(hipEventCreate(&e1));
    float best_ms = INFINITY;
    dr_status st = DR_OK;
    uint32_t k = 0;
    for (const dr_batch_options &c : cand) {
        std::vector<float> t;
        for (uint32_t r = 0; r <= reps && st == DR_OK; ++r) {  // the first run sizes buffers, untimed
            hipError_t e = hipEventRecord(e0, nullptr);
            if (e == hipSuccess && (st = dr_batch_render(bt, tv.data(), n_views, params, &c, nullptr)) != DR_OK) break;
            if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float m = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&m, e0, e1);
            if (e != hipSuccess) st = fail(DR_E_DEVICE, std::string("autotune timing: ") + hipGetErrorString(e));
            if (r > 0) t.push_back(m);
        }
        if (st) break;
        std::sort(t.begin(), t.end());
        const float med = t[t.size() / 2];
        if (k < max_tried) {
            tried[k] = c;
            ms[k] = med;
        }
        ++k;
        if (med < best_ms) {
            best_ms = med;
            *best = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (n_tried) *n_tried = k;
    return st;
}

}  // extern "C"
