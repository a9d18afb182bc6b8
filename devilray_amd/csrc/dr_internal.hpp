// Internal declarations shared by the HIP translation units of libdevilray.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dr_api.h"

namespace dr {

// ---- error plumbing --------------------------------------------------------
void set_error(const std::string &msg);
dr_status fail(dr_status s, const std::string &msg);

#define DR_HIP_TRY(expr)                                                              \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ::dr::fail(e_ == hipErrorOutOfMemory ? DR_E_OOM : DR_E_DEVICE,     \
                              std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

// ---- device buffer ---------------------------------------------------------
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    dr_status ensure(size_t n);  // grow-only, contents not preserved
    void release();
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// ---- wide BVH node (128 B): SoA child boxes, child references --------------
// child ref >= 0: index of an interior wide node
// child ref <  0: leaf, bit 31 set, bits 28..30 = count-1, bits 0..27 = first
// (first + count - 1 <= kMaxLeafRefs keeps every leaf distinct from DR_EMPTY_REF)
// empty slot: DR_EMPTY_REF with a box at +inf in all axes (never hit)
struct alignas(16) WideNode {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t child[4];
    int32_t pad[4];
};
static_assert(sizeof(WideNode) == 128, "wide node is 128 B");

// The traversal's copy of a wide node: the child boxes quantised to one byte
// per plane on a grid per axis (node_quant.hpp). Axis a's planes are
// origin[a] + q * step[a], q = 0..255; every plane is an exact float, and a
// child's quantised box contains its WideNode box. A child slot with an empty
// box (or none) has q = 255 below and 0 above.
struct alignas(64) QNode {
    float origin[3];
    float step[3];
    uint32_t qlo[3];   // per axis x, y, z: child c's lower plane index in byte c
    uint32_t qhi[3];   // per axis: child c's upper plane index in byte c
    int32_t child[4];
};
static_assert(sizeof(QNode) == 64, "quantised node is 64 B");

constexpr int32_t DR_EMPTY_REF = -1;  // 0xFFFFFFFF
constexpr int kLeafMaxCap = 8;
constexpr uint32_t kMaxLeafRefs = (1u << 28) - 2;  // triangle references per scene or batch
// wide BVH levels a build may produce: traversal keeps up to 3 entries per
// level on its stack, and a deeper tree is refused at build time
constexpr uint32_t kMaxWideDepth = 256;

__host__ __device__ inline int32_t make_leaf_ref(uint32_t first, uint32_t count)
{
    return (int32_t)(0x80000000u | ((count - 1u) << 28) | first);
}
__host__ __device__ inline uint32_t leaf_first(int32_t ref) { return (uint32_t)ref & 0x0FFFFFFFu; }
__host__ __device__ inline uint32_t leaf_count(int32_t ref) { return (((uint32_t)ref >> 28) & 0x7u) + 1u; }

// One instance of a two-level forest: the root of its scene's tree and the
// inverse of its scale-and-translate transform. A ray enters model space as
// ((o - t) * inv_s, d * inv_s), so its t stays world t; hits report
// id_delta + the scene-local triangle id.
struct InstRec {
    int32_t root;
    float isx, isy, isz;
    float tx, ty, tz;
    uint32_t id_delta;
};
static_assert(sizeof(InstRec) == 32, "instance record is 32 B");

// Shading view of one instance: its transform forward (world = model * s +
// t, rounded per operation as the flattened copy is) and where its scene's
// triangles start in the forest's concatenated scene arrays; id_delta as in
// InstRec.
struct InstShade {
    float sx, sy, sz;
    float tx, ty, tz;
    uint32_t tri_off;
    uint32_t id_delta;
};
static_assert(sizeof(InstShade) == 32, "instance shading record is 32 B");

struct Context {
    int device = 0;
    int cu_count = 0;
    hipStream_t stream = nullptr;  // owned helper stream (unused by the async API)
    DevBuf scratch_sort[6];         // radix-sort / scan temporaries
    DevBuf display_thr;             // 2 x 255 FP64 decision thresholds of the 8-bit display encodings
    // this device's copy of the zero-allocation traversal's static scratch
    // array (its address from hipGetSymbolAddress at context creation)
    int32_t *noalloc_scratch = nullptr;
};

// A lane of the wavefront render: a share of each chunk's paths with its own
// queues, shadow rays and traversal spill region, on its own stream, so the
// lanes' launches run side by side and one lane's launch tails are filled by
// another's work. Lane 0 uses the scene's rs_* queues and the caller's
// stream; lanes 1.. own these buffers, a non-blocking stream of default
// priority and an event that ends their part of a chunk.
struct RenderLane {
    DevBuf list_a, list_b, counters, srays, socc, scon, sroot, ovf;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
};
constexpr uint32_t kMaxRenderLanes = 4;

struct Scene {
    Context *ctx = nullptr;
    uint32_t n_tris = 0, n_verts = 0, n_mats = 0;
    // input data on device (input order)
    DevBuf verts;     // float3 * n_verts
    DevBuf indices;   // uint3 * n_tris
    DevBuf tri_mat;   // uint32 * n_tris
    DevBuf materials; // dr_material * n_mats
    std::vector<dr_material> host_materials;
    // acceleration structure
    DevBuf tri_v;     // 3 float4 per triangle reference, leaf order; w of the first = id bits
    DevBuf nodes;     // QNode: what the traversal reads
    DevBuf bnodes;    // WideNode: the same nodes with float boxes (build and refit)
    uint32_t n_refs = 0;  // tri_v entries: triangles plus split pieces
    // two-level forest (instance.hip): nodes [0, n_top) are a tree over the
    // instances whose leaf references name an InstRec, then the instanced
    // scenes' trees; 0 for an ordinary scene
    uint32_t n_inst = 0, n_top = 0;
    DevBuf inst_recs;  // InstRec * n_inst
    // a forest's shading data: verts / indices / tri_mat / materials hold the
    // instanced scenes' own arrays concatenated (n_cat_tris triangles, indices
    // and material ids rebased), inst_shade one InstShade per instance, and
    // the light table names emitters by index: emit_tri holds their instanced
    // triangle ids (ascending), emit_pdf their selection pdf / world area
    uint32_t n_cat_tris = 0;
    DevBuf inst_shade;
    DevBuf emit_pdf;
    std::vector<uint32_t> cat_tri_off;  // host: scene k's triangles are [off[k], off[k + 1])
    std::vector<uint32_t> inst_scene;   // host: the scene of each instance
    uint32_t n_nodes = 0, n_leaves = 0, depth = 0;
    // wide nodes by depth: level d is [level_ofs[d], level_ofs[d + 1]) (refit)
    std::vector<uint32_t> level_ofs;
    DevBuf refit_tmp;  // refit: ordered-int vertex bounds
    double build_ms = 0.0, sah_cost = 0.0;
    bool built = false;
    // traversal overflow stack + flag
    DevBuf ovf_stack;
    DevBuf flags;     // uint32[512]: [0] overflow flag, [16..271] per-launch ray counters
    uint64_t launch_seq = 0;
    uint32_t sched_closest = DR_SCHED_PERSISTENT_SPECULATIVE;  // traversal schedules
    uint32_t sched_any = DR_SCHED_PERSISTENT_SPECULATIVE;
    uint32_t refill_min = 16;  // persistent waves refill when this many lanes are idle
    // Persistent grid: resident blocks per CU (0 = occupancy limit). On
    // incoherent rays over the split 1.3M-triangle icosphere (closest hit,
    // 68 registers, 7 blocks fit): 1.377, 1.344 and 1.347 ms for 5, 6 and 7.
    uint32_t blocks_per_cu = 6;
    uint32_t sort_key = 0;       // ray reorder key: 0 origin, 1 direction + origin
    float bounds[6] = {0, 0, 0, 0, 0, 0};  // scene box (lo, hi), set by the BVH build
    DevBuf ray_keys, ray_perm;             // ray reordering scratch
    // light sampler (light.hip)
    bool lights_built = false;
    uint32_t n_emit = 0;
    uint64_t light_q_total = 0;
    DevBuf tri_area;        // double per triangle
    DevBuf light_pdf_area;  // float per triangle: selection pdf / area (0: not an emitter)
    DevBuf emit_tri;        // u32 per emitter
    DevBuf alias;           // 20-byte alias entries, one per emitter
    DevBuf light_scalars;
    DevBuf debug_colors;    // float4 per material for the UV-checker view (default: the albedo)
    bool debug_colors_set = false;
    DevBuf dbg_rays, dbg_hits;  // debug views: one sample of rays and hits
    // wavefront render state (render.hip)
    DevBuf rs_rays, rs_hits, rs_beta, rs_radiance, rs_list_a, rs_list_b, rs_counters;
    DevBuf rs_shadow_rays, rs_shadow_occ, rs_shadow_contrib;
    DevBuf rs_view, rs_root, rs_sroot, rs_chunk, rs_views;  // per-path view / BVH root, chunk cursor, view table
    DevBuf rs_guide_head;   // DR_GUIDE_HEAD_FLOATS floats per path (guided renders)
    DevBuf rs_grid;         // persistent render: grid barrier state and trace ray counters
    DevBuf rs_pending;      // u32 per path: training sample awaiting its radiance snapshot
    // u64: extension rays (camera and bounce rays, not shadow rays) traced by
    // the renders of this scene since it was created or the count was reset
    DevBuf ray_casts;
    uint32_t rs_capacity = 0;
    // lanes of the launch sequence (RenderLane): how many a render uses and
    // the event that starts them
    uint32_t render_lanes = 4;
    RenderLane lanes[kMaxRenderLanes];
    hipEvent_t lane_fork = nullptr;
    void release_lanes();
    // The recorded bounce loop of the last graph render and an event after its
    // last replay: an executable graph is freed only once its launches are
    // done (at the next graph render, or with the scene).
    hipGraphExec_t rs_exec = nullptr;
    hipEvent_t rs_exec_done = nullptr;
    void release_graph();
    Scene() = default;
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;
    ~Scene()
    {
        release_graph();
        release_lanes();
    }
};

// ---- learned path guide (guide.hip) ----------------------------------------
constexpr uint32_t kGuideNoSample = 0xFFFFFFFFu;

// Smallest b with 2^b >= v (0 for v <= 1).
__host__ __device__ inline uint32_t ceil_log2_u64(uint64_t v)
{
    uint32_t b = 0;
    while (b < 64 && (1ull << b) < v) ++b;
    return b;
}

// Order key of a staged training sample: a bijection of the path item on
// [0, 2^item_bits) (odd multipliers and xor-shifts, each invertible modulo
// 2^item_bits), then the vertex in the low vertex_bits. Keys are unique, and
// the leading samples in key order are spread over the image rather than
// taken from its first rows.
__host__ __device__ inline uint64_t guide_sample_key(uint64_t item, uint32_t vertex, uint32_t item_bits,
                                                     uint32_t vertex_bits)
{
    const uint64_t mask = item_bits >= 64 ? ~0ull : (1ull << item_bits) - 1ull;
    const uint32_t sh = (item_bits + 1u) / 2u;
    uint64_t x = (item * 0x9E3779B97F4A7C15ull) & mask;
    x ^= x >> sh;
    x = (x * 0xBF58476D1CE4E5B9ull) & mask;
    x ^= x >> sh;
    return (x << vertex_bits) | vertex;
}
constexpr int kGuideTrainBlocks = 256;  // partial-gradient slots (deterministic reduction)

// Per training sample, what the render needs to complete it: the path slot
// (kGuideNoSample once finalised), the pdf the direction was drawn with, the
// throughput after the bounce and the path radiance before the next vertex.
struct alignas(16) GuideAux {
    uint32_t path;
    float pdf;
    float beta[3];
    float lsnap[3];
};

struct Guide {
    Context *ctx = nullptr;
    dr_guide_params prm{};
    DevBuf master;   // FP32 parameters
    DevBuf m1, m2;   // Adam moments
    DevBuf frag_f;   // BF16 weight fragments for the forward products
    DevBuf frag_b;   // BF16 transposed fragments for the backward products
    DevBuf partials; // kGuideTrainBlocks rows of N_PARAMS + 1 FP32, padded to 64
    DevBuf samples;  // staged samples of the last collecting render: dr_guide_sample x stage_cap
    DevBuf aux;      // GuideAux x stage_cap
    DevBuf ring;     // training ring: the latest `capacity` samples of non-zero weight
    DevBuf flags;    // u32 x (stage_cap + 1): compaction ranks
    DevBuf keys;     // u32 x 2 x stage_cap: each staged sample's order key, low words then high words
    DevBuf perm;     // u32 x stage_cap: stage slots in key order (valid after a collecting render)
    DevBuf counters; // u32[4]: [0] samples offered to the stage, [1] ring head, [2] ring fill
    uint32_t stage_cap = 0;
    uint32_t step = 0;
    float collect_fraction = 1.0f;
    uint32_t guided_vertices = 64;  // leading path vertices that sample from the guide
    uint32_t train_batch = 0;       // ring samples per step (0: the whole ring), a rotating window
    uint32_t window = 0;
    uint32_t sel_variance = 0;      // selection logit trained on the second moment instead of the KL
};

// Network inference for the paths of a wavefront queue: for each listed path
// whose hit is a Lambert surface, head[DR_GUIDE_HEAD_FLOATS * path] receives
// the decoded mixture (other paths are left untouched).
dr_status guide_infer_paths(Guide *g, const Scene *s, const dr_ray *rays, const dr_hit *hits, const uint32_t *list,
                            const uint32_t *count, uint32_t n_max, float *head, hipStream_t stream);
// Complete the training samples of a finished chunk from the path radiance.
dr_status guide_finalize(Guide *g, const float4 *radiance, hipStream_t stream);
// Append the staged samples of non-zero weight to the training ring in the
// order of their keys (guide_sample_key, unique per sample), so the ring does
// not depend on the order in which the stage's slots were taken. key_bits:
// item_bits + vertex_bits of the render's keys.
dr_status guide_compact(Guide *g, uint32_t key_bits, hipStream_t stream);
dr_status guide_reset_counters(Guide *g, hipStream_t stream);

// ---- rendering (render.hip) ------------------------------------------------
struct CamParams {
    float m[12];
    float f, cx, cy, psx, psy;
    uint32_t w, h, sampling;
};

// One view of a render call: its items are item_off + sample * npix + pixel
// for sample in [0, spp); its samples carry indices sample_offset + sample.
struct ViewInfo {
    CamParams cam;
    uint64_t item_off;
    uint64_t seed;
    double *accum;
    uint32_t npix, spp, sample_offset, scene;
};

// Per scene of a batch: BVH root and light-table range.
struct SceneRange {
    uint32_t root, alias_off, n_emit, pad;
};

// default paths per chunk: a larger wavefront hides the traversal's latency
// and its launch tails better (many-lights at 1920x1080 x 8 spp, one lane:
// 27.8 ms at 2^22 paths per chunk, 23.4 at 2^23, 21.3 at 2^24); about 150 B
// of path state and queues per path (2.5 GB at 2^24)
constexpr uint32_t kPathCap = 1u << 24;
constexpr uint32_t kMaxChunkPaths = 1u << 24;

// A render call: views (device table) over one scene, or over a batch's
// merged scene with a per-scene range table.
struct RenderJob {
    Scene *s = nullptr;
    const SceneRange *scenes = nullptr;
    const ViewInfo *views = nullptr;
    uint32_t n_views = 0;
    uint64_t total = 0;  // items over all views
    const dr_render_params *prm = nullptr;
    Guide *guide = nullptr;
    uint32_t guide_flags = 0;
    uint32_t cap = kPathCap;  // paths per chunk
    uint32_t sched = DR_SCHED_PER_RAY_SPECULATIVE;
    bool graph = false;  // record the bounce loop once, replay it per chunk
    // run every chunk and bounce in one persistent launch (plain renders),
    // with this many resident blocks per CU (0: the occupancy limit)
    bool persistent = false;
    uint32_t blocks_per_cu = 0;
};

CamParams cam_params(const dr_camera &c);
dr_status check_render_args(const dr_camera *cam, const dr_render_params *prm, const double *accum);
dr_status render_job(const RenderJob &j, hipStream_t stream);
// allocates and zeroes the scene's ray-cast counter on first use, then adds
// n (a count known on the host), or sets the count to n when `set`
dr_status count_ray_casts(Scene *s, uint64_t n, hipStream_t stream, bool set = false);

// light.hip
dr_status build_lights(Scene *s, hipStream_t stream);
// the light table of an instance forest: every emissive triangle of every
// instance, weighted by its world area
dr_status build_instance_lights(Scene *forest, hipStream_t stream);

// render.hip
dr_status render(Scene *s, const dr_camera *cam, const dr_render_params *prm, Guide *guide, uint32_t guide_flags,
                 double *accum, hipStream_t stream);
// render of an instance forest (plain sampler, the launch sequence)
dr_status render_instances(Scene *forest, const dr_camera *cam, const dr_render_params *prm, double *accum,
                           hipStream_t stream);

// batch.hip: copy n wide nodes, moving interior references by node_off and
// leaf references by ref_off
dr_status rebase_nodes(const QNode *src, uint32_t n, uint32_t node_off, uint32_t ref_off, QNode *dst,
                       hipStream_t stream);

// bvh_build.hip
dr_status build_bvh(Scene *s, const dr_bvh_params *p, hipStream_t stream);
// new boxes for the built tree over the scene's current vertices
dr_status refit_bvh(Scene *s, hipStream_t stream);

// trace.hip
// ray_list (optional): indices of the rays to trace (results go to those
// indices); n_dev (optional): device-resident count, at most n_max.
// ray_root (optional): per ray, the BVH node its traversal starts at (a
// batch's merged forest); NULL starts every ray at node 0.
// ovf (optional): the spill region to use instead of the scene's (a render
// lane's own, so lanes on several streams never share one).
dr_status launch_trace_ex(Scene *s, const dr_ray *rays, const uint32_t *ray_list, uint32_t n_max,
                          const uint32_t *n_dev, dr_hit *hits, uint32_t *occluded, uint32_t *box_tests,
                          uint32_t *tri_tests, bool any_hit, uint32_t sched, hipStream_t stream,
                          const uint32_t *ray_root = nullptr, DevBuf *ovf = nullptr);
dr_status launch_trace(Scene *s, const dr_ray *rays, uint32_t n, dr_hit *hits, uint32_t *occluded,
                       uint32_t *box_tests, uint32_t *tri_tests, bool any_hit, hipStream_t stream);
// an any-hit queue (rays_a, device count na_dev) and a closest-hit queue
// (rays_c through list_c, device count nc_dev) traced in one launch, each
// ray exactly as launch_trace_ex would; both counts at most n_max. Only for
// the per-ray schedules (trace_pair_fits); instances run the speculative one.
bool trace_pair_fits(const Scene *s, uint32_t sched);
dr_status launch_trace_pair(Scene *s, const dr_ray *rays_a, const uint32_t *na_dev, uint32_t *occluded,
                            const uint32_t *roots_a, const dr_ray *rays_c, const uint32_t *list_c,
                            const uint32_t *nc_dev, dr_hit *hits, const uint32_t *roots_c, uint32_t n_max,
                            uint32_t sched, hipStream_t stream, DevBuf *ovf = nullptr);
// closest hit with every scratch buffer in the static device array (no
// allocation); noalloc_init looks up this device's copy of it
dr_status noalloc_init(Context *ctx);
dr_status launch_trace_noalloc(Scene *s, const dr_ray *rays, uint32_t n, dr_hit *hits, hipStream_t stream);

// sort.hip
dr_status radix_sort_pairs(Context *ctx, uint32_t *keys, uint32_t *vals, uint32_t n, int key_bits,
                           hipStream_t stream);
dr_status exclusive_scan_u32(Context *ctx, uint32_t *data, uint32_t n, hipStream_t stream);
dr_status exclusive_scan_u64(Context *ctx, uint64_t *data, uint32_t n, hipStream_t stream);

}  // namespace dr
