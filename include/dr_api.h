/*
 * DevilRay MI355X render core -- C ABI.
 *
 * Plain C types and pointers only (no HIP or torch types): a caller binds it
 * with ctypes, cgo, JNI or a C++ facade. hipStream_t is passed as an opaque
 * pointer (`dr_stream`, NULL = the default stream).
 *
 * Conventions
 *   - Every entry point returns a dr_status; dr_last_error() gives the message
 *     of the last failure on the calling thread.
 *   - Scene inputs are host pointers and are copied. Ray, hit, occlusion and
 *     image buffers are device pointers (allocate them with dr_device_alloc or
 *     any HIP allocator on the context's device).
 *   - All trace/render entry points are asynchronous on the given stream;
 *     *_create / *_destroy / dr_scene_build_bvh synchronise.
 *   - A context or scene is externally synchronised (not thread-safe).
 *   - Triangle id = index into the caller's input triangle array.
 *
 * Reference counterparts (Frontier789/DevilRay) are cited per entry point.
 */
#ifndef DEVILRAY_DR_API_H
#define DEVILRAY_DR_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    DR_OK = 0,
    DR_E_INVALID = 1,     /* bad argument */
    DR_E_OOM = 2,         /* device allocation failed */
    DR_E_DEVICE = 3,      /* HIP runtime error (no device, launch failure) */
    DR_E_UNSUPPORTED = 4, /* valid request the library does not implement */
    DR_E_OVERFLOW = 5     /* a bounded structure (traversal stack) overflowed */
} dr_status;

typedef struct dr_context dr_context;
typedef struct dr_scene dr_scene;
typedef void *dr_stream;

/* 32-byte ray: origin, tmin, direction, tmax. A hit is reported for
 * tmin <= t <= tmax; tmin must be >= 0. The direction need not be unit. */
typedef struct { float ox, oy, oz, tmin, dx, dy, dz, tmax; } dr_ray;

/* 16-byte hit: triangle id (-1 on miss), distance t (+inf on miss; a hit at
 * distance zero reports +0), barycentric weights of the triangle's 2nd (u)
 * and 3rd (v) vertex. Ties on t resolve to the lower id. */
typedef struct { int32_t id; float t, u, v; } dr_hit;

/* Materials: the reference's Material variant (include/tracing/Material.hpp:9-25):
 * DiffuseMaterial {emission, diffuse_reflectance} and TransparentMaterial {ior}. */
enum { DR_MAT_LAMBERT = 0, DR_MAT_DIELECTRIC = 1 };
typedef struct {
    int32_t kind;
    float albedo[3];
    float emission[3];
    float ior;
} dr_material;

/* Scene description (host memory, copied). The reference builds one
 * TriangleMesh per object (include/tracing/TriangleMesh.hpp:9-27) with one
 * material each; here geometry is flattened to world space with a material
 * index per triangle (tri_material == NULL means material 0 everywhere). */
typedef struct {
    const float *vertices;        /* 3 floats per vertex */
    uint32_t n_vertices;
    const uint32_t *indices;      /* 3 vertex indices per triangle */
    uint32_t n_triangles;
    const uint32_t *tri_material; /* n_triangles entries or NULL */
    const dr_material *materials;
    uint32_t n_materials;
} dr_scene_desc;

/* BVH build parameters: max triangles per leaf (1..8) and DR_BVH_* flags.
 * The binary tree is built by PLOC clustering over the Morton order (default)
 * or as a Karras radix tree over the Morton codes (DR_BVH_LBVH, faster to
 * build, lower quality); both then go through the same SAH leaf collapse and
 * 4-wide collapse. Bits 8-15 set the PLOC search radius (0: 16, max 64).
 * Triangle splitting: a triangle whose box half area is `ratio` times
 * A (|nx| + |ny| + |nz|) (2 to 3 for well-shaped triangles, far more for long
 * thin ones at an angle to the axes) is cut into floor(ratio / r0) pieces,
 * 1 to 16, each bounded by the box of its clipped part. Bits 24-31 set r0 in
 * quarters (0: default 8, r0 = 2); bits 16-23 cap the extra references in
 * quarters of the triangle count (0: default 32, i.e. 8 per triangle; never
 * more than 2^28 - 2 references in all); DR_BVH_NO_SPLIT turns it off. Hits
 * do not depend on any of these choices. */
typedef struct {
    uint32_t leaf_max;
    uint32_t flags;
} dr_bvh_params;
#define DR_BVH_LBVH 1u
#define DR_BVH_NO_SPLIT 2u
#define DR_BVH_RADIUS_MASK 0xFF00u
#define DR_BVH_SPLIT_SHIFT 16
#define DR_BVH_SPLIT_MASK 0xFF0000u
#define DR_BVH_SPLIT_RATIO_SHIFT 24
#define DR_BVH_SPLIT_RATIO_MASK 0xFF000000u

typedef struct {
    uint32_t n_triangles;
    uint32_t n_nodes;        /* wide (4-ary) nodes */
    uint32_t n_leaves;
    uint32_t depth;          /* wide levels (at most 256: a deeper build fails with DR_E_OVERFLOW) */
    uint32_t overflow_flag;  /* non-zero if a traversal overflowed its stack; traversal sizes its
                              * stack to 3 x depth entries, which bounds every walk, so this
                              * stays 0 (a guard that is checked by the tests) */
    /* last counted traversal (dr_trace_closest_counted): wave-level node-loop
     * iterations, node visits, wave-level triangle iterations, triangle tests;
     * visits / (64 * iterations) is the SIMD utilisation of each loop */
    uint32_t simd_stats[4];
    double build_ms;         /* device time of the last build */
    double sah_cost;         /* SAH cost of the binary tree before collapse */
    uint32_t n_references;   /* leaf entries: triangles plus split pieces */
    uint32_t stack_high_water; /* last counted traversal: deepest stack of any ray (<= 3 x depth) */
} dr_scene_info;

const char *dr_status_string(dr_status s);
const char *dr_last_error(void);

/* Device / context. Replaces the reference's device runtime
 * (src/device/DeviceMemory.cu:7-69, src/device/DevUtils.cu:12-74). */
dr_status dr_device_count(int32_t *count);
dr_status dr_context_create(int32_t device, dr_context **out);
dr_status dr_context_destroy(dr_context *ctx);
dr_status dr_device_alloc(dr_context *ctx, size_t bytes, void **out);
dr_status dr_device_free(dr_context *ctx, void *ptr);
dr_status dr_memcpy_htod(dr_context *ctx, void *dst, const void *src, size_t bytes);
dr_status dr_memcpy_dtoh(dr_context *ctx, void *dst, const void *src, size_t bytes);
/* Device-to-device copy, asynchronous on `stream` (also the known-size
 * kernel against which memory-traffic counters are calibrated). */
dr_status dr_memcpy_dtod(dr_context *ctx, void *dst, const void *src, size_t bytes, dr_stream stream);
dr_status dr_synchronize(dr_context *ctx);
/* Zero bytes of device memory with a kernel (not a memset node, so callers
 * may record it into a graph), asynchronous on `stream`; dst and bytes must
 * be multiples of 16. Replaces the device side of Buffers::reset
 * (src/Renderer.cpp:17-21), which clears the image between frames. */
dr_status dr_fill_zero(dr_context *ctx, void *dst, size_t bytes, dr_stream stream);

/* Scene upload. Replaces convertMeshToTris / viewGpuTris
 * (src/GpuTris.cpp:5-85) and Scene (include/tracing/Scene.hpp:15-26). */
dr_status dr_scene_create(dr_context *ctx, const dr_scene_desc *desc, dr_scene **out);
dr_status dr_scene_destroy(dr_scene *scene);
dr_status dr_scene_get_info(dr_scene *scene, dr_scene_info *info);

/* On-device BVH build: Morton-code LBVH + SAH leaf collapse + 4-wide nodes.
 * Replaces generateSimpleBBH (src/models/BBH.cpp:101-115), which is a host
 * median split with one triangle per leaf. */
dr_status dr_scene_build_bvh(dr_scene *scene, const dr_bvh_params *params, dr_stream stream);
/* Moves the scene's vertices (host array, the same n_vertices x 3 floats and
 * the same triangles) and refits the built BVH to them: the tree and its
 * triangle pieces are kept and every box is recomputed bottom-up, level by
 * level, so hits stay exact while the tree's quality follows how far the
 * geometry moved. The light table is rebuilt at the next render. The
 * reference moves objects by their transform (TriangleMesh::setPosition /
 * setScale, src/tracing/TriangleMesh.cpp:3-15) over per-object trees that it
 * never rebuilds; the flattened scene here refits instead. Batches made from
 * the scene earlier keep their own copy. Synchronises. */
dr_status dr_scene_refit(dr_scene *scene, const float *vertices, dr_stream stream);

/* Closest hit over a ray batch. Replaces getIntersection / cast
 * (include/tracing/IntersectionTestsImpl.hpp:222-251) for n rays at once. */
dr_status dr_trace_closest(dr_scene *scene, const dr_ray *rays, uint32_t n, dr_hit *hits,
                           dr_stream stream);

/* dr_trace_closest without any allocation. No counterpart in the reference:
 * for callers that cannot call hipMalloc (HIP graph capture, tight serving
 * loops, harnesses that forbid allocation). Same hits as dr_trace_closest,
 * bit for bit. Its scratch (ray counter and traversal stack spill) is one
 * static device array of 48 MiB + 256 B that every GPU running the library
 * holds until the process ends; it covers the resident grid (at most 1024
 * blocks of 256 lanes) over trees of at most 21 wide levels, and a deeper
 * tree returns DR_E_OVERFLOW (dr_scene_info.depth tells). Calls on one GPU
 * share that array: calls that overlap in time on one GPU (different
 * streams, threads or scenes) race and must be serialised by the caller. */
dr_status dr_trace_closest_noalloc(dr_scene *scene, const dr_ray *rays, uint32_t n, dr_hit *hits,
                                   dr_stream stream);

/* Any hit (occlusion): occluded[i] = 1 if some triangle has tmin <= t <= tmax.
 * Replaces visible() (include/tracing/PathGeneration.hpp:213-224), which the
 * reference answers with a closest-hit cast. */
dr_status dr_trace_any(dr_scene *scene, const dr_ray *rays, uint32_t n, uint32_t *occluded,
                       dr_stream stream);

/* Traversal schedules (how lanes take rays and order interior/leaf work):
 * PER_RAY: one lane per ray for the launch; SPECULATIVE: leaves are parked
 * until every lane of the wave holds one (Aila & Laine 2009); PERSISTENT:
 * resident waves refill finished lanes from a global counter. The default,
 * chosen by measurement, is PERSISTENT_SPECULATIVE without REORDER (over split
 * triangle references the reorder's sort costs more than it saves);
 * dr_scene_set_trace_schedule overrides it. */
enum {
    DR_SCHED_PER_RAY = 0,
    DR_SCHED_PER_RAY_SPECULATIVE = 1,
    DR_SCHED_PERSISTENT = 2,
    DR_SCHED_PERSISTENT_SPECULATIVE = 3,
    /* flag: sort the batch by (dominant axis, origin Morton code) first */
    DR_SCHED_REORDER = 4,
    /* flag (persistent schedules): split the ray range into 8 contiguous
     * slices with one counter each, one per XCD (its own L2); waves move on
     * to the next slice when theirs is drained. Measured on the traversal
     * benchmark: within noise for incoherent rays, about 10% slower for
     * coherent camera rays, so not in the default. */
    DR_SCHED_XCD_SLICES = 8
};
dr_status dr_scene_set_trace_schedule(dr_scene *scene, uint32_t closest_sched, uint32_t any_sched);
/* Tuning knobs (chosen by the autotuner): persistent waves refill their
 * finished lanes once at least refill_min (1..64, default 16) are idle;
 * blocks_per_cu caps the resident grid (default 6; 0 = the occupancy limit);
 * sort_key selects the reorder
 * key (0: dominant axis + origin Morton code, 1: dominant axis + quantised
 * direction + origin). */
dr_status dr_scene_set_trace_tuning(dr_scene *scene, uint32_t refill_min, uint32_t blocks_per_cu, uint32_t sort_key);
/* Render lanes (1..4, default 4): dr_render and dr_batch_render deal each
 * chunk's paths to this many lanes, each running its bounce loop on its own
 * stream (the caller's and streams the scene owns), joined before the
 * resolve, so one lane's launch tails are filled by another's work. Images
 * do not depend on it. A recorded bounce loop (graph replay) uses one lane,
 * and so does a chunk of fewer than 2^17 paths. No counterpart in the
 * reference (its megakernel renders a frame in one launch,
 * src/device/RendererImpl.hpp:49-102). */
dr_status dr_scene_set_render_lanes(dr_scene *scene, uint32_t lanes);

/* Closest hit with the reference's HitTests counters
 * (include/tracing/Benchmark.hpp:5-14, IntersectionTestsImpl.hpp:228-231):
 * per ray, the number of AABB tests and triangle tests. */
dr_status dr_trace_closest_counted(dr_scene *scene, const dr_ray *rays, uint32_t n, dr_hit *hits,
                                   uint32_t *box_tests, uint32_t *tri_tests, dr_stream stream);

/* ---- rendering -------------------------------------------------------- */

/* Pinhole camera of the reference (include/tracing/Camera.hpp:6-19,
 * include/tracing/CameraRay.hpp:6-34): the camera-space direction of pixel
 * (x, y) with sub-pixel offset s is normalize(((x+s.x)*pixel_size.x -
 * center.x) / f, ((y+s.y)*pixel_size.y - center.y) / f, 1), rotated by the
 * 3x3 part of `transform` (row-major rows 0..2 of the reference's Matrix4x4);
 * the origin is its translation column. Pixel y grows with row index. */
enum { DR_PIXEL_CENTER = 0, DR_PIXEL_UNIFORM_RANDOM = 1 };
typedef struct {
    float transform[12];
    float focal_length;
    float center[2];
    float pixel_size[2];
    uint32_t width, height;
    uint32_t pixel_sampling;
} dr_camera;

/* Fresnel model of dielectrics: exact unpolarised dielectric Fresnel, or
 * Schlick's approximation as the reference uses (PathGeneration.hpp:75-86). */
enum { DR_FRESNEL_EXACT = 0, DR_FRESNEL_SCHLICK = 1 };

/* spp samples per pixel are added to the accumulation buffer; sample indices
 * start at sample_offset (progressive rendering: the counter-based RNG makes
 * the result independent of how samples are split over calls). max_depth
 * counts path vertices (the reference uses 10, no Russian roulette);
 * Russian roulette applies from vertex rr_depth on (rr_depth >= max_depth
 * disables it). */
typedef struct {
    uint32_t spp;
    uint32_t sample_offset;
    uint32_t max_depth;
    uint32_t rr_depth;
    uint64_t seed;
    uint32_t fresnel;
    uint32_t flags;
} dr_render_params;

/* Build the emissive-triangle alias table (weights area x luminance(E),
 * exact integer sweep construction). dr_render builds it on first use. */
dr_status dr_scene_build_lights(dr_scene *scene, dr_stream stream);
/* Host copies of the light table: n_emit entries of 20 B
 * {float p_A, pdf_A, pdf_B; int32 A, B} (the reference's AliasEntry,
 * include/tracing/DistributionSamplers.hpp:44-51, with triangle ids as items)
 * and per-triangle area pdf (selection probability / area). Either pointer
 * may be NULL. */
dr_status dr_scene_light_count(dr_scene *scene, uint32_t *n_emit);
dr_status dr_scene_copy_lights(dr_scene *scene, void *entries, float *pdf_area);

/* Path tracing with NEE + MIS (power heuristic) into an FP64 accumulation
 * buffer accum[4 * pixel] = {sum r, sum g, sum b, sample count} (device
 * memory, width*height*32 bytes, caller-owned). Replaces Renderer::render's
 * kernel (src/device/RendererImpl.hpp:22-102, PathGeneration.hpp:363-515):
 * wavefront kernels (camera, extend, shade, shadow, resolve) on one stream. */
dr_status dr_render(dr_scene *scene, const dr_camera *camera, const dr_render_params *params, double *accum,
                    dr_stream stream);
/* Extension rays (camera and bounce rays; shadow rays are not counted, as in
 * the reference's SampleStats::ray_casts, PathGeneration.hpp:88-111) traced
 * by dr_render, dr_render_guided and dr_render_debug on this scene since it
 * was created or last reset. Replaces the per-pixel casts buffer behind
 * Renderer::getTotalCasts (src/Renderer.cpp:26-33, include/Renderer.hpp:51).
 * Synchronises `stream` before reading; reset != 0 then zeroes the count,
 * ordered on `stream` before later work. */
dr_status dr_scene_ray_casts(dr_scene *scene, uint64_t *casts, int32_t reset, dr_stream stream);

/* ---- display and debug views --------------------------------------------- */

/* 8-bit display pixels from an FP64 accumulation image (width x height,
 * device memory): rgb x (1 / w) (black where w = 0), clamped to [0, 1], encoded,
 * rounded to 8 bits and packed R | G << 8 | B << 16 | 255 << 24 into rgba
 * (device memory, width*height u32) with the rows flipped. Replaces
 * Renderer::createPixels (src/Renderer.cpp:51-90); the rounding is exact
 * (decision thresholds computed once in FP64), so every device and the CPU
 * oracle give the same bytes. Asynchronous on `stream`. */
enum { DR_OUTPUT_LINEAR = 0, DR_OUTPUT_SRGB = 1 };
dr_status dr_resolve_rgba8(dr_context *ctx, const double *accum, uint32_t width, uint32_t height, uint32_t output,
                           uint32_t *rgba, dr_stream stream);

/* Debug views of the camera ray's first hit (DebugOptions,
 * include/DebugOptions.hpp:3-8; sampleColorDebug and the debug branch of
 * sampleColor, include/tracing/PathGeneration.hpp:274-299,377-429):
 * UV_CHECKER adds 0.5 x the material's debug colour (the reference's uv is
 * always (0, 0)) and skips misses entirely; BARYCENTRIC adds (1-u-v, u, v);
 * WINDING adds (1, 0.73, 0.47) where ((A-B) x (A-C)) . d > 0, else
 * (0.53, 0.82, 1), and black for a miss. params: spp, sample_offset, seed and
 * the camera's pixel sampling, as dr_render; accumulation as dr_render. */
enum { DR_DEBUG_UV_CHECKER = 1, DR_DEBUG_BARYCENTRIC = 2, DR_DEBUG_WINDING = 3 };
dr_status dr_render_debug(dr_scene *scene, const dr_camera *camera, const dr_render_params *params, uint32_t mode,
                          double *accum, dr_stream stream);
/* The materials' debug colours (Material::debug_color,
 * include/tracing/Material.hpp:9-12), 4 floats per material (rgba); the
 * default is each material's albedo. */
dr_status dr_scene_set_debug_colors(dr_scene *scene, const float *rgba, uint32_t n_materials);

/* ---- batched rendering (no reference counterpart: a reference Renderer
 * renders one scene, src/device/RendererImpl.hpp:49-102) ----------------
 *
 * A batch merges built scenes into one device forest (each scene keeps its
 * own BVH root, materials and light table), so that views of many scenes of
 * uneven size render in one sequence of wavefront launches whose bounce loop
 * can be recorded once as a HIP graph and replayed per chunk of paths. */
typedef struct dr_batch dr_batch;

typedef struct {
    uint32_t scene;          /* index into the batch's scenes */
    uint32_t spp;            /* samples per pixel added by this call */
    uint32_t sample_offset;  /* first sample index (progressive rendering) */
    uint32_t pad;
    dr_camera camera;
    double *accum;           /* device FP64 RGBW, width*height*32 bytes */
} dr_batch_view;

/* Launch shape (the autotuner's knobs). NULL = defaults {1 << 21,
 * DR_SCHED_PER_RAY_SPECULATIVE, 16, 1, DR_BATCH_LAUNCH_SEQUENCE, 0}.
 *
 * DR_BATCH_LAUNCH_SEQUENCE: per chunk of paths, one launch per wavefront
 * stage (camera, then per bounce trace, shade, shadow trace, shadow apply,
 * then resolve), the bounce loop optionally recorded as a HIP graph.
 * DR_BATCH_LAUNCH_PERSISTENT: the whole render in one cooperative launch of
 * resident blocks (blocks_per_cu per CU, 0 = the occupancy limit) that runs
 * the same stages in the same order, separated by grid-wide barriers; the
 * traversal uses its persistent schedule (speculative if trace_sched has
 * bit 0). Same images as the sequence. The call then waits for the launch,
 * and fails with DR_E_DEVICE if a barrier wait ran out (it is bounded). */
#define DR_BATCH_LAUNCH_SEQUENCE 0u
#define DR_BATCH_LAUNCH_PERSISTENT 1u
typedef struct {
    uint32_t chunk_paths;   /* paths in flight per chunk (4096 .. 1 << 24) */
    uint32_t trace_sched;   /* DR_SCHED_* for the extension and shadow traces */
    uint32_t refill_min;    /* persistent traversal refill threshold (1..64) */
    uint32_t use_graph;     /* sequence: record the bounce loop once, replay it per chunk */
    uint32_t launch;        /* DR_BATCH_LAUNCH_* */
    uint32_t blocks_per_cu; /* persistent launch: resident blocks per CU (0 = occupancy limit, at most 16) */
} dr_batch_options;

/* Scenes must have their BVH built (their light tables are built here if
 * needed). The batch copies what it needs; the scenes may be destroyed. */
dr_status dr_batch_create(dr_context *ctx, dr_scene *const *scenes, uint32_t n_scenes, dr_batch **out);
dr_status dr_batch_destroy(dr_batch *batch);
/* Render every view; view v uses the random stream of a dr_render call
 * with seed params->seed ^ ((uint64_t)v << 32), so each view's result
 * equals the corresponding single-scene render. params->spp and
 * sample_offset are ignored (per view). Waits for earlier work on `stream`
 * at entry (the view table is rewritten), then runs asynchronously. */
dr_status dr_batch_render(dr_batch *batch, const dr_batch_view *views, uint32_t n_views,
                          const dr_render_params *params, const dr_batch_options *options, dr_stream stream);
/* Launch-shape autotuner: renders the views with each view's spp cut to
 * budget_spp into a scratch accumulation buffer (the views' own images are
 * not touched), reps + 1 times per candidate (the first untimed), and writes
 * the candidate of least median time to *best. Candidates: the launch
 * sequence over chunk_paths 2^21 .. 2^24, the per-ray and the resident-grid
 * speculative schedules (refill_min 8, 16, 32 for the latter) and graph
 * recording off and on; the persistent launch over chunk_paths 2^22 and
 * 2^23, the resident-grid speculative schedule with one ray counter and with
 * one per XCD (DR_SCHED_XCD_SLICES), refill_min 32 and 64 and blocks_per_cu
 * 1 and 2. If max_tried > 0,
 * the first max_tried candidates and their median milliseconds go to tried
 * and ms; *n_tried (may be NULL) is the number of candidates timed.
 * Synchronous. */
dr_status dr_batch_autotune(dr_batch *batch, const dr_batch_view *views, uint32_t n_views,
                            const dr_render_params *params, uint32_t budget_spp, uint32_t reps,
                            dr_batch_options *best, dr_batch_options *tried, float *ms, uint32_t max_tried,
                            uint32_t *n_tried);

/* ---- instancing: one scene placed several times ----------------------------
 *
 * Replaces the reference's objects that share one GpuTris storage with a
 * transform each (viewGpuTris, src/GpuTris.cpp:61-85; TriangleMesh's
 * model_to_world, include/tracing/TriangleMesh.hpp:9-27) and its per-object
 * loop in cast (include/tracing/IntersectionTestsImpl.hpp:234-251). A ray is
 * tested against an instance in model space, as getIntersectionImpl does
 * (IntersectionTestsImpl.hpp:143-151): o' = (o - translate) * inv_s,
 * d' = d * inv_s with inv_s = 1 / scale per axis (Transform::applyInverse,
 * include/Transform.hpp:14-21), d' not renormalised, so t is world t. A
 * top-level 4-wide tree over the instances' world boxes selects the instances
 * a ray enters; each scene's tree is stored once.
 *
 * Hit ids number the instances' triangles in instance order: instance i's
 * triangle j is (triangles of instances 0 .. i-1) + j, as in a scene that
 * concatenates the transformed copies in that order. Ties on t go to the
 * lower id. */
typedef struct dr_instances dr_instances;

typedef struct {
    uint32_t scene;      /* index into the scenes passed to dr_instances_create */
    float scale[3];      /* per axis, finite and non-zero (negative mirrors) */
    float translate[3];  /* world = model * scale + translate */
} dr_instance;

/* Scenes must have their BVH built; the instance set copies their trees, so
 * they may be destroyed afterwards (a later rebuild or refit of a scene does
 * not reach the set). At most 2^28 - 2 instances, 2^31 - 1 instanced
 * triangles and 2^28 - 2 triangle references in the scenes. */
dr_status dr_instances_create(dr_context *ctx, dr_scene *const *scenes, uint32_t n_scenes,
                              const dr_instance *instances, uint32_t n_instances, dr_instances **out);
dr_status dr_instances_destroy(dr_instances *inst);
/* Closest hit and occlusion as dr_trace_closest / dr_trace_any. Instances
 * run the speculative schedules: DR_SCHED_PER_RAY becomes
 * DR_SCHED_PER_RAY_SPECULATIVE and DR_SCHED_PERSISTENT becomes
 * DR_SCHED_PERSISTENT_SPECULATIVE (default: persistent speculative). */
dr_status dr_instances_trace_closest(dr_instances *inst, const dr_ray *rays, uint32_t n, dr_hit *hits,
                                     dr_stream stream);
dr_status dr_instances_trace_any(dr_instances *inst, const dr_ray *rays, uint32_t n, uint32_t *occluded,
                                 dr_stream stream);
dr_status dr_instances_set_trace_schedule(dr_instances *inst, uint32_t closest_sched, uint32_t any_sched);
/* n_triangles: instanced triangles; n_nodes: top-level plus scene nodes;
 * depth: top-level levels plus the deepest scene's levels (the traversal
 * stack holds 3 x depth + 1 entries); overflow_flag as for scenes. */
dr_status dr_instances_get_info(dr_instances *inst, dr_scene_info *info);
/* Path-traced render of the instance set as dr_render renders a scene (the
 * reference renders its objects, each a TriangleMesh view with its own
 * model_to_world, Renderer::render, src/device/RendererImpl.hpp:22-102):
 * each instanced triangle keeps its scene's material; hit points, normals
 * and lights are in world space (model * scale + translate, per axis), and
 * every emissive triangle of every instance is a light weighted by its world
 * area (the reference weights an object's lights by its scaled surface area,
 * src/Renderer.cpp:113-152). Traversal runs in model space through the
 * two-level forest, so nothing is copied per instance. With exact transforms
 * the image equals dr_render of the flattened copy bit for bit. Plain
 * sampler only (no guide); the light table is built on first use. */
dr_status dr_instances_render(dr_instances *inst, const dr_camera *camera, const dr_render_params *params,
                              double *accum, dr_stream stream);
/* The instance set's light table (built if needed): n_emit entries as
 * dr_scene_copy_lights but naming emitters by index, their instanced
 * triangle ids (ascending) and their area pdfs; any pointer may be NULL. */
dr_status dr_instances_light_count(dr_instances *inst, uint32_t *n_emit);
dr_status dr_instances_copy_lights(dr_instances *inst, void *entries, uint32_t *emit_tri, float *emit_pdf);

/* ---- learned path guide (no reference counterpart: the reference samples
 * Lambert bounces from the BSDF only, PathGeneration.hpp:406-420) ----------
 *
 * Network: 32 input features (frequency encoding of the position within the
 * guide's box: per axis sin and cos of 2^(k+1) pi u, k = 0..3, for u in [0, 1]
 * across the box; the shading normal and the outgoing direction) -> 4 hidden
 * ReLU layers of 64 -> 48 outputs, of which 41 are used: an 8-lobe von
 * Mises-Fisher mixture (per lobe a mean direction, log concentration and
 * mixture logit) and one selection logit alpha = 0.1 + 0.8 sigmoid(s).
 * Weights and activations are BF16 on the matrix cores, accumulation FP32;
 * the FP32 master parameters and Adam moments stay on the device.
 *
 * Parameter vector (DR_GUIDE_N_PARAMS floats): for layer l = 0..4 with
 * in = {32,64,64,64,64}, out = {64,64,64,64,48}: W_l row-major [out][in]
 * then b_l [out]. Output row r of the last layer means: with g = (r % 16)/4
 * and q = 4*(r/16) + r%4, lobe 2g+q/5 component q%5 (mu x, mu y, mu z,
 * log kappa, logit) for q < 10; the selection logit for g = 0, q = 10; unused
 * otherwise. */
#define DR_GUIDE_N_PARAMS 17712
#define DR_GUIDE_N_LOBES 8
#define DR_GUIDE_HEAD_FLOATS 44 /* 8 x {mu x, mu y, mu z, kappa, weight}, alpha, 3 pad */

typedef struct dr_guide dr_guide;

typedef struct {
    float bounds[6];     /* box the position encoding is normalised to: lo xyz, hi xyz */
    uint32_t capacity;   /* training ring: latest samples of non-zero weight (e.g. 1 << 18);
                          * a collecting render stages up to 4 x capacity vertices */
    float lr;            /* Adam step size (1e-2) */
    float beta1, beta2;  /* Adam moment decay (0.9, 0.99) */
    float eps;           /* Adam epsilon (1e-8) */
    uint64_t seed;       /* parameter initialisation */
} dr_guide_params;

/* One training sample (64 B): a Lambert vertex at pos with shading normal,
 * outgoing direction wo (towards the previous vertex) and sampled incident
 * direction wi; weight = luminance(L_i(wi)) * max(0, n . wi) / pdf(wi), the
 * one-sample estimate of the target product over the pdf it was drawn from,
 * and pdf that density (the training loss does not read it). */
typedef struct {
    float pos[3], weight;
    float normal[3], pdf;
    float wo[3], pad1;
    float wi[3], pad2;
} dr_guide_sample;

dr_status dr_guide_create(dr_context *ctx, const dr_guide_params *params, dr_guide **out);
dr_status dr_guide_destroy(dr_guide *guide);
/* FP32 master parameters (host arrays of DR_GUIDE_N_PARAMS floats). Setting
 * them resets the Adam moments and step count. */
dr_status dr_guide_get_params(dr_guide *guide, float *params);
dr_status dr_guide_set_params(dr_guide *guide, const float *params);
/* Network evaluation for n samples (device arrays; only pos, normal and wo
 * are read): head[DR_GUIDE_HEAD_FLOATS * i] = the decoded mixture. */
dr_status dr_guide_eval(dr_guide *guide, const dr_guide_sample *samples, uint32_t n, float *head, dr_stream stream);
/* Gradient of the training loss over n samples: with q the mixture density
 * at wi and b = max(0, n . wi) / pi, the lobes descend
 * sum_i weight_i * -log(q/2 + b/2) and the selection logit descends
 * sum_i weight_i * -log(alpha q + (1 - alpha) b) with q held fixed. grad
 * receives DR_GUIDE_N_PARAMS + 1 floats (device), the last sum_i weight_i. */
dr_status dr_guide_gradient(dr_guide *guide, const dr_guide_sample *samples, uint32_t n, float *grad,
                            dr_stream stream);
/* One Adam step on the weight-normalised loss. samples == NULL trains on the
 * guide's ring of the latest non-zero-weight samples from collecting renders
 * (fill read on the device). */
dr_status dr_guide_train_step(dr_guide *guide, const dr_guide_sample *samples, uint32_t n, dr_stream stream);
/* Samples staged by the last collecting render, weights included (each such
 * render starts a new stage of at most 4 x capacity; *n_total counts all
 * offered). host may be NULL to query the counts. Synchronises.
 * Order: by a key of (path sample, vertex) that hashes the path sample, the
 * order in which they also enter the training ring (when a render offers
 * more than the ring holds, the ring keeps the first in this order, spread
 * over the image). So a collecting render whose stage did not fill
 * (*n_total <= 4 x capacity) stages and trains the same way on every run.
 * When it fills, which of the offered samples it kept depends on the
 * device's scheduling. */
dr_status dr_guide_copy_samples(dr_guide *guide, dr_guide_sample *host, uint32_t max_n, uint32_t *n_kept,
                                uint32_t *n_total);
/* Fraction of eligible vertices a collecting render records (default 1);
 * the number of leading path vertices that sample from the guide (default
 * 64 = all; later vertices sample the BSDF alone); samples per training step
 * from the ring (0 = all of it; otherwise consecutive steps take successive
 * windows of that size, wrapping around the ring). */
dr_status dr_guide_set_options(dr_guide *guide, float collect_fraction, uint32_t guided_vertices,
                               uint32_t train_batch);
/* Adam step size for the following training steps (> 0), for schedules that
 * decay it as the guide converges. */
dr_status dr_guide_set_learning_rate(dr_guide *guide, float lr);
/* Training objective of the selection logit (no reference counterpart).
 * DR_GUIDE_SELECT_KL (default) descends sum_i weight_i * -log(alpha q + (1 -
 * alpha) b); DR_GUIDE_SELECT_VARIANCE descends the one-sample estimator's
 * second moment sum_i weight_i^2 pdf_i / (alpha q + (1 - alpha) b) (Rath et
 * al. 2020), with pdf_i the density the sample was drawn from, and trains
 * only the output row of the selection logit. The lobes keep their loss. */
enum { DR_GUIDE_SELECT_KL = 0, DR_GUIDE_SELECT_VARIANCE = 1 };
dr_status dr_guide_set_selection_objective(dr_guide *guide, int32_t objective);

/* dr_render with the guide: DR_GUIDE_USE samples Lambert bounces from the
 * one-sample MIS of the guide mixture and the cosine BSDF (with probability
 * alpha from the guide), and NEE weights against that combined pdf;
 * DR_GUIDE_COLLECT gathers training samples. guide_flags == 0 is dr_render. */
enum { DR_GUIDE_USE = 1, DR_GUIDE_COLLECT = 2 };
dr_status dr_render_guided(dr_scene *scene, const dr_camera *camera, const dr_render_params *params,
                           dr_guide *guide, uint32_t guide_flags, double *accum, dr_stream stream);

/* ---- workload generators (host, deterministic; no reference counterpart:
 * the reference's tests build their meshes by hand) ---- */
/* Icosphere with 20*4^subdiv triangles, shared vertices, outward winding;
 * vertex radius 1 + sigma*N(0,1) from Philox(seed, vertex index). Query sizes
 * with NULL outputs. */
dr_status dr_gen_icosphere(uint32_t subdiv, float sigma, uint64_t seed, float *vertices,
                           uint32_t *n_vertices, uint32_t *indices, uint32_t *n_triangles);
/* Incoherent rays: origins area-uniform on the mesh, directions cosine-
 * distributed about the inward (inward != 0) or outward face normal, origin
 * offset by `offset` along the direction; tmin 0, tmax inf. */
dr_status dr_gen_rays_surface(const float *vertices, const uint32_t *indices, uint32_t n_triangles,
                              uint32_t n_rays, uint64_t seed, int32_t inward, float offset,
                              dr_ray *rays);
/* Pinhole primary rays, width x height, pixel centres, row-major from the top. */
dr_status dr_gen_rays_camera(const float eye[3], const float target[3], const float up[3],
                             float fov_y_deg, uint32_t width, uint32_t height, dr_ray *rays);

#ifdef __cplusplus
}
#endif
#endif
