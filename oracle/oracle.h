/*
 * DevilRay CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C, the reference's render-path algorithms
 * (Frontier789/DevilRay) and the build's own well-defined variants of them, so
 * that the HIP product path can be checked against it. Only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark may load it; the
 * product library never links, loads or calls it.
 *
 * Two semantics live here:
 *   - "reference semantics": restatements of the reference's functions
 *     (plane-based triangle test, slab box test, median-split BBH with the
 *     stackless bit-trail walk, heap-paired alias table). Cited per function.
 *   - "build semantics": the exact arithmetic the GPU kernels use (watertight
 *     triangle test, robust slab test, (t, id) tie-break, Philox RNG), so that
 *     hit ids are bit-exact and t/u/v identical between CPU and GPU.
 *
 * Every function is pinned by a test under tests/ (known-answer vectors taken
 * from the reference's own GoogleTest cases, published Philox vectors,
 * closed forms, brute force on small inputs). A function without such a check
 * would say "parity unpinned" here; none is at present except where noted.
 */
#ifndef DEVILRAY_ORACLE_H
#define DEVILRAY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 32-byte ray and 16-byte hit, same layout as the product's dr_ray / dr_hit. */
typedef struct { float ox, oy, oz, tmin, dx, dy, dz, tmax; } or_ray;
typedef struct { int32_t id; float t, u, v; } or_hit;

/* ---- RNG ---------------------------------------------------------------- */
/* Philox-4x32-10 (Salmon et al., SC'11). */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* u32 -> float in [0,1): top 24 bits * 2^-24. */
float or_u01(uint32_t x);

/* ---- triangle / box tests ----------------------------------------------- */
/* Watertight test (Woop, Benthin, Wald, JCGT 2013), two-sided, build semantics.
 * Returns 1 and t, u (weight of v1), v (weight of v2) when the ray's line
 * crosses the triangle with det != 0; the caller applies the [tmin, tmax]
 * window. Arithmetic order is fixed and must match the device kernel. */
int or_tri_watertight(const or_ray *r, const float *v0, const float *v1, const float *v2,
                      float *t, float *u, float *v);

/* Reference semantics: testTriangleIntersection
 * (include/tracing/IntersectionTestsImpl.hpp:253-297). bari = (wA, wB, wC). */
int or_tri_reference(const float p[3], const float d[3], const float a[3], const float b[3],
                     const float c[3], float *t, float bari[3]);

/* Reference semantics: findBoxInterval / testBoxIntersection
 * (include/tracing/IntersectionTestsImpl.hpp:80-105). */
void or_box_interval_ref(float lo, float hi, float p, float v, float *enter, float *exit_);
int or_box_ref(const float lo[3], const float hi[3], const float p[3], const float d[3], float *t);

/* ---- canonical binary SAH BVH (build semantics) -------------------------- */
/* Binned SAH, `bins` bins, C_trav = C_isect = 1, at most `max_leaf` triangles
 * per leaf. Deterministic. Returns an opaque handle. */
void *or_bvh_build(const float *verts, int32_t nverts, const int32_t *tris, int32_t ntris,
                   int32_t max_leaf, int32_t bins);
void or_bvh_free(void *bvh);
int32_t or_bvh_node_count(const void *bvh);
/* SAH cost (C_trav = C_isect = 1) normalised by the root box area. */
double or_bvh_sah_cost(const void *bvh);
/* Checks structural invariants (leaves tile [0,n), child boxes inside parent,
 * triangles inside leaf boxes). Returns 0 when all hold. */
int32_t or_bvh_validate(const void *bvh, const float *verts, const int32_t *tris);

/* Closest-hit (any_hit = 0) or any-hit (any_hit = 1) over a batch, threaded.
 * Closest: hits[i] = min (t, id) over triangles with tmin <= t <= tmax, id -1
 * on miss. Any-hit: occluded[i] = 1 if such a triangle exists.
 * node_tests / tri_tests (may be NULL): per-ray counts of AABB tests and
 * triangle tests of the front-to-back traversal (32 B and 40 B units). */
void or_trace_bvh(const void *bvh, const float *verts, const int32_t *tris, const or_ray *rays,
                  int32_t n, int32_t any_hit, or_hit *hits, uint32_t *occluded,
                  int32_t *node_tests, int32_t *tri_tests, int32_t threads);

/* Brute force over all triangles (small inputs only). */
void or_trace_brute(const float *verts, const int32_t *tris, int32_t ntris, const or_ray *rays,
                    int32_t n, int32_t any_hit, or_hit *hits, uint32_t *occluded, int32_t threads);

/* Brute force over instances, each a scene placed by p * scale + translate.
 * A ray is tested in each instance's model space as the reference's
 * getIntersectionImpl does (include/tracing/IntersectionTestsImpl.hpp:143-151
 * with Transform::applyInverse, include/Transform.hpp:14-21):
 * o' = (o - translate) * (1 / scale), d' = d * (1 / scale), per axis, t
 * unchanged. Scene k's triangles are tris[scene_tofs[k] .. + scene_ntris[k])
 * indexing verts from scene_vofs[k]. Hit ids number the instances'
 * triangles in instance order; ties on t go to the lower id. */
typedef struct { uint32_t scene; float scale[3]; float translate[3]; } or_instance;  /* = dr_instance */
void or_trace_instances_brute(const float *verts, const int32_t *tris, const int32_t *scene_vofs,
                              const int32_t *scene_tofs, const int32_t *scene_ntris,
                              const or_instance *inst, int32_t n_inst, const or_ray *rays, int32_t n,
                              int32_t any_hit, or_hit *hits, uint32_t *occluded, int32_t threads);

/* ---- reference-semantics BBH (median split, 1 tri / leaf) ----------------- */
/* generateSimpleBBH / generateBBHLayer (src/models/BBH.cpp:38-115) and the
 * stackless bit-trail walk getIntersectionImpl
 * (include/tracing/IntersectionTestsImpl.hpp:141-220) with the reference
 * triangle test. The triangle permutation is returned so that ids can be
 * reported in input order. */
void *or_refbbh_build(const float *verts, int32_t nverts, const int32_t *tris, int32_t ntris);
void or_refbbh_free(void *bbh);
int32_t or_refbbh_depth(const void *bbh);
/* The nodes in the order generateSimpleBBH creates them (depth first, root
 * 0): lo xyz, hi xyz per node, and left, right (-1: none), parent and the
 * triangle range [tb, te) over the reordered triangles; order[i] is the
 * caller's id of reordered triangle i. */
int32_t or_refbbh_node_count(const void *bbh);
void or_refbbh_nodes(const void *bbh, float *lo_hi, int32_t *links);
void or_refbbh_order(const void *bbh, int32_t ntris, int32_t *order);
/* hits: id = input triangle index or -1, t, u = wB, v = wC. box_tests /
 * tri_tests: the reference HitTests counters (include/tracing/Benchmark.hpp). */
void or_refbbh_trace(const void *bbh, const float *verts, const int32_t *tris, const or_ray *rays,
                     int32_t n, or_hit *hits, int64_t *box_tests, int64_t *tri_tests,
                     int32_t threads);

/* ---- light sampler and path tracer (build semantics) --------------------- */
typedef struct { int32_t kind; float albedo[3]; float emission[3]; float ior; } or_material;  /* = dr_material */
typedef struct { float p_a, pdf_a, pdf_b; int32_t a, b; } or_alias_entry;                   /* 20 B */
typedef struct {
    float transform[12];
    float focal_length;
    float center[2];
    float pixel_size[2];
    uint32_t width, height, pixel_sampling;
} or_camera;                                                                               /* = dr_camera */
/* Pinhole ray through sensor point (x + sx, y + sy), sx, sy in [0, 1]
 * (Center sampling uses 0.5, 0.5). */
void or_camera_ray(const or_camera *cam, uint32_t x, uint32_t y, float sx, float sy, or_ray *out);
typedef struct {
    uint32_t spp, sample_offset, max_depth, rr_depth;
    uint64_t seed;
    uint32_t fresnel, flags;
} or_render_params;                                                                        /* = dr_render_params */
/* flags bits read by the oracle only: each plants a known mistake, so that
 * the tests can show their checks would catch it. RR_NO_REWEIGHT: roulette
 * survivors keep their throughput; TRANSMIT_WEIGHT: refracted throughput is
 * multiplied by 1 - F although refraction was already chosen with that
 * probability. */
#define OR_PLANT_RR_NO_REWEIGHT 0x10000u
#define OR_PLANT_TRANSMIT_WEIGHT 0x20000u

/* Alias table over emissive triangles, weights area x luminance(E) (FP64),
 * built by a sequential sweep in exact integer units. entries: one per
 * emitter (index order); pdf_area: per triangle (0 for non-emitters).
 * Returns the emitter count. */
int32_t or_lights_build(const float *verts, const int32_t *tris, int32_t ntris, const int32_t *tri_mat,
                        const or_material *mats, or_alias_entry *entries, float *pdf_area);

/* sin and cos of 2 pi u (the build's own polynomial). */
void or_sincos_2pi(float u, float *s, float *c);
/* The path tracer's samplers, exported for their own tests: a direction
 * uniform on the sphere, the cosine-weighted direction about n (normalised
 * sphere sample + n, the reference's PathGeneration.hpp:406-420), and the
 * area-uniform point of a triangle. r1, r2 in [0, 1). */
void or_sample_sphere(float r1, float r2, float out[3]);
/* Lighting terms of the integrator, exported for their own tests: the power
 * heuristic, the MIS weight of emission reached by a bounce, and one
 * next-event estimate (unoccluded, unweighted; returns 0 when the light is
 * behind the surface). */
float or_power_heuristic(float a, float b);
float or_emission_weight(int spec_prev, float pdf_bsdf, float pdf_light);
int or_direct_light(const float p[3], const float n[3], const float albedo[3], const float xl[3], const float nl[3],
                    const float emission[3], float pdf_area, float ld[3], float *pdf_light, float *pdf_bsdf);
void or_sample_cosine(const float n[3], float r1, float r2, float out[3]);
/* Scattering steps of the integrator, exported for their own tests: the
 * Fresnel reflectance (exact, or the reference's Schlick), one dielectric
 * scattering event (returns 1 for reflection, out = the new direction), and
 * Russian roulette after a vertex (returns whether the path goes on; beta is
 * divided by the survival probability). */
float or_fresnel(float cos_i, float n1, float n2, float cos_t, int32_t schlick);
int32_t or_dielectric_scatter(const float d[3], const float n[3], int32_t entering, float ior, float u,
                              int32_t schlick, float out[3]);
int32_t or_roulette(float beta[3], float u, uint32_t vertex, uint32_t rr_depth, uint32_t max_depth);
void or_sample_triangle(const float a[3], const float b[3], const float c[3], float r1, float r2, float out[3]);

/* Path-trace pixels [pixel_begin, pixel_end) adding spp samples each to
 * accum (4 doubles per pixel, in sample order). vertex_hist (optional,
 * max_depth ints) counts paths reaching each vertex. counts (optional,
 * int64[6], added to): extension rays, their interior-node visits and
 * triangle tests in the canonical BVH, then the same for shadow rays. */
void or_render(const void *bvh, const float *verts, const int32_t *tris, const int32_t *tri_mat,
               const or_material *mats, const or_alias_entry *alias, const float *pdf_area, int32_t n_emit,
               const or_camera *cam, const or_render_params *prm, int32_t pixel_begin, int32_t pixel_end,
               double *accum, int32_t *vertex_hist, int64_t *counts, int32_t threads);

/* ---- the reference's integrator (reference semantics, or_refrender.c) ------ */
/* An object: a contiguous range of triangles (the reference's TriangleMesh,
 * one material, its own BBH). */
typedef struct { int32_t first, count; } or_object;
/* Per-object median-split BBHs over verts / tris (kept by the caller). */
void *or_refscene_build(const float *verts, int32_t nverts, const int32_t *tris, int32_t ntris,
                        const or_object *objects, int32_t n_objects);
void or_refscene_free(void *scene);
/* sampleColor (PathGeneration.hpp:363-515) for pixels [pixel_begin,
 * pixel_end), prm->spp samples each (prm->max_depth vertices, 10 in the
 * reference; rr_depth and fresnel are ignored: no roulette, Schlick), added
 * to accum in sample order. alias / n_emit: the emissive-triangle table of
 * or_lights_build. counts (optional, int64[4], added to): extension casts,
 * shadow casts, BBH box tests, triangle tests. */
void or_render_ref(const void *scene, const int32_t *tri_mat, const or_material *mats,
                   const or_alias_entry *alias, int32_t n_emit, const or_camera *cam, const or_render_params *prm,
                   int32_t pixel_begin, int32_t pixel_end, double *accum, int64_t *counts, int32_t threads);

/* ---- display and debug views ------------------------------------------------ */
/* 8-bit decision thresholds of the linear (srgb = 0) or sRGB encoding. */
void or_display_thresholds(int32_t srgb, double thr[255]);
/* Renderer::createPixels (src/Renderer.cpp:51-90) from FP64 RGBW: rows flipped, RGBA8 packed. */
void or_resolve_rgba8(const double *accum, uint32_t width, uint32_t height, int32_t srgb, uint32_t *out);
/* Debug views of the first hit; mode 1 UV checker, 2 barycentric, 3 winding. */
void or_debug_aov(const void *bvh, const float *verts, const int32_t *tris, const int32_t *tri_mat,
                  const float *debug_rgba, const or_camera *cam, const or_render_params *prm, int32_t mode,
                  double *accum);

/* ---- learned path guide: FP64 network, head, loss and gradient ----------- */
#define OR_GUIDE_NP 17712
typedef struct {
    float pos[3], weight;
    float normal[3], pdf;
    float wo[3], pad1;
    float wi[3], pad2;
} or_guide_sample; /* = dr_guide_sample */

/* 32 input features of a sample (positions normalised to bounds lo..hi). */
void or_guide_encode(const float bounds[6], const or_guide_sample *s, double feat[32]);
/* Network outputs (48 raw) and decoded head (44: 8 x {mu xyz, kappa,
 * weight}, alpha, pad) for n samples. emulate_bf16 != 0 rounds weights,
 * features and activations to BF16 (round to nearest even) where the device
 * kernels do; 0 evaluates the real-valued network. */
void or_guide_forward(const double *params, const float bounds[6], const or_guide_sample *s, int32_t n,
                      int32_t emulate_bf16, double *raw, double *head);
/* Loss sum_i w_i * [-log(q/2 + cos+/(2 pi)) - log(alpha q0 + (1 - alpha) cos+/pi)]
 * with q, alpha from params and q0 the mixture of `frozen` (NULL: params). */
double or_guide_loss(const double *params, const double *frozen, const float bounds[6], const or_guide_sample *s,
                     int32_t n, int32_t emulate_bf16);
/* Gradient of the loss (OR_GUIDE_NP doubles); returns sum_i w_i. With
 * emulate_bf16 the backward pass also rounds the output and hidden deltas
 * to BF16 as the device does. */
double or_guide_grad(const double *params, const float bounds[6], const or_guide_sample *s, int32_t n,
                     int32_t emulate_bf16, double *grad);
/* The same with the selection term chosen by `objective`: 0 the KL above,
 * 1 the one-sample second moment sum_i w_i^2 pdf_i / (alpha q0 + (1 - alpha)
 * cos+/pi) (Rath et al. 2020), 2 none (lobe term alone). The gradient of
 * objective 1 trains only the selection row of the output layer: no delta
 * of that row reaches the hidden layers (objective 2 has no gradient form). */
double or_guide_loss_sel(const double *params, const double *frozen, const float bounds[6],
                         const or_guide_sample *s, int32_t n, int32_t emulate_bf16, int32_t objective);
double or_guide_grad_sel(const double *params, const float bounds[6], const or_guide_sample *s, int32_t n,
                         int32_t emulate_bf16, int32_t objective, double *grad);
/* vMF density kappa / (2 pi (1 - e^-2kappa)) e^(kappa (mu.w - 1)) and
 * W. Jakob's stable sampler (2012): w = 1 + log(x1 + (1 - x1) e^-2kappa) / kappa. */
double or_vmf_pdf(const double mu[3], double kappa, const double w[3]);
void or_vmf_sample(const double mu[3], double kappa, double x1, double x2, double out[3]);

#ifdef __cplusplus
}
#endif
#endif
