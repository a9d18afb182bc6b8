/*
 * DevilRay CPU oracle: the reference's integrator, in reference semantics.
 * TEST INFRASTRUCTURE ONLY -- see oracle.h. Compiled with -ffp-contract=off.
 *
 * Restates sampleColor (include/tracing/PathGeneration.hpp:363-515) as the
 * reference defines it, over the reference's scene model:
 *   - objects, each with its own median-split BBH (src/models/BBH.cpp:38-115);
 *     cast() tries every object and keeps the first strictly nearest hit
 *     (IntersectionTestsImpl.hpp:234-251), with the plane-based triangle test
 *     (:253-297); the hit point is p + v t and the normal the unit geometric
 *     normal (A-B)x(A-C) turned against the ray (:38-66);
 *   - the whole path is generated first (at most max_depth vertices, 10 in
 *     the reference, no Russian roulette, a miss ends it), then next-event
 *     estimation and emission MIS run over the stored vertices (:455-510);
 *   - Lambert bounces: normalize(uniform sphere + n), throughput x albedo;
 *     dielectrics: the IOR stack of object identities with capacity 3 whose
 *     top() pops (:42-68, 139-176), total internal reflection, Schlick's
 *     Fresnel on the incident cosine (:75-86), reflect / refract (:178-207);
 *     every new ray starts 1e-5 along the normal of the side it leaves from
 *     (:210);
 *   - emission weighted by the power heuristic of (bsdf pdf, light pdf) of
 *     the previous vertex, 1 after a dielectric or the camera
 *     (misWeightedEmission :306-318, computeNextBounceMisPdfs :320-334);
 *   - one light sample per Lambert vertex, visible() through a closest-hit
 *     cast (:213-224), evaluateDirectLighting (:336-361).
 *
 * Where this restatement differs from the reference's text (each by a
 * definition the reference leaves open, not by a changed formula):
 *   - random numbers: Philox-4x32-10 streams keyed by (pixel, seed) with
 *     counters (sample, vertex, stream), the build's own, instead of
 *     cuRAND XORWOW; so images agree with the reference statistically;
 *   - light selection uses the flattened emissive-triangle table. Its
 *     probability per triangle, area x luminance(E) / sum, equals the
 *     reference's object table (area x luminance(pi E)) times its
 *     per-object triangle table (area); the alias pairing is not restated;
 *   - an object's identity for the IOR stack is its index in the object
 *     list, and its material that of its first triangle; a push onto a full
 *     stack (more than 3 nested dielectrics, undefined in the reference) is
 *     dropped.
 */
#include "oracle.h"
#include "or_internal.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y, z; } v3;
static const float kPi = 3.14159274f;

static inline v3 mk(float x, float y, float z) { v3 r = { x, y, z }; return r; }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 vdiv(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline v3 vmulv(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float vlen(v3 a) { return sqrtf(vdot(a, a)); }
static inline v3 vcross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline v3 vnorm(v3 a) { return vdiv(a, vlen(a)); }

typedef struct {
    const float *verts;
    const int32_t *tris;
    int32_t n_obj;
    or_object *objects;
    void **bbh;
} refscene;

void *or_refscene_build(const float *verts, int32_t nverts, const int32_t *tris, int32_t ntris,
                        const or_object *objects, int32_t n_objects)
{
    refscene *s = (refscene *)calloc(1, sizeof(refscene));
    s->verts = verts;
    s->tris = tris;
    s->n_obj = n_objects;
    s->objects = (or_object *)malloc(sizeof(or_object) * (size_t)(n_objects > 0 ? n_objects : 1));
    s->bbh = (void **)calloc((size_t)(n_objects > 0 ? n_objects : 1), sizeof(void *));
    for (int32_t o = 0; o < n_objects; ++o) {
        s->objects[o] = objects[o];
        if (objects[o].first < 0 || objects[o].count < 0 || objects[o].first + objects[o].count > ntris) continue;
        s->bbh[o] = or_refbbh_build(verts, nverts, tris + 3 * (size_t)objects[o].first, objects[o].count);
    }
    return s;
}

void or_refscene_free(void *h)
{
    refscene *s = (refscene *)h;
    if (!s) return;
    for (int32_t o = 0; o < s->n_obj; ++o) or_refbbh_free(s->bbh[o]);
    free(s->bbh);
    free(s->objects);
    free(s);
}

typedef struct {
    int32_t obj, tri;
    float t;
} refhit;

/* cast (IntersectionTestsImpl.hpp:234-251): every object, first strictly
 * nearest hit kept */
static int cast(const refscene *s, v3 p, v3 d, refhit *best, int64_t *counts)
{
    const float pp[3] = { p.x, p.y, p.z }, dd[3] = { d.x, d.y, d.z };
    int has = 0;
    for (int32_t o = 0; o < s->n_obj; ++o) {
        if (!s->bbh[o]) continue;
        int32_t id;
        float t, bari[3];
        if (!or_refbbh_closest_one(s->bbh[o], s->verts, s->tris + 3 * (size_t)s->objects[o].first, pp, dd, &id, &t,
                                   bari, counts ? &counts[2] : NULL, counts ? &counts[3] : NULL))
            continue;
        if (!has || best->t > t) {
            has = 1;
            best->obj = o;
            best->tri = s->objects[o].first + id;
            best->t = t;
        }
    }
    return has;
}

static inline v3 vtx(const refscene *s, int32_t tri, int k)
{
    const float *v = s->verts + 3 * (size_t)s->tris[3 * (size_t)tri + k];
    return mk(v[0], v[1], v[2]);
}

/* visible (PathGeneration.hpp:213-224) */
static int visible(const refscene *s, v3 p0, v3 p1, int64_t *counts)
{
    const float distance = vlen(vsub(p1, p0));
    const v3 v = vdiv(vsub(p1, p0), distance);
    refhit h;
    if (counts) ++counts[1];
    if (!cast(s, vadd(p0, vmul(v, 1e-5f)), v, &h, counts)) return 1;
    return h.t > distance - 1e-5f * 2.0f;
}

static inline float luminance(const float *c) { return 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]; }

static inline float power_heuristic(float a, float b) { return a * a / (a * a + b * b); }

/* areaToSolidAngle, Utils.hpp:236-247 */
static inline float area_to_solid_angle(v3 from, v3 to, v3 to_n)
{
    const v3 diff = vsub(to, from);
    const float d2 = vdot(diff, diff);
    const v3 dir = vnorm(diff);
    const float c = fabsf(vdot(dir, to_n));
    if (c < 1e-6f) return 0.0f;
    return d2 / c;
}

/* cosineWeightedHemispherePdf, Utils.hpp:249-254 */
static inline float cosine_pdf(v3 from, v3 to, v3 n)
{
    const v3 dir = vnorm(vsub(to, from));
    const float c = vdot(dir, n);
    return c > 0.0f ? c / kPi : 0.0f;
}

/* schlick, PathGeneration.hpp:75-86 */
static inline float schlick(float dotp, float n1, float n2)
{
    return or_fresnel(dotp, n1, n2, 0.0f, 1);
}

typedef struct {
    v3 p, n, thr;
    int32_t mat;
} entry;

typedef struct {
    int32_t arr[3];
    int32_t size;
} ior_stack;

static inline int st_is_top(const ior_stack *s, int32_t v) { return s->size > 0 && s->arr[s->size - 1] == v; }
static inline void st_push(ior_stack *s, int32_t v) { if (s->size < 3) s->arr[s->size++] = v; }
static inline void st_pop(ior_stack *s) { if (s->size > 0) --s->size; }
/* the reference's top() returns the top element and pops it */
static inline int32_t st_top(ior_stack *s) { --s->size; return s->arr[s->size]; }

typedef struct {
    const refscene *s;
    const int32_t *tri_mat;
    const or_material *mats;
    const or_alias_entry *alias;
    int32_t n_emit;
    double total_power;  /* sum over emissive triangles of area x luminance(E) */
    const or_camera *cam;
    const or_render_params *prm;
    double *accum;
} rctx;

static inline const or_material *tri_material(const rctx *c, int32_t tri)
{
    return &c->mats[c->tri_mat ? c->tri_mat[tri] : 0];
}

static inline float object_ior(const rctx *c, int32_t obj)
{
    return tri_material(c, c->s->objects[obj].first)->ior;
}

static v3 sample_color(const rctx *c, uint32_t pixel, uint32_t sample, int64_t *counts)
{
    const or_render_params *prm = c->prm;
    const refscene *s = c->s;
    entry path[64];
    const uint32_t max_depth = prm->max_depth < 64 ? prm->max_depth : 64;
    int n = 0;
    or_ray r;
    or_camera_sample(c->cam, prm->seed, pixel, sample, &r);
    v3 ro = mk(r.ox, r.oy, r.oz), rv = mk(r.dx, r.dy, r.dz);
    v3 thr = mk(1.0f, 1.0f, 1.0f);
    ior_stack stack;
    stack.size = 0;
    /* ---- path generation (PathGeneration.hpp:432-453) */
    for (uint32_t depth = 0; depth < max_depth; ++depth) {
        refhit h;
        if (counts) ++counts[0];
        if (!cast(s, ro, rv, &h, counts)) break;
        const v3 A = vtx(s, h.tri, 0), B = vtx(s, h.tri, 1), C = vtx(s, h.tri, 2);
        v3 nrm = vnorm(vcross(vsub(A, B), vsub(A, C)));
        if (vdot(nrm, rv) > 0.0f) nrm = vmul(nrm, -1.0f);
        nrm = vnorm(nrm);  /* (n * inv_s).normalized() with unit scale */
        const v3 p = vadd(ro, vmul(rv, h.t));
        const or_material *m = tri_material(c, h.tri);
        path[n].p = p;
        path[n].n = nrm;
        path[n].thr = thr;
        path[n].mat = c->tri_mat ? c->tri_mat[h.tri] : 0;
        ++n;
        uint32_t r0[4];
        or_path_rng(pixel, sample, depth, 0u, prm->seed, r0);
        v3 normal = nrm;
        if (m->kind == 0) {
            float w[3];
            const float nn[3] = { nrm.x, nrm.y, nrm.z };
            or_sample_cosine(nn, or_u01(r0[0]), or_u01(r0[1]), w);
            thr = vmulv(thr, mk(m->albedo[0], m->albedo[1], m->albedo[2]));
            ro = p;
            rv = mk(w[0], w[1], w[2]);
        } else {
            float n1, n2;
            if (st_is_top(&stack, h.obj)) {  /* leaving the last object entered */
                n1 = m->ior;
                st_pop(&stack);
                n2 = stack.size == 0 ? 1.0f : object_ior(c, st_top(&stack));
                st_push(&stack, h.obj);
            } else {
                n2 = m->ior;
                n1 = stack.size == 0 ? 1.0f : object_ior(c, st_top(&stack));
            }
            const float eta = n1 / n2;
            const float d = vdot(rv, normal);
            const float k = 1.0f - eta * eta * (1.0f - d * d);
            v3 v;
            if (n1 > n2 && k < 0.0f) {
                v = vsub(rv, vmul(normal, 2.0f * d));
            } else {
                const float F = schlick(-d, n1, n2);
                if (or_u01(r0[3]) < F) {
                    v = vsub(rv, vmul(normal, 2.0f * d));
                } else {
                    if (st_is_top(&stack, h.obj)) st_pop(&stack);
                    else st_push(&stack, h.obj);
                    normal = vmul(normal, -1.0f);
                    const float dotp = -d;
                    v = vadd(vmul(normal, sqrtf(k) - dotp * eta), vmul(rv, eta));
                }
            }
            ro = p;
            rv = v;
        }
        ro = vadd(ro, vmul(normal, 1e-5f));
    }
    /* ---- emission and next-event estimation over the stored path (:455-510) */
    v3 color = mk(0.0f, 0.0f, 0.0f);
    int prev_specular = 1;
    float pdf_bsdf = 0.0f, pdf_nee = 0.0f;
    for (int i = 0; i < n; ++i) {
        const or_material *m = &c->mats[path[i].mat];
        if (m->kind != 0) {
            prev_specular = 1;
            continue;
        }
        const v3 E = mk(m->emission[0], m->emission[1], m->emission[2]);
        const v3 th = path[i].thr;
        if (prev_specular) {
            color = vadd(color, vmulv(E, th));
        } else if (pdf_bsdf + pdf_nee > 0.0f) {
            color = vadd(color, vmul(vmulv(E, th), power_heuristic(pdf_bsdf, pdf_nee)));
        }
        pdf_bsdf = 0.0f;
        pdf_nee = 0.0f;
        if (i + 1 < n) {
            const or_material *mn = &c->mats[path[i + 1].mat];
            pdf_bsdf = cosine_pdf(path[i].p, path[i + 1].p, path[i].n);
            const float exit_lum = mn->kind == 0 ? luminance(mn->emission) : 0.0f;
            const float pdf_area = (float)((double)exit_lum / c->total_power);
            pdf_nee = pdf_area * area_to_solid_angle(path[i].p, path[i + 1].p, path[i + 1].n);
        }
        if (c->n_emit > 0) {
            /* samplePointOnLights (:234-272) over the flattened table */
            uint32_t r1[4];
            or_path_rng(pixel, sample, (uint32_t)i, 1u, prm->seed, r1);
            float findex = or_u01(r1[0]) * (float)c->n_emit;
            if (findex >= (float)c->n_emit) findex = (float)(c->n_emit - 1);
            const or_alias_entry *en = &c->alias[(int32_t)findex];
            const int take_a = or_u01(r1[1]) <= en->p_a;
            const int32_t lt = take_a ? en->a : en->b;
            const float psel = take_a ? en->pdf_a : en->pdf_b;
            if (lt >= 0) {
                const v3 LA = vtx(s, lt, 0), LB = vtx(s, lt, 1), LC = vtx(s, lt, 2);
                float lp[3];
                const float a3[3] = { LA.x, LA.y, LA.z }, b3[3] = { LB.x, LB.y, LB.z }, c3[3] = { LC.x, LC.y, LC.z };
                or_sample_triangle(a3, b3, c3, or_u01(r1[2]), or_u01(r1[3]), lp);
                const v3 perp = vcross(vsub(LA, LB), vsub(LA, LC));
                const float plen = vlen(perp);
                const v3 ln = vdiv(perp, plen);
                const float lpdf = psel / (plen / 2.0f);
                const v3 xl = mk(lp[0], lp[1], lp[2]);
                const or_material *lm = tri_material(c, lt);
                const float nb = cosine_pdf(path[i].p, xl, path[i].n);
                const float nn = lpdf * area_to_solid_angle(path[i].p, xl, ln);
                /* evaluateDirectLighting (:336-361) */
                v3 Ld = mk(0.0f, 0.0f, 0.0f);
                if (visible(s, path[i].p, xl, counts)) {
                    const float distance = vlen(vsub(xl, path[i].p));
                    const v3 wl = vdiv(vsub(xl, path[i].p), distance);
                    const float cs = vdot(wl, path[i].n);
                    const float cl = vdot(wl, ln);
                    if (cs > 0.0f) {
                        const float g = cs * fabsf(cl) / (distance * distance);
                        const v3 brdf = vdiv(mk(m->albedo[0], m->albedo[1], m->albedo[2]), kPi);
                        Ld = vdiv(vmul(vmulv(mk(lm->emission[0], lm->emission[1], lm->emission[2]), brdf), g), lpdf);
                    }
                }
                if (nn + nb > 0.0f) color = vadd(color, vmul(vmulv(Ld, th), power_heuristic(nn, nb)));
            }
        }
        prev_specular = 0;
    }
    return color;
}

typedef struct {
    const rctx *c;
    int32_t begin, end;
    int64_t counts[4];
} rjob;

static void *ref_thread(void *arg)
{
    rjob *j = (rjob *)arg;
    const rctx *c = j->c;
    for (int32_t pixel = j->begin; pixel < j->end; ++pixel) {
        double *a = c->accum + 4 * (size_t)pixel;
        for (uint32_t k = 0; k < c->prm->spp; ++k) {
            const v3 L = sample_color(c, (uint32_t)pixel, c->prm->sample_offset + k, j->counts);
            a[0] = a[0] + (double)L.x;
            a[1] = a[1] + (double)L.y;
            a[2] = a[2] + (double)L.z;
            a[3] = a[3] + 1.0;
        }
    }
    return NULL;
}

void or_render_ref(const void *refscene_h, const int32_t *tri_mat, const or_material *mats,
                   const or_alias_entry *alias, int32_t n_emit, const or_camera *cam, const or_render_params *prm,
                   int32_t pixel_begin, int32_t pixel_end, double *accum, int64_t *counts, int32_t threads)
{
    const refscene *s = (const refscene *)refscene_h;
    double total = 0.0;
    for (int32_t o = 0; o < s->n_obj; ++o)
        for (int32_t t = s->objects[o].first; t < s->objects[o].first + s->objects[o].count; ++t) {
            const or_material *m = &mats[tri_mat ? tri_mat[t] : 0];
            if (m->kind != 0) continue;
            const float lum = luminance(m->emission);
            if (!(lum > 0.0f)) continue;
            const v3 A = vtx(s, t, 0), B = vtx(s, t, 1), C = vtx(s, t, 2);
            total += (double)(vlen(vcross(vsub(A, B), vsub(A, C))) / 2.0f) * (double)lum;
        }
    rctx c = { s, tri_mat, mats, alias, n_emit, total > 0.0 ? total : 1.0, cam, prm, accum };
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    const int32_t n = pixel_end - pixel_begin;
    pthread_t th[256];
    rjob *jobs = (rjob *)calloc(256, sizeof(rjob));
    int started = 0;
    const int32_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        jobs[t].c = &c;
        jobs[t].begin = pixel_begin + t * chunk;
        jobs[t].end = pixel_begin + (t + 1) * chunk < pixel_end ? pixel_begin + (t + 1) * chunk : pixel_end;
        if (jobs[t].begin >= jobs[t].end) break;
        pthread_create(&th[t], NULL, ref_thread, &jobs[t]);
        ++started;
    }
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
    if (counts)
        for (int t = 0; t < started; ++t)
            for (int k = 0; k < 4; ++k) counts[k] += jobs[t].counts[k];
    free(jobs);
}
