/*
 * DevilRay CPU oracle: emissive-triangle alias table and the path tracer.
 * TEST INFRASTRUCTURE ONLY -- see oracle.h. Compiled with -ffp-contract=off.
 *
 * The integrator restates the reference's sampleColor
 * (include/tracing/PathGeneration.hpp:363-515) as the build defines it:
 * emission with power-heuristic MIS at every Lambert vertex (weight 1 after
 * the camera or a specular bounce), NEE with one light sample per Lambert
 * vertex (misWeightedEmission :306-318, computeNextBounceMisPdfs :320-334,
 * evaluateDirectLighting :336-361, samplePointOnLights :234-272), cosine
 * bounces (cosineWeightedHemisphereSample, DistributionSamplers.hpp:33-42),
 * dielectric reflect / refract (generateNewRay :114-211), optional Russian
 * roulette, per-pixel FP64 accumulation in sample order.
 */
#include "oracle.h"
#include "or_internal.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Alias table: a sequential sweep over lights and heavies in index order,   */
/* in exact integer units (bucket capacity Q).                              */
/* ------------------------------------------------------------------------ */
int32_t or_lights_build(const float *verts, const int32_t *tris, int32_t ntris, const int32_t *tri_mat,
                        const or_material *mats, or_alias_entry *entries, float *pdf_area)
{
    double *w = (double *)malloc(sizeof(double) * (size_t)(ntris > 0 ? ntris : 1));
    double *area = (double *)malloc(sizeof(double) * (size_t)(ntris > 0 ? ntris : 1));
    int32_t *emit = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ntris > 0 ? ntris : 1));
    int32_t ne = 0;
    double wmax = 0.0;
    for (int32_t t = 0; t < ntris; ++t) {
        const or_material *m = &mats[tri_mat ? tri_mat[t] : 0];
        const double lum = 0.2126 * (double)m->emission[0] + 0.7152 * (double)m->emission[1] +
                           0.0722 * (double)m->emission[2];
        double a3[3][3];
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) a3[k][c] = (double)verts[3 * (size_t)tris[3 * (size_t)t + k] + c];
        const double e1[3] = { a3[1][0] - a3[0][0], a3[1][1] - a3[0][1], a3[1][2] - a3[0][2] };
        const double e2[3] = { a3[2][0] - a3[0][0], a3[2][1] - a3[0][1], a3[2][2] - a3[0][2] };
        const double cx = e1[1] * e2[2] - e1[2] * e2[1];
        const double cy = e1[2] * e2[0] - e1[0] * e2[2];
        const double cz = e1[0] * e2[1] - e1[1] * e2[0];
        const double ar = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
        area[t] = ar;
        if (pdf_area) pdf_area[t] = 0.0f;
        if (m->kind == 0 && lum > 0.0 && ar > 0.0) {
            w[ne] = ar * lum;
            if (w[ne] > wmax) wmax = w[ne];
            emit[ne++] = t;
        }
    }
    if (ne > 0) {
        int L = 1;
        while ((1ull << L) < (uint64_t)ne) ++L;
        const int S = 62 - 2 * L < 52 ? 62 - 2 * L : 52;
        const double scale = ldexp(1.0, S);
        uint64_t *q = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)ne);
        uint64_t Q = 0;
        for (int32_t e = 0; e < ne; ++e) { q[e] = (uint64_t)floor(w[e] / wmax * scale); Q += q[e]; }
        int32_t *lights = (int32_t *)malloc(sizeof(int32_t) * (size_t)ne);
        int32_t *heavies = (int32_t *)malloc(sizeof(int32_t) * (size_t)ne);
        int32_t na = 0, nb = 0;
        for (int32_t e = 0; e < ne; ++e) {
            if ((uint64_t)ne * q[e] < Q) lights[na++] = e;
            else heavies[nb++] = e;
        }
        for (int32_t e = 0; e < ne; ++e) {  /* default: full own bucket */
            entries[e].p_a = 1.0f;
            entries[e].pdf_a = (float)((double)q[e] / (double)Q);
            entries[e].pdf_b = 0.0f;
            entries[e].a = emit[e];
            entries[e].b = -1;
        }
        int32_t j = 0;
        uint64_t r = (uint64_t)ne * q[heavies[0]];  /* remaining mass of the current heavy */
        for (int32_t i = 0; i < na; ++i) {
            const int32_t e = lights[i], h = heavies[j];
            const uint64_t m = (uint64_t)ne * q[e];
            entries[e].p_a = (float)((double)m / (double)Q);
            entries[e].pdf_b = (float)((double)q[h] / (double)Q);
            entries[e].b = emit[h];
            r -= Q - m;
            while (r < Q && j < nb - 1) {
                const int32_t hj = heavies[j], hn = heavies[j + 1];
                entries[hj].p_a = (float)((double)r / (double)Q);
                entries[hj].pdf_b = (float)((double)q[hn] / (double)Q);
                entries[hj].b = emit[hn];
                ++j;
                r = (uint64_t)ne * q[hn] - (Q - r);
            }
        }
        if (pdf_area)
            for (int32_t e = 0; e < ne; ++e) pdf_area[emit[e]] = (float)((double)q[e] / (double)Q / area[emit[e]]);
        free(q); free(lights); free(heavies);
    }
    free(w); free(area); free(emit);
    return ne;
}

/* ------------------------------------------------------------------------ */
/* Shading math (same evaluation order as the device kernels).              */
/* ------------------------------------------------------------------------ */
typedef struct { float x, y, z; } v3;
static const float kPi = 3.14159274f;

static inline v3 mk(float x, float y, float z) { v3 r = { x, y, z }; return r; }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 vdiv(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline v3 vnorm(v3 a) { return vdiv(a, sqrtf(vdot(a, a))); }

/* sin / cos of 2 pi u: quadrant reduction, Taylor to x^11 / x^12 (Horner). */
void or_sincos_2pi(float u, float *s, float *c)
{
    const float q4 = u * 4.0f;
    int q = (int)q4;
    if (q > 3) q = 3;
    const float f = q4 - (float)q;
    const float x = f * 1.57079637f;
    const float x2 = x * x;
    float ps = -2.50521084e-8f;
    ps = ps * x2 + 2.75573192e-6f;
    ps = ps * x2 + -1.98412698e-4f;
    ps = ps * x2 + 8.33333333e-3f;
    ps = ps * x2 + -1.66666667e-1f;
    ps = ps * x2 + 1.0f;
    const float sn = ps * x;
    float pc = 2.08767570e-9f;
    pc = pc * x2 + -2.75573192e-7f;
    pc = pc * x2 + 2.48015873e-5f;
    pc = pc * x2 + -1.38888889e-3f;
    pc = pc * x2 + 4.16666667e-2f;
    pc = pc * x2 + -0.5f;
    pc = pc * x2 + 1.0f;
    const float cs = pc;
    if (q == 0) { *s = sn; *c = cs; }
    else if (q == 1) { *s = cs; *c = -sn; }
    else if (q == 2) { *s = -sn; *c = -cs; }
    else { *s = -cs; *c = sn; }
}

static inline v3 uniform_sphere(float r1, float r2)
{
    const float z = 1.0f - 2.0f * r2;
    const float st = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float s0, c0;
    or_sincos_2pi(r1, &s0, &c0);
    return mk(st * s0, st * c0, z);
}

static inline v3 cosine_dir(v3 n, float r1, float r2)
{
    const v3 w = vadd(uniform_sphere(r1, r2), n);
    const float l2 = vdot(w, w);
    if (!(l2 > 1e-20f)) return n;
    return vdiv(w, sqrtf(l2));
}

/* area-uniform point of triangle ABC from two uniforms, u + v > 1 folded back
 * (the reference's uniformTriangleSample, DistributionSamplers.hpp:92-104) */
static inline v3 triangle_point(v3 a, v3 b, v3 c, float u, float v)
{
    if (u + v > 1.0f) { u = 1.0f - u; v = 1.0f - v; }
    return vadd(vadd(a, vmul(vsub(b, a), u)), vmul(vsub(c, a), v));
}

void or_sample_sphere(float r1, float r2, float out[3])
{
    const v3 w = uniform_sphere(r1, r2);
    out[0] = w.x; out[1] = w.y; out[2] = w.z;
}

void or_sample_cosine(const float n[3], float r1, float r2, float out[3])
{
    const v3 w = cosine_dir(mk(n[0], n[1], n[2]), r1, r2);
    out[0] = w.x; out[1] = w.y; out[2] = w.z;
}

void or_sample_triangle(const float a[3], const float b[3], const float c[3], float r1, float r2, float out[3])
{
    const v3 p = triangle_point(mk(a[0], a[1], a[2]), mk(b[0], b[1], b[2]), mk(c[0], c[1], c[2]), r1, r2);
    out[0] = p.x; out[1] = p.y; out[2] = p.z;
}

static inline float max_abs(v3 a) { return fmaxf(fabsf(a.x), fmaxf(fabsf(a.y), fabsf(a.z))); }

/* offset scale: 1e-5 times the largest vertex coordinate of the triangle */
static inline float origin_eps(v3 a, v3 b, v3 c)
{
    return 1e-5f * fmaxf(1.0f, fmaxf(max_abs(a), fmaxf(max_abs(b), max_abs(c))));
}

/* hit point from barycentrics: A + (B - A) u + (C - A) v */
static inline v3 bary_point(v3 a, v3 b, v3 c, float u, float v)
{
    return vadd(vadd(a, vmul(vsub(b, a), u)), vmul(vsub(c, a), v));
}

/* power heuristic (beta = 2); 0 when both densities are 0 */
static inline float power_heuristic(float a, float b)
{
    const float a2 = a * a, b2 = b * b;
    const float s = a2 + b2;
    return s > 0.0f ? a2 / s : 0.0f;
}

float or_power_heuristic(float a, float b) { return power_heuristic(a, b); }

/* MIS weight of emission reached by a bounce: 1 after a specular vertex or
 * the camera, else the power heuristic of the bounce's pdf against the
 * light sampler's solid-angle pdf (misWeightedEmission,
 * PathGeneration.hpp:306-318) */
float or_emission_weight(int spec_prev, float pdf_bsdf, float pdf_light)
{
    return spec_prev ? 1.0f : power_heuristic(pdf_bsdf, pdf_light);
}

/* Next-event estimate of one light point xl (normal nl, emission E, area pdf
 * pa) seen from a Lambert point p (normal n, albedo alb), unoccluded and
 * unweighted: Ld = E (alb / pi) cos_s |cos_l| / d^2 / pa, with the light's
 * solid-angle pdf pn = pa d^2 / |cos_l| and the cosine bounce's pb = cos_s / pi
 * (evaluateDirectLighting, PathGeneration.hpp:336-361). Returns 0 (and no
 * estimate) when the light is behind the surface. */
static inline int nee_terms(v3 p, v3 n, v3 alb, v3 xl, v3 nl, const float *E, float pa, v3 *Ld, float *pn,
                            float *pb, float *dist, v3 *om)
{
    const v3 toL = vsub(xl, p);
    const float dist2 = vdot(toL, toL);
    *dist = sqrtf(dist2);
    *om = vdiv(toL, *dist);
    const float cs = vdot(*om, n);
    const float acl = fabsf(vdot(*om, nl));
    if (!(cs > 0.0f && pa > 0.0f && *dist > 0.0f)) return 0;
    const float G = cs * acl / dist2;
    *pn = acl < 1e-6f ? 0.0f : pa * dist2 / acl;
    *pb = cs / kPi;
    *Ld = mk(E[0] * (alb.x / kPi) * G / pa, E[1] * (alb.y / kPi) * G / pa, E[2] * (alb.z / kPi) * G / pa);
    return 1;
}

int or_direct_light(const float p[3], const float n[3], const float albedo[3], const float xl[3], const float nl[3],
                    const float emission[3], float pdf_area, float ld[3], float *pdf_light, float *pdf_bsdf)
{
    v3 L = mk(0.0f, 0.0f, 0.0f), om;
    float pn = 0.0f, pb = 0.0f, dist;
    const int ok = nee_terms(mk(p[0], p[1], p[2]), mk(n[0], n[1], n[2]), mk(albedo[0], albedo[1], albedo[2]),
                             mk(xl[0], xl[1], xl[2]), mk(nl[0], nl[1], nl[2]), emission, pdf_area, &L, &pn, &pb,
                             &dist, &om);
    ld[0] = L.x; ld[1] = L.y; ld[2] = L.z;
    *pdf_light = pn;
    *pdf_bsdf = pb;
    return ok;
}

/* Fresnel reflectance of a ray arriving with cosine cos_i from index n1 into
 * n2 (cos_t: cosine of the transmitted direction). Exact: the mean of the s-
 * and p-polarised reflectances. Schlick: the reference's schlick
 * (PathGeneration.hpp:75-86) with its types: float powers of 1 - cos_i, then
 * R0 + (1.0 - R0) c^5 in double (its 1.0 literal), rounded once to float. */
float or_fresnel(float cos_i, float n1, float n2, float cos_t, int32_t schlick)
{
    if (schlick) {
        const float r = (n1 - n2) / (n1 + n2);
        const float r0 = r * r;
        const float c = 1.0f - cos_i;
        const float c2 = c * c;
        const float c4 = c2 * c2;
        const float c5 = c4 * c;
        return (float)((double)r0 + (1.0 - (double)r0) * (double)c5);
    }
    const float rs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t);
    const float rp = (n2 * cos_i - n1 * cos_t) / (n2 * cos_i + n1 * cos_t);
    return 0.5f * (rs * rs + rp * rp);
}

/* Dielectric scattering at a hit (generateNewRay's transparent branch,
 * PathGeneration.hpp:139-207): d is the ray direction, n the geometric
 * normal turned against d, `entering` whether d arrived on the front side
 * (then n1 = 1, n2 = ior, else the reverse; winding replaces the reference's
 * IOR stack). Total internal reflection when n1 > n2 and
 * k = 1 - eta^2 (1 - (d.n)^2) < 0; otherwise reflection with probability F
 * (u < F), else refraction n'(sqrt k - cos_i eta) + eta d with n' = -n.
 * Returns 1 for reflection; *side is the normal on the outgoing side (the
 * ray origin moves along it). */
static inline int dielectric_scatter(v3 d, v3 n, int entering, float ior, float u, int schlick, v3 *nd, v3 *side,
                                     float *fres)
{
    const float n1 = entering ? 1.0f : ior;
    const float n2 = entering ? ior : 1.0f;
    const float dd = vdot(d, n);
    const float eta = n1 / n2;
    const float k = 1.0f - eta * eta * (1.0f - dd * dd);
    int reflect_ray = 1;
    *fres = 1.0f;
    if (!(n1 > n2 && k < 0.0f)) {
        const float ci = -dd;
        const float F = or_fresnel(ci, n1, n2, sqrtf(fmaxf(k, 0.0f)), schlick);
        *fres = F;
        reflect_ray = u < F;
    }
    if (reflect_ray) {
        *nd = mk(d.x - n.x * 2.0f * dd, d.y - n.y * 2.0f * dd, d.z - n.z * 2.0f * dd);
        *side = n;
    } else {
        const v3 nn = vmul(n, -1.0f);
        const float ci = -dd;
        const float sk = sqrtf(k);
        *nd = vadd(vmul(nn, sk - ci * eta), vmul(d, eta));
        *side = nn;
    }
    return reflect_ray;
}

int32_t or_dielectric_scatter(const float d[3], const float n[3], int32_t entering, float ior, float u,
                              int32_t schlick, float out[3])
{
    v3 nd, side;
    float F;
    const int r = dielectric_scatter(mk(d[0], d[1], d[2]), mk(n[0], n[1], n[2]), entering, ior, u, schlick, &nd,
                                     &side, &F);
    out[0] = nd.x; out[1] = nd.y; out[2] = nd.z;
    return r;
}

/* Russian roulette after vertex `vertex` (0-based): from vertex rr_depth - 1
 * on, a path whose largest throughput component q is below 1 survives with
 * probability q (u < q) and is divided by q; at max_depth vertices, or with
 * zero throughput, it ends. Returns whether the path goes on. */
static inline int roulette(v3 *beta, float u, uint32_t vertex, uint32_t rr_depth, uint32_t max_depth,
                           int planted_no_reweight)
{
    int alive = vertex + 1u < max_depth;
    if (alive && vertex + 1u >= rr_depth) {
        const float q = fmaxf(beta->x, fmaxf(beta->y, beta->z));
        if (q < 1.0f) {
            if (u >= q) alive = 0;
            else if (!planted_no_reweight) *beta = mk(beta->x / q, beta->y / q, beta->z / q);
        }
    }
    if (!(beta->x > 0.0f || beta->y > 0.0f || beta->z > 0.0f)) alive = 0;
    return alive;
}

int32_t or_roulette(float beta[3], float u, uint32_t vertex, uint32_t rr_depth, uint32_t max_depth)
{
    v3 b = mk(beta[0], beta[1], beta[2]);
    const int alive = roulette(&b, u, vertex, rr_depth, max_depth, 0);
    beta[0] = b.x; beta[1] = b.y; beta[2] = b.z;
    return alive;
}

static inline void path_rng(uint32_t pixel, uint32_t sample, uint32_t vertex, uint32_t stream, uint64_t seed,
                            uint32_t out[4])
{
    const uint32_t ctr[4] = { sample, vertex, stream, (uint32_t)(seed >> 32) };
    const uint32_t key[2] = { pixel, (uint32_t)seed };
    or_philox4x32_10(ctr, key, out);
}

static inline v3 vtx(const float *verts, int32_t vi)
{
    return mk(verts[3 * (size_t)vi], verts[3 * (size_t)vi + 1], verts[3 * (size_t)vi + 2]);
}

typedef struct {
    const void *bvh;
    const float *verts;
    const int32_t *tris;
    const int32_t *tri_mat;
    const or_material *mats;
    const or_alias_entry *alias;
    const float *pdf_area;
    int32_t n_emit;
    const or_camera *cam;
    const or_render_params *prm;
    double *accum;
    int32_t *vertex_hist;  /* optional: paths reaching each vertex count */
    int64_t *counts;       /* optional: traversal work, see or_render */
} render_ctx;

/* one traversal, its work added to c[0..2] (rays, interior-node visits,
 * triangle tests) when c is given */
static inline void trace_counted(const render_ctx *rc, const or_ray *ray, int any, or_hit *h, uint32_t *occ,
                                 int64_t *c)
{
    if (!c) {
        or_trace_one(rc->bvh, rc->verts, rc->tris, ray, any, h, occ);
        return;
    }
    int32_t nb = 0, nt = 0;
    or_trace_one_counted(rc->bvh, rc->verts, rc->tris, ray, any, h, occ, &nb, &nt);
    c[0] += 1;
    c[1] += nb > 0 ? (nb - 1) / 2 : 0;
    c[2] += nt;
}

/* Pinhole camera ray of a pixel sample (cameraRay,
 * include/tracing/CameraRay.hpp:6-34, with the build's Philox jitter). */
/* pinhole ray through (x + sx, y + sy) of the sensor (cameraRay,
 * include/tracing/CameraRay.hpp:6-34): origin = translation column,
 * direction = rotation x normalize(((p * pixel_size - center) / f, 1)) */
void or_camera_ray(const or_camera *cam, uint32_t x, uint32_t y, float sx, float sy, or_ray *out)
{
    const float pcx = ((float)x + sx) * cam->pixel_size[0] - cam->center[0];
    const float pcy = ((float)y + sy) * cam->pixel_size[1] - cam->center[1];
    const v3 cv = vnorm(mk(pcx / cam->focal_length, pcy / cam->focal_length, 1.0f));
    const float *m = cam->transform;
    or_ray ray;
    ray.ox = m[3]; ray.oy = m[7]; ray.oz = m[11];
    ray.dx = m[0] * cv.x + m[1] * cv.y + m[2] * cv.z;
    ray.dy = m[4] * cv.x + m[5] * cv.y + m[6] * cv.z;
    ray.dz = m[8] * cv.x + m[9] * cv.y + m[10] * cv.z;
    ray.tmin = 0.0f;
    ray.tmax = INFINITY;
    *out = ray;
}

static or_ray camera_ray(const or_camera *cam, uint64_t seed, uint32_t pixel, uint32_t sample)
{
    const uint32_t x = pixel % cam->width, y = pixel / cam->width;
    float sx = 0.5f, sy = 0.5f;
    if (cam->pixel_sampling == 1) {
        uint32_t r[4];
        path_rng(pixel, sample, 0xFFFFFFFFu, 0u, seed, r);
        sx = or_u01(r[0]);
        sy = or_u01(r[1]);
    }
    or_ray ray;
    or_camera_ray(cam, x, y, sx, sy, &ray);
    return ray;
}

void or_path_rng(uint32_t pixel, uint32_t sample, uint32_t vertex, uint32_t stream, uint64_t seed, uint32_t out[4])
{
    path_rng(pixel, sample, vertex, stream, seed, out);
}

void or_camera_sample(const or_camera *cam, uint64_t seed, uint32_t pixel, uint32_t sample, or_ray *out)
{
    *out = camera_ray(cam, seed, pixel, sample);
}

/* One path: returns its radiance. */
static v3 trace_path(const render_ctx *rc, uint32_t pixel, uint32_t sample, int64_t *counts)
{
    const or_render_params *prm = rc->prm;
    or_ray ray = camera_ray(rc->cam, prm->seed, pixel, sample);
    v3 beta = mk(1.0f, 1.0f, 1.0f);
    v3 L = mk(0.0f, 0.0f, 0.0f);
    int spec_prev = 1;
    float pdf_prev = 0.0f;
    for (uint32_t vertex = 0; vertex < prm->max_depth; ++vertex) {
        or_hit h;
        trace_counted(rc, &ray, 0, &h, NULL, counts);
        if (h.id < 0) break;
        if (rc->vertex_hist) __atomic_fetch_add(&rc->vertex_hist[vertex], 1, __ATOMIC_RELAXED);
        const int32_t tri = h.id;
        const v3 A = vtx(rc->verts, rc->tris[3 * (size_t)tri]);
        const v3 B = vtx(rc->verts, rc->tris[3 * (size_t)tri + 1]);
        const v3 C = vtx(rc->verts, rc->tris[3 * (size_t)tri + 2]);
        const or_material mat = rc->mats[rc->tri_mat ? rc->tri_mat[tri] : 0];
        const v3 d = mk(ray.dx, ray.dy, ray.dz);
        const v3 p = bary_point(A, B, C, h.u, h.v);
        const v3 ng = vnorm(vcross(vsub(B, A), vsub(C, A)));
        const int back = vdot(ng, d) > 0.0f;
        const v3 n = back ? vmul(ng, -1.0f) : ng;
        const float eps = origin_eps(A, B, C);
        uint32_t r0[4], r1[4];
        path_rng(pixel, sample, vertex, 0u, prm->seed, r0);
        path_rng(pixel, sample, vertex, 1u, prm->seed, r1);
        v3 nd, no;
        int spec;
        float pdf;
        if (mat.kind == 0) {
            const v3 E = mk(mat.emission[0], mat.emission[1], mat.emission[2]);
            if (E.x != 0.0f || E.y != 0.0f || E.z != 0.0f) {
                float pl = 0.0f;
                if (!spec_prev) {
                    const float cl = fabsf(vdot(d, ng));
                    pl = cl < 1e-6f ? 0.0f : rc->pdf_area[tri] * (h.t * h.t) / cl;
                }
                const float w = or_emission_weight(spec_prev, pdf_prev, pl);
                L = vadd(L, mk(beta.x * E.x * w, beta.y * E.y * w, beta.z * E.z * w));
            }
            const v3 alb = mk(mat.albedo[0], mat.albedo[1], mat.albedo[2]);
            if (rc->n_emit > 0) {
                const uint32_t s = (uint32_t)(((uint64_t)r1[0] * (uint64_t)rc->n_emit) >> 32);
                const or_alias_entry en = rc->alias[s];
                const int take_a = or_u01(r1[1]) < en.p_a;
                const int32_t lt = take_a ? en.a : en.b;
                const float psel = take_a ? en.pdf_a : en.pdf_b;
                if (lt >= 0 && psel > 0.0f) {
                    const v3 LA = vtx(rc->verts, rc->tris[3 * (size_t)lt]);
                    const v3 LB = vtx(rc->verts, rc->tris[3 * (size_t)lt + 1]);
                    const v3 LC = vtx(rc->verts, rc->tris[3 * (size_t)lt + 2]);
                    const v3 xl = triangle_point(LA, LB, LC, or_u01(r1[2]), or_u01(r1[3]));
                    const v3 nl = vnorm(vcross(vsub(LB, LA), vsub(LC, LA)));
                    const float pa = rc->pdf_area[lt];
                    const or_material lm = rc->mats[rc->tri_mat ? rc->tri_mat[lt] : 0];
                    v3 Ld, om;
                    float pn = 0.0f, pb = 0.0f, dist;
                    if (nee_terms(p, n, alb, xl, nl, lm.emission, pa, &Ld, &pn, &pb, &dist, &om)) {
                        const float w = power_heuristic(pn, pb);
                        const float tmax = dist - 2.0f * fmaxf(eps, origin_eps(LA, LB, LC));
                        if (pn > 0.0f && tmax > 0.0f) {
                            const v3 c = mk(beta.x * Ld.x * w, beta.y * Ld.y * w, beta.z * Ld.z * w);
                            const v3 so = vadd(p, vmul(om, eps));
                            or_ray sr = { so.x, so.y, so.z, 0.0f, om.x, om.y, om.z, tmax };
                            uint32_t occ = 0;
                            trace_counted(rc, &sr, 1, NULL, &occ, counts ? counts + 3 : NULL);
                            if (!occ) L = vadd(L, c);
                        }
                    }
                }
            }
            nd = cosine_dir(n, or_u01(r0[0]), or_u01(r0[1]));
            const float cn = vdot(nd, n);
            pdf = cn > 0.0f ? cn / kPi : 0.0f;
            beta = mk(beta.x * alb.x, beta.y * alb.y, beta.z * alb.z);
            no = vadd(p, vmul(n, eps));
            spec = 0;
        } else {
            v3 side;
            float F;
            const int refl = dielectric_scatter(d, n, !back, mat.ior, or_u01(r0[3]), prm->fresnel == 1, &nd, &side, &F);
            if (!refl && (prm->flags & OR_PLANT_TRANSMIT_WEIGHT)) beta = vmul(beta, 1.0f - F);
            no = vadd(p, vmul(side, eps));
            pdf = 0.0f;
            spec = 1;
        }
        if (!roulette(&beta, or_u01(r0[2]), vertex, prm->rr_depth, prm->max_depth,
                      (prm->flags & OR_PLANT_RR_NO_REWEIGHT) != 0))
            break;
        spec_prev = spec;
        pdf_prev = pdf;
        ray.ox = no.x; ray.oy = no.y; ray.oz = no.z; ray.tmin = 0.0f;
        ray.dx = nd.x; ray.dy = nd.y; ray.dz = nd.z; ray.tmax = INFINITY;
    }
    return L;
}

typedef struct { const render_ctx *rc; int32_t begin, end; int64_t counts[6]; } render_job;

static void *render_thread(void *arg)
{
    render_job *j = (render_job *)arg;
    const render_ctx *rc = j->rc;
    for (int32_t pixel = j->begin; pixel < j->end; ++pixel) {
        double *a = rc->accum + 4 * (size_t)pixel;
        for (uint32_t s = 0; s < rc->prm->spp; ++s) {
            const v3 L = trace_path(rc, (uint32_t)pixel, rc->prm->sample_offset + s, rc->counts ? j->counts : NULL);
            a[0] = a[0] + (double)L.x;
            a[1] = a[1] + (double)L.y;
            a[2] = a[2] + (double)L.z;
            a[3] = a[3] + 1.0;
        }
    }
    return NULL;
}

void or_render(const void *bvh, const float *verts, const int32_t *tris, const int32_t *tri_mat,
               const or_material *mats, const or_alias_entry *alias, const float *pdf_area, int32_t n_emit,
               const or_camera *cam, const or_render_params *prm, int32_t pixel_begin, int32_t pixel_end,
               double *accum, int32_t *vertex_hist, int64_t *counts, int32_t threads)
{
    render_ctx rc = { bvh, verts, tris, tri_mat, mats, alias, pdf_area, n_emit, cam, prm, accum, vertex_hist, counts };
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    const int32_t n = pixel_end - pixel_begin;
    pthread_t th[256];
    render_job jobs[256];
    memset(jobs, 0, sizeof(jobs));
    int started = 0;
    /* interleaved row blocks keep the threads' loads even */
    const int32_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        jobs[t].rc = &rc;
        jobs[t].begin = pixel_begin + t * chunk;
        jobs[t].end = pixel_begin + (t + 1) * chunk < pixel_end ? pixel_begin + (t + 1) * chunk : pixel_end;
        if (jobs[t].begin >= jobs[t].end) break;
        pthread_create(&th[t], NULL, render_thread, &jobs[t]);
        ++started;
    }
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
    if (counts)
        for (int t = 0; t < started; ++t)
            for (int k = 0; k < 6; ++k) counts[k] += jobs[t].counts[k];
}

/* ---- display and debug views --------------------------------------------- */

/* The reference's display encoding (Renderer::createPixels,
 * src/Renderer.cpp:51-90): toSRGB(c) = 12.92 c below c = 0.0031308, else
 * 1.055 c^(1/2.4) - 0.055, then round(255 x). The 8-bit value of c is the
 * number of thresholds f^-1((b - 1/2) / 255), b = 1..255, that c reaches. */
void or_display_thresholds(int32_t srgb, double thr[255])
{
    const double knee = 0.0031308 * 12.92;
    for (int b = 1; b <= 255; ++b) {
        const double y = (b - 0.5) / 255.0;
        thr[b - 1] = !srgb ? y : (y < knee ? y / 12.92 : pow((y + 0.055) / 1.055, 2.4));
    }
}

static uint32_t quantize8(double c, const double thr[255])
{
    if (!(c > 0.0)) return 0u;
    if (c >= 1.0) return 255u;
    uint32_t n = 0;
    while (n < 255u && c >= thr[n]) ++n;
    return n;
}

void or_resolve_rgba8(const double *accum, uint32_t width, uint32_t height, int32_t srgb, uint32_t *out)
{
    double thr[255];
    or_display_thresholds(srgb, thr);
    for (uint32_t y = 0; y < height; ++y)
        for (uint32_t x = 0; x < width; ++x) {
            const double *a = accum + 4 * ((size_t)y * width + x);
            uint32_t r = 0, g = 0, b = 0;
            if (a[3] != 0.0) {
                const double inv = 1.0 / a[3];
                r = quantize8(a[0] * inv, thr);
                g = quantize8(a[1] * inv, thr);
                b = quantize8(a[2] * inv, thr);
            }
            out[(size_t)(height - 1 - y) * width + x] = r | (g << 8) | (b << 16) | (255u << 24);
        }
}

/* Debug views of the first hit (sampleColorDebug and the debug branch of
 * sampleColor, include/tracing/PathGeneration.hpp:274-299,377-429). mode 1:
 * UV checker (uv is (0, 0) in the reference: 0.5 x debug colour; misses are
 * not counted), 2: barycentric (1-u-v, u, v), 3: winding ((A-B) x (A-C)) . d
 * > 0 -> (1, 0.73, 0.47) else (0.53, 0.82, 1). Samples added in order. */
void or_debug_aov(const void *bvh, const float *verts, const int32_t *tris, const int32_t *tri_mat,
                  const float *debug_rgba, const or_camera *cam, const or_render_params *prm, int32_t mode,
                  double *accum)
{
    const uint32_t npix = cam->width * cam->height;
    for (uint32_t k = 0; k < prm->spp; ++k)
        for (uint32_t p = 0; p < npix; ++p) {
            or_ray ray = camera_ray(cam, prm->seed, p, prm->sample_offset + k);
            or_hit h;
            or_trace_one(bvh, verts, tris, &ray, 0, &h, NULL);
            float c[3] = {0.0f, 0.0f, 0.0f};
            if (h.id >= 0) {
                const int32_t t = h.id;
                if (mode == 1) {
                    const float *dc = debug_rgba + 4 * (size_t)(tri_mat ? tri_mat[t] : 0);
                    c[0] = 0.5f * dc[0]; c[1] = 0.5f * dc[1]; c[2] = 0.5f * dc[2];
                } else if (mode == 2) {
                    c[0] = 1.0f - h.u - h.v; c[1] = h.u; c[2] = h.v;
                } else {
                    const v3 A = vtx(verts, tris[3 * (size_t)t]);
                    const v3 B = vtx(verts, tris[3 * (size_t)t + 1]);
                    const v3 C = vtx(verts, tris[3 * (size_t)t + 2]);
                    const v3 n = vcross(vsub(A, B), vsub(A, C));
                    const int ccw = n.x * ray.dx + n.y * ray.dy + n.z * ray.dz > 0.0f;
                    c[0] = ccw ? 1.0f : 0.53f; c[1] = ccw ? 0.73f : 0.82f; c[2] = ccw ? 0.47f : 1.0f;
                }
            } else if (mode == 1) {
                continue;
            }
            double *a = accum + 4 * (size_t)p;
            a[0] += (double)c[0]; a[1] += (double)c[1]; a[2] += (double)c[2]; a[3] += 1.0;
        }
}
