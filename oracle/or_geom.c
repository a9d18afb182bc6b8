/*
 * DevilRay CPU oracle: RNG, triangle/box tests, BVHs and batch traversal.
 * TEST INFRASTRUCTURE ONLY -- see oracle.h. Compiled with -ffp-contract=off so
 * that every product and sum below rounds exactly as written.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

/* ------------------------------------------------------------------------ */
/* Philox-4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As */
/* Easy as 1, 2, 3", SC'11. Multipliers and Weyl constants from the paper.  */
/* ------------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

float or_u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

/* ------------------------------------------------------------------------ */
/* Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013).      */
/* ------------------------------------------------------------------------ */
typedef struct {
    float o[3];
    float tmin, tmax;
    int kx, ky, kz;
    float sx, sy, sz;
    float invd[3];
} ray_pre;

static void ray_prepare(const or_ray *r, ray_pre *p)
{
    const float d[3] = { r->dx, r->dy, r->dz };
    p->o[0] = r->ox; p->o[1] = r->oy; p->o[2] = r->oz;
    p->tmin = r->tmin; p->tmax = r->tmax;
    const float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < 0.0f) { const int tmp = kx; kx = ky; ky = tmp; }
    p->kx = kx; p->ky = ky; p->kz = kz;
    p->sx = d[kx] / d[kz];
    p->sy = d[ky] / d[kz];
    p->sz = 1.0f / d[kz];
    for (int a = 0; a < 3; ++a) p->invd[a] = 1.0f / d[a];
}

/* Returns 1 with t, u, v when the line crosses the triangle (det != 0). */
static int tri_wt(const ray_pre *p, const float *v0, const float *v1, const float *v2,
                  float *t, float *u, float *v)
{
    const int kx = p->kx, ky = p->ky, kz = p->kz;
    const float A[3] = { v0[0] - p->o[0], v0[1] - p->o[1], v0[2] - p->o[2] };
    const float B[3] = { v1[0] - p->o[0], v1[1] - p->o[1], v1[2] - p->o[2] };
    const float C[3] = { v2[0] - p->o[0], v2[1] - p->o[1], v2[2] - p->o[2] };
    const float Ax = A[kx] - p->sx * A[kz];
    const float Ay = A[ky] - p->sy * A[kz];
    const float Bx = B[kx] - p->sx * B[kz];
    const float By = B[ky] - p->sy * B[kz];
    const float Cx = C[kx] - p->sx * C[kz];
    const float Cy = C[ky] - p->sy * C[kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        /* products of two floats are exact in double */
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
    const float det = U + V + W;
    if (det == 0.0f || det != det) return 0;
    const float Az = p->sz * A[kz];
    const float Bz = p->sz * B[kz];
    const float Cz = p->sz * C[kz];
    const float T = U * Az + V * Bz + W * Cz;
    const float rcp = 1.0f / det;
    *t = T * rcp;
    *u = V * rcp;
    *v = W * rcp;
    return 1;
}

int or_tri_watertight(const or_ray *r, const float *v0, const float *v1, const float *v2,
                      float *t, float *u, float *v)
{
    ray_pre p;
    ray_prepare(r, &p);
    return tri_wt(&p, v0, v1, v2, t, u, v);
}

/* ------------------------------------------------------------------------ */
/* Reference semantics.                                                     */
/* ------------------------------------------------------------------------ */
static inline float dot3(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void sub3(const float *a, const float *b, float *o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static inline void cross3(const float *a, const float *b, float *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* testTriangleIntersection, include/tracing/IntersectionTestsImpl.hpp:253-297:
 * NaN reject, n_f = (A-B)x(A-C), |n_f|^2 < 1e-14 reject, plane distance along
 * the side-flipped unit normal, d < 1e-7 reject, t = dp/d, barycentrics from
 * the two sub-triangle cross products over |n_f|^2, reject any < 0. */
int or_tri_reference(const float p[3], const float d[3], const float a[3], const float b[3],
                     const float c[3], float *t, float bari[3])
{
    for (int i = 0; i < 3; ++i)
        if (p[i] != p[i] || d[i] != d[i]) return 0;
    float ab[3], ac[3], nf[3];
    sub3(a, b, ab); sub3(a, c, ac);
    cross3(ab, ac, nf);
    const float area2 = dot3(nf, nf);
    if (area2 < 1e-14f) return 0;
    const float len = sqrtf(area2);
    float n[3] = { nf[0] / len, nf[1] / len, nf[2] / len };
    float pa[3]; sub3(p, a, pa);
    float dp = dot3(pa, n);
    if (dp < 0) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; dp = -dp; }
    const float dd = -dot3(d, n);
    if (dd < 1e-7f) return 0;
    const float tt = dp / dd;
    const float hp[3] = { p[0] + d[0] * tt, p[1] + d[1] * tt, p[2] + d[2] * tt };
    float hpa[3], ca[3], ba[3], n1[3], n2[3];
    sub3(hp, a, hpa); sub3(c, a, ca); sub3(b, a, ba);
    cross3(hpa, ca, n1);
    cross3(ba, hpa, n2);
    const float wB = dot3(nf, n1) / area2;
    const float wC = dot3(nf, n2) / area2;
    const float wA = 1 - wB - wC;
    if (wA < 0 || wB < 0 || wC < 0) return 0;
    *t = tt;
    bari[0] = wA; bari[1] = wB; bari[2] = wC;
    return 1;
}

/* findBoxInterval, include/tracing/IntersectionTestsImpl.hpp:80-89. */
void or_box_interval_ref(float lo, float hi, float p, float v, float *enter, float *exit_)
{
    const float t0 = (lo - p) / v;
    const float t1 = (hi - p) / v;
    /* std::min / std::max(a, b) return a when the comparison is false */
    *enter = (t1 < t0) ? t1 : t0;
    *exit_ = (t0 < t1) ? t1 : t0;
}

/* testBoxIntersection, include/tracing/IntersectionTestsImpl.hpp:91-105. */
int or_box_ref(const float lo[3], const float hi[3], const float p[3], const float d[3], float *t)
{
    float en[3], ex[3];
    for (int i = 0; i < 3; ++i) or_box_interval_ref(lo[i], hi[i], p[i], d[i], &en[i], &ex[i]);
    /* std::max(std::max(x, y), z) and std::min(std::min(x, y), z) */
    float enter = (en[0] < en[1]) ? en[1] : en[0];
    enter = (enter < en[2]) ? en[2] : enter;
    float exit_ = (ex[1] < ex[0]) ? ex[1] : ex[0];
    exit_ = (ex[2] < exit_) ? ex[2] : exit_;
    if (exit_ < 0 || enter > exit_) return 0;
    *t = (enter < 0.0f) ? 0.0f : enter;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Threading helper.                                                        */
/* ------------------------------------------------------------------------ */
typedef void (*range_fn)(void *ctx, int32_t begin, int32_t end);
typedef struct { range_fn fn; void *ctx; int32_t begin, end; } range_job;

static void *range_thread(void *arg)
{
    range_job *j = (range_job *)arg;
    j->fn(j->ctx, j->begin, j->end);
    return NULL;
}

static void parallel_for(int32_t n, int32_t threads, range_fn fn, void *ctx)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if (threads == 1 || n < 64) { fn(ctx, 0, n); return; }
    pthread_t th[256];
    range_job jobs[256];
    /* interleave chunks so that uneven rays spread over threads */
    const int32_t chunk = (n + threads - 1) / threads;
    int32_t started = 0;
    for (int i = 0; i < threads; ++i) {
        jobs[i].fn = fn; jobs[i].ctx = ctx;
        jobs[i].begin = i * chunk;
        jobs[i].end = (i + 1) * chunk < n ? (i + 1) * chunk : n;
        if (jobs[i].begin >= jobs[i].end) break;
        pthread_create(&th[i], NULL, range_thread, &jobs[i]);
        ++started;
    }
    for (int i = 0; i < started; ++i) pthread_join(th[i], NULL);
}

/* ------------------------------------------------------------------------ */
/* Canonical binary binned-SAH BVH (build semantics).                       */
/* ------------------------------------------------------------------------ */
typedef struct {
    float lo[3], hi[3];
    int32_t first;   /* interior: index of left child (right = first + 1); leaf: first prim */
    int32_t count;   /* 0 for interior nodes */
} bnode;

typedef struct {
    bnode *nodes;
    int32_t n_nodes, cap;
    int32_t *prim;   /* leaf-ordered triangle ids */
    int32_t ntris;
} obvh;

typedef struct { float lo[3], hi[3]; } box3;

static inline void box_empty(box3 *b)
{
    for (int i = 0; i < 3; ++i) { b->lo[i] = FLT_MAX; b->hi[i] = -FLT_MAX; }
}
static inline void box_grow(box3 *b, const box3 *o)
{
    for (int i = 0; i < 3; ++i) {
        if (o->lo[i] < b->lo[i]) b->lo[i] = o->lo[i];
        if (o->hi[i] > b->hi[i]) b->hi[i] = o->hi[i];
    }
}
static inline double box_half_area(const box3 *b)
{
    const double dx = (double)b->hi[0] - b->lo[0], dy = (double)b->hi[1] - b->lo[1], dz = (double)b->hi[2] - b->lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0;
    return dx * dy + dy * dz + dz * dx;
}

typedef struct {
    obvh *b;
    const box3 *pbox;
    const float *cen;  /* 3 per prim */
    int32_t max_leaf, bins;
    int32_t *tmp;
} build_ctx;

static int32_t alloc_nodes(obvh *b, int32_t k)
{
    if (b->n_nodes + k > b->cap) {
        b->cap = (b->cap * 2 > b->n_nodes + k) ? b->cap * 2 : b->n_nodes + k + 16;
        b->nodes = (bnode *)realloc(b->nodes, sizeof(bnode) * (size_t)b->cap);
    }
    const int32_t at = b->n_nodes;
    b->n_nodes += k;
    return at;
}

/* Outward rounding by one ulp: every vertex coordinate is then strictly
 * inside the box, so the slab test never multiplies 0 by inf on a plane that
 * carries geometry. */
static void box_inflate(box3 *b)
{
    for (int i = 0; i < 3; ++i) {
        b->lo[i] = nextafterf(b->lo[i], -INFINITY);
        b->hi[i] = nextafterf(b->hi[i], INFINITY);
    }
}

static void build_rec(build_ctx *c, int32_t node, int32_t begin, int32_t end)
{
    obvh *b = c->b;
    box3 nb, cb;
    box_empty(&nb); box_empty(&cb);
    for (int32_t i = begin; i < end; ++i) {
        const int32_t p = b->prim[i];
        box_grow(&nb, &c->pbox[p]);
        box3 cp = { { c->cen[3 * p], c->cen[3 * p + 1], c->cen[3 * p + 2] },
                    { c->cen[3 * p], c->cen[3 * p + 1], c->cen[3 * p + 2] } };
        box_grow(&cb, &cp);
    }
    box3 inb = nb; box_inflate(&inb);
    memcpy(b->nodes[node].lo, inb.lo, sizeof(inb.lo));
    memcpy(b->nodes[node].hi, inb.hi, sizeof(inb.hi));
    const int32_t n = end - begin;
    if (n == 1) { b->nodes[node].first = begin; b->nodes[node].count = 1; return; }

    const double parent_area = box_half_area(&nb);
    double best_cost = INFINITY;
    int best_axis = -1, best_split = -1;
    enum { MAXB = 64 };
    int32_t bins = c->bins > MAXB ? MAXB : c->bins;
    for (int a = 0; a < 3; ++a) {
        const float ext = cb.hi[a] - cb.lo[a];
        if (!(ext > 0.0f)) continue;
        box3 bb[MAXB]; int32_t cnt[MAXB];
        for (int k = 0; k < bins; ++k) { box_empty(&bb[k]); cnt[k] = 0; }
        const float scale = (float)bins / ext;
        for (int32_t i = begin; i < end; ++i) {
            const int32_t p = b->prim[i];
            int k = (int)((c->cen[3 * p + a] - cb.lo[a]) * scale);
            if (k >= bins) k = bins - 1;
            if (k < 0) k = 0;
            box_grow(&bb[k], &c->pbox[p]);
            cnt[k]++;
        }
        double rarea[MAXB]; int32_t rcnt[MAXB];
        box3 acc; box_empty(&acc); int32_t ac = 0;
        for (int k = bins - 1; k >= 1; --k) {
            box_grow(&acc, &bb[k]); ac += cnt[k];
            rarea[k] = box_half_area(&acc); rcnt[k] = ac;
        }
        box_empty(&acc); ac = 0;
        for (int k = 0; k < bins - 1; ++k) {
            box_grow(&acc, &bb[k]); ac += cnt[k];
            if (ac == 0 || rcnt[k + 1] == 0) continue;
            const double cost = 1.0 + (box_half_area(&acc) * ac + rarea[k + 1] * rcnt[k + 1]) /
                                      (parent_area > 0 ? parent_area : 1.0);
            if (cost < best_cost) { best_cost = cost; best_axis = a; best_split = k; }
        }
    }
    if (n <= c->max_leaf && (double)n <= best_cost) {
        b->nodes[node].first = begin; b->nodes[node].count = n; return;
    }
    int32_t mid;
    if (best_axis < 0) {
        mid = begin + n / 2;   /* all centroids coincide: split the current order */
    } else {
        /* stable partition by bin < best_split + 1 */
        const float ext = cb.hi[best_axis] - cb.lo[best_axis];
        const float scale = (float)bins / ext;
        int32_t nl = 0, nr = 0;
        for (int32_t i = begin; i < end; ++i) {
            const int32_t p = b->prim[i];
            int k = (int)((c->cen[3 * p + best_axis] - cb.lo[best_axis]) * scale);
            if (k >= bins) k = bins - 1;
            if (k < 0) k = 0;
            if (k <= best_split) b->prim[begin + nl++] = p;
            else c->tmp[nr++] = p;
        }
        memcpy(b->prim + begin + nl, c->tmp, sizeof(int32_t) * (size_t)nr);
        mid = begin + nl;
    }
    const int32_t left = alloc_nodes(b, 2);
    b->nodes[node].first = left;
    b->nodes[node].count = 0;
    build_rec(c, left, begin, mid);
    build_rec(c, left + 1, mid, end);
}

static void tri_box(const float *verts, const int32_t *tris, int32_t p, box3 *bx)
{
    box_empty(bx);
    for (int k = 0; k < 3; ++k) {
        const float *v = verts + 3 * (size_t)tris[3 * (size_t)p + k];
        box3 pv = { { v[0], v[1], v[2] }, { v[0], v[1], v[2] } };
        box_grow(bx, &pv);
    }
}

void *or_bvh_build(const float *verts, int32_t nverts, const int32_t *tris, int32_t ntris,
                   int32_t max_leaf, int32_t bins)
{
    (void)nverts;
    obvh *b = (obvh *)calloc(1, sizeof(obvh));
    b->ntris = ntris;
    b->prim = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ntris > 0 ? ntris : 1));
    box3 *pbox = (box3 *)malloc(sizeof(box3) * (size_t)(ntris > 0 ? ntris : 1));
    float *cen = (float *)malloc(sizeof(float) * 3 * (size_t)(ntris > 0 ? ntris : 1));
    for (int32_t i = 0; i < ntris; ++i) {
        b->prim[i] = i;
        tri_box(verts, tris, i, &pbox[i]);
        for (int a = 0; a < 3; ++a) cen[3 * i + a] = 0.5f * (pbox[i].lo[a] + pbox[i].hi[a]);
    }
    build_ctx c = { b, pbox, cen, max_leaf < 1 ? 1 : max_leaf, bins < 2 ? 2 : bins,
                    (int32_t *)malloc(sizeof(int32_t) * (size_t)(ntris > 0 ? ntris : 1)) };
    if (ntris > 0) {
        alloc_nodes(b, 1);
        build_rec(&c, 0, 0, ntris);
    }
    free(c.tmp); free(pbox); free(cen);
    return b;
}

void or_bvh_free(void *h)
{
    obvh *b = (obvh *)h;
    if (!b) return;
    free(b->nodes); free(b->prim); free(b);
}

int32_t or_bvh_node_count(const void *h) { return ((const obvh *)h)->n_nodes; }

/* SAH cost with C_trav = C_isect = 1, normalised by the root's area:
 * sum over interior nodes of A(n)/A(root) + sum over leaves of N A(l)/A(root). */
double or_bvh_sah_cost(const void *h)
{
    const obvh *b = (const obvh *)h;
    if (b->n_nodes == 0) return 0.0;
    box3 root;
    memcpy(root.lo, b->nodes[0].lo, sizeof(root.lo));
    memcpy(root.hi, b->nodes[0].hi, sizeof(root.hi));
    const double ra = box_half_area(&root);
    double cost = 0.0;
    for (int32_t i = 0; i < b->n_nodes; ++i) {
        box3 bx;
        memcpy(bx.lo, b->nodes[i].lo, sizeof(bx.lo));
        memcpy(bx.hi, b->nodes[i].hi, sizeof(bx.hi));
        const double a = box_half_area(&bx) / ra;
        cost += b->nodes[i].count > 0 ? a * b->nodes[i].count : a;
    }
    return cost;
}

int32_t or_bvh_validate(const void *h, const float *verts, const int32_t *tris)
{
    const obvh *b = (const obvh *)h;
    if (b->ntris == 0) return b->n_nodes == 0 ? 0 : 1;
    char *seen = (char *)calloc((size_t)b->ntris, 1);
    int32_t bad = 0;
    for (int32_t i = 0; i < b->n_nodes && !bad; ++i) {
        const bnode *nd = &b->nodes[i];
        if (nd->count == 0) {
            if (nd->first <= i || nd->first + 1 >= b->n_nodes) { bad = 2; break; }
            for (int k = 0; k < 2; ++k) {
                const bnode *ch = &b->nodes[nd->first + k];
                for (int a = 0; a < 3; ++a)
                    if (ch->lo[a] < nd->lo[a] || ch->hi[a] > nd->hi[a]) bad = 3;
            }
        } else {
            for (int32_t j = nd->first; j < nd->first + nd->count; ++j) {
                if (j < 0 || j >= b->ntris || seen[j]) { bad = 4; break; }
                seen[j] = 1;
                box3 tb; tri_box(verts, tris, b->prim[j], &tb);
                for (int a = 0; a < 3; ++a)
                    if (!(tb.lo[a] > nd->lo[a] && tb.hi[a] < nd->hi[a])) bad = 5;
            }
        }
    }
    for (int32_t j = 0; j < b->ntris && !bad; ++j) if (!seen[j]) bad = 6;
    free(seen);
    return bad;
}

/* Robust slab test (Ize, "Robust BVH Ray Traversal", JCGT 2(2) 2013): the
 * far distance is scaled by 1 + 2*gamma(3) so rounding never loses a box the
 * exact ray touches. NaN slabs (0 * inf on an empty plane) leave the interval
 * unchanged. `tfar_limit` is the current best distance times TBEST_SLACK. */
#define GAMMA3_X2 (2.0f * (3.0f * 0x1p-24f) / (1.0f - 3.0f * 0x1p-24f))
#define TBEST_SLACK (1.0f + 0x1p-18f)

static inline int slab(const float *lo, const float *hi, const ray_pre *p, float tnear0,
                       float tfar_limit, float *tentry)
{
    float tn = tnear0, tf = tfar_limit;
    for (int a = 0; a < 3; ++a) {
        float t0 = (lo[a] - p->o[a]) * p->invd[a];
        float t1 = (hi[a] - p->o[a]) * p->invd[a];
        if (t0 > t1) { const float x = t0; t0 = t1; t1 = x; }
        t1 = t1 * (1.0f + GAMMA3_X2);
        tn = (t0 > tn) ? t0 : tn;
        tf = (t1 < tf) ? t1 : tf;
    }
    *tentry = tn;
    return tn <= tf;
}

/* Candidate acceptance: tmin <= t <= tbest, ties on t go to the lower id. */
static inline int accept_hit(float t, int32_t id, float tmin, float tbest, int32_t best_id)
{
    if (!(t >= tmin)) return 0;
    if (t < tbest) return 1;
    return t == tbest && (uint32_t)id < (uint32_t)best_id;
}

typedef struct {
    const obvh *b;
    const float *verts;
    const int32_t *tris;
    const or_ray *rays;
    int32_t any;
    or_hit *hits;
    uint32_t *occ;
    int32_t *ntest, *ttest;
} trace_ctx;

/* Front-to-back traversal of one ray; counts AABB tests (nt) and triangle
 * tests (tt). */
static void trace_ray(const obvh *b, const float *verts, const int32_t *tris, const or_ray *ray, int any,
                      or_hit *hit_out, uint32_t *occ_out, int32_t *nt_out, int32_t *tt_out)
{
    typedef struct { int32_t node; float t; } entry;
    entry stack[128];
    ray_pre p;
    ray_prepare(ray, &p);
    float tbest = p.tmax;
    int32_t best = -1;
    float bu = 0, bv = 0;
    int32_t nt = 0, tt = 0;
    int occluded = 0;
    int sp = 0;
    if (b->n_nodes > 0) {
        float te;
        ++nt;
        if (slab(b->nodes[0].lo, b->nodes[0].hi, &p, p.tmin, p.tmax * TBEST_SLACK, &te)) {
            stack[sp].node = 0; stack[sp].t = te; ++sp;
        }
    }
    while (sp > 0) {
        --sp;
        const int32_t ni = stack[sp].node;
        if (stack[sp].t > tbest * TBEST_SLACK) continue;
        const bnode *nd = &b->nodes[ni];
        if (nd->count > 0) {
            for (int32_t j = nd->first; j < nd->first + nd->count; ++j) {
                const int32_t id = b->prim[j];
                const int32_t *tri = tris + 3 * (size_t)id;
                float t, u, v;
                ++tt;
                if (!tri_wt(&p, verts + 3 * (size_t)tri[0], verts + 3 * (size_t)tri[1],
                            verts + 3 * (size_t)tri[2], &t, &u, &v)) continue;
                if (any) {
                    if (t >= p.tmin && t <= p.tmax) { occluded = 1; break; }
                } else if (accept_hit(t, id, p.tmin, tbest, best)) {
                    tbest = t; best = id; bu = u; bv = v;
                }
            }
            if (occluded) break;
            continue;
        }
        float t0, t1;
        const float lim = (any ? p.tmax : tbest) * TBEST_SLACK;
        nt += 2;
        const int h0 = slab(b->nodes[nd->first].lo, b->nodes[nd->first].hi, &p, p.tmin, lim, &t0);
        const int h1 = slab(b->nodes[nd->first + 1].lo, b->nodes[nd->first + 1].hi, &p, p.tmin, lim, &t1);
        if (h0 && h1) {
            /* push far first so that the near child is popped next */
            if (t1 < t0) {
                stack[sp].node = nd->first; stack[sp].t = t0; ++sp;
                stack[sp].node = nd->first + 1; stack[sp].t = t1; ++sp;
            } else {
                stack[sp].node = nd->first + 1; stack[sp].t = t1; ++sp;
                stack[sp].node = nd->first; stack[sp].t = t0; ++sp;
            }
        } else if (h0) {
            stack[sp].node = nd->first; stack[sp].t = t0; ++sp;
        } else if (h1) {
            stack[sp].node = nd->first + 1; stack[sp].t = t1; ++sp;
        }
    }
    if (any) {
        if (occ_out) *occ_out = (uint32_t)occluded;
    } else if (hit_out) {
        hit_out->id = best;
        hit_out->t = best >= 0 ? tbest + 0.0f : INFINITY;  /* a zero distance as +0 */
        hit_out->u = best >= 0 ? bu : 0.0f;
        hit_out->v = best >= 0 ? bv : 0.0f;
    }
    if (nt_out) *nt_out = nt;
    if (tt_out) *tt_out = tt;
}

void or_trace_one(const void *bvh, const float *verts, const int32_t *tris, const or_ray *ray, int any,
                  or_hit *hit, uint32_t *occluded)
{
    trace_ray((const obvh *)bvh, verts, tris, ray, any, hit, occluded, NULL, NULL);
}

void or_trace_one_counted(const void *bvh, const float *verts, const int32_t *tris, const or_ray *ray, int any,
                          or_hit *hit, uint32_t *occluded, int32_t *box_tests, int32_t *tri_tests)
{
    trace_ray((const obvh *)bvh, verts, tris, ray, any, hit, occluded, box_tests, tri_tests);
}

static void trace_range(void *vc, int32_t begin, int32_t end)
{
    trace_ctx *c = (trace_ctx *)vc;
    for (int32_t i = begin; i < end; ++i)
        trace_ray(c->b, c->verts, c->tris, &c->rays[i], c->any, c->hits ? &c->hits[i] : NULL,
                  c->occ ? &c->occ[i] : NULL, c->ntest ? &c->ntest[i] : NULL, c->ttest ? &c->ttest[i] : NULL);
}

void or_trace_bvh(const void *bvh, const float *verts, const int32_t *tris, const or_ray *rays,
                  int32_t n, int32_t any_hit, or_hit *hits, uint32_t *occluded,
                  int32_t *node_tests, int32_t *tri_tests, int32_t threads)
{
    trace_ctx c = { (const obvh *)bvh, verts, tris, rays, any_hit, hits, occluded, node_tests, tri_tests };
    parallel_for(n, threads, trace_range, &c);
}

typedef struct {
    const float *verts;
    const int32_t *tris;
    int32_t ntris;
    const or_ray *rays;
    int32_t any;
    or_hit *hits;
    uint32_t *occ;
} brute_ctx;

static void brute_range(void *vc, int32_t begin, int32_t end)
{
    brute_ctx *c = (brute_ctx *)vc;
    for (int32_t i = begin; i < end; ++i) {
        ray_pre p;
        ray_prepare(&c->rays[i], &p);
        float tbest = p.tmax, bu = 0, bv = 0;
        int32_t best = -1;
        int occ = 0;
        for (int32_t id = 0; id < c->ntris; ++id) {
            const int32_t *tri = c->tris + 3 * (size_t)id;
            float t, u, v;
            if (!tri_wt(&p, c->verts + 3 * (size_t)tri[0], c->verts + 3 * (size_t)tri[1],
                        c->verts + 3 * (size_t)tri[2], &t, &u, &v)) continue;
            if (c->any) {
                if (t >= p.tmin && t <= p.tmax) { occ = 1; break; }
            } else if (accept_hit(t, id, p.tmin, tbest, best)) {
                tbest = t; best = id; bu = u; bv = v;
            }
        }
        if (c->any) c->occ[i] = (uint32_t)occ;
        else {
            c->hits[i].id = best;
            c->hits[i].t = best >= 0 ? tbest + 0.0f : INFINITY;  /* a zero distance as +0 */
            c->hits[i].u = best >= 0 ? bu : 0.0f;
            c->hits[i].v = best >= 0 ? bv : 0.0f;
        }
    }
}

void or_trace_brute(const float *verts, const int32_t *tris, int32_t ntris, const or_ray *rays,
                    int32_t n, int32_t any_hit, or_hit *hits, uint32_t *occluded, int32_t threads)
{
    brute_ctx c = { verts, tris, ntris, rays, any_hit, hits, occluded };
    parallel_for(n, threads, brute_range, &c);
}

typedef struct {
    const float *verts;
    const int32_t *tris, *vofs, *tofs, *ntris;
    const or_instance *inst;
    int32_t n_inst;
    const or_ray *rays;
    int32_t any;
    or_hit *hits;
    uint32_t *occ;
} inst_ctx;

/* include/tracing/IntersectionTestsImpl.hpp:143-151 (ray to model space per
 * object) and :234-251 (nearest over all objects) */
static void inst_range(void *vc, int32_t begin, int32_t end)
{
    inst_ctx *c = (inst_ctx *)vc;
    for (int32_t i = begin; i < end; ++i) {
        const or_ray *w = &c->rays[i];
        float tbest = w->tmax, bu = 0, bv = 0;
        int32_t best = -1;
        int occ = 0;
        uint32_t id_off = 0;
        for (int32_t k = 0; k < c->n_inst && !occ; ++k) {
            const or_instance *in = &c->inst[k];
            const float isx = 1.0f / in->scale[0], isy = 1.0f / in->scale[1], isz = 1.0f / in->scale[2];
            const or_ray m = { (w->ox - in->translate[0]) * isx, (w->oy - in->translate[1]) * isy,
                               (w->oz - in->translate[2]) * isz, w->tmin,
                               w->dx * isx, w->dy * isy, w->dz * isz, w->tmax };
            ray_pre p;
            ray_prepare(&m, &p);
            const int32_t s = (int32_t)in->scene;
            const float *vb = c->verts + 3 * (size_t)c->vofs[s];
            const int32_t *tb = c->tris + 3 * (size_t)c->tofs[s];
            for (int32_t j = 0; j < c->ntris[s]; ++j) {
                const int32_t *tri = tb + 3 * (size_t)j;
                float t, u, v;
                if (!tri_wt(&p, vb + 3 * (size_t)tri[0], vb + 3 * (size_t)tri[1], vb + 3 * (size_t)tri[2],
                            &t, &u, &v)) continue;
                const int32_t id = (int32_t)(id_off + (uint32_t)j);
                if (c->any) {
                    if (t >= p.tmin && t <= p.tmax) { occ = 1; break; }
                } else if (accept_hit(t, id, p.tmin, tbest, best)) {
                    tbest = t; best = id; bu = u; bv = v;
                }
            }
            id_off += (uint32_t)c->ntris[s];
        }
        if (c->any) c->occ[i] = (uint32_t)occ;
        else {
            c->hits[i].id = best;
            c->hits[i].t = best >= 0 ? tbest + 0.0f : INFINITY;  /* a zero distance as +0 */
            c->hits[i].u = best >= 0 ? bu : 0.0f;
            c->hits[i].v = best >= 0 ? bv : 0.0f;
        }
    }
}

void or_trace_instances_brute(const float *verts, const int32_t *tris, const int32_t *scene_vofs,
                              const int32_t *scene_tofs, const int32_t *scene_ntris,
                              const or_instance *inst, int32_t n_inst, const or_ray *rays, int32_t n,
                              int32_t any_hit, or_hit *hits, uint32_t *occluded, int32_t threads)
{
    inst_ctx c = { verts, tris, scene_vofs, scene_tofs, scene_ntris, inst, n_inst, rays, any_hit, hits, occluded };
    parallel_for(n, threads, inst_range, &c);
}

/* ------------------------------------------------------------------------ */
/* Reference-semantics BBH: generateSimpleBBH (src/models/BBH.cpp:38-115).   */
/* ------------------------------------------------------------------------ */
typedef struct {
    float lo[3], hi[3];
    int32_t left, right, parent, tb, te;
} rnode;

typedef struct {
    rnode *nodes;
    int32_t n, cap;
    int32_t *order;   /* permuted triangle ids (the reference reorders in place) */
    int32_t depth;
} rbbh;

typedef struct {
    const float *verts;
    const int32_t *tris;
    int axis;
} sort_ctx;

static const sort_ctx *g_sort;   /* qsort has no context argument */
static pthread_mutex_t g_sort_lock = PTHREAD_MUTEX_INITIALIZER;

static float max_extent(const sort_ctx *s, int32_t id)
{
    const int32_t *t = s->tris + 3 * (size_t)id;
    float m = -INFINITY;
    for (int k = 0; k < 3; ++k) {
        const float x = s->verts[3 * (size_t)t[k] + s->axis];
        if (x > m) m = x;
    }
    return m;
}

/* The reference sorts with std::sort (unstable); ties here fall back to the
 * triangle id so the layout is deterministic. Hits differ only on exact ties. */
static int cmp_extent(const void *a, const void *b)
{
    const int32_t ia = *(const int32_t *)a, ib = *(const int32_t *)b;
    const float ea = max_extent(g_sort, ia), eb = max_extent(g_sort, ib);
    if (ea < eb) return -1;
    if (ea > eb) return 1;
    return (ia > ib) - (ia < ib);
}

static int32_t ref_layer(rbbh *b, const float *verts, const int32_t *tris, int32_t tb, int32_t te,
                         int32_t depth, int32_t parent)
{
    if (te - tb < 1) return -1;
    if (b->n == b->cap) {
        b->cap = b->cap ? b->cap * 2 : 64;
        b->nodes = (rnode *)realloc(b->nodes, sizeof(rnode) * (size_t)b->cap);
    }
    const int32_t idx = b->n++;
    rnode nd;
    for (int a = 0; a < 3; ++a) { nd.lo[a] = INFINITY; nd.hi[a] = -INFINITY; }
    for (int32_t i = tb; i < te; ++i)
        for (int k = 0; k < 3; ++k) {
            const float *v = verts + 3 * (size_t)tris[3 * (size_t)b->order[i] + k];
            for (int a = 0; a < 3; ++a) {
                if (v[a] < nd.lo[a]) nd.lo[a] = v[a];
                if (v[a] > nd.hi[a]) nd.hi[a] = v[a];
            }
        }
    nd.left = -1; nd.right = -1; nd.parent = parent; nd.tb = tb; nd.te = te;
    b->nodes[idx] = nd;
    if (depth + 1 > b->depth) b->depth = depth + 1;
    if (te - tb > 1) {
        sort_ctx s = { verts, tris, depth % 3 };
        g_sort = &s;
        qsort(b->order + tb, (size_t)(te - tb), sizeof(int32_t), cmp_extent);
        const int32_t mid = (tb + te) / 2;
        if (mid - tb >= 1) {
            const int32_t l = ref_layer(b, verts, tris, tb, mid, depth + 1, idx);
            b->nodes[idx].left = l;
        }
        if (te - mid >= 1) {
            const int32_t r = ref_layer(b, verts, tris, mid, te, depth + 1, idx);
            b->nodes[idx].right = r;
        }
    }
    return idx;
}

void *or_refbbh_build(const float *verts, int32_t nverts, const int32_t *tris, int32_t ntris)
{
    (void)nverts;
    rbbh *b = (rbbh *)calloc(1, sizeof(rbbh));
    b->order = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ntris > 0 ? ntris : 1));
    for (int32_t i = 0; i < ntris; ++i) b->order[i] = i;
    pthread_mutex_lock(&g_sort_lock);
    ref_layer(b, verts, tris, 0, ntris, 0, -1);
    pthread_mutex_unlock(&g_sort_lock);
    return b;
}

void or_refbbh_free(void *h)
{
    rbbh *b = (rbbh *)h;
    if (!b) return;
    free(b->nodes); free(b->order); free(b);
}

int32_t or_refbbh_depth(const void *h) { return ((const rbbh *)h)->depth; }

int32_t or_refbbh_node_count(const void *h) { return ((const rbbh *)h)->n; }

void or_refbbh_nodes(const void *h, float *lo_hi, int32_t *links)
{
    const rbbh *b = (const rbbh *)h;
    for (int32_t i = 0; i < b->n; ++i) {
        const rnode *nd = b->nodes + i;
        for (int a = 0; a < 3; ++a) {
            lo_hi[6 * (size_t)i + a] = nd->lo[a];
            lo_hi[6 * (size_t)i + 3 + a] = nd->hi[a];
        }
        links[5 * (size_t)i] = nd->left;
        links[5 * (size_t)i + 1] = nd->right;
        links[5 * (size_t)i + 2] = nd->parent;
        links[5 * (size_t)i + 3] = nd->tb;
        links[5 * (size_t)i + 4] = nd->te;
    }
}

void or_refbbh_order(const void *h, int32_t ntris, int32_t *order)
{
    const rbbh *b = (const rbbh *)h;
    for (int32_t i = 0; i < ntris; ++i) order[i] = b->order[i];
}

typedef struct {
    const rbbh *b;
    const float *verts;
    const int32_t *tris;
    const or_ray *rays;
    or_hit *hits;
    int64_t *box_tests, *tri_tests;
} rtrace_ctx;

/* getIntersectionImpl, include/tracing/IntersectionTestsImpl.hpp:141-220:
 * descend to the near child (left is taken when the ray points to +axis on
 * axis depth % 3) and record the branch in bit `depth` of a 32-bit trail;
 * on a miss or a leaf climb through parent links until an unvisited far
 * child appears. Boxes are skipped when their entry distance is not below
 * the best hit. Leaf triangles keep the first strictly nearer hit. */
int or_refbbh_closest_one(const void *bbh, const float *verts, const int32_t *tris, const float p[3],
                          const float d[3], int32_t *id_out, float *t_out, float bari_out[3], int64_t *box_tests,
                          int64_t *tri_tests)
{
    const rbbh *b = (const rbbh *)bbh;
    int has = 0;
    float bt = 0, bb[3] = { 0, 0, 0 };
    int32_t bid = -1;
    int64_t nbox = 0, ntri = 0;
    uint32_t trail = 0, depth = 0;
    int32_t cur = b->n > 0 ? 0 : -1;
    while (cur >= 0 && cur < b->n) {
        const rnode *nd = &b->nodes[cur];
        ++nbox;
        float tb;
        if (or_box_ref(nd->lo, nd->hi, p, d, &tb)) {
            if (!has || tb < bt) {
                if (nd->left == -1 && nd->right == -1) {
                    for (int32_t j = nd->tb; j < nd->te; ++j) {
                        const int32_t id = b->order[j];
                        const int32_t *tri = tris + 3 * (size_t)id;
                        float t, bari[3];
                        ++ntri;
                        if (!or_tri_reference(p, d, verts + 3 * (size_t)tri[0], verts + 3 * (size_t)tri[1],
                                              verts + 3 * (size_t)tri[2], &t, bari)) continue;
                        if (!has || bt > t) {
                            has = 1; bt = t; bid = id;
                            bb[0] = bari[0]; bb[1] = bari[1]; bb[2] = bari[2];
                        }
                    }
                } else {
                    const uint32_t axis = depth % 3;
                    const int right = d[axis] > 0.0f;
                    const int32_t nearc = right ? cur + 1 : nd->right;
                    trail &= ~(1u << depth);
                    cur = nearc;
                    ++depth;
                    continue;
                }
            }
        }
        for (;;) {
            if (cur == 0) { cur = -1; break; }
            const int32_t pi = b->nodes[cur].parent;
            const rnode *pn = &b->nodes[pi];
            --depth;
            if (!(trail & (1u << depth))) {
                const uint32_t axis = depth % 3;
                const int right = d[axis] > 0.0f;
                const int32_t farc = right ? pn->right : pi + 1;
                trail |= (1u << depth);
                cur = farc;
                ++depth;
                break;
            }
            cur = pi;
        }
    }
    *id_out = has ? bid : -1;
    *t_out = has ? bt : INFINITY;
    bari_out[0] = bb[0]; bari_out[1] = bb[1]; bari_out[2] = bb[2];
    if (box_tests) *box_tests += nbox;
    if (tri_tests) *tri_tests += ntri;
    return has;
}

static void rtrace_range(void *vc, int32_t begin, int32_t end)
{
    rtrace_ctx *c = (rtrace_ctx *)vc;
    for (int32_t i = begin; i < end; ++i) {
        const or_ray *r = &c->rays[i];
        const float p[3] = { r->ox, r->oy, r->oz };
        const float d[3] = { r->dx, r->dy, r->dz };
        int32_t id;
        float t, bari[3];
        int64_t nbox = 0, ntri = 0;
        const int has = or_refbbh_closest_one(c->b, c->verts, c->tris, p, d, &id, &t, bari, &nbox, &ntri);
        c->hits[i].id = id;
        c->hits[i].t = t;
        c->hits[i].u = has ? bari[1] : 0.0f;
        c->hits[i].v = has ? bari[2] : 0.0f;
        if (c->box_tests) c->box_tests[i] = nbox;
        if (c->tri_tests) c->tri_tests[i] = ntri;
    }
}

void or_refbbh_trace(const void *bbh, const float *verts, const int32_t *tris, const or_ray *rays,
                     int32_t n, or_hit *hits, int64_t *box_tests, int64_t *tri_tests,
                     int32_t threads)
{
    rtrace_ctx c = { (const rbbh *)bbh, verts, tris, rays, hits, box_tests, tri_tests };
    parallel_for(n, threads, rtrace_range, &c);
}
