// Wavefront path tracer: camera -> [extend -> shade -> shadow -> apply] x
// max_depth -> resolve, all on one stream with device-resident queue counts.
//
// Estimator (the reference's sampleColor, include/tracing/PathGeneration.hpp:
// 363-515, computed incrementally instead of after the whole path): at every
// Lambert vertex, emission weighted by the power heuristic against the light
// pdf (weight 1 after the camera or a specular bounce), then next-event
// estimation with one alias-table light sample and a shadow ray, weighted by
// the power heuristic against the BSDF pdf; then a cosine-weighted bounce.
// Dielectric vertices reflect or refract by Fresnel (exact or Schlick).
// Russian roulette from vertex rr_depth. Paths are independent work items
// (pixel, sample); per-path radiance is resolved into the FP64 accumulation
// buffer in sample order, so results do not depend on queue order.
//
// With a learned guide, Lambert bounces are drawn from the one-sample MIS of
// the guide's vMF mixture (probability alpha, from the network) and the
// cosine BSDF, weighted by f cos / (alpha q + (1 - alpha) cos/pi); next-event
// estimation and emission hits are MIS-weighted against that combined pdf,
// so the estimator stays unbiased for any network output.
#include "dr_internal.hpp"
#include "guide_device.hpp"
#include "shade.hpp"
#include "trace_device.hpp"

#include <algorithm>
#include <string>

namespace dr {

namespace {

constexpr int kT = 256;

struct AliasEntry {
    float p_a, pdf_a, pdf_b;
    int32_t a, b;
};



// Device-resident chunk cursor: every kernel reads the chunk's first item and
// size from here, so one recorded bounce loop serves every chunk.
struct ChunkState {
    uint64_t item0, total;
    uint32_t n, next, cap, pad;
};

struct PathParams {
    const ViewInfo *views;
    const SceneRange *scenes;  // NULL: one scene, root 0, the whole light table
    const ChunkState *chunk;
    uint32_t n_views;
    uint32_t max_depth, rr_depth, schlick;
};

// the last view whose first item is <= item (empty views are skipped)
__device__ __forceinline__ uint32_t find_view(const ViewInfo *__restrict__ v, uint32_t nv, uint64_t item)
{
    uint32_t lo = 0, hi = nv - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) / 2u;
        if (v[mid].item_off <= item) lo = mid;
        else hi = mid - 1u;
    }
    return lo;
}

__device__ __forceinline__ void item_of(const ViewInfo &V, uint64_t item, uint32_t &pixel, uint32_t &sample)
{
    const uint64_t local = item - V.item_off;
    pixel = (uint32_t)(local % V.npix);
    sample = V.sample_offset + (uint32_t)(local / V.npix);
}

__global__ void k_chunk_init(ChunkState *cs, uint64_t total, uint32_t cap)
{
    if (threadIdx.x != 0) return;
    cs->item0 = 0;
    cs->total = total;
    cs->n = 0;
    cs->next = 0;
    cs->cap = cap;
}

__device__ __forceinline__ void next_chunk(ChunkState *cs)
{
    const uint64_t item0 = (uint64_t)cs->next * cs->cap;
    cs->item0 = item0;
    const uint64_t left = item0 < cs->total ? cs->total - item0 : 0ull;
    cs->n = left < cs->cap ? (uint32_t)left : cs->cap;
    cs->next = cs->next + 1u;
}
__global__ void k_next_chunk(ChunkState *cs)
{
    if (threadIdx.x != 0) return;
    next_chunk(cs);
}

__global__ void k_set_view(ViewInfo *dst, ViewInfo v)
{
    if (threadIdx.x == 0) *dst = v;
}

// Each stage is written for one item (path slot, queue slot or shadow ray)
// and run either as its own kernel, one thread per item, or as a
// grid-stride loop of the persistent render kernel.
// The lanes a chunk's paths are dealt to: lane L takes paths [L h, (L + 1) h)
// with h = ceil(n / n_lanes), each into its own first queue and counters.
struct CamLanes {
    uint32_t *list[kMaxRenderLanes];
    uint32_t *counters[kMaxRenderLanes];
    uint32_t n_lanes;
};

__device__ __forceinline__ void camera_item(uint32_t p, const PathParams &pp, dr_ray *__restrict__ rays,
                                            float4 *__restrict__ beta, float4 *__restrict__ rad,
                                            const CamLanes &cl, uint32_t *__restrict__ pending,
                                            uint32_t *__restrict__ pview, uint32_t *__restrict__ proot)
{
    const uint32_t n = pp.chunk->n;
    const uint32_t h = (n + cl.n_lanes - 1u) / cl.n_lanes;
    if (p < cl.n_lanes) {
        const uint32_t lo = min(p * h, n);
        uint32_t *c = cl.counters[p];
        c[0] = min(lo + h, n) - lo;
        c[1] = 0u;
        c[2] = 0u;
        c[3] = 0u;
    }
    if (p >= n) return;
    const uint64_t item = pp.chunk->item0 + p;
    const uint32_t view = find_view(pp.views, pp.n_views, item);
    const ViewInfo &V = pp.views[view];
    const CamParams &cam = V.cam;
    uint32_t pixel, sample;
    item_of(V, item, pixel, sample);
    pview[p] = view;
    if (proot) proot[p] = pp.scenes[V.scene].root;
    rays[p] = camera_ray(cam, pixel, sample, V.seed);
    beta[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);       // w: pdf of the last BSDF sample
    rad[p] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);        // w: 1 = last bounce specular / camera
    cl.list[p / h][p % h] = p;
    if (pending) pending[p] = kGuideNoSample;
}

__global__ __launch_bounds__(kT) void k_camera(PathParams pp, dr_ray *__restrict__ rays,
                                               float4 *__restrict__ beta, float4 *__restrict__ rad, CamLanes cl,
                                               uint32_t *__restrict__ pending, uint32_t *__restrict__ pview,
                                               uint32_t *__restrict__ proot)
{
    camera_item(blockIdx.x * kT + threadIdx.x, pp, rays, beta, rad, cl, pending, pview, proot);
}

// also adds this vertex's extension rays (the input queue) to the scene's
// ray-cast count (atomically: render lanes run on several streams)
__device__ __forceinline__ void reset_counts(uint32_t *out_count, uint32_t *shadow_count, const uint32_t *in_count,
                                             unsigned long long *casts)
{
    *out_count = 0u;
    *shadow_count = 0u;
    atomicAdd(casts, (unsigned long long)*in_count);
}
__global__ void k_reset_counts(uint32_t *out_count, uint32_t *shadow_count, const uint32_t *in_count,
                               unsigned long long *casts)
{
    reset_counts(out_count, shadow_count, in_count, casts);
}

__global__ void k_add_casts(unsigned long long *casts, unsigned long long n, int zero)
{
    *casts = (zero ? 0ull : *casts) + n;
}

struct SceneView {
    const float *verts;
    const uint32_t *idx;
    const uint32_t *tri_mat;
    const dr_material *mats;
    const AliasEntry *alias;
    const float *pdf_area;
    uint32_t n_emit;
    // an instance forest (shaded by shade_item<., true>): hit ids name
    // instanced triangles, the arrays above are the scenes' own, the alias
    // entries name emitters, emit_tri their triangle ids (ascending) and
    // emit_pdf their selection pdf / world area; pdf_area is unused
    const InstShade *inst;
    uint32_t n_inst;
    const uint32_t *emit_tri;
    const float *emit_pdf;
};

// Guide state seen by the shading kernel: head != NULL samples from the
// guide (records written by the inference kernel for this vertex);
// smp != NULL collects training samples, and key[k] / key[cap + k] receive
// the low / high word of slot k's order key.
struct GuideView {
    const float *head;
    dr_guide_sample *smp;
    GuideAux *aux;
    uint32_t *key;
    uint32_t *counter;
    uint32_t *pending;
    uint32_t cap;
    float frac;
    uint32_t item_bits, vertex_bits;
};

__device__ __forceinline__ V3 vtx(const SceneView &sv, uint32_t vi)
{
    return V3{sv.verts[3 * (size_t)vi], sv.verts[3 * (size_t)vi + 1], sv.verts[3 * (size_t)vi + 2]};
}

// World vertices and material id of triangle `tri` (an instanced triangle id
// when INST: found by the instance whose id range holds it, then moved to
// world space as p * s + t, one rounding per operation, as a flattened copy
// of the instance is)
template <bool INST>
__device__ __forceinline__ void tri_world(const SceneView &sv, uint32_t tri, V3 &A, V3 &B, V3 &C, uint32_t &mat)
{
    uint32_t t = tri;
    InstShade r{};
    if (INST) {
        uint32_t lo = 0, hi = sv.n_inst;  // the last instance whose first id is <= tri
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sv.inst[mid].id_delta <= tri) lo = mid; else hi = mid;
        }
        r = sv.inst[lo];
        t = r.tri_off + (tri - r.id_delta);
    }
    A = vtx(sv, sv.idx[3 * (size_t)t]);
    B = vtx(sv, sv.idx[3 * (size_t)t + 1]);
    C = vtx(sv, sv.idx[3 * (size_t)t + 2]);
    if (INST) {
        A = V3{A.x * r.sx + r.tx, A.y * r.sy + r.ty, A.z * r.sz + r.tz};
        B = V3{B.x * r.sx + r.tx, B.y * r.sy + r.ty, B.z * r.sz + r.tz};
        C = V3{C.x * r.sx + r.tx, C.y * r.sy + r.ty, C.z * r.sz + r.tz};
    }
    mat = sv.tri_mat[t];
}

// selection pdf / area of an emitting triangle hit by a BSDF sample (0 for a
// triangle outside the light table)
template <bool INST>
__device__ __forceinline__ float hit_light_pdf(const SceneView &sv, uint32_t tri)
{
    if (!INST) return sv.pdf_area[tri];
    uint32_t lo = 0, hi = sv.n_emit;  // the first emitter whose triangle id is >= tri
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sv.emit_tri[mid] < tri) lo = mid + 1; else hi = mid;
    }
    return (lo < sv.n_emit && sv.emit_tri[lo] == tri) ? sv.emit_pdf[lo] : 0.0f;
}

// GUIDE = false compiles the plain shader (no guide reads, no collection):
// the same arithmetic in fewer registers. INST shades an instance forest.
template <bool GUIDE, bool INST = false>
__device__ __forceinline__ void shade_item(uint32_t slot, const SceneView &sv, const PathParams &pp, uint32_t vertex,
                                           const uint32_t *__restrict__ in_list,
                                           const uint32_t *__restrict__ in_count, uint32_t *__restrict__ out_list,
                                           uint32_t *__restrict__ out_count, dr_ray *__restrict__ rays,
                                           const dr_hit *__restrict__ hits, float4 *__restrict__ beta,
                                           float4 *__restrict__ rad, dr_ray *__restrict__ srays,
                                           float4 *__restrict__ scontrib, uint32_t *__restrict__ scount,
                                           const GuideView &gv, const uint32_t *__restrict__ pview,
                                           uint32_t *__restrict__ sroot)
{
    if (slot >= *in_count) return;
    const uint32_t path = in_list[slot];
    const dr_hit h = hits[path];
    float4 l4 = rad[path];
    if (GUIDE && gv.smp) {
        // the previous bounce's training sample: radiance before this vertex
        const uint32_t pk = gv.pending[path];
        if (pk != kGuideNoSample) {
            GuideAux &x = gv.aux[pk];
            x.lsnap[0] = l4.x;
            x.lsnap[1] = l4.y;
            x.lsnap[2] = l4.z;
            gv.pending[path] = kGuideNoSample;
        }
    }
    if (h.id < 0) return;  // escaped: no environment light
    const ViewInfo &V = pp.views[pview[path]];
    const uint64_t seed = V.seed;
    uint32_t pixel, sample;
    item_of(V, pp.chunk->item0 + path, pixel, sample);
    // the path's scene: light-table range and BVH root for shadow rays
    const SceneRange S = pp.scenes ? pp.scenes[V.scene] : SceneRange{0u, 0u, sv.n_emit, 0u};
    const dr_ray ray = rays[path];
    float4 b4 = beta[path];
    V3 beta3 = V3{b4.x, b4.y, b4.z};
    const bool spec_prev = l4.w != 0.0f;
    const float pdf_prev = b4.w;

    const uint32_t tri = (uint32_t)h.id;
    V3 A, B, C;
    uint32_t mi;
    tri_world<INST>(sv, tri, A, B, C, mi);
    const dr_material mat = sv.mats[mi];
    const V3 d = V3{ray.dx, ray.dy, ray.dz};
    const V3 p = bary_point(A, B, C, h.u, h.v);
    const V3 ng = normalize(cross(sub(B, A), sub(C, A)));
    const bool back = dot(ng, d) > 0.0f;
    const V3 n = back ? mul(ng, -1.0f) : ng;
    const float eps = origin_eps(A, B, C);
    const U4 r0 = path_rng(pixel, sample, vertex, 0u, seed);
    const U4 r1 = path_rng(pixel, sample, vertex, 1u, seed);

    V3 L = V3{l4.x, l4.y, l4.z};
    V3 nd;
    V3 no;
    bool spec;
    float pdf;
    float hd[kHeadFloats];
    float alpha = 0.0f;
    const bool guided = GUIDE && gv.head != nullptr && mat.kind == DR_MAT_LAMBERT;
    if (guided) {
        const float4 *H = reinterpret_cast<const float4 *>(gv.head + (size_t)kHeadFloats * path);
#pragma unroll
        for (int k = 0; k < kHeadFloats / 4; ++k) {
            const float4 v = H[k];
            hd[4 * k] = v.x;
            hd[4 * k + 1] = v.y;
            hd[4 * k + 2] = v.z;
            hd[4 * k + 3] = v.w;
        }
        alpha = hd[40];
    }
    if (mat.kind == DR_MAT_LAMBERT) {
        // emission, MIS against light sampling
        const V3 E = V3{mat.emission[0], mat.emission[1], mat.emission[2]};
        if (E.x != 0.0f || E.y != 0.0f || E.z != 0.0f) {
            float w = 1.0f;
            if (!spec_prev) {
                const float cl = fabsf(dot(d, ng));
                const float pl = cl < 1e-6f ? 0.0f : hit_light_pdf<INST>(sv, tri) * (h.t * h.t) / cl;
                w = power_heuristic(pdf_prev, pl);
            }
            L = add(L, V3{beta3.x * E.x * w, beta3.y * E.y * w, beta3.z * E.z * w});
        }
        // next-event estimation
        const V3 alb = V3{mat.albedo[0], mat.albedo[1], mat.albedo[2]};
        if (S.n_emit > 0) {
            const uint32_t s = (uint32_t)(((uint64_t)r1.x * S.n_emit) >> 32);
            const AliasEntry en = sv.alias[S.alias_off + s];
            const bool take_a = u01(r1.y) < en.p_a;
            const int32_t lt = take_a ? en.a : en.b;
            const float psel = take_a ? en.pdf_a : en.pdf_b;
            if (lt >= 0 && psel > 0.0f) {
                float a1 = u01(r1.z), a2 = u01(r1.w);
                if (a1 + a2 > 1.0f) { a1 = 1.0f - a1; a2 = 1.0f - a2; }
                // (an instance forest's entries name emitters)
                V3 LA, LB, LC;
                uint32_t lmi;
                tri_world<INST>(sv, INST ? sv.emit_tri[lt] : (uint32_t)lt, LA, LB, LC, lmi);
                const V3 x = add(add(LA, mul(sub(LB, LA), a1)), mul(sub(LC, LA), a2));
                const V3 nl = normalize(cross(sub(LB, LA), sub(LC, LA)));
                const float pa = INST ? sv.emit_pdf[lt] : sv.pdf_area[lt];
                const V3 toL = sub(x, p);
                const float dist2 = dot(toL, toL);
                const float dist = sqrtf(dist2);
                const V3 om = divs(toL, dist);
                const float cs = dot(om, n);
                const float acl = fabsf(dot(om, nl));
                if (cs > 0.0f && pa > 0.0f && dist > 0.0f) {
                    const dr_material lm = sv.mats[lmi];
                    const float G = cs * acl / dist2;
                    const float pn = acl < 1e-6f ? 0.0f : pa * dist2 / acl;
                    float pb = cs / kPi;
                    if (guided) pb = alpha * head_mixture(hd, om) + (1.0f - alpha) * pb;
                    const float w = power_heuristic(pn, pb);
                    const float tmax = dist - 2.0f * fmaxf(eps, origin_eps(LA, LB, LC));
                    if (pn > 0.0f && tmax > 0.0f) {
                        const V3 Ld = V3{lm.emission[0] * (alb.x / kPi) * G / pa, lm.emission[1] * (alb.y / kPi) * G / pa,
                                         lm.emission[2] * (alb.z / kPi) * G / pa};
                        const V3 c = V3{beta3.x * Ld.x * w, beta3.y * Ld.y * w, beta3.z * Ld.z * w};
                        const uint32_t k = atomicAdd(scount, 1u);
                        const V3 so = add(p, mul(om, eps));
                        srays[k] = dr_ray{so.x, so.y, so.z, 0.0f, om.x, om.y, om.z, tmax};
                        if (sroot) sroot[k] = S.root;
                        scontrib[k] = make_float4(c.x, c.y, c.z, __int_as_float((int32_t)path));
                    }
                }
            }
        }
        if (guided) {
            // one-sample MIS: a guide lobe with probability alpha, else the BSDF
            const U4 r2 = path_rng(pixel, sample, vertex, 2u, seed);
            if (u01(r2.x) < alpha) {
                const float u = u01(r2.y);
                // the first lobe whose cumulative weight exceeds u (else the
                // last), picked by selects so the head stays in registers
                float cdf = 0.0f;
                bool found = false;
                float mx = hd[5 * (kGuideLobes - 1)], my = hd[5 * (kGuideLobes - 1) + 1];
                float mz = hd[5 * (kGuideLobes - 1) + 2], kr = hd[5 * (kGuideLobes - 1) + 3];
#pragma unroll
                for (int k = 0; k < kGuideLobes - 1; ++k) {
                    cdf += hd[5 * k + 4];
                    const bool take = u < cdf && !found;
                    mx = take ? hd[5 * k] : mx;
                    my = take ? hd[5 * k + 1] : my;
                    mz = take ? hd[5 * k + 2] : mz;
                    kr = take ? hd[5 * k + 3] : kr;
                    found = found || take;
                }
                nd = vmf_sample(V3{mx, my, mz}, kr, u01(r2.z), u01(r2.w));
            } else {
                nd = cosine_dir(n, u01(r0.x), u01(r0.y));
            }
            const float cn = dot(nd, n);
            const float pbsdf = cn > 0.0f ? cn / kPi : 0.0f;
            pdf = alpha * head_mixture(hd, nd) + (1.0f - alpha) * pbsdf;
            const float fw = (cn > 0.0f && pdf > 0.0f) ? pbsdf / pdf : 0.0f;
            beta3 = V3{beta3.x * alb.x * fw, beta3.y * alb.y * fw, beta3.z * alb.z * fw};
        } else {
            // cosine-weighted bounce
            nd = cosine_dir(n, u01(r0.x), u01(r0.y));
            const float cn = dot(nd, n);
            pdf = cn > 0.0f ? cn / kPi : 0.0f;
            beta3 = V3{beta3.x * alb.x, beta3.y * alb.y, beta3.z * alb.z};
        }
        no = add(p, mul(n, eps));
        spec = false;
    } else {
        // dielectric: n faces the incoming ray; ng (winding) points outside
        const bool entering = !back;
        const float n1 = entering ? 1.0f : mat.ior;
        const float n2 = entering ? mat.ior : 1.0f;
        const float dd = dot(d, n);
        const float eta = n1 / n2;
        const float k = 1.0f - eta * eta * (1.0f - dd * dd);
        bool reflect_ray = true;
        if (!(n1 > n2 && k < 0.0f)) {
            const float ci = -dd;
            const float F = fresnel(ci, n1, n2, sqrtf(fmaxf(k, 0.0f)), pp.schlick != 0u);
            reflect_ray = u01(r0.w) < F;
        }
        if (reflect_ray) {
            nd = V3{d.x - n.x * 2.0f * dd, d.y - n.y * 2.0f * dd, d.z - n.z * 2.0f * dd};
            no = add(p, mul(n, eps));
        } else {
            const V3 nn = mul(n, -1.0f);
            const float ci = -dd;
            const float sk = sqrtf(k);
            nd = add(mul(nn, sk - ci * eta), mul(d, eta));
            no = add(p, mul(nn, eps));
        }
        pdf = 0.0f;
        spec = true;
    }
    rad[path] = make_float4(L.x, L.y, L.z, spec ? 1.0f : 0.0f);
    // continue?
    bool alive = vertex + 1u < pp.max_depth;
    if (alive && vertex + 1u >= pp.rr_depth) {
        const float q = fmaxf(beta3.x, fmaxf(beta3.y, beta3.z));
        if (q < 1.0f) {
            if (u01(r0.z) >= q) alive = false;
            else beta3 = V3{beta3.x / q, beta3.y / q, beta3.z / q};
        }
    }
    if (!(beta3.x > 0.0f || beta3.y > 0.0f || beta3.z > 0.0f)) alive = false;
    if (!alive) return;
    if (GUIDE && gv.smp && !spec && u01(r0.w) < gv.frac) {
        const uint32_t k = atomicAdd(gv.counter, 1u);
        if (k < gv.cap) {
            dr_guide_sample q;
            q.pos[0] = p.x; q.pos[1] = p.y; q.pos[2] = p.z; q.weight = 0.0f;
            q.normal[0] = n.x; q.normal[1] = n.y; q.normal[2] = n.z; q.pdf = pdf;
            q.wo[0] = -d.x; q.wo[1] = -d.y; q.wo[2] = -d.z; q.pad1 = 0.0f;
            q.wi[0] = nd.x; q.wi[1] = nd.y; q.wi[2] = nd.z; q.pad2 = 0.0f;
            gv.smp[k] = q;
            GuideAux x;
            x.path = path;
            x.pdf = pdf;
            x.beta[0] = beta3.x; x.beta[1] = beta3.y; x.beta[2] = beta3.z;
            x.lsnap[0] = x.lsnap[1] = x.lsnap[2] = 0.0f;
            gv.aux[k] = x;
            // (path item, vertex) names the sample whatever slot it took
            const uint64_t key = guide_sample_key(pp.chunk->item0 + path, vertex, gv.item_bits, gv.vertex_bits);
            gv.key[k] = (uint32_t)key;
            gv.key[gv.cap + k] = (uint32_t)(key >> 32);
            gv.pending[path] = k;
        }
    }
    beta[path] = make_float4(beta3.x, beta3.y, beta3.z, pdf);
    rays[path] = dr_ray{no.x, no.y, no.z, 0.0f, nd.x, nd.y, nd.z, INFINITY};
    out_list[atomicAdd(out_count, 1u)] = path;
}

template <bool GUIDE, bool INST = false>
__global__ __launch_bounds__(kT) void k_shade(SceneView sv, PathParams pp, uint32_t vertex,
                                              const uint32_t *__restrict__ in_list,
                                              const uint32_t *__restrict__ in_count, uint32_t *__restrict__ out_list,
                                              uint32_t *__restrict__ out_count, dr_ray *__restrict__ rays,
                                              const dr_hit *__restrict__ hits, float4 *__restrict__ beta,
                                              float4 *__restrict__ rad, dr_ray *__restrict__ srays,
                                              float4 *__restrict__ scontrib, uint32_t *__restrict__ scount,
                                              GuideView gv, const uint32_t *__restrict__ pview,
                                              uint32_t *__restrict__ sroot)
{
    shade_item<GUIDE, INST>(blockIdx.x * kT + threadIdx.x, sv, pp, vertex, in_list, in_count, out_list, out_count,
                            rays, hits, beta, rad, srays, scontrib, scount, gv, pview, sroot);
}

__device__ __forceinline__ void shadow_apply_item(uint32_t i, const uint32_t *__restrict__ scount,
                                                  const uint32_t *__restrict__ occ,
                                                  const float4 *__restrict__ scontrib, float4 *__restrict__ rad)
{
    if (i >= *scount || occ[i]) return;
    const float4 c = scontrib[i];
    const uint32_t path = (uint32_t)__float_as_int(c.w);
    float4 l = rad[path];
    l.x = l.x + c.x;
    l.y = l.y + c.y;
    l.z = l.z + c.z;
    rad[path] = l;
}
__global__ __launch_bounds__(kT) void k_shadow_apply(const uint32_t *__restrict__ scount,
                                                     const uint32_t *__restrict__ occ,
                                                     const float4 *__restrict__ scontrib, float4 *__restrict__ rad)
{
    shadow_apply_item(blockIdx.x * kT + threadIdx.x, scount, occ, scontrib, rad);
}

// Lane p resolves path p if it is the chunk's first sample of its pixel and
// adds that pixel's samples of the chunk in increasing sample order (the
// next sample of a pixel is npix items later, within the same view).
__device__ __forceinline__ void resolve_item(uint32_t p, const PathParams &pp, const float4 *__restrict__ rad,
                                             const uint32_t *__restrict__ pview)
{
    const uint32_t n = pp.chunk->n;
    if (p >= n) return;
    const ViewInfo &V = pp.views[pview[p]];
    const uint64_t local = pp.chunk->item0 + p - V.item_off;
    if (local >= V.npix && p >= V.npix) return;  // an earlier sample of this pixel is in the chunk
    const uint64_t view_end = (uint64_t)V.npix * V.spp;
    double *a = V.accum + 4 * (size_t)(local % V.npix);
    double r = a[0], g = a[1], b = a[2], w = a[3];
    uint64_t l = local;
    for (uint32_t path = p; path < n && l < view_end; path += V.npix, l += V.npix) {
        const float4 c = rad[path];
        r = r + (double)c.x;
        g = g + (double)c.y;
        b = b + (double)c.z;
        w = w + 1.0;
    }
    a[0] = r; a[1] = g; a[2] = b; a[3] = w;
}
__global__ __launch_bounds__(kT) void k_resolve(PathParams pp, const float4 *__restrict__ rad,
                                                const uint32_t *__restrict__ pview)
{
    resolve_item(blockIdx.x * kT + threadIdx.x, pp, rad, pview);
}

inline uint32_t blocks(uint32_t n) { return (n + kT - 1) / kT; }

// One render lane's view of its buffers (RenderLane; lane 0 holds the
// scene's rs_* queues) and the stream its launches go to.
struct LaneView {
    uint32_t *list_a, *list_b, *cnt;
    dr_ray *srays;
    uint32_t *socc;
    float4 *scon;
    uint32_t *sroot;
    DevBuf *ovf;
    hipStream_t stream;
};

// The bounce loop of one lane's share of a chunk: vertices 0 .. max_depth - 1
// of its paths, every queue of at most `cap` entries.
dr_status bounce_loop(const RenderJob &j, const PathParams &pp, const SceneView &sv, GuideView gv,
                      const LaneView &L, uint32_t cap)
{
    Scene *s = j.s;
    const hipStream_t stream = L.stream;
    const bool use_guide = (j.guide_flags & DR_GUIDE_USE) != 0;
    uint32_t *cnt = L.cnt;
    auto *rays = s->rs_rays.as<dr_ray>();
    auto *hits = s->rs_hits.as<dr_hit>();
    auto *beta = s->rs_beta.as<float4>();
    auto *rad = s->rs_radiance.as<float4>();
    auto *pview = s->rs_view.as<uint32_t>();
    uint32_t *proot = j.scenes ? s->rs_root.as<uint32_t>() : nullptr;
    // With a per-ray schedule, the shadow rays of vertex v - 1 are traced in
    // the same launch as the extension rays of vertex v, and applied before
    // vertex v is shaded: each radiance gets its terms in the same order as
    // with separate launches, and two short queues fill one grid.
    const bool paired = s->n_emit && trace_pair_fits(s, j.sched);
    auto trace_shadows = [&]() -> dr_status {
        if (dr_status st = launch_trace_ex(s, L.srays, nullptr, cap, cnt + 2, nullptr, L.socc, nullptr, nullptr, true,
                                           j.sched, stream, L.sroot, L.ovf))
            return st;
        k_shadow_apply<<<blocks(cap), kT, 0, stream>>>(cnt + 2, L.socc, L.scon, rad);
        DR_HIP_TRY(hipGetLastError());
        return DR_OK;
    };
    for (uint32_t v = 0; v < j.prm->max_depth; ++v) {
        uint32_t *lin = (v & 1u) ? L.list_b : L.list_a;
        uint32_t *lout = (v & 1u) ? L.list_a : L.list_b;
        uint32_t *cin = cnt + (v & 1u);
        uint32_t *cout = cnt + ((v + 1u) & 1u);
        const bool pair_v = paired && v > 0;
        if (pair_v) {
            if (dr_status st = launch_trace_pair(s, L.srays, cnt + 2, L.socc, L.sroot, rays, lin, cin, hits, proot,
                                                 cap, j.sched, stream, L.ovf))
                return st;
        } else if (dr_status st = launch_trace_ex(s, rays, lin, cap, cin, hits, nullptr, nullptr, nullptr, false,
                                                  j.sched, stream, proot, L.ovf)) {
            return st;
        }
        const bool guided_v = use_guide && v < j.guide->guided_vertices;
        if (guided_v)
            if (dr_status st = guide_infer_paths(j.guide, s, rays, hits, lin, cin, cap, s->rs_guide_head.as<float>(),
                                                 stream))
                return st;
        gv.head = guided_v ? s->rs_guide_head.as<float>() : nullptr;
        if (pair_v) {
            k_shadow_apply<<<blocks(cap), kT, 0, stream>>>(cnt + 2, L.socc, L.scon, rad);
            DR_HIP_TRY(hipGetLastError());
        }
        k_reset_counts<<<1, 1, 0, stream>>>(cout, cnt + 2, cin, s->ray_casts.as<unsigned long long>());
        const auto shade = sv.inst ? k_shade<false, true> : (gv.head || gv.smp) ? k_shade<true> : k_shade<false>;
        shade<<<blocks(cap), kT, 0, stream>>>(sv, pp, v, lin, cin, lout, cout, rays, hits, beta, rad, L.srays,
                                              L.scon, cnt + 2, gv, pview, L.sroot);
        DR_HIP_TRY(hipGetLastError());
        if (s->n_emit && !paired)
            if (dr_status st = trace_shadows()) return st;
    }
    // the last vertex's shadow rays
    if (paired && j.prm->max_depth > 0)
        if (dr_status st = trace_shadows()) return st;
    return DR_OK;
}

// One chunk: camera rays, each lane's bounce loop on its own stream (lanes
// 1.. start after the camera rays and end before the resolve, through
// events), then the resolve. Every argument is fixed for the whole call
// (chunk position and queue counts live in device memory), so with one lane
// the same launches can be recorded once and replayed for every chunk.
dr_status run_chunk(const RenderJob &j, const PathParams &pp, const SceneView &sv, const GuideView &gv,
                    hipStream_t stream, uint32_t n_lanes)
{
    Scene *s = j.s;
    const uint32_t cap = j.cap;
    const bool collect = (j.guide_flags & DR_GUIDE_COLLECT) != 0;
    const uint32_t lane_cap = (cap + n_lanes - 1u) / n_lanes;
    LaneView lv[kMaxRenderLanes];
    CamLanes cl{};
    cl.n_lanes = n_lanes;
    for (uint32_t k = 0; k < n_lanes; ++k) {
        LaneView &L = lv[k];
        if (k == 0) {
            L = LaneView{s->rs_list_a.as<uint32_t>(), s->rs_list_b.as<uint32_t>(), s->rs_counters.as<uint32_t>(),
                         s->rs_shadow_rays.as<dr_ray>(), s->rs_shadow_occ.as<uint32_t>(),
                         s->rs_shadow_contrib.as<float4>(), j.scenes ? s->rs_sroot.as<uint32_t>() : nullptr,
                         nullptr, stream};
        } else {
            RenderLane &R = s->lanes[k];
            L = LaneView{R.list_a.as<uint32_t>(), R.list_b.as<uint32_t>(), R.counters.as<uint32_t>(),
                         R.srays.as<dr_ray>(), R.socc.as<uint32_t>(), R.scon.as<float4>(),
                         j.scenes ? R.sroot.as<uint32_t>() : nullptr, &R.ovf, R.stream};
        }
        cl.list[k] = L.list_a;
        cl.counters[k] = L.cnt;
    }
    auto *rad = s->rs_radiance.as<float4>();
    auto *pview = s->rs_view.as<uint32_t>();
    k_next_chunk<<<1, 64, 0, stream>>>(s->rs_chunk.as<ChunkState>());
    DR_HIP_TRY(hipGetLastError());
    k_camera<<<blocks(cap), kT, 0, stream>>>(pp, s->rs_rays.as<dr_ray>(), s->rs_beta.as<float4>(), rad, cl,
                                             collect ? gv.pending : nullptr, pview,
                                             j.scenes ? s->rs_root.as<uint32_t>() : nullptr);
    DR_HIP_TRY(hipGetLastError());
    if (n_lanes > 1) {
        DR_HIP_TRY(hipEventRecord(s->lane_fork, stream));
        for (uint32_t k = 1; k < n_lanes; ++k) DR_HIP_TRY(hipStreamWaitEvent(lv[k].stream, s->lane_fork, 0));
    }
    for (uint32_t k = 0; k < n_lanes; ++k)
        if (dr_status st = bounce_loop(j, pp, sv, gv, lv[k], lane_cap)) return st;
    for (uint32_t k = 1; k < n_lanes; ++k) {
        DR_HIP_TRY(hipEventRecord(s->lanes[k].done, lv[k].stream));
        DR_HIP_TRY(hipStreamWaitEvent(stream, s->lanes[k].done, 0));
    }
    if (collect)
        if (dr_status st = guide_finalize(j.guide, rad, stream)) return st;
    k_resolve<<<blocks(cap), kT, 0, stream>>>(pp, rad, pview);
    DR_HIP_TRY(hipGetLastError());
    return DR_OK;
}

// ---- one persistent launch ------------------------------------------------
// The whole render (every chunk, its camera rays and every bounce) in one
// launch of resident blocks: each wavefront stage is a grid-stride loop (the
// traversal stages its persistent schedule) and the stages are separated by
// grid-wide barriers. The stages and their order are those of the launch
// sequence, so the images are the same.
//
// Grid state: [0] top-level arrivals, [1] barrier generation, [2] abort
// flag; [4, 12) the closest-hit stage's ray counters, [12, 20) the any-hit
// stage's; from word 32 the barrier groups, 16 words (one 64-byte line) each:
// [0] arrivals, [1] generation.
constexpr uint32_t kBarrierGroups = 32;
constexpr uint32_t kGridWords = 32 + 16 * kBarrierGroups;
// barrier wait bound: about 2^23 x 512 cycles, a few seconds
constexpr uint32_t kBarrierSpins = 1u << 23;

struct PersistArgs {
    PathParams pp;
    SceneView sv;
    dr_ray *rays;
    dr_hit *hits;
    float4 *beta, *rad;
    uint32_t *list_a, *list_b, *cnt;
    dr_ray *srays;
    uint32_t *socc;
    float4 *scon;
    uint32_t *pview, *proot, *sroot;
    unsigned long long *casts;
    ChunkState *chunk;
    const QNode *nodes;
    const float4 *tri_v;
    int32_t *ovf;
    uint32_t *flags;
    uint32_t *grid;
    uint32_t refill_min, slices, max_stack, cap, has_emit;
};

__device__ __forceinline__ uint32_t load_agent(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until *gen differs from `seen`; bounded: on time-out or an abort by
// another block it sets the abort flag and returns false.
__device__ __forceinline__ bool wait_generation(const uint32_t *gen, uint32_t seen, uint32_t *abort_flag)
{
    uint32_t spins = 0;
    while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == seen) {
        if (load_agent(abort_flag) != 0u || ++spins > kBarrierSpins) {
            atomicOr(abort_flag, 1u);
            return false;
        }
        __builtin_amdgcn_s_sleep(8);
    }
    return true;
}

// Arrival at a counter of `size` participants: true for the last one, which
// resets the counter; the others see false. `gen` is the generation read
// before arriving.
__device__ __forceinline__ bool arrive(uint32_t *count, uint32_t size)
{
    if (atomicAdd(count, 1u) + 1u != size) return false;
    __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// Grid-wide barrier for a launch whose blocks are all resident, in two
// levels: block b arrives at group b % 32 (at most ceil(grid / 32)
// arrivals on one cache line), the last block of each group at the top
// counter, and the release runs back down through the groups' generations,
// so no address sees more than a few dozen atomics or pollers. One thread
// per block arrives and fences (the block's barriers order its other threads'
// accesses around it). Every wait is bounded: a block that waits too long,
// or sees the abort flag, sets the flag and the barrier returns false (the
// same value for every thread of the block), so a launch whose blocks
// disagree ends instead of hanging.
__device__ bool grid_barrier(uint32_t *g)
{
    __shared__ uint32_t ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();  // this block's writes, device-wide
        const uint32_t nb = gridDim.x;
        const uint32_t ng = nb < kBarrierGroups ? nb : kBarrierGroups;
        const uint32_t k = blockIdx.x % kBarrierGroups;
        const uint32_t gsize = (nb - k + kBarrierGroups - 1u) / kBarrierGroups;
        uint32_t *grp = g + 32 + 16 * k;
        uint32_t good = 1u;
        const uint32_t ggen = __hip_atomic_load(grp + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (arrive(grp, gsize)) {
            const uint32_t tgen = __hip_atomic_load(g + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            if (arrive(g, ng))
                __hip_atomic_store(g + 1, tgen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            else
                good = wait_generation(g + 1, tgen, g + 2) ? 1u : 0u;
            __hip_atomic_store(grp + 1, ggen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            good = wait_generation(grp + 1, ggen, g + 2) ? 1u : 0u;
        }
        if (load_agent(g + 2) != 0u) good = 0u;
        __threadfence();  // the other blocks' writes
        ok = good;
    }
    __syncthreads();
    return ok != 0u;
}

// Block size of the persistent launch. The traversal stages treat each
// kBlock-thread part of a block as a block of their own. A grid barrier
// costs about 0.13 us per resident block (each block's agent-scope fences),
// so waves are cheapest in few large blocks: on the many-lights view at
// 1920 x 1080 x 2 spp, one 512-thread block per CU takes 14.3 ms, against
// 17.3 ms for two 256-thread blocks (the same waves) and 14.5 ms for one
// 1024-thread block (occupancy 4 waves per SIMD gains nothing over 2).
constexpr int kPersistThreads = 512;
constexpr int kPersistParts = kPersistThreads / kBlock;

template <bool SPEC>
__global__ __launch_bounds__(kPersistThreads) void k_render_persistent(PersistArgs a)
{
    __shared__ int32_t stacks[kPersistParts][kLdsStack][kBlock];
    const uint32_t gtid = blockIdx.x * kPersistThreads + threadIdx.x, nthr = gridDim.x * kPersistThreads;
    const uint32_t part = threadIdx.x / kBlock, ptid = threadIdx.x % kBlock;
    const uint32_t vbid = blockIdx.x * kPersistParts + part, vnblk = gridDim.x * kPersistParts;
    int32_t(*stack)[kBlock] = stacks[part];
    uint32_t *ctr_closest = a.grid + 4, *ctr_any = a.grid + 12;
    const GuideView gv{};
    for (;;) {
        if (gtid == 0) {
            next_chunk(a.chunk);
            for (int k = 0; k < 8; ++k) {
                ctr_any[k] = 0u;
                ctr_closest[k] = 0u;
            }
        }
        if (!grid_barrier(a.grid)) return;
        const uint32_t n = load_agent(&a.chunk->n);
        if (n == 0u) return;  // every block reads the same count
        for (uint32_t p = gtid; p < n; p += nthr)
            camera_item(p, a.pp, a.rays, a.beta, a.rad, CamLanes{{a.list_a}, {a.cnt}, 1u}, nullptr, a.pview,
                        a.proot);
        if (!grid_barrier(a.grid)) return;
        // Vertex v's stage traces the shadow rays of vertex v - 1 and the
        // extension rays of vertex v (each block drains the shadow queue's
        // share first, then takes extension rays), applies the shadow terms,
        // then shades vertex v; a radiance gets its terms in the order of the
        // launch sequence. Shadow queue counts alternate between words 2 and
        // 3 by vertex parity, so the shade's count restarts while the
        // previous vertex's is still read. Vertex max_depth only traces and
        // applies the last shadow rays.
        for (uint32_t v = 0; v <= a.pp.max_depth; ++v) {
            uint32_t *lin = (v & 1u) ? a.list_b : a.list_a;
            uint32_t *lout = (v & 1u) ? a.list_a : a.list_b;
            uint32_t *cin = a.cnt + (v & 1u);
            uint32_t *cout = a.cnt + ((v + 1u) & 1u);
            uint32_t *scur = a.cnt + 2u + (v & 1u), *sprev = a.cnt + 2u + ((v + 1u) & 1u);
            // the same counts in every block (read after a barrier)
            const uint32_t nin = v < a.pp.max_depth ? load_agent(cin) : 0u;
            const uint32_t ns = (v > 0u && a.has_emit) ? load_agent(sprev) : 0u;
            if (nin == 0u && ns == 0u) break;  // no path and no shadow ray left in this chunk
            // the output queue and this vertex's shadow queue restart (the
            // former was the previous vertex's input, the latter held the
            // shadow rays of vertex v - 2, applied at v - 1); the extension
            // rays traced here count as ray casts
            if (gtid == 0) {
                *cout = 0u;
                *scur = 0u;
                *a.casts += nin;
            }
            if (ns)
                trace_body<true, false, true, SPEC, false>(stack, ptid, vbid, vnblk, a.nodes, a.tri_v,
                                                           reinterpret_cast<const float4 *>(a.srays), nullptr, a.cap,
                                                           sprev, nullptr, a.socc, nullptr, nullptr, a.ovf, a.flags,
                                                           ctr_any, a.refill_min, a.sroot, a.slices, a.max_stack,
                                                           nullptr);
            if (nin)
                trace_body<false, false, true, SPEC, false>(stack, ptid, vbid, vnblk, a.nodes, a.tri_v,
                                                            reinterpret_cast<const float4 *>(a.rays), lin, a.cap, cin,
                                                            a.hits, nullptr, nullptr, nullptr, a.ovf, a.flags,
                                                            ctr_closest, a.refill_min, a.proot, a.slices, a.max_stack,
                                                            nullptr);
            if (!grid_barrier(a.grid)) return;
            if (gtid == 0)
                for (int k = 0; k < 8; ++k) {
                    ctr_any[k] = 0u;
                    ctr_closest[k] = 0u;
                }
            if (ns) {
                for (uint32_t i = gtid; i < ns; i += nthr) shadow_apply_item(i, sprev, a.socc, a.scon, a.rad);
                if (!grid_barrier(a.grid)) return;
            }
            if (nin == 0u) break;
            for (uint32_t slot = gtid; slot < nin; slot += nthr)
                shade_item<false>(slot, a.sv, a.pp, v, lin, cin, lout, cout, a.rays, a.hits, a.beta, a.rad, a.srays,
                                  a.scon, scur, gv, a.pview, a.sroot);
            if (!grid_barrier(a.grid)) return;
        }
        for (uint32_t p = gtid; p < n; p += nthr) resolve_item(p, a.pp, a.rad, a.pview);
        // the next chunk rewrites the path state
        if (!grid_barrier(a.grid)) return;
    }
}

__global__ void k_grid_init(uint32_t *g)
{
    for (uint32_t k = threadIdx.x; k < kGridWords; k += blockDim.x) g[k] = 0u;
}

// One cooperative launch of resident blocks for the whole job (plain renders;
// the traversal runs its persistent schedule, speculative or not as j.sched
// says). Waits for the launch and reports an aborted barrier as an error.
dr_status run_persistent(const RenderJob &j, const PathParams &pp, const SceneView &sv, hipStream_t stream)
{
    Scene *s = j.s;
    if (s->depth > kMaxWideDepth)
        return fail(DR_E_OVERFLOW, "BVH deeper than the traversal stack bound of " + std::to_string(kMaxWideDepth) +
                                       " wide levels");
    const bool spec = (j.sched & 1u) != 0u;
    const void *kern = spec ? reinterpret_cast<const void *>(&k_render_persistent<true>)
                            : reinterpret_cast<const void *>(&k_render_persistent<false>);
    int occ = 0;
    DR_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kPersistThreads, 0));
    if (occ < 1) return fail(DR_E_DEVICE, "persistent render kernel does not fit on a CU");
    const int bpc = j.blocks_per_cu ? std::min<int>((int)j.blocks_per_cu, occ) : occ;
    const uint32_t grid = (uint32_t)(s->ctx->cu_count * bpc);
    const uint32_t max_stack = std::max<uint32_t>((uint32_t)kLdsStack, 3u * s->depth + 1u);
    if (dr_status st = s->ovf_stack.ensure(sizeof(int32_t) * (size_t)(max_stack - kLdsStack) * grid * kPersistThreads + 4))
        return st;
    if (dr_status st = s->rs_grid.ensure(sizeof(uint32_t) * kGridWords)) return st;
    k_grid_init<<<1, 64, 0, stream>>>(s->rs_grid.as<uint32_t>());
    DR_HIP_TRY(hipGetLastError());
    PersistArgs a;
    a.pp = pp;
    a.sv = sv;
    a.rays = s->rs_rays.as<dr_ray>();
    a.hits = s->rs_hits.as<dr_hit>();
    a.beta = s->rs_beta.as<float4>();
    a.rad = s->rs_radiance.as<float4>();
    a.list_a = s->rs_list_a.as<uint32_t>();
    a.list_b = s->rs_list_b.as<uint32_t>();
    a.cnt = s->rs_counters.as<uint32_t>();
    a.srays = s->rs_shadow_rays.as<dr_ray>();
    a.socc = s->rs_shadow_occ.as<uint32_t>();
    a.scon = s->rs_shadow_contrib.as<float4>();
    a.pview = s->rs_view.as<uint32_t>();
    a.proot = j.scenes ? s->rs_root.as<uint32_t>() : nullptr;
    a.sroot = j.scenes ? s->rs_sroot.as<uint32_t>() : nullptr;
    a.casts = s->ray_casts.as<unsigned long long>();
    a.chunk = s->rs_chunk.as<ChunkState>();
    a.nodes = s->nodes.as<QNode>();
    a.tri_v = s->tri_v.as<float4>();
    a.ovf = s->ovf_stack.as<int32_t>();
    a.flags = s->flags.as<uint32_t>();
    a.grid = s->rs_grid.as<uint32_t>();
    a.refill_min = s->refill_min;
    a.slices = ((j.sched & DR_SCHED_XCD_SLICES) && grid * kPersistParts >= 8u) ? 8u : 1u;
    a.max_stack = max_stack;
    a.cap = j.cap;
    a.has_emit = s->n_emit ? 1u : 0u;
    void *args[] = {&a};
    DR_HIP_TRY(hipLaunchCooperativeKernel(kern, dim3(grid), dim3(kPersistThreads), args, 0, stream));
    uint32_t abort_flag = 0;
    DR_HIP_TRY(hipMemcpyAsync(&abort_flag, s->rs_grid.as<uint32_t>() + 2, 4, hipMemcpyDeviceToHost, stream));
    DR_HIP_TRY(hipStreamSynchronize(stream));
    if (abort_flag) return fail(DR_E_DEVICE, "persistent render: a grid barrier timed out (the launch was stopped)");
    return DR_OK;
}

}  // namespace

dr_status count_ray_casts(Scene *s, uint64_t n, hipStream_t stream, bool set)
{
    const bool fresh = s->ray_casts.ptr == nullptr;
    if (fresh)
        if (dr_status st = s->ray_casts.ensure(sizeof(unsigned long long))) return st;
    if (fresh || set || n) {
        k_add_casts<<<1, 1, 0, stream>>>(s->ray_casts.as<unsigned long long>(), n, (fresh || set) ? 1 : 0);
        DR_HIP_TRY(hipGetLastError());
    }
    return DR_OK;
}

CamParams cam_params(const dr_camera &c)
{
    CamParams cp;
    for (int i = 0; i < 12; ++i) cp.m[i] = c.transform[i];
    cp.f = c.focal_length;
    cp.cx = c.center[0];
    cp.cy = c.center[1];
    cp.psx = c.pixel_size[0];
    cp.psy = c.pixel_size[1];
    cp.w = c.width;
    cp.h = c.height;
    cp.sampling = c.pixel_sampling;
    return cp;
}

dr_status check_render_args(const dr_camera *cam, const dr_render_params *prm, const double *accum)
{
    if (!cam || !prm || !accum) return fail(DR_E_INVALID, "null argument");
    if (cam->width == 0 || cam->height == 0) return fail(DR_E_INVALID, "empty image");
    if ((uint64_t)cam->width * cam->height > 0xFFFFFFFFull) return fail(DR_E_INVALID, "image too large");
    if (prm->max_depth == 0 || prm->max_depth > 64) return fail(DR_E_INVALID, "max_depth must be 1..64");
    return DR_OK;
}

dr_status render_job(const RenderJob &j, hipStream_t stream)
{
    Scene *s = j.s;
    if (j.guide_flags & ~(uint32_t)(DR_GUIDE_USE | DR_GUIDE_COLLECT)) return fail(DR_E_INVALID, "unknown guide flags");
    if (j.guide_flags && !j.guide) return fail(DR_E_INVALID, "guide flags without a guide");
    const bool use_guide = (j.guide_flags & DR_GUIDE_USE) != 0;
    const bool collect = (j.guide_flags & DR_GUIDE_COLLECT) != 0;
    if (j.total == 0) return DR_OK;
    if (collect && j.total > (1ull << 58)) return fail(DR_E_INVALID, "a collecting render takes at most 2^58 samples");
    const uint32_t cap = j.cap;
    if (s->rs_capacity < cap) {
        dr_status st = DR_OK;
        auto need = [&](DevBuf &b, size_t bytes) { if (st == DR_OK) st = b.ensure(bytes); };
        need(s->rs_rays, sizeof(dr_ray) * (size_t)cap);
        need(s->rs_hits, sizeof(dr_hit) * (size_t)cap);
        need(s->rs_beta, sizeof(float4) * (size_t)cap);
        need(s->rs_radiance, sizeof(float4) * (size_t)cap);
        need(s->rs_list_a, 4ull * cap);
        need(s->rs_list_b, 4ull * cap);
        need(s->rs_view, 4ull * cap);
        need(s->rs_counters, 64);
        need(s->rs_chunk, sizeof(ChunkState));
        need(s->rs_shadow_rays, sizeof(dr_ray) * (size_t)cap);
        need(s->rs_shadow_occ, 4ull * cap);
        need(s->rs_shadow_contrib, sizeof(float4) * (size_t)cap);
        if (st) return st;
        s->rs_capacity = cap;
    }
    // render lanes: one for a recorded bounce loop, else the scene's setting
    // (never more lanes than 2^16 paths each)
    const uint64_t n_chunks_all = (j.total + cap - 1) / cap;
    uint32_t n_lanes = (j.graph && n_chunks_all > 1) ? 1u : std::max(1u, std::min(s->render_lanes, kMaxRenderLanes));
    while (n_lanes > 1 && cap / n_lanes < 65536u) --n_lanes;  // (2^17 paths for two lanes)
    if (n_lanes > 1) {
        const uint32_t lane_cap = (cap + n_lanes - 1u) / n_lanes;
        if (!s->lane_fork) DR_HIP_TRY(hipEventCreateWithFlags(&s->lane_fork, hipEventDisableTiming));
        for (uint32_t k = 1; k < n_lanes; ++k) {
            RenderLane &R = s->lanes[k];
            if (!R.stream) DR_HIP_TRY(hipStreamCreateWithFlags(&R.stream, hipStreamNonBlocking));
            if (!R.done) DR_HIP_TRY(hipEventCreateWithFlags(&R.done, hipEventDisableTiming));
            dr_status st = DR_OK;
            auto need = [&](DevBuf &b, size_t bytes) { if (st == DR_OK) st = b.ensure(bytes); };
            need(R.list_a, 4ull * lane_cap);
            need(R.list_b, 4ull * lane_cap);
            need(R.counters, 64);
            need(R.srays, sizeof(dr_ray) * (size_t)lane_cap);
            need(R.socc, 4ull * lane_cap);
            need(R.scon, sizeof(float4) * (size_t)lane_cap);
            if (j.scenes) need(R.sroot, 4ull * lane_cap);
            if (st) return st;
        }
    }
    if (dr_status st = count_ray_casts(s, 0, stream)) return st;
    if (j.scenes) {
        if (dr_status st = s->rs_root.ensure(4ull * s->rs_capacity)) return st;
        if (dr_status st = s->rs_sroot.ensure(4ull * s->rs_capacity)) return st;
    }
    if (use_guide)
        if (dr_status st = s->rs_guide_head.ensure(sizeof(float) * kHeadFloats * (size_t)s->rs_capacity)) return st;
    if (collect) {
        if (dr_status st = s->rs_pending.ensure(4ull * s->rs_capacity)) return st;
        if (dr_status st = guide_reset_counters(j.guide, stream)) return st;
    }
    GuideView gv{};
    if (collect) {
        gv.smp = j.guide->samples.as<dr_guide_sample>();
        gv.aux = j.guide->aux.as<GuideAux>();
        gv.key = j.guide->keys.as<uint32_t>();
        gv.counter = j.guide->counters.as<uint32_t>();
        gv.pending = s->rs_pending.as<uint32_t>();
        gv.cap = j.guide->stage_cap;
        gv.frac = j.guide->collect_fraction;
        gv.item_bits = ceil_log2_u64(j.total);
        gv.vertex_bits = ceil_log2_u64(j.prm->max_depth);
    }
    SceneView sv;
    sv.verts = s->verts.as<float>();
    sv.idx = s->indices.as<uint32_t>();
    sv.tri_mat = s->tri_mat.as<uint32_t>();
    sv.mats = s->materials.as<dr_material>();
    sv.alias = s->alias.as<AliasEntry>();
    sv.pdf_area = s->light_pdf_area.as<float>();
    sv.n_emit = s->n_emit;
    sv.inst = s->n_inst ? s->inst_shade.as<InstShade>() : nullptr;
    sv.n_inst = s->n_inst;
    sv.emit_tri = s->emit_tri.as<uint32_t>();
    sv.emit_pdf = s->emit_pdf.as<float>();
    if (sv.inst && (use_guide || collect || j.persistent || j.scenes))
        return fail(DR_E_UNSUPPORTED, "an instance set renders with the plain sampler and the launch sequence");
    PathParams pp;
    pp.views = j.views;
    pp.scenes = j.scenes;
    pp.chunk = s->rs_chunk.as<ChunkState>();
    pp.n_views = j.n_views;
    pp.max_depth = j.prm->max_depth;
    pp.rr_depth = j.prm->rr_depth;
    pp.schlick = j.prm->fresnel == DR_FRESNEL_SCHLICK ? 1u : 0u;
    k_chunk_init<<<1, 64, 0, stream>>>(s->rs_chunk.as<ChunkState>(), j.total, cap);
    DR_HIP_TRY(hipGetLastError());
    const uint64_t n_chunks = (j.total + cap - 1) / cap;
    if (j.persistent) {
        if (use_guide || collect) return fail(DR_E_UNSUPPORTED, "the persistent launch renders without a guide");
        if (dr_status st = run_persistent(j, pp, sv, stream)) return st;
    } else if (!j.graph || n_chunks == 1) {
        for (uint64_t c = 0; c < n_chunks; ++c)
            if (dr_status st = run_chunk(j, pp, sv, gv, stream, n_lanes)) return st;
    } else {
        // the first chunk runs eagerly (it sizes every buffer and caches the
        // launch shapes); the loop is then recorded once and replayed
        if (dr_status st = run_chunk(j, pp, sv, gv, stream, 1u)) return st;
        DR_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        dr_status cst = run_chunk(j, pp, sv, gv, stream, 1u);
        hipGraph_t graph = nullptr;
        const hipError_t ce = hipStreamEndCapture(stream, &graph);
        if (cst) {
            if (graph) (void)hipGraphDestroy(graph);
            return cst;
        }
        DR_HIP_TRY(ce);
        hipGraphExec_t exec = nullptr;
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        DR_HIP_TRY(ie);
        hipError_t le = hipSuccess;
        for (uint64_t c = 1; c < n_chunks && le == hipSuccess; ++c) le = hipGraphLaunch(exec, stream);
        // the replays may still be queued: keep the executable graph until
        // an event after them has completed
        j.s->release_graph();
        j.s->rs_exec = exec;
        hipError_t ee = hipEventCreateWithFlags(&j.s->rs_exec_done, hipEventDisableTiming);
        if (ee == hipSuccess) ee = hipEventRecord(j.s->rs_exec_done, stream);
        DR_HIP_TRY(le);
        DR_HIP_TRY(ee);
    }
    if (collect)
        if (dr_status st = guide_compact(j.guide, gv.item_bits + gv.vertex_bits, stream)) return st;
    return DR_OK;
}

void Scene::release_lanes()
{
    for (RenderLane &R : lanes) {
        if (R.stream) {
            (void)hipStreamSynchronize(R.stream);
            (void)hipStreamDestroy(R.stream);
            R.stream = nullptr;
        }
        if (R.done) {
            (void)hipEventDestroy(R.done);
            R.done = nullptr;
        }
    }
    if (lane_fork) {
        (void)hipEventDestroy(lane_fork);
        lane_fork = nullptr;
    }
}

void Scene::release_graph()
{
    if (rs_exec) {
        // with no completion event (its creation failed) wait for the device
        if (rs_exec_done) (void)hipEventSynchronize(rs_exec_done);
        else (void)hipDeviceSynchronize();
        (void)hipGraphExecDestroy(rs_exec);
        rs_exec = nullptr;
    }
    if (rs_exec_done) {
        (void)hipEventDestroy(rs_exec_done);
        rs_exec_done = nullptr;
    }
}

dr_status render(Scene *s, const dr_camera *cam, const dr_render_params *prm, Guide *guide, uint32_t guide_flags,
                 double *accum, hipStream_t stream)
{
    if (!s->built) return fail(DR_E_INVALID, "scene BVH not built (call dr_scene_build_bvh)");
    if (dr_status st = check_render_args(cam, prm, accum)) return st;
    if (!s->lights_built)
        if (dr_status st = s->n_inst ? build_instance_lights(s, stream) : build_lights(s, stream)) return st;
    const uint32_t npix = cam->width * cam->height;
    const uint64_t items = (uint64_t)npix * prm->spp;
    if (items == 0) return DR_OK;
    if (dr_status st = s->rs_views.ensure(sizeof(ViewInfo))) return st;
    ViewInfo v;
    v.cam = cam_params(*cam);
    v.item_off = 0;
    v.seed = prm->seed;
    v.accum = accum;
    v.npix = npix;
    v.spp = prm->spp;
    v.sample_offset = prm->sample_offset;
    v.scene = 0;
    k_set_view<<<1, 64, 0, stream>>>(s->rs_views.as<ViewInfo>(), v);
    DR_HIP_TRY(hipGetLastError());
    RenderJob j;
    j.s = s;
    j.scenes = nullptr;
    j.views = s->rs_views.as<ViewInfo>();
    j.n_views = 1;
    j.total = items;
    j.prm = prm;
    j.guide = guide;
    j.guide_flags = guide_flags;
    j.cap = (uint32_t)std::min<uint64_t>(items, kPathCap);
    j.sched = DR_SCHED_PER_RAY_SPECULATIVE;
    j.graph = false;
    return render_job(j, stream);
}

dr_status render_instances(Scene *f, const dr_camera *cam, const dr_render_params *prm, double *accum,
                           hipStream_t stream)
{
    if (!f->n_inst) return fail(DR_E_INVALID, "not an instance set");
    return render(f, cam, prm, nullptr, 0u, accum, stream);
}

}  // namespace dr
