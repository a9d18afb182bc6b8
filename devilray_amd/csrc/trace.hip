// Closest-hit and any-hit traversal of the 4-wide BVH.
//
// One lane per ray. Each lane keeps a short stack of child references in LDS
// (kLdsStack entries, laid out [entry][lane] so a wave's pushes at equal
// depth hit distinct banks); deeper entries spill to a per-lane region in
// global memory, up to kMaxStack in total. At an interior node all four child
// boxes are tested with the robust slab test, the hit children are sorted by
// entry distance, the nearest is taken next and the rest are pushed far to
// near. Leaves hold up to 8 triangles stored contiguously as 3 float4.
#include "dr_device.hpp"
#include "dr_internal.hpp"

namespace dr {

namespace {

constexpr int kBlock = 256;
constexpr int kLdsStack = 16;
constexpr int kMaxStack = 64;

struct Best {
    float t, u, v;
    uint32_t id;
};

__device__ __forceinline__ void cswap(float &ta, int32_t &ra, float &tb, int32_t &rb)
{
    if (tb < ta) {
        const float t = ta; ta = tb; tb = t;
        const int32_t r = ra; ra = rb; rb = r;
    }
}

template <bool ANY, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_trace(const WideNode *__restrict__ nodes,
                                                  const float4 *__restrict__ tri_v,
                                                  const float4 *__restrict__ rays, uint32_t n,
                                                  dr_hit *__restrict__ hits, uint32_t *__restrict__ occ,
                                                  uint32_t *__restrict__ box_tests,
                                                  uint32_t *__restrict__ tri_tests,
                                                  int32_t *__restrict__ ovf, uint32_t *__restrict__ flags)
{
    __shared__ int32_t stack[kLdsStack][kBlock];
    const uint32_t tid = threadIdx.x;
    const uint32_t i = blockIdx.x * kBlock + tid;
    if (i >= n) return;
    const float4 r0 = rays[2 * (size_t)i];
    const float4 r1 = rays[2 * (size_t)i + 1];
    RayPre p;
    ray_prepare(r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, p);

    Best best{p.tmax, 0.0f, 0.0f, 0xFFFFFFFFu};
    bool occluded = false;
    uint32_t nbox = 0, ntri = 0;
    int sp = 0;
    int32_t ref = 0;  // the root is always an interior wide node
    bool overflow = false;
    // A NaN anywhere in the ray hits nothing (the triangle test rejects it
    // too); skipping traversal also keeps NaN slabs from admitting empty slots.
    const bool nan_ray = (r0.x != r0.x) || (r0.y != r0.y) || (r0.z != r0.z) || (r0.w != r0.w) ||
                         (r1.x != r1.x) || (r1.y != r1.y) || (r1.z != r1.z) || (r1.w != r1.w);
    if (nan_ray) ref = DR_EMPTY_REF;

    auto push = [&](int32_t r) {
        if (sp < kLdsStack) stack[sp][tid] = r;
        else if (sp < kMaxStack) ovf[(size_t)(sp - kLdsStack) * n + i] = r;
        else { overflow = true; return; }
        ++sp;
    };

    for (;;) {
        if (ref == DR_EMPTY_REF) {
            // nothing to visit (NaN ray)
        } else if (ref >= 0) {
            const float4 *np = reinterpret_cast<const float4 *>(nodes + ref);
            const float4 lox = np[0], hix = np[1], loy = np[2], hiy = np[3], loz = np[4], hiz = np[5];
            const int4 ch = reinterpret_cast<const int4 *>(np)[6];
            const float lim = (ANY ? p.tmax : best.t) * kBestSlack;
            float tn[4];
            const float lx[4] = {lox.x, lox.y, lox.z, lox.w}, hx[4] = {hix.x, hix.y, hix.z, hix.w};
            const float ly[4] = {loy.x, loy.y, loy.z, loy.w}, hy[4] = {hiy.x, hiy.y, hiy.z, hiy.w};
            const float lz[4] = {loz.x, loz.y, loz.z, loz.w}, hz[4] = {hiz.x, hiz.y, hiz.z, hiz.w};
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float t0x = (lx[c] - p.ox) * p.ix, t1x = (hx[c] - p.ox) * p.ix;
                const float t0y = (ly[c] - p.oy) * p.iy, t1y = (hy[c] - p.oy) * p.iy;
                const float t0z = (lz[c] - p.oz) * p.iz, t1z = (hz[c] - p.oz) * p.iz;
                const float near = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), p.tmin));
                float far = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z)) * kFarScale;
                far = fminf(far, lim);
                tn[c] = (near <= far) ? near : INFINITY;
            }
            int32_t r[4] = {ch.x, ch.y, ch.z, ch.w};
            if (COUNT) nbox += (ch.x != DR_EMPTY_REF) + (ch.y != DR_EMPTY_REF) + (ch.z != DR_EMPTY_REF) + (ch.w != DR_EMPTY_REF);
            // sorting network on (distance, ref)
            cswap(tn[0], r[0], tn[1], r[1]);
            cswap(tn[2], r[2], tn[3], r[3]);
            cswap(tn[0], r[0], tn[2], r[2]);
            cswap(tn[1], r[1], tn[3], r[3]);
            cswap(tn[1], r[1], tn[2], r[2]);
            // empty slots carry boxes at +inf and never pass the slab test;
            // the ref check keeps that true for any ray
            if (tn[3] != INFINITY && r[3] != DR_EMPTY_REF) push(r[3]);
            if (tn[2] != INFINITY && r[2] != DR_EMPTY_REF) push(r[2]);
            if (tn[1] != INFINITY && r[1] != DR_EMPTY_REF) push(r[1]);
            if (overflow) break;
            if (tn[0] != INFINITY && r[0] != DR_EMPTY_REF) { ref = r[0]; continue; }
        } else {
            const uint32_t first = leaf_first(ref), cnt = leaf_count(ref);
            for (uint32_t k = 0; k < cnt; ++k) {
                const float4 *tv = tri_v + 3 * (size_t)(first + k);
                const float4 a = tv[0], b = tv[1], c = tv[2];
                float t, u, v;
                if (COUNT) ++ntri;
                if (!tri_watertight(p, a, b, c, t, u, v)) continue;
                const uint32_t id = (uint32_t)__float_as_int(a.w);
                if (ANY) {
                    if (t >= p.tmin && t <= p.tmax) { occluded = true; break; }
                } else if (accept_hit(t, id, p.tmin, best.t, best.id)) {
                    best.t = t; best.u = u; best.v = v; best.id = id;
                }
            }
            if (ANY && occluded) break;
        }
        if (sp == 0) break;
        --sp;
        ref = sp < kLdsStack ? stack[sp][tid] : ovf[(size_t)(sp - kLdsStack) * n + i];
    }
    if (overflow) atomicOr(&flags[0], 1u);
    if (ANY) {
        occ[i] = occluded ? 1u : 0u;
    } else {
        dr_hit h;
        const bool hit = best.id != 0xFFFFFFFFu;
        h.id = hit ? (int32_t)best.id : -1;
        h.t = hit ? best.t : INFINITY;
        h.u = hit ? best.u : 0.0f;
        h.v = hit ? best.v : 0.0f;
        hits[i] = h;
    }
    if (COUNT) {
        box_tests[i] = nbox;
        tri_tests[i] = ntri;
    }
}

}  // namespace

dr_status launch_trace(Scene *s, const dr_ray *rays, uint32_t n, dr_hit *hits, uint32_t *occluded,
                       uint32_t *box_tests, uint32_t *tri_tests, bool any_hit, hipStream_t stream)
{
    if (!s->built) return fail(DR_E_INVALID, "scene BVH not built (call dr_scene_build_bvh)");
    if (n == 0) return DR_OK;
    if (!rays || (any_hit ? !occluded : !hits)) return fail(DR_E_INVALID, "null ray or output buffer");
    if (dr_status st = s->ovf_stack.ensure(sizeof(int32_t) * (size_t)(kMaxStack - kLdsStack) * n)) return st;
    if (dr_status st = s->flags.ensure(16)) return st;
    const dim3 grid((n + kBlock - 1) / kBlock);
    const auto *nodes = s->nodes.as<WideNode>();
    const auto *tv = s->tri_v.as<float4>();
    const auto *r4 = reinterpret_cast<const float4 *>(rays);
    int32_t *ovf = s->ovf_stack.as<int32_t>();
    uint32_t *flags = s->flags.as<uint32_t>();
    const bool count = box_tests && tri_tests;
    if (any_hit)
        k_trace<true, false><<<grid, kBlock, 0, stream>>>(nodes, tv, r4, n, nullptr, occluded, nullptr, nullptr, ovf, flags);
    else if (count)
        k_trace<false, true><<<grid, kBlock, 0, stream>>>(nodes, tv, r4, n, hits, nullptr, box_tests, tri_tests, ovf, flags);
    else
        k_trace<false, false><<<grid, kBlock, 0, stream>>>(nodes, tv, r4, n, hits, nullptr, nullptr, nullptr, ovf, flags);
    DR_HIP_TRY(hipGetLastError());
    return DR_OK;
}

}  // namespace dr
