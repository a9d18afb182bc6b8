// Device side of the 4-wide BVH traversal, shared by the trace kernels
// (trace.hip) and the persistent render kernel (render.hip).
//
// Ray order. Optionally (DR_SCHED_REORDER) the batch is first sorted by the
// key (dominant direction axis, 30-bit Morton code of the origin inside the
// scene box): neighbouring lanes then start in the same part of the tree and
// a wave's rays share the dominant axis, so the triangle test runs a
// compile-time specialisation without per-lane component selects. Results are
// written back in input order.
//
// Schedules. PER_RAY: one lane per ray. SPECULATIVE: the "while-while" loop
// with postponed leaves (Aila & Laine, "Understanding the Efficiency of Ray
// Traversal on GPUs", HPG 2009): a lane that reaches a leaf parks it and keeps
// descending interior nodes until every lane of the wave holds a leaf, then
// all lanes intersect their leaves together. PERSISTENT: resident waves
// refill finished lanes from a global counter (one atomic per wave).
//
// Each lane keeps a short stack of child references in LDS (kLdsStack
// entries, laid out [entry][lane]); deeper entries spill to a per-lane region
// in global memory. A visit pushes at most 3 children, and every stacked entry
// is a child of a node on the path from the root to the node being visited,
// so a tree of D wide levels never needs more than 3 D entries: the host sizes
// the spill region from the built depth, and no ray is ever cut short. At an
// interior node (a 64-byte QNode) the four child boxes are decoded from their
// plane bytes (one exact fma per plane) and tested with the robust slab test
// (node_quant.hpp); near and far planes are picked by the ray's octant so no
// per-slab min/max is needed. The quantised boxes
// contain the exact ones, so the hits are those of the float boxes; the node
// is half the bytes of a float node. Hit children
// are sorted by entry distance, the nearest is taken next and the rest are
// pushed far to near. Leaves hold up to 8 triangles stored as 3 float4.
//
// Instances (INST). A two-level forest starts with a tree over instance
// boxes in world space whose leaves name an instance. A lane that reaches
// one enters it at once: it pushes an exit marker, moves its ray into the
// instance's model space ((o - t) / s, d / s, so t stays world t) and
// descends the instance's tree; popping the marker restores the world ray. A
// model-space leaf parked by the speculative loop is intersected before the
// marker is taken, so a parked leaf always belongs to the current instance.
// The stack bound is 3 entries per level of the top tree and of the deepest
// instanced tree, plus the marker.
#pragma once

#include "dr_device.hpp"
#include "dr_internal.hpp"
#include "node_quant.hpp"

namespace dr {

namespace {

constexpr int kBlock = 256;
constexpr int kLdsStack = 16;
constexpr int32_t kNone = DR_EMPTY_REF;
// stack marker of an instance's model space (a leaf reference of count 8 at
// first 2^28 - 2, beyond every valid leaf)
constexpr int32_t kInstExit = (int32_t)0xFFFFFFFEu;

__device__ __forceinline__ void cswap(float &ta, int32_t &ra, float &tb, int32_t &rb)
{
    if (tb < ta) {
        const float t = ta; ta = tb; tb = t;
        const int32_t r = ra; ra = rb; rb = r;
    }
}

template <bool INST = false>
__device__ __forceinline__ bool is_leaf(int32_t r)
{
    return r < 0 && r != kNone && (!INST || r != kInstExit);
}

struct LeafState {
    float bt, bu, bv;
    uint32_t bid;
    bool occluded;
    uint32_t ntri;
    uint32_t wave_tri_iters;  // counted variant: triangle-loop iterations issued while this lane led the wave
};

// Counted variant only: true on one active lane per wave, so a lane-local sum
// of these events over the wave counts wave-level loop iterations (SIMD slots).
__device__ __forceinline__ bool leads_wave()
{
    return (uint32_t)(__ffsll((unsigned long long)__ballot(true)) - 1) == (threadIdx.x & 63u);
}

template <bool ANY, bool COUNT, int KZ>
__device__ __forceinline__ void intersect_leaf(const float4 *__restrict__ tri_v, int32_t leaf, const RayPre &p,
                                               LeafState &st, uint32_t id_add = 0u)
{
    const uint32_t first = leaf_first(leaf), cnt = leaf_count(leaf);
    for (uint32_t k = 0; k < cnt; ++k) {
        const float4 *tv = tri_v + 3 * (size_t)(first + k);
        const float4 a = tv[0], b = tv[1], c = tv[2];
        float t, u, v;
        if (COUNT) {
            ++st.ntri;
            if (leads_wave()) ++st.wave_tri_iters;
        }
        if (!tri_watertight<KZ>(p, a, b, c, t, u, v)) continue;
        const uint32_t id = (uint32_t)__float_as_int(a.w) + id_add;
        if (ANY) {
            if (t >= p.tmin && t <= p.tmax) { st.occluded = true; return; }
        } else if (accept_hit(t, id, p.tmin, st.bt, st.bid)) {
            st.bt = t; st.bu = u; st.bv = v; st.bid = id;
        }
    }
}

// The traversal of one launch, as thread `tid` (0 .. kBlock - 1) of block
// `bid` of `nblk`: a kernel passes its own thread and block indices and grid
// size; the persistent render kernel calls it as one of its stages, each
// kBlock-thread part of its blocks acting as one block. `stack` is that
// block's LDS short stack. Every thread of the block must call it; it returns
// when no ray is left.
template <bool ANY, bool COUNT, bool PERSISTENT, bool SPECULATIVE, bool INST = false>
__device__ __forceinline__ void trace_body(int32_t (*__restrict__ stack)[kBlock], uint32_t tid, uint32_t bid,
                                           uint32_t nblk,
                                           const QNode *__restrict__ nodes, const float4 *__restrict__ tri_v,
                                           const float4 *__restrict__ rays, const uint32_t *__restrict__ perm,
                                           uint32_t n_max, const uint32_t *__restrict__ n_dev,
                                           dr_hit *__restrict__ hits, uint32_t *__restrict__ occ,
                                           uint32_t *__restrict__ box_tests, uint32_t *__restrict__ tri_tests,
                                           int32_t *__restrict__ ovf, uint32_t *__restrict__ flags,
                                           uint32_t *__restrict__ ray_counter, uint32_t refill_min,
                                           const uint32_t *__restrict__ roots, uint32_t slices, uint32_t max_stack,
                                           const InstRec *__restrict__ inst)
{
    static_assert(!INST || (SPECULATIVE && !COUNT), "instances run the speculative loop, uncounted");
    // ray count: fixed at launch, or read from device memory (on-device queues)
    const uint32_t n = n_dev ? min(*n_dev, n_max) : n_max;
    const uint32_t lane = tid & 63u;
    const size_t gl = (size_t)bid * kBlock + tid;
    const size_t nlanes = (size_t)nblk * kBlock;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64u - lane));
    // persistent: the (sorted) ray range splits into `slices` contiguous
    // parts with a counter each; workgroups are dispatched to the XCDs in
    // turn, so with 8 slices each XCD's L2 serves one region of the batch. A
    // wave whose part is drained moves on to the next one (wave-uniform
    // state), so uneven parts do not leave an XCD idle.
    uint32_t part = bid % slices, parts_seen = 1u;

    int32_t ray = -1;
    bool exhausted = false;
    RayPre p{};
    LeafState st{0.0f, 0.0f, 0.0f, 0xFFFFFFFFu, false, 0u, 0u};
    uint32_t wave_node_iters = 0, node_visits = 0, tri_total = 0;  // counted variant
    uint32_t tri_cur = 0, tri_end = 0;  // single-step schedule: triangles of the current leaf
    int sp = 0;
    int32_t node = kNone, leaf = kNone;
    uint32_t nbox = 0;
    bool negx = false, negy = false, negz = false;  // the ray's octant: the upper plane of the axis is near
    bool overflow = false, any_overflow = false;
    uint32_t high_water = 0;  // counted variant: deepest stack of the lane
    const char *nbase = reinterpret_cast<const char *>(nodes);
    // instances: the world ray, whether the lane is in an instance's model
    // space, and that instance's triangle id offset
    float wox = 0.0f, woy = 0.0f, woz = 0.0f, wdx = 0.0f, wdy = 0.0f, wdz = 0.0f;
    bool in_inst = false;
    uint32_t id_add = 0;
    auto set_octant = [&]() {
        negx = signbit(p.ix);
        negy = signbit(p.iy);
        negz = signbit(p.iz);
    };

    auto pop = [&]() -> int32_t {
        if (sp == 0) return kNone;
        --sp;
        return sp < kLdsStack ? stack[sp][tid] : ovf[(size_t)(sp - kLdsStack) * nlanes + gl];
    };
    // max_stack >= 3 x the tree's wide levels, so the last branch is a
    // guard that never runs; if it did, the ray would end early and the
    // scene's overflow flag would report it
    auto push = [&](int32_t r) {
        if (sp < kLdsStack) stack[sp][tid] = r;
        else if ((uint32_t)sp < max_stack) ovf[(size_t)(sp - kLdsStack) * nlanes + gl] = r;
        else { overflow = true; return; }
        ++sp;
        if (COUNT) high_water = max(high_water, (uint32_t)sp);
    };

    for (;;) {
        // ---- fetch: the lane's own slot, or (persistent) the next free slots
        const bool need = ray < 0 && !exhausted;
        const uint64_t m = __ballot(need);
        // persistent waves refill in batches (each refill is an atomic and a
        // ray load on the wave's critical path): when refill_min lanes are
        // idle, or when none is busy
        const bool refill = PERSISTENT ? ((uint32_t)__popcll(m) >= refill_min || __ballot(ray >= 0) == 0ull) : true;
        if (m && refill) {
            uint32_t base = 0;
            const uint32_t part_lo = PERSISTENT ? (uint32_t)(((uint64_t)n * part) / slices) : 0u;
            const uint32_t part_hi = PERSISTENT ? (uint32_t)(((uint64_t)n * (part + 1u)) / slices) : n;
            if (PERSISTENT) {
                const int leader = __ffsll((unsigned long long)m) - 1;
                if ((int)lane == leader) base = atomicAdd(ray_counter + part, (uint32_t)__popcll(m));
                base = __shfl(base, leader, 64);
            }
            const bool drained = PERSISTENT && part_lo + base + (uint32_t)__popcll(m) >= part_hi;
            if (need) {
                const uint32_t slot = PERSISTENT ? part_lo + base + (uint32_t)__popcll(m & lt) : (uint32_t)gl;
                if (!PERSISTENT) exhausted = true;
                if (slot < part_hi) {
                    const uint32_t r = perm ? perm[slot] : slot;
                    ray = (int32_t)r;
                    const float4 r0 = rays[2 * (size_t)r];
                    const float4 r1 = rays[2 * (size_t)r + 1];
                    ray_prepare(r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, p);
                    st.bt = p.tmax; st.bu = 0.0f; st.bv = 0.0f; st.bid = 0xFFFFFFFFu;
                    st.occluded = false;
                    tri_total += st.ntri;
                    st.ntri = 0;
                    sp = 0; leaf = kNone; nbox = 0;
                    tri_cur = 0; tri_end = 0;
                    overflow = false;
                    // a NaN anywhere in the ray hits nothing (the triangle test
                    // rejects it too), and NaN slabs would admit empty slots
                    const bool nan_ray = (r0.x != r0.x) || (r0.y != r0.y) || (r0.z != r0.z) || (r0.w != r0.w) ||
                                         (r1.x != r1.x) || (r1.y != r1.y) || (r1.z != r1.z) || (r1.w != r1.w);
                    // the root is always an interior node
                    node = nan_ray ? kNone : (roots ? (int32_t)roots[r] : 0);
                    set_octant();
                    if (INST) {
                        wox = r0.x; woy = r0.y; woz = r0.z;
                        wdx = r1.x; wdy = r1.y; wdz = r1.z;
                        in_inst = false;
                        id_add = 0u;
                    }
                } else if (!PERSISTENT || parts_seen >= slices) {
                    exhausted = true;
                }
            }
            if (drained && parts_seen < slices) {
                part = part + 1u == slices ? 0u : part + 1u;
                ++parts_seen;
            }
        }
        // done when no lane holds a ray and none can fetch one (every part
        // visited and drained); each pass either fetches or visits a part
        if (__ballot(ray >= 0) == 0ull && __ballot(!exhausted) == 0ull) break;

        if (ray >= 0) {
            // ---- interior nodes, leaves postponed
            // one interior node: returns the next reference (nearest hit child
            // or the top of the stack) after pushing the other hit children
            auto visit = [&](int32_t cur) -> int32_t {
                if (COUNT) {
                    ++node_visits;
                    if (leads_wave()) ++wave_node_iters;
                }
                // a QNode: origin xyz and step x | steps y z, lower plane
                // bytes x y | lower z, upper x y z | the children
                const char *nd = nbase + (size_t)(uint32_t)cur * sizeof(QNode);
                const float4 g0 = *reinterpret_cast<const float4 *>(nd);
                const float4 g1 = *reinterpret_cast<const float4 *>(nd + 16);
                const uint4 g2 = *reinterpret_cast<const uint4 *>(nd + 32);
                const int4 ch = *reinterpret_cast<const int4 *>(nd + 48);
                const float lim = (ANY ? p.tmax : st.bt) * kBestSlack;
                const QNodeWords nw{g0.x, g0.y, g0.z, g0.w, g1.x, g1.y,
                                    {__float_as_uint(g1.z), __float_as_uint(g1.w), g2.x}, {g2.y, g2.z, g2.w}};
                const SlabRay sr{p.ox, p.oy, p.oz, p.ix, p.iy, p.iz, p.tmin, negx, negy, negz};
                float nr[4], fr[4];
                qnode_slabs(nw, sr, lim, nr, fr);
                float tn[4];
                int32_t r[4] = {ch.x, ch.y, ch.z, ch.w};
#pragma unroll
                for (int c = 0; c < 4; ++c) tn[c] = (nr[c] <= fr[c] && r[c] != kNone) ? nr[c] : INFINITY;  // empty slots never pass
                if (COUNT) nbox += (ch.x != kNone) + (ch.y != kNone) + (ch.z != kNone) + (ch.w != kNone);
                if (!ANY) {  // nearest child first
                    cswap(tn[0], r[0], tn[1], r[1]);
                    cswap(tn[2], r[2], tn[3], r[3]);
                    cswap(tn[0], r[0], tn[2], r[2]);
                    cswap(tn[1], r[1], tn[3], r[3]);
                    cswap(tn[1], r[1], tn[2], r[2]);
                } else {
                    // any hit ends the query: only the nearest hit child goes first
                    // (6% faster than a full sort on incoherent occlusion rays)
                    cswap(tn[0], r[0], tn[1], r[1]);
                    cswap(tn[0], r[0], tn[2], r[2]);
                    cswap(tn[0], r[0], tn[3], r[3]);
                }
                if (tn[3] != INFINITY) push(r[3]);
                if (tn[2] != INFINITY) push(r[2]);
                if (tn[1] != INFINITY) push(r[1]);
                return (tn[0] != INFINITY) ? r[0] : pop();
            };
            if (SPECULATIVE) {
                // ---- interior nodes, leaves postponed
                for (;;) {
                    if (INST) {
                        // (a failed push leaves the stack without its marker:
                        // stop here and let the overflow path end the ray)
                        if (overflow) break;
                        // leave the instance once its parked leaf is done
                        if (node == kInstExit && leaf == kNone) {
                            ray_prepare(wox, woy, woz, p.tmin, wdx, wdy, wdz, p.tmax, p);
                            set_octant();
                            in_inst = false;
                            id_add = 0u;
                            node = pop();
                        }
                        // an instance leaf of the top tree: enter it
                        if (!in_inst && is_leaf<true>(node)) {
                            const InstRec ir = inst[leaf_first(node)];
                            push(kInstExit);
                            ray_prepare((wox - ir.tx) * ir.isx, (woy - ir.ty) * ir.isy, (woz - ir.tz) * ir.isz,
                                        p.tmin, wdx * ir.isx, wdy * ir.isy, wdz * ir.isz, p.tmax, p);
                            set_octant();
                            in_inst = true;
                            id_add = ir.id_delta;
                            node = ir.root;
                        }
                    }
                    if (node < 0) break;
                    node = visit(node);
                    if (is_leaf<INST>(node) && leaf == kNone && (!INST || in_inst)) {
                        leaf = node;
                        node = pop();
                    }
                    if (__ballot(leaf == kNone) == 0ull) break;  // every active lane holds a leaf
                }
                // ---- leaves; a wave whose lanes share the dominant axis runs
                // the specialised triangle test
                const int kz0 = __builtin_amdgcn_readfirstlane(p.kz);
                const bool uniform = __ballot(p.kz != kz0) == 0ull;
                // (a parked leaf is in the model space of the lane's instance,
                // and so is a leaf taken from its stack below)
                while (is_leaf<INST>(leaf) && !(INST && overflow)) {
                    if (uniform) {
                        if (kz0 == 0) intersect_leaf<ANY, COUNT, 0>(tri_v, leaf, p, st, id_add);
                        else if (kz0 == 1) intersect_leaf<ANY, COUNT, 1>(tri_v, leaf, p, st, id_add);
                        else intersect_leaf<ANY, COUNT, 2>(tri_v, leaf, p, st, id_add);
                    } else {
                        intersect_leaf<ANY, COUNT, -1>(tri_v, leaf, p, st, id_add);
                    }
                    if (ANY && st.occluded) { leaf = kNone; node = kNone; sp = 0; break; }
                    leaf = kNone;
                    if (is_leaf<INST>(node)) {
                        leaf = node;
                        node = pop();
                    }
                }
            } else {
                // ---- single step: one triangle test or one node visit per
                // iteration, so lanes in either phase make progress together
                if (tri_cur < tri_end) {
                    const int kz0 = __builtin_amdgcn_readfirstlane(p.kz);
                    const bool uniform = __ballot(p.kz != kz0) == 0ull;
                    const int32_t one = make_leaf_ref(tri_cur, 1u);
                    if (uniform) {
                        if (kz0 == 0) intersect_leaf<ANY, COUNT, 0>(tri_v, one, p, st);
                        else if (kz0 == 1) intersect_leaf<ANY, COUNT, 1>(tri_v, one, p, st);
                        else intersect_leaf<ANY, COUNT, 2>(tri_v, one, p, st);
                    } else {
                        intersect_leaf<ANY, COUNT, -1>(tri_v, one, p, st);
                    }
                    ++tri_cur;
                    if (ANY && st.occluded) { tri_cur = tri_end; node = kNone; sp = 0; }
                } else if (node >= 0) {
                    node = visit(node);
                }
                if (tri_cur >= tri_end && is_leaf(node)) {
                    tri_cur = leaf_first(node);
                    tri_end = tri_cur + leaf_count(node);
                    node = pop();
                }
            }
            if (overflow) { any_overflow = true; node = kNone; leaf = kNone; sp = 0; tri_cur = tri_end; }
            // ---- finished: write the result, free the lane
            if (node == kNone && leaf == kNone && tri_cur >= tri_end) {
                const uint32_t rr = (uint32_t)ray;
                if (ANY) {
                    occ[rr] = st.occluded ? 1u : 0u;
                } else {
                    const bool hit = st.bid != 0xFFFFFFFFu;
                    dr_hit h;
                    h.id = hit ? (int32_t)st.bid : -1;
                    // + 0 reports a hit at distance zero as +0 whatever the
                    // sign of the test's determinant
                    h.t = hit ? st.bt + 0.0f : INFINITY;
                    h.u = hit ? st.bu : 0.0f;
                    h.v = hit ? st.bv : 0.0f;
                    hits[rr] = h;
                }
                if (COUNT) { box_tests[rr] = nbox; tri_tests[rr] = st.ntri; }
                ray = -1;
            }
        }
    }
    if (any_overflow) atomicOr(&flags[0], 1u);
    if (COUNT) {
        // SIMD utilisation statistics: [2] wave-level node-loop iterations,
        // [3] node visits, [4] wave-level triangle-loop iterations, [5] tests;
        // [6] the deepest traversal stack of any ray
        atomicAdd(&flags[2], wave_node_iters);
        atomicAdd(&flags[3], node_visits);
        atomicAdd(&flags[4], st.wave_tri_iters);
        atomicAdd(&flags[5], tri_total + st.ntri);
        atomicMax(&flags[6], high_water);
    }
}

}  // namespace

}  // namespace dr
