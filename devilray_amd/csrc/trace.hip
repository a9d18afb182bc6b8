// Closest-hit and any-hit traversal launches of the 4-wide BVH (the
// traversal itself is in trace_device.hpp): ray reordering, launch
// shapes, the zero-allocation entry point.
#include "trace_device.hpp"

namespace dr {

namespace {

// the spill region covers schedules up to this many entries per lane at the
// full per-ray grid; deeper trees run on the resident (persistent) grid
constexpr uint32_t kPerRaySpillMax = 48;
constexpr uint32_t kCounterSlots = 256;

template <bool ANY, bool COUNT, bool PERSISTENT, bool SPECULATIVE, bool INST = false>
__global__ __launch_bounds__(kBlock) void k_trace(const QNode *__restrict__ nodes,
                                                  const float4 *__restrict__ tri_v,
                                                  const float4 *__restrict__ rays, const uint32_t *__restrict__ perm,
                                                  uint32_t n_max, const uint32_t *__restrict__ n_dev,
                                                  dr_hit *__restrict__ hits, uint32_t *__restrict__ occ,
                                                  uint32_t *__restrict__ box_tests,
                                                  uint32_t *__restrict__ tri_tests,
                                                  int32_t *__restrict__ ovf, uint32_t *__restrict__ flags,
                                                  uint32_t *__restrict__ ray_counter, uint32_t refill_min,
                                                  const uint32_t *__restrict__ roots, uint32_t slices,
                                                  uint32_t max_stack, const InstRec *__restrict__ inst)
{
    __shared__ int32_t stack[kLdsStack][kBlock];
    trace_body<ANY, COUNT, PERSISTENT, SPECULATIVE, INST>(stack, threadIdx.x, blockIdx.x, gridDim.x, nodes, tri_v, rays,
                                                          perm, n_max,
                                                          n_dev, hits, occ, box_tests, tri_tests, ovf, flags,
                                                          ray_counter, refill_min, roots, slices, max_stack, inst);
}

// An any-hit queue and a closest-hit queue in one launch of per-ray blocks:
// blocks [0, nblk_a) take the any-hit rays, the rest the closest-hit rays,
// each part with its own block indices, grid size and spill region, so every
// ray is traced exactly as in a launch of its own. Both counts live in device
// memory; blocks past a count exit at once.
template <bool SPECULATIVE, bool INST = false>
__global__ __launch_bounds__(kBlock) void k_trace_pair(const QNode *__restrict__ nodes,
                                                       const float4 *__restrict__ tri_v,
                                                       const float4 *__restrict__ rays_a, const uint32_t *__restrict__ na,
                                                       uint32_t *__restrict__ occ, const uint32_t *__restrict__ roots_a,
                                                       uint32_t nblk_a, int32_t *__restrict__ ovf_a,
                                                       const float4 *__restrict__ rays_c,
                                                       const uint32_t *__restrict__ list_c,
                                                       const uint32_t *__restrict__ nc, dr_hit *__restrict__ hits,
                                                       const uint32_t *__restrict__ roots_c,
                                                       int32_t *__restrict__ ovf_c, uint32_t n_max,
                                                       uint32_t *__restrict__ flags, uint32_t max_stack,
                                                       const InstRec *__restrict__ inst)
{
    __shared__ int32_t stack[kLdsStack][kBlock];
    if (blockIdx.x < nblk_a)
        trace_body<true, false, false, SPECULATIVE, INST>(stack, threadIdx.x, blockIdx.x, nblk_a, nodes, tri_v,
                                                          rays_a, nullptr, n_max, na, nullptr, occ, nullptr, nullptr,
                                                          ovf_a, flags, nullptr, 0u, roots_a, 1u, max_stack, inst);
    else
        trace_body<false, false, false, SPECULATIVE, INST>(stack, threadIdx.x, blockIdx.x - nblk_a,
                                                           gridDim.x - nblk_a, nodes, tri_v, rays_c, list_c, n_max,
                                                           nc, hits, nullptr, nullptr, nullptr, ovf_c, flags, nullptr,
                                                           0u, roots_c, 1u, max_stack, inst);
}

__global__ void k_zero_counter(uint32_t *c) { *c = 0u; }
__global__ void k_zero_counters8(uint32_t *c)
{
    if (threadIdx.x < 8) c[threadIdx.x] = 0u;
}

__device__ __forceinline__ uint32_t expand10(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// key = dominant axis (2 bits) | Morton code of the origin (30 bits)
__device__ __forceinline__ uint32_t spread2(uint32_t v)  // 16 bits -> even bits
{
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// key mode 0: dominant axis (2 bits) | origin Morton code (30 bits)
// key mode 1: dominant axis | direction (octahedral map, 3+3 bits) | origin Morton (24 bits)
__global__ __launch_bounds__(kBlock) void k_ray_keys(const float4 *__restrict__ rays, uint32_t n, float lx, float ly,
                                                     float lz, float sx, float sy, float sz, uint32_t mode,
                                                     uint32_t *__restrict__ keys, uint32_t *__restrict__ idx)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 o = rays[2 * (size_t)i];
    const float4 d = rays[2 * (size_t)i + 1];
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    const uint32_t kz = (ax >= ay && ax >= az) ? 0u : (ay >= az ? 1u : 2u);
    // NaN origins map to cell 0 (fmaxf drops the NaN)
    const float qx = fminf(fmaxf((o.x - lx) * sx, 0.0f), 1023.0f);
    const float qy = fminf(fmaxf((o.y - ly) * sy, 0.0f), 1023.0f);
    const float qz = fminf(fmaxf((o.z - lz) * sz, 0.0f), 1023.0f);
    if (mode == 0) {
        const uint32_t code = (expand10((uint32_t)qx) << 2) | (expand10((uint32_t)qy) << 1) | expand10((uint32_t)qz);
        keys[i] = (kz << 30) | code;
    } else {
        const float l1 = ax + ay + az;
        float u = l1 > 0.0f ? d.x / l1 : 0.0f, v = l1 > 0.0f ? d.y / l1 : 0.0f;
        if (d.z < 0.0f) {
            const float uu = (1.0f - fabsf(v)) * (u < 0.0f ? -1.0f : 1.0f);
            const float vv = (1.0f - fabsf(u)) * (v < 0.0f ? -1.0f : 1.0f);
            u = uu;
            v = vv;
        }
        const uint32_t du = (uint32_t)fminf(fmaxf((u + 1.0f) * 4.0f, 0.0f), 7.0f);
        const uint32_t dv = (uint32_t)fminf(fmaxf((v + 1.0f) * 4.0f, 0.0f), 7.0f);
        const uint32_t dir = (spread2(du) << 1) | spread2(dv);
        const uint32_t code = (expand10((uint32_t)qx >> 2) << 2) | (expand10((uint32_t)qy >> 2) << 1) |
                              expand10((uint32_t)qz >> 2);
        keys[i] = (kz << 30) | (dir << 24) | code;
    }
    idx[i] = i;
}

using TraceKernel = void (*)(const QNode *, const float4 *, const float4 *, const uint32_t *, uint32_t,
                            const uint32_t *, dr_hit *,
                            uint32_t *, uint32_t *, uint32_t *, int32_t *, uint32_t *, uint32_t *, uint32_t,
                            const uint32_t *, uint32_t, uint32_t, const InstRec *);

template <bool ANY, bool COUNT, bool INST = false>
TraceKernel pick(uint32_t sched)
{
    if (INST)  // the speculative loops only (the launcher promotes the others)
        return (sched & 2u) ? k_trace<ANY, false, true, true, true> : k_trace<ANY, false, false, true, true>;
    switch (sched & 3u) {
    case DR_SCHED_PER_RAY: return k_trace<ANY, COUNT, false, false>;
    case DR_SCHED_PER_RAY_SPECULATIVE: return k_trace<ANY, COUNT, false, true>;
    case DR_SCHED_PERSISTENT: return k_trace<ANY, COUNT, true, false>;
    default: return k_trace<ANY, COUNT, true, true>;
    }
}

// Blocks per CU of a kernel variant at kBlock threads (occupancy limit),
// looked up once per variant.
template <bool ANY, bool COUNT, bool INST = false>
dr_status blocks_per_cu(uint32_t loop, TraceKernel kern, int &bpc)
{
    static int cache[5][4] = {};
    int &c = cache[INST ? 3 + ANY : (ANY ? 0 : (COUNT ? 1 : 2))][loop & 3u];
    if (!c) {
        int b = 0;
        DR_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kern, kBlock, 0));
        c = b > 0 ? b : 1;
    }
    bpc = c;
    return DR_OK;
}

// ---- zero-allocation closest hit ------------------------------------------
// The scratch of a persistent closest-hit launch (the overflow flag, the ray
// counter and the per-lane stack spill region) in one static device array,
// for callers that cannot allocate. Its launch shape is capped so the array
// covers it: at most kNoAllocBlocks resident blocks, and trees of at most
// kNoAllocMaxDepth wide levels (3 x 21 + 1 = 64 stack entries, 16 of them in
// LDS). Words [0, 64): [0] overflow flag, [8, 16) ray counters; then the
// spill region, kNoAllocSpill entries x kNoAllocBlocks x kBlock lanes.
constexpr uint32_t kNoAllocBlocks = 1024;
constexpr uint32_t kNoAllocMaxDepth = 21;
constexpr uint32_t kNoAllocSpill = 3u * kNoAllocMaxDepth + 1u - (uint32_t)kLdsStack;
constexpr size_t kNoAllocHeadWords = 64;
constexpr size_t kScratchWords = kNoAllocHeadWords + (size_t)kNoAllocSpill * kNoAllocBlocks * kBlock;

}  // namespace

static __device__ __align__(256) int32_t scratch[kScratchWords];  // 48 MiB + 256 B per device

dr_status noalloc_init(Context *ctx)
{
    // the array's name is a host-side stand-in: kernels get this device
    // address instead
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(DR_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    ctx->noalloc_scratch = static_cast<int32_t *>(p);
    return DR_OK;
}

dr_status launch_trace_noalloc(Scene *s, const dr_ray *rays, uint32_t n, dr_hit *hits, hipStream_t stream)
{
    if (!s->built) return fail(DR_E_INVALID, "scene BVH not built (call dr_scene_build_bvh)");
    if (n == 0) return DR_OK;
    if (!rays || !hits) return fail(DR_E_INVALID, "null ray or output buffer");
    if (s->depth > kNoAllocMaxDepth)
        return fail(DR_E_OVERFLOW, "dr_trace_closest_noalloc: the BVH has " + std::to_string(s->depth) +
                                       " wide levels, the static scratch covers at most " +
                                       std::to_string(kNoAllocMaxDepth) + " (use dr_trace_closest)");
    int32_t *base = s->ctx->noalloc_scratch;
    if (!base) return fail(DR_E_DEVICE, "zero-allocation scratch not set up on this context");
    const uint32_t max_stack = std::max<uint32_t>((uint32_t)kLdsStack, 3u * s->depth + 1u);
    // the resident-grid form of the scene's closest-hit schedule, no reordering
    const uint32_t loop = (s->sched_closest & 1u) | 2u;
    const TraceKernel kern = pick<false, false>(loop);
    int occ = 1;
    if (dr_status st = blocks_per_cu<false, false>(loop, kern, occ)) return st;
    const int bpc = s->blocks_per_cu ? std::min<int>((int)s->blocks_per_cu, occ) : occ;
    const uint32_t need = (n + kBlock - 1) / kBlock;
    const uint32_t grid = std::min<uint32_t>({need, (uint32_t)(s->ctx->cu_count * bpc), kNoAllocBlocks});
    uint32_t *flags = reinterpret_cast<uint32_t *>(base);
    uint32_t *counter = flags + 8;
    k_zero_counters8<<<1, 64, 0, stream>>>(counter);
    kern<<<grid, kBlock, 0, stream>>>(s->nodes.as<QNode>(), s->tri_v.as<float4>(),
                                      reinterpret_cast<const float4 *>(rays), nullptr, n, nullptr, hits, nullptr,
                                      nullptr, nullptr, base + kNoAllocHeadWords, flags, counter, s->refill_min,
                                      nullptr, 1u, max_stack, nullptr);
    DR_HIP_TRY(hipGetLastError());
    return DR_OK;
}

dr_status launch_trace(Scene *s, const dr_ray *rays, uint32_t n, dr_hit *hits, uint32_t *occluded,
                       uint32_t *box_tests, uint32_t *tri_tests, bool any_hit, hipStream_t stream)
{
    return launch_trace_ex(s, rays, nullptr, n, nullptr, hits, occluded, box_tests, tri_tests, any_hit,
                           any_hit ? s->sched_any : s->sched_closest, stream);
}

dr_status launch_trace_ex(Scene *s, const dr_ray *rays, const uint32_t *ray_list, uint32_t n_max,
                          const uint32_t *n_dev, dr_hit *hits, uint32_t *occluded, uint32_t *box_tests,
                          uint32_t *tri_tests, bool any_hit, uint32_t sched, hipStream_t stream,
                          const uint32_t *ray_root, DevBuf *ovf_buf)
{
    const uint32_t n = n_max;
    if (!s->built) return fail(DR_E_INVALID, "scene BVH not built (call dr_scene_build_bvh)");
    if (n == 0) return DR_OK;
    if (!rays || (any_hit ? !occluded : !hits)) return fail(DR_E_INVALID, "null ray or output buffer");
    const bool count = box_tests && tri_tests;
    const bool inst = s->n_inst != 0;
    if (inst && (count || ray_root))
        return fail(DR_E_UNSUPPORTED, "instances: no counted traversal and no per-ray roots");
    // stack capacity from the built depth (3 entries per wide level, one slot
    // of slack; for instances the depth counts both levels and the slot holds
    // the marker); a deep tree moves per-ray schedules onto the resident grid so
    // the spill region stays bounded by the number of resident lanes
    const uint32_t max_stack = std::max<uint32_t>((uint32_t)kLdsStack, 3u * s->depth + 1u);
    if (s->depth > kMaxWideDepth)
        return fail(DR_E_OVERFLOW, "BVH deeper than the traversal stack bound of " + std::to_string(kMaxWideDepth) +
                                       " wide levels");
    uint32_t loop = sched & 3u;
    if (max_stack - kLdsStack > kPerRaySpillMax) loop |= 2u;  // PER_RAY -> PERSISTENT, keeping speculation
    if (inst) loop |= 1u;  // instances run the speculative loops
    const TraceKernel kern = inst ? (any_hit ? pick<true, false, true>(loop) : pick<false, false, true>(loop))
                             : any_hit ? pick<true, false>(loop)
                             : (count ? pick<false, true>(loop) : pick<false, false>(loop));
    const bool persistent = loop == DR_SCHED_PERSISTENT || loop == DR_SCHED_PERSISTENT_SPECULATIVE;
    int occ = 1;
    if (dr_status st = inst ? (any_hit ? blocks_per_cu<true, false, true>(loop, kern, occ)
                                       : blocks_per_cu<false, false, true>(loop, kern, occ))
                       : any_hit ? blocks_per_cu<true, false>(loop, kern, occ)
                       : (count ? blocks_per_cu<false, true>(loop, kern, occ)
                                : blocks_per_cu<false, false>(loop, kern, occ)))
        return st;
    const uint32_t need = (n + kBlock - 1) / kBlock;
    const int bpc = s->blocks_per_cu ? std::min<int>((int)s->blocks_per_cu, occ) : occ;
    const uint32_t resident = (uint32_t)(s->ctx->cu_count * bpc);
    const uint32_t grid = persistent ? (need < resident ? need : resident) : need;
    const size_t nlanes = (size_t)grid * kBlock;
    DevBuf &spill = ovf_buf ? *ovf_buf : s->ovf_stack;
    if (dr_status st = spill.ensure(sizeof(int32_t) * (size_t)(max_stack - kLdsStack) * nlanes + 4)) return st;
    const auto *r4 = reinterpret_cast<const float4 *>(rays);
    const uint32_t *perm = ray_list;
    if (!ray_list && !n_dev && !ray_root && (sched & DR_SCHED_REORDER) && n >= 4096) {
        if (dr_status st = s->ray_keys.ensure(sizeof(uint32_t) * (size_t)n)) return st;
        if (dr_status st = s->ray_perm.ensure(sizeof(uint32_t) * (size_t)n)) return st;
        const float *lo = s->bounds, *hi = s->bounds + 3;
        float sc[3];
        for (int a = 0; a < 3; ++a) sc[a] = hi[a] > lo[a] ? 1024.0f / (hi[a] - lo[a]) : 0.0f;
        k_ray_keys<<<need, kBlock, 0, stream>>>(r4, n, lo[0], lo[1], lo[2], sc[0], sc[1], sc[2], s->sort_key,
                                                s->ray_keys.as<uint32_t>(), s->ray_perm.as<uint32_t>());
        DR_HIP_TRY(hipGetLastError());
        if (dr_status st = radix_sort_pairs(s->ctx, s->ray_keys.as<uint32_t>(), s->ray_perm.as<uint32_t>(), n, 32,
                                            stream))
            return st;
        perm = s->ray_perm.as<uint32_t>();
    }
    // flags[0]: overflow; flags[16 ..]: per-launch ray counters (ring)
    uint32_t *flags = s->flags.as<uint32_t>();
    uint32_t *counter = flags + 16 + 8 * (s->launch_seq++ % (kCounterSlots / 8));
    if (persistent) k_zero_counters8<<<1, 64, 0, stream>>>(counter);
    // every slice needs at least one workgroup to drain it
    const uint32_t slices = (persistent && (sched & DR_SCHED_XCD_SLICES) && grid >= 8u) ? 8u : 1u;
    if (count)
        for (int k = 2; k <= 6; ++k) k_zero_counter<<<1, 1, 0, stream>>>(flags + k);
    int32_t *ovf = spill.as<int32_t>();
    kern<<<grid, kBlock, 0, stream>>>(s->nodes.as<QNode>(), s->tri_v.as<float4>(), r4, perm, n, n_dev, hits, occluded,
                                      box_tests, tri_tests, ovf, flags, counter, s->refill_min, ray_root, slices,
                                      max_stack, inst ? s->inst_recs.as<InstRec>() : nullptr);
    DR_HIP_TRY(hipGetLastError());
    return DR_OK;
}

bool trace_pair_fits(const Scene *s, uint32_t sched)
{
    const uint32_t max_stack = std::max<uint32_t>((uint32_t)kLdsStack, 3u * s->depth + 1u);
    return s->built && (sched & 2u) == 0 && s->depth <= kMaxWideDepth && max_stack - kLdsStack <= kPerRaySpillMax;
}

dr_status launch_trace_pair(Scene *s, const dr_ray *rays_a, const uint32_t *na_dev, uint32_t *occluded,
                            const uint32_t *roots_a, const dr_ray *rays_c, const uint32_t *list_c,
                            const uint32_t *nc_dev, dr_hit *hits, const uint32_t *roots_c, uint32_t n_max,
                            uint32_t sched, hipStream_t stream, DevBuf *ovf_buf)
{
    if (!trace_pair_fits(s, sched)) return fail(DR_E_INVALID, "paired trace launch: needs a per-ray schedule");
    if (n_max == 0) return DR_OK;
    if (!rays_a || !na_dev || !occluded || !rays_c || !nc_dev || !hits)
        return fail(DR_E_INVALID, "paired trace launch: null ray, count or output buffer");
    const bool inst = s->n_inst != 0;
    if (inst && (roots_a || roots_c)) return fail(DR_E_UNSUPPORTED, "instances: no per-ray roots");
    const uint32_t max_stack = std::max<uint32_t>((uint32_t)kLdsStack, 3u * s->depth + 1u);
    const uint32_t need = (n_max + kBlock - 1) / kBlock;
    const size_t part = (size_t)(max_stack - kLdsStack) * need * kBlock;  // spill entries of one part
    DevBuf &spill = ovf_buf ? *ovf_buf : s->ovf_stack;
    if (dr_status st = spill.ensure(sizeof(int32_t) * 2 * part + 4)) return st;
    int32_t *ovf = spill.as<int32_t>();
    // instances run the speculative loop (as launch_trace_ex promotes them)
    const auto kern = inst ? k_trace_pair<true, true> : (sched & 1u) ? k_trace_pair<true> : k_trace_pair<false>;
    kern<<<2 * need, kBlock, 0, stream>>>(s->nodes.as<QNode>(), s->tri_v.as<float4>(),
                                          reinterpret_cast<const float4 *>(rays_a), na_dev, occluded, roots_a, need,
                                          ovf, reinterpret_cast<const float4 *>(rays_c), list_c, nc_dev, hits, roots_c,
                                          ovf + part, n_max, s->flags.as<uint32_t>(), max_stack,
                                          inst ? s->inst_recs.as<InstRec>() : nullptr);
    DR_HIP_TRY(hipGetLastError());
    return DR_OK;
}

}  // namespace dr
