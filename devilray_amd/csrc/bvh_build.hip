// On-device BVH build: Morton-code LBVH, SAH leaf collapse, 4-wide nodes.
//
//  1. per-triangle boxes and the centroid bounds (one atomic per block);
//     optionally, triangles with large empty boxes are split into several
//     references with clipped boxes (k_split below)
//  2. 30-bit Morton codes of the reference centroids, radix-sorted with their ids
//  3. binary tree: PLOC clustering of the Morton-ordered leaves (default,
//     k_ploc_* below, leaves renumbered depth-first), or the binary radix tree
//     over the sorted codes (Karras, "Maximizing Parallelism in the
//     Construction of BVHs, Octrees, and k-d Trees", HPG 2012; equal codes are
//     ordered by their position)
//  4. node boxes, reference counts and the SAH cost
//     C(n) = min(Ci*A(n)*N(n), Ct*A(n) + C(l) + C(r)), at each PLOC merge or,
//     for the radix tree, in a bottom-up pass with atomic arrival flags; a subtree is
//     collapsed into one leaf when the leaf term wins and N(n) <= leaf_max
//     (the collapse rule of Karras & Aila, "Fast Parallel Construction of
//     High-Quality BVHs", HPG 2013)
//  5. top-down level-synchronous collapse into 4-wide nodes: each wide node
//     opens its largest-area interior child until it has four children
//  6. triangles are written leaf-ordered as 3 float4 with the
//     caller's triangle id in the first w.
// Child boxes are rounded outward by one ulp so every vertex lies strictly
// inside its leaf box.
#include "dr_internal.hpp"
#include "node_quant.hpp"

#include <cfloat>
#include <cmath>
#include <cstring>

namespace dr {

namespace {

constexpr int kT = 256;

struct BuildTmp {
    DevBuf tri_lo, tri_hi;          // float4 per triangle (input order)
    DevBuf ref_lo, ref_hi, ref_tri; // split references: box and triangle
    DevBuf offs;                    // split: reference offsets per triangle
    const float4 *rlo = nullptr, *rhi = nullptr;  // reference boxes (the triangle boxes without splitting)
    const uint32_t *rtri = nullptr;               // reference -> triangle (null: identity)
    DevBuf keys, vals;              // sorted codes / ids
    DevBuf leaf_lo, leaf_hi;        // float4 per leaf (sorted order)
    DevBuf child_l, child_r;        // int per internal node (<0: ~leaf)
    DevBuf parent_i, parent_l;      // parent of internal nodes / leaves
    DevBuf range_first, count;      // per internal node
    DevBuf node_lo, node_hi;        // float4 per internal node
    DevBuf cost;                    // float per internal node
    DevBuf visit;                   // u32 arrival flags
    DevBuf collapse;                // u8-in-u32 per internal node
    DevBuf queue_a, queue_b;        // int2 (binary node, wide node)
    DevBuf counters;                // u32[8] + ordered-int centroid bounds[6] at 8, of references at 16
};

__device__ __forceinline__ int32_t f2ord(float f)
{
    const int32_t i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float ord2f(int32_t i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

__device__ __forceinline__ int longest_axis(const float (&tl)[3], const float (&th)[3])
{
    int ax = 0;
    for (int a = 1; a < 3; ++a)
        if (th[a] - tl[a] > th[ax] - tl[ax]) ax = a;
    return ax;
}

// Block-wide min (c) and max (C) of a centroid box, then one atomic per
// component per block: same-address atomics serialise at the L2, so one per
// wave over a million triangles cost milliseconds. Every thread of the block
// must call it.
__device__ __forceinline__ void block_bounds_atomic(float (&c)[3], float (&C)[3], int32_t *__restrict__ bounds)
{
    __shared__ float red[6][kT / 64];
    for (int a = 0; a < 3; ++a) {
        for (int off = 32; off >= 1; off >>= 1) {
            c[a] = fminf(c[a], __shfl_xor(c[a], off, 64));
            C[a] = fmaxf(C[a], __shfl_xor(C[a], off, 64));
        }
    }
    const int w = (int)(threadIdx.x >> 6);
    if ((threadIdx.x & 63u) == 0) {
        for (int a = 0; a < 3; ++a) {
            red[a][w] = c[a];
            red[3 + a][w] = C[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = (int)threadIdx.x;
        float v = red[a][0];
        for (int k = 1; k < kT / 64; ++k) v = a < 3 ? fminf(v, red[a][k]) : fmaxf(v, red[a][k]);
        if (a < 3) atomicMin(&bounds[a], f2ord(v));
        else atomicMax(&bounds[a], f2ord(v));
    }
}

// grid of the bounds kernels: a few blocks per CU, each thread striding over
// the triangles
constexpr uint32_t kBoundsBlocks = 2048;

__global__ __launch_bounds__(kT) void k_tri_bounds(const float *__restrict__ verts,
                                                   const uint32_t *__restrict__ idx, uint32_t n,
                                                   float4 *__restrict__ lo, float4 *__restrict__ hi,
                                                   int32_t *__restrict__ cbounds)
{
    float c[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, C[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
        float l[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, h[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int k = 0; k < 3; ++k) {
            const uint32_t vi = idx[3 * (size_t)i + k];
            for (int a = 0; a < 3; ++a) {
                const float x = verts[3 * (size_t)vi + a];
                l[a] = fminf(l[a], x);
                h[a] = fmaxf(h[a], x);
            }
        }
        // A triangle with a non-finite coordinate can never be hit (the
        // triangle test rejects it): it gets an empty box, which no ray
        // enters and which merges away, and stays out of the centroid bounds.
        bool finite = true;
        for (int a = 0; a < 3; ++a) finite = finite && isfinite(l[a]) && isfinite(h[a]);
        if (!finite)
            for (int a = 0; a < 3; ++a) { l[a] = INFINITY; h[a] = -INFINITY; }
        lo[i] = make_float4(l[0], l[1], l[2], 0.0f);
        hi[i] = make_float4(h[0], h[1], h[2], 0.0f);
        if (finite)
            for (int a = 0; a < 3; ++a) {
                const float m = 0.5f * (l[a] + h[a]);
                c[a] = fminf(c[a], m);
                C[a] = fmaxf(C[a], m);
            }
    }
    block_bounds_atomic(c, C, cbounds);
}

__device__ __forceinline__ uint32_t expand10(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ __launch_bounds__(kT) void k_morton(const float4 *__restrict__ lo, const float4 *__restrict__ hi,
                                               uint32_t n, const int32_t *__restrict__ cbounds,
                                               uint32_t *__restrict__ keys, uint32_t *__restrict__ vals)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const float4 l = lo[i], h = hi[i];
    const float c[3] = {0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z)};
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
        const float bl = ord2f(cbounds[a]), bh = ord2f(cbounds[3 + a]);
        const float ext = bh - bl;
        float x = ext > 0.0f ? (c[a] - bl) / ext : 0.0f;
        x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
        q[a] = (uint32_t)x;
    }
    keys[i] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
    vals[i] = i;
}

// leaf boxes and leaf-ordered triangle vertices; vals[k] is a reference,
// whose triangle is ref_tri[vals[k]] (or vals[k] itself without a map). The w
// components carry the triangle id and, for a refit, which piece of how many
// the reference is, with the axis of its slabs in bits 16..17 of the piece
// (offs: the split's reference offsets per triangle).
__global__ __launch_bounds__(kT) void k_gather(const float *__restrict__ verts, const uint32_t *__restrict__ idx,
                                               const uint32_t *__restrict__ vals, uint32_t n,
                                               const float4 *__restrict__ tlo, const float4 *__restrict__ thi,
                                               const uint32_t *__restrict__ ref_tri, const uint32_t *__restrict__ offs,
                                               float4 *__restrict__ llo, float4 *__restrict__ lhi,
                                               float4 *__restrict__ tri_v)
{
    const uint32_t k = blockIdx.x * kT + threadIdx.x;
    if (k >= n) return;
    const uint32_t r = vals[k];
    const uint32_t id = ref_tri ? ref_tri[r] : r;
    const uint32_t pieces = (ref_tri && offs) ? offs[id + 1] - offs[id] : 1u;
    uint32_t piece = 0;
    if (pieces > 1) {  // piece index and the axis its slabs cut (the split's longest axis)
        float tl[3] = {INFINITY, INFINITY, INFINITY}, th[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int j = 0; j < 3; ++j) {
            const uint32_t vi = idx[3 * (size_t)id + j];
            for (int a = 0; a < 3; ++a) {
                tl[a] = fminf(tl[a], verts[3 * (size_t)vi + a]);
                th[a] = fmaxf(th[a], verts[3 * (size_t)vi + a]);
            }
        }
        piece = (r - offs[id]) | ((uint32_t)longest_axis(tl, th) << 16);
    }
    llo[k] = tlo[r];
    lhi[k] = thi[r];
    for (int j = 0; j < 3; ++j) {
        const uint32_t vi = idx[3 * (size_t)id + j];
        const float w = __int_as_float((int32_t)(j == 0 ? id : (j == 1 ? piece : pieces)));
        tri_v[3 * (size_t)k + j] = make_float4(verts[3 * (size_t)vi], verts[3 * (size_t)vi + 1],
                                               verts[3 * (size_t)vi + 2], w);
    }
}

__device__ __forceinline__ int delta(const uint32_t *__restrict__ codes, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    const uint32_t a = codes[i], b = codes[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

__global__ __launch_bounds__(kT) void k_karras(const uint32_t *__restrict__ codes, int n,
                                               int32_t *__restrict__ cl, int32_t *__restrict__ cr,
                                               int32_t *__restrict__ pi, int32_t *__restrict__ pl,
                                               uint32_t *__restrict__ rfirst, uint32_t *__restrict__ cnt)
{
    const int i = (int)(blockIdx.x * kT + threadIdx.x);
    if (i >= n - 1) return;
    const int d = (delta(codes, n, i, i + 1) - delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(codes, n, i, i - d);
    int lmax = 2;
    while (delta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(codes, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int first = min(i, j), last = max(i, j);
    const int32_t left = (first == gamma) ? ~gamma : gamma;
    const int32_t right = (last == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    cl[i] = left;
    cr[i] = right;
    if (left < 0) pl[~left] = i; else pi[left] = i;
    if (right < 0) pl[~right] = i; else pi[right] = i;
    rfirst[i] = (uint32_t)first;
    cnt[i] = (uint32_t)(last - first + 1);
    if (i == 0) pi[0] = -1;
}

__device__ __forceinline__ float half_area(float4 l, float4 h)
{
    const float dx = h.x - l.x, dy = h.y - l.y, dz = h.z - l.z;
    return dx * dy + dy * dz + dz * dx;
}

__global__ __launch_bounds__(kT) void k_fit(int n, const int32_t *__restrict__ cl, const int32_t *__restrict__ cr,
                                            const int32_t *__restrict__ pi, const int32_t *__restrict__ pl,
                                            const uint32_t *__restrict__ cnt, const float4 *__restrict__ llo,
                                            const float4 *__restrict__ lhi, float4 *node_lo, float4 *node_hi,
                                            float *cost, uint32_t *visit, uint32_t *collapse,
                                            uint32_t leaf_max, float ct, float ci)
{
    const int k = (int)(blockIdx.x * kT + threadIdx.x);
    if (k >= n) return;
    int node = pl[k];
    while (node >= 0) {
        __threadfence();
        if (atomicAdd(&visit[node], 1u) == 0u) return;  // first arrival: sibling not ready
        __threadfence();
        const int32_t a = cl[node], b = cr[node];
        float4 alo, ahi, blo, bhi;
        float ac, bc;
        if (a < 0) { alo = llo[~a]; ahi = lhi[~a]; ac = ci * half_area(alo, ahi); }
        else { alo = node_lo[a]; ahi = node_hi[a]; ac = cost[a]; }
        if (b < 0) { blo = llo[~b]; bhi = lhi[~b]; bc = ci * half_area(blo, bhi); }
        else { blo = node_lo[b]; bhi = node_hi[b]; bc = cost[b]; }
        const float4 lo = make_float4(fminf(alo.x, blo.x), fminf(alo.y, blo.y), fminf(alo.z, blo.z), 0.0f);
        const float4 hi = make_float4(fmaxf(ahi.x, bhi.x), fmaxf(ahi.y, bhi.y), fmaxf(ahi.z, bhi.z), 0.0f);
        const float area = half_area(lo, hi);
        const uint32_t N = cnt[node];
        const float c_split = ct * area + ac + bc;
        const float c_leaf = (N <= leaf_max) ? ci * area * (float)N : INFINITY;
        // N <= leaf_max as well: with overflowing areas both costs are infinite
        const bool col = N <= leaf_max && c_leaf <= c_split;
        node_lo[node] = lo;
        node_hi[node] = hi;
        cost[node] = col ? c_leaf : c_split;
        collapse[node] = col ? 1u : 0u;
        node = pi[node];
    }
}

__device__ __forceinline__ float down(float x) { return nextafterf(x, -INFINITY); }
__device__ __forceinline__ float up(float x) { return nextafterf(x, INFINITY); }

__global__ __launch_bounds__(kT) void k_collapse(
    const int2 *__restrict__ qin, uint32_t nin, int2 *__restrict__ qout, uint32_t *counters,
    const int32_t *__restrict__ cl, const int32_t *__restrict__ cr, const uint32_t *__restrict__ rfirst,
    const uint32_t *__restrict__ cnt, const float4 *__restrict__ node_lo, const float4 *__restrict__ node_hi,
    const float4 *__restrict__ llo, const float4 *__restrict__ lhi, const uint32_t *__restrict__ collapse,
    WideNode *__restrict__ wide, uint32_t wide_cap)
{
    const uint32_t qi = blockIdx.x * kT + threadIdx.x;
    const bool valid = qi < nin;  // every lane stays for the wave-wide allocation below
    int2 e = make_int2(0, 0);
    int32_t kids[4] = {0, 0, 0, 0};
    int nk = 0;
    if (valid) {
        e = qin[qi];
        kids[0] = cl[e.x];
        kids[1] = cr[e.x];
        nk = 2;
    }
    while (valid && nk < 4) {
        int best = -1;
        float best_a = -1.0f;
        for (int c = 0; c < nk; ++c) {
            const int32_t k = kids[c];
            if (k < 0 || collapse[k]) continue;
            const float a = half_area(node_lo[k], node_hi[k]);
            if (a > best_a) { best_a = a; best = c; }
        }
        if (best < 0) break;
        const int32_t k = kids[best];
        kids[best] = cl[k];
        kids[nk++] = cr[k];
    }
    // Wide-node slots and queue entries for the interior children, leaf
    // counts: one atomic per wave and counter (a prefix sum over the lanes
    // hands out the slots); per-child atomics on the same three counters
    // serialised at the L2 and made the widest levels take milliseconds.
    uint32_t n_int = 0, n_leaf = 0;
    for (int c = 0; c < nk; ++c) {
        const int32_t k = kids[c];
        if (k < 0 || collapse[k]) ++n_leaf;
        else ++n_int;
    }
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t incl = n_int;
    for (uint32_t d = 1; d < 64u; d <<= 1) {
        const uint32_t v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
    }
    const uint32_t total = __shfl(incl, 63, 64);
    uint32_t leaves = n_leaf;
    for (int off = 32; off >= 1; off >>= 1) leaves += __shfl_xor(leaves, off, 64);
    uint32_t bw = 0, bq = 0;
    if (lane == 0) {
        if (total) {
            bw = atomicAdd(&counters[0], total);
            bq = atomicAdd(&counters[1], total);
        }
        if (leaves) atomicAdd(&counters[2], leaves);
    }
    uint32_t w = __shfl(bw, 0, 64) + incl - n_int;
    uint32_t q = __shfl(bq, 0, 64) + incl - n_int;
    if (!valid) return;
    WideNode wn;
    for (int c = 0; c < 4; ++c) {
        if (c >= nk) {
            wn.lox[c] = wn.hix[c] = wn.loy[c] = wn.hiy[c] = wn.loz[c] = wn.hiz[c] = INFINITY;
            wn.child[c] = DR_EMPTY_REF;
            continue;
        }
        const int32_t k = kids[c];
        float4 lo, hi;
        int32_t ref;
        if (k < 0) {
            lo = llo[~k]; hi = lhi[~k];
            ref = make_leaf_ref((uint32_t)~k, 1u);
        } else {
            lo = node_lo[k]; hi = node_hi[k];
            if (collapse[k]) {
                ref = make_leaf_ref(rfirst[k], cnt[k]);
            } else {
                if (w >= wide_cap) {  // cannot happen by the bound above; flag instead of writing
                    atomicOr(&counters[3], 1u);
                    ref = DR_EMPTY_REF;
                } else {
                    qout[q] = make_int2(k, (int)w);
                    ref = (int32_t)w;
                }
                ++w;
                ++q;
            }
        }
        wn.lox[c] = down(lo.x); wn.loy[c] = down(lo.y); wn.loz[c] = down(lo.z);
        wn.hix[c] = up(hi.x); wn.hiy[c] = up(hi.y); wn.hiz[c] = up(hi.z);
        wn.child[c] = ref;
    }
    wn.pad[0] = nk; wn.pad[1] = wn.pad[2] = wn.pad[3] = 0;
    wide[e.y] = wn;
}

// Root wide node with a single leaf child covering all triangles (n <= leaf
// cap, or a collapsed root).
__global__ void k_single_leaf_root(const float4 *__restrict__ lo, const float4 *__restrict__ hi, uint32_t n,
                                   WideNode *__restrict__ wide, const float4 *__restrict__ root_lo,
                                   const float4 *__restrict__ root_hi)
{
    WideNode wn;
    const float4 l = root_lo ? *root_lo : lo[0];
    const float4 h = root_hi ? *root_hi : hi[0];
    for (int c = 0; c < 4; ++c) {
        wn.lox[c] = wn.hix[c] = wn.loy[c] = wn.hiy[c] = wn.loz[c] = wn.hiz[c] = INFINITY;
        wn.child[c] = DR_EMPTY_REF;
        wn.pad[c] = 0;
    }
    wn.lox[0] = down(l.x); wn.loy[0] = down(l.y); wn.loz[0] = down(l.z);
    wn.hix[0] = up(h.x); wn.hiy[0] = up(h.y); wn.hiz[0] = up(h.z);
    wn.child[0] = make_leaf_ref(0u, n);
    wn.pad[0] = 1;
    *wide = wn;
}

// QNode copies of the wide nodes; *err is set if a node's planes are out of
// range (a vertex coordinate of magnitude 2^126 or more).
__global__ __launch_bounds__(kT) void k_quantize(const WideNode *__restrict__ src, uint32_t n, QNode *__restrict__ dst,
                                                 uint32_t *__restrict__ err)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const WideNode w = src[i];
    QNode q;
    bool use[4];
    for (int c = 0; c < 4; ++c)
        use[c] = w.child[c] != DR_EMPTY_REF && w.lox[c] <= w.hix[c] && w.loy[c] <= w.hiy[c] && w.loz[c] <= w.hiz[c];
    bool ok = quantize_axis(w.lox, w.hix, use, q.origin[0], q.step[0], q.qlo[0], q.qhi[0]);
    ok = quantize_axis(w.loy, w.hiy, use, q.origin[1], q.step[1], q.qlo[1], q.qhi[1]) && ok;
    ok = quantize_axis(w.loz, w.hiz, use, q.origin[2], q.step[2], q.qlo[2], q.qhi[2]) && ok;
    for (int c = 0; c < 4; ++c) q.child[c] = w.child[c];
    dst[i] = q;
    if (!ok) atomicOr(err, 1u);
}

__global__ void k_init_counters(uint32_t *c, int32_t *cb, int32_t *rb)
{
    const int t = threadIdx.x;
    if (t < 8) c[t] = 0;
    if (t < 3) {
        cb[t] = 0x7FFFFFFF; cb[3 + t] = (int32_t)0x80000000;
        rb[t] = 0x7FFFFFFF; rb[3 + t] = (int32_t)0x80000000;
    }
}

// counters: [0] wide nodes allocated (the root), [1] queue out, [2] leaves, [3] error
__global__ void k_collapse_init(uint32_t *c, int2 *queue)
{
    c[0] = 1u; c[1] = 0u; c[2] = 0u; c[3] = 0u;
    queue[0] = make_int2(0, 0);
}

__global__ void k_reset_queue_count(uint32_t *c) { c[1] = 0u; }

__global__ __launch_bounds__(kT) void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i < n) p[i] = v;
}

inline uint32_t blocks(uint32_t n) { return (n + kT - 1) / kT; }

// ---- Triangle reference splitting: a long thin triangle at an angle to the
// axes has a box that is mostly empty, so rays that pass near it test it. Such
// triangles are cut into pieces before the build, each piece bounded by the box
// of the triangle clipped to its cell ("early split clipping": Ernst &
// Greiner, RT 2007). The piece count follows the triangle's shape: its box's
// half area over the half area that boxes of ever finer pieces tend to,
// A (|nx| + |ny| + |nz|). That ratio is 2 to 3 for well-shaped triangles and 5
// to 50 for the slivers of a displaced fine mesh; a triangle gets
// floor(ratio / r0) pieces, 1 to kSplitMax, and a cap on the extra references
// scales the counts down if needed. Every piece refers to the whole triangle,
// so hits are those of the unsplit scene.
constexpr int kSplitMax = 16;  // pieces per triangle
// Defaults: r0 = 2 (8 quarters) and a cap of 8 extra references per triangle
// (32 quarters). On the displaced 1.3M-triangle icosphere this makes about 7.7
// references per triangle and cuts triangle tests per incoherent ray from 163
// to about 16; scenes of well-shaped triangles are left unsplit.
constexpr uint32_t kSplitRatioDefaultQuarters = 8;
constexpr uint32_t kSplitCapDefaultQuarters = 32;
constexpr uint32_t kMaxRefs = kMaxLeafRefs;

// pieces per triangle; cnt[n] = 0 (the scan then gives offsets and the total)
__global__ __launch_bounds__(kT) void k_split_counts(const float *__restrict__ verts, const uint32_t *__restrict__ idx,
                                                     uint32_t n, const float4 *__restrict__ tlo,
                                                     const float4 *__restrict__ thi, float r0,
                                                     uint32_t *__restrict__ cnt)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i > n) return;
    if (i == n) { cnt[n] = 0u; return; }
    float v[3][3];
    for (int k = 0; k < 3; ++k) {
        const uint32_t vi = idx[3 * (size_t)i + k];
        for (int a = 0; a < 3; ++a) v[k][a] = verts[3 * (size_t)vi + a];
    }
    const float e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
    const float e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
    const float cx = e1[1] * e2[2] - e1[2] * e2[1];
    const float cy = e1[2] * e2[0] - e1[0] * e2[2];
    const float cz = e1[0] * e2[1] - e1[1] * e2[0];
    const float ideal = 0.5f * (fabsf(cx) + fabsf(cy) + fabsf(cz));
    const float ratio = half_area(tlo[i], thi[i]) / ideal;  // inf or NaN for degenerate triangles
    uint32_t k = 1u;
    if (ratio >= 2.0f * r0 && ratio <= FLT_MAX) k = (uint32_t)fminf(floorf(ratio / r0), (float)kSplitMax);
    cnt[i] = k;
}

// cap: k -> 1 + floor((k - 1) * cap / extra) when the extra pieces exceed the cap
__global__ __launch_bounds__(kT) void k_split_rescale(uint32_t *__restrict__ cnt, uint32_t n, double scale)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    cnt[i] = 1u + (uint32_t)floor((double)(cnt[i] - 1u) * scale);
}

__device__ __forceinline__ float round_down(double d)
{
    float f = (float)d;
    if ((double)f > d) f = nextafterf(f, -INFINITY);
    return f;
}
__device__ __forceinline__ float round_up(double d)
{
    float f = (float)d;
    if ((double)f < d) f = nextafterf(f, INFINITY);
    return f;
}

// Cuts triangle i into k = offs[i+1] - offs[i] pieces: k slabs of equal
// width across the longest axis of its box (for the slivers that get split,
// the cut that shrinks the pieces' boxes most). The part of a triangle inside
// a closed slab is a convex polygon whose corners are the triangle's vertices
// inside the slab and the crossings of its edges with the two planes, so the
// box of those points is the piece's box; crossings are computed in double and
// rounded outward to float. The box is then widened by a margin of 2^-17 of the
// scene's largest coordinate (so rays that meet the triangle on a cut are
// inside the boxes on both sides whatever the rounding of the triangle test)
// and cut back to the triangle's own box. Triangles left whole keep their
// exact box.
__device__ __forceinline__ float slab_cut(float lo, float hi, uint32_t j, uint32_t k)
{
    return j == 0 ? lo : (j == k ? hi : lo + (hi - lo) * ((float)j / (float)k));
}

// Box of piece j of k of a triangle (vertices v in FP64, box [tl, th]): slab
// j across axis ax, as described above.
__device__ void piece_box(const double (&v)[3][3], const float (&tl)[3], const float (&th)[3], int ax, uint32_t j,
                          uint32_t k, float margin, float (&plo)[3], float (&phi)[3])
{
    const double s0 = slab_cut(tl[ax], th[ax], j, k), s1 = slab_cut(tl[ax], th[ax], j + 1, k);
    double l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
    auto add = [&](const double *p) {
        for (int a = 0; a < 3; ++a) { l[a] = fmin(l[a], p[a]); h[a] = fmax(h[a], p[a]); }
    };
    for (int q = 0; q < 3; ++q)
        if (v[q][ax] >= s0 && v[q][ax] <= s1) add(v[q]);
    for (int q = 0; q < 3; ++q) {
        const double *a0 = v[q], *b0 = v[q == 2 ? 0 : q + 1];
        for (int e = 0; e < 2; ++e) {
            const double sp = e ? s1 : s0;
            const double da = a0[ax] - sp, db = b0[ax] - sp;
            if ((da < 0.0 && db > 0.0) || (da > 0.0 && db < 0.0)) {
                const double f = da / (da - db);
                double p[3];
                for (int a = 0; a < 3; ++a) p[a] = a0[a] + f * (b0[a] - a0[a]);
                p[ax] = sp;
                add(p);
            }
        }
    }
    for (int a = 0; a < 3; ++a) {
        if (l[a] <= h[a]) {
            plo[a] = fmaxf(round_down(l[a]) - margin, tl[a]);
            phi[a] = fminf(round_up(h[a]) + margin, th[a]);
        } else {  // no point found (cannot happen for a slab inside the box): the whole box
            plo[a] = tl[a];
            phi[a] = th[a];
        }
    }
}

__global__ __launch_bounds__(kT) void k_split(const float *__restrict__ verts, const uint32_t *__restrict__ idx,
                                              uint32_t n, const float4 *__restrict__ tlo,
                                              const float4 *__restrict__ thi, const uint32_t *__restrict__ offs,
                                              const int32_t *__restrict__ cbounds, float4 *__restrict__ rlo,
                                              float4 *__restrict__ rhi, uint32_t *__restrict__ rtri,
                                              int32_t *__restrict__ rbounds)
{
    float c[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, C[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = blockIdx.x * kT + threadIdx.x; i < n; i += gridDim.x * kT) {
        const uint32_t base = offs[i], k = offs[i + 1] - base;
        const float4 L = tlo[i], H = thi[i];
        const float tl[3] = {L.x, L.y, L.z}, th[3] = {H.x, H.y, H.z};
        if (k == 1) {
            rlo[base] = L;
            rhi[base] = H;
            rtri[base] = i;
            for (int a = 0; a < 3; ++a) {
                const float m = 0.5f * (tl[a] + th[a]);
                c[a] = fminf(c[a], m);
                C[a] = fmaxf(C[a], m);
            }
        } else {
            double v[3][3];
            for (int q = 0; q < 3; ++q) {
                const uint32_t vi = idx[3 * (size_t)i + q];
                for (int a = 0; a < 3; ++a) v[q][a] = verts[3 * (size_t)vi + a];
            }
            float mag = 0.0f;
            for (int a = 0; a < 6; ++a) mag = fmaxf(mag, fabsf(ord2f(cbounds[a])));
            for (int a = 0; a < 3; ++a) mag = fmaxf(mag, fmaxf(fabsf(tl[a]), fabsf(th[a])));
            const float margin = mag * 0x1p-17f;
            const int ax = longest_axis(tl, th);
            for (uint32_t j = 0; j < k; ++j) {
                float plo[3], phi[3];
                piece_box(v, tl, th, ax, j, k, margin, plo, phi);
                rlo[base + j] = make_float4(plo[0], plo[1], plo[2], 0.0f);
                rhi[base + j] = make_float4(phi[0], phi[1], phi[2], 0.0f);
                rtri[base + j] = i;
                for (int a = 0; a < 3; ++a) {
                    const float m = 0.5f * (plo[a] + phi[a]);
                    c[a] = fminf(c[a], m);
                    C[a] = fmaxf(C[a], m);
                }
            }
        }
    }
    block_bounds_atomic(c, C, rbounds);
}

// ---- PLOC binary tree (Meister & Bittner, "Parallel Locally-Ordered
// Clustering for Bounding Volume Hierarchy Construction", IEEE TVCG 2018).
// Clusters stay in Morton order. Each round every cluster finds the cluster
// within +-r positions whose merged box has the least surface area; mutual
// nearest neighbours merge, the rest carry over, and a scan compacts the
// list. Pairs are strictly ordered (area first, ties broken by the indices),
// so the smallest pair of a round is always mutual and every round merges at
// least one pair. Internal node ids are handed out from n-2 downwards, so the last
// merge is node 0 (the root) and a child always has a larger id than its
// parent. Leaves are then renumbered in depth-first order so each subtree
// covers a contiguous range, which the SAH collapse needs.
constexpr int kPlocRMax = 64;

__global__ __launch_bounds__(kT) void k_ploc_init(uint32_t n, int32_t *__restrict__ ref, uint32_t *__restrict__ cnt,
                                                  int32_t *__restrict__ pi)
{
    const uint32_t k = blockIdx.x * kT + threadIdx.x;
    if (k >= n) return;
    ref[k] = ~(int32_t)k;
    cnt[k] = 1u;
    if (k == 0) pi[0] = -1;
}

__global__ __launch_bounds__(kT) void k_ploc_nn(const float4 *__restrict__ lo, const float4 *__restrict__ hi,
                                                uint32_t m, int r, int32_t *__restrict__ nn)
{
    __shared__ float4 slo[kT + 2 * kPlocRMax], shi[kT + 2 * kPlocRMax];
    const int base = (int)(blockIdx.x * kT) - r;
    for (int t = threadIdx.x; t < kT + 2 * r; t += kT) {
        const int g = base + t;
        if (g >= 0 && g < (int)m) { slo[t] = lo[g]; shi[t] = hi[g]; }
    }
    __syncthreads();
    const int i = (int)(blockIdx.x * kT + threadIdx.x);
    if (i >= (int)m) return;
    const float4 a = slo[i - base], b = shi[i - base];
    const int j0 = i - r > 0 ? i - r : 0;
    const int j1 = i + r < (int)m - 1 ? i + r : (int)m - 1;
    // Pair order: (area, index distance, parity of the lower index, lower
    // index). On exact ties this pairs (0,1), (2,3), ... instead of chaining
    // every cluster to its left neighbour, which would merge one pair a round.
    float best = 0.0f;
    int bj = -1, bdist = 0, blow = 0;
    for (int j = j0; j <= j1; ++j) {
        if (j == i) continue;
        const float4 c = slo[j - base], d = shi[j - base];
        const float4 l = make_float4(fminf(a.x, c.x), fminf(a.y, c.y), fminf(a.z, c.z), 0.0f);
        const float4 h = make_float4(fmaxf(b.x, d.x), fmaxf(b.y, d.y), fmaxf(b.z, d.z), 0.0f);
        const float area = half_area(l, h);
        const int dist = j > i ? j - i : i - j;
        const int low = j < i ? j : i;
        const bool better =
            bj < 0 || area < best ||
            (area == best && (dist < bdist || (dist == bdist && ((low & 1) < (blow & 1) ||
                                                                 ((low & 1) == (blow & 1) && low < blow)))));
        if (better) { best = area; bj = j; bdist = dist; blow = low; }
    }
    nn[i] = bj;
}

// f[i] = (merges here << 32) | (cluster i survives); f[m] = 0 for the totals
__global__ __launch_bounds__(kT) void k_ploc_flags(const int32_t *__restrict__ nn, uint32_t m,
                                                   uint64_t *__restrict__ f)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i > m) return;
    if (i == m) { f[m] = 0; return; }
    const int32_t j = nn[i];
    const bool mutual = nn[j] == (int32_t)i;
    const bool merge = mutual && (int32_t)i < j;
    const bool keep = !mutual || (int32_t)i < j;
    f[i] = ((uint64_t)(merge ? 1u : 0u) << 32) | (keep ? 1u : 0u);
}

__global__ __launch_bounds__(kT) void k_ploc_emit(
    uint32_t m, const int32_t *__restrict__ nn, const uint64_t *__restrict__ f, const int32_t *__restrict__ ref_in,
    const float4 *__restrict__ lo_in, const float4 *__restrict__ hi_in, const uint32_t *__restrict__ cnt_in,
    int32_t *__restrict__ ref_out, float4 *__restrict__ lo_out, float4 *__restrict__ hi_out,
    uint32_t *__restrict__ cnt_out, int32_t *__restrict__ cl, int32_t *__restrict__ cr, int32_t *__restrict__ pi,
    int32_t *__restrict__ pl, uint32_t *__restrict__ ncnt, uint32_t id_base, float4 *__restrict__ node_lo,
    float4 *__restrict__ node_hi, float *__restrict__ cost, uint32_t *__restrict__ collapse, uint32_t leaf_max,
    float ct, float ci)
{
    const int32_t i = (int32_t)(blockIdx.x * kT + threadIdx.x);
    if (i >= (int32_t)m) return;
    const int32_t j = nn[i];
    const bool mutual = nn[j] == i;
    if (mutual && i > j) return;
    const uint64_t fi = f[i];
    const uint32_t p = (uint32_t)fi;
    if (!mutual) {
        ref_out[p] = ref_in[i];
        lo_out[p] = lo_in[i];
        hi_out[p] = hi_in[i];
        cnt_out[p] = cnt_in[i];
        return;
    }
    const int32_t id = (int32_t)(id_base - (uint32_t)(fi >> 32));
    const int32_t a = ref_in[i], b = ref_in[j];
    cl[id] = a;
    cr[id] = b;
    if (a < 0) pl[~a] = id; else pi[a] = id;
    if (b < 0) pl[~b] = id; else pi[b] = id;
    const float4 al = lo_in[i], ah = hi_in[i], bl = lo_in[j], bh = hi_in[j];
    const uint32_t c = cnt_in[i] + cnt_in[j];
    const float4 lo = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.0f);
    const float4 hi = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.0f);
    // The SAH fit of k_fit, done here: both children are finished subtrees.
    const float ac = a < 0 ? ci * half_area(al, ah) : cost[a];
    const float bc = b < 0 ? ci * half_area(bl, bh) : cost[b];
    const float area = half_area(lo, hi);
    const float c_split = ct * area + ac + bc;
    const float c_leaf = (c <= leaf_max) ? ci * area * (float)c : INFINITY;
    const bool col = c <= leaf_max && c_leaf <= c_split;
    node_lo[id] = lo;
    node_hi[id] = hi;
    cost[id] = col ? c_leaf : c_split;
    collapse[id] = col ? 1u : 0u;
    ncnt[id] = c;
    ref_out[p] = id;
    lo_out[p] = lo;
    hi_out[p] = hi;
    cnt_out[p] = c;
}

// Depth-first first-leaf position of every node: the counts of the left
// siblings on the way up to the root. Internal nodes x < n-1 get rfirst[x],
// leaf k (x = n-1+k, Morton position) gets pos[k].
__global__ __launch_bounds__(kT) void k_ploc_first(uint32_t n, const int32_t *__restrict__ cl,
                                                   const int32_t *__restrict__ cr, const int32_t *__restrict__ pi,
                                                   const int32_t *__restrict__ pl, const uint32_t *__restrict__ cnt,
                                                   uint32_t *__restrict__ rfirst, uint32_t *__restrict__ pos)
{
    const uint32_t x = blockIdx.x * kT + threadIdx.x;
    if (x >= 2 * n - 1) return;
    const bool leaf = x >= n - 1;
    int32_t me = leaf ? ~(int32_t)(x - (n - 1)) : (int32_t)x;
    int32_t p = leaf ? pl[x - (n - 1)] : pi[x];
    uint32_t first = 0;
    while (p >= 0) {
        if (cr[p] == me) {
            const int32_t l = cl[p];
            first += l < 0 ? 1u : cnt[l];
        }
        me = p;
        p = pi[p];
    }
    if (leaf) pos[x - (n - 1)] = first; else rfirst[x] = first;
}

__global__ __launch_bounds__(kT) void k_ploc_relabel_nodes(uint32_t ni, int32_t *__restrict__ cl,
                                                           int32_t *__restrict__ cr, const uint32_t *__restrict__ pos)
{
    const uint32_t x = blockIdx.x * kT + threadIdx.x;
    if (x >= ni) return;
    const int32_t a = cl[x], b = cr[x];
    if (a < 0) cl[x] = ~(int32_t)pos[~a];
    if (b < 0) cr[x] = ~(int32_t)pos[~b];
}

__global__ __launch_bounds__(kT) void k_ploc_relabel_leaves(uint32_t n, const uint32_t *__restrict__ vals,
                                                            const int32_t *__restrict__ pl_morton,
                                                            const uint32_t *__restrict__ pos,
                                                            uint32_t *__restrict__ vals_dfs,
                                                            int32_t *__restrict__ pl_dfs)
{
    const uint32_t k = blockIdx.x * kT + threadIdx.x;
    if (k >= n) return;
    const uint32_t p = pos[k];
    vals_dfs[p] = vals[k];
    pl_dfs[p] = pl_morton[k];
}

// Builds the binary tree over the n >= 2 leaves whose boxes t.leaf_lo/hi hold
// in Morton order, into the arrays the collapse reads (child_l/r, parent_i/l,
// range_first, count, and the boxes, SAH costs and collapse flags that k_fit
// computes for the radix tree), and re-gathers the leaf boxes and triangles
// in depth-first order.
dr_status build_ploc(Scene *s, BuildTmp &t, uint32_t n, int radius, uint32_t leaf_max, float ct, float ci,
                     hipStream_t stream)
{
    Context *ctx = s->ctx;
    DevBuf ref[2], lo1, hi1, cnt[2], nn, f, pl_morton, vals_dfs;
    dr_status st = DR_OK;
    auto need = [&](DevBuf &b, size_t bytes) { if (st == DR_OK) st = b.ensure(bytes); };
    for (int b = 0; b < 2; ++b) {
        need(ref[b], sizeof(int32_t) * n);
        need(cnt[b], sizeof(uint32_t) * n);
    }
    need(lo1, sizeof(float4) * n);
    need(hi1, sizeof(float4) * n);
    need(nn, sizeof(int32_t) * n);
    need(f, sizeof(uint64_t) * (n + 1));
    need(pl_morton, sizeof(int32_t) * n);
    need(vals_dfs, sizeof(uint32_t) * n);
    if (st) return st;
    float4 *lo[2] = {t.leaf_lo.as<float4>(), lo1.as<float4>()};
    float4 *hi[2] = {t.leaf_hi.as<float4>(), hi1.as<float4>()};
    int32_t *cl = t.child_l.as<int32_t>(), *cr = t.child_r.as<int32_t>(), *pi = t.parent_i.as<int32_t>();
    uint32_t *ncnt = t.count.as<uint32_t>();
    k_ploc_init<<<blocks(n), kT, 0, stream>>>(n, ref[0].as<int32_t>(), cnt[0].as<uint32_t>(), pi);
    DR_HIP_TRY(hipGetLastError());
    uint32_t m = n, made = 0;
    int cur = 0;
    while (m > 1) {
        k_ploc_nn<<<blocks(m), kT, 0, stream>>>(lo[cur], hi[cur], m, radius, nn.as<int32_t>());
        k_ploc_flags<<<blocks(m + 1), kT, 0, stream>>>(nn.as<int32_t>(), m, f.as<uint64_t>());
        DR_HIP_TRY(hipGetLastError());
        if ((st = exclusive_scan_u64(ctx, f.as<uint64_t>(), m + 1, stream))) return st;
        k_ploc_emit<<<blocks(m), kT, 0, stream>>>(m, nn.as<int32_t>(), f.as<uint64_t>(), ref[cur].as<int32_t>(),
                                                  lo[cur], hi[cur], cnt[cur].as<uint32_t>(), ref[1 - cur].as<int32_t>(),
                                                  lo[1 - cur], hi[1 - cur], cnt[1 - cur].as<uint32_t>(), cl, cr, pi,
                                                  pl_morton.as<int32_t>(), ncnt, n - 2 - made, t.node_lo.as<float4>(),
                                                  t.node_hi.as<float4>(), t.cost.as<float>(), t.collapse.as<uint32_t>(),
                                                  leaf_max, ct, ci);
        DR_HIP_TRY(hipGetLastError());
        uint64_t tot = 0;
        DR_HIP_TRY(hipMemcpyAsync(&tot, f.as<uint64_t>() + m, 8, hipMemcpyDeviceToHost, stream));
        DR_HIP_TRY(hipStreamSynchronize(stream));
        const uint32_t merges = (uint32_t)(tot >> 32), kept = (uint32_t)tot;
        // every round merges at least one pair (the smallest pair is mutual)
        if (merges == 0 || kept != m - merges || made + merges > n - 1)
            return fail(DR_E_DEVICE, "BVH build: clustering made no progress");
        made += merges;
        m = kept;
        cur = 1 - cur;
    }
    k_ploc_first<<<blocks(2 * n - 1), kT, 0, stream>>>(n, cl, cr, pi, pl_morton.as<int32_t>(), ncnt,
                                                       t.range_first.as<uint32_t>(), t.keys.as<uint32_t>());
    k_ploc_relabel_nodes<<<blocks(n - 1), kT, 0, stream>>>(n - 1, cl, cr, t.keys.as<uint32_t>());
    k_ploc_relabel_leaves<<<blocks(n), kT, 0, stream>>>(n, t.vals.as<uint32_t>(), pl_morton.as<int32_t>(),
                                                        t.keys.as<uint32_t>(), vals_dfs.as<uint32_t>(),
                                                        t.parent_l.as<int32_t>());
    k_gather<<<blocks(n), kT, 0, stream>>>(s->verts.as<float>(), s->indices.as<uint32_t>(), vals_dfs.as<uint32_t>(), n,
                                           t.rlo, t.rhi, t.rtri, t.rtri ? t.offs.as<uint32_t>() : nullptr,
                                           t.leaf_lo.as<float4>(), t.leaf_hi.as<float4>(), s->tri_v.as<float4>());
    DR_HIP_TRY(hipGetLastError());
    DR_HIP_TRY(hipStreamSynchronize(stream));  // the temporaries are freed on return
    return DR_OK;
}

// ---- refit: the built tree over moved vertices ------------------------------
__global__ void k_init_bounds(int32_t *b)
{
    const int t = threadIdx.x;
    if (t < 3) {
        b[t] = 0x7FFFFFFF;
        b[3 + t] = (int32_t)0x80000000;
    }
    if (t == 6) b[6] = 0;  // refit: quantisation error flag
}

// box of the finite vertex coordinates (ordered ints)
__global__ __launch_bounds__(kT) void k_vertex_bounds(const float *__restrict__ verts, uint32_t nv,
                                                      int32_t *__restrict__ vb)
{
    float c[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, C[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = blockIdx.x * kT + threadIdx.x; i < nv; i += gridDim.x * kT) {
        for (int a = 0; a < 3; ++a) {
            const float x = verts[3 * (size_t)i + a];
            if (isfinite(x)) {
                c[a] = fminf(c[a], x);
                C[a] = fmaxf(C[a], x);
            }
        }
    }
    block_bounds_atomic(c, C, vb);
}

// One level of wide nodes [lo, hi), deepest level first. An interior child's
// box becomes the union of its node's (already refitted) child boxes. A leaf's
// references take their triangle's new vertices and, for a split triangle,
// the box of their piece: slab j of k of the triangle's new box across the
// axis the build cut (the k slabs still cover the triangle, and an axis-
// aligned scale and shift maps each piece onto the same part as before). A
// triangle with a non-finite coordinate gets no box, as in the build. Boxes
// are widened by one ulp outward as in the collapse.
__global__ __launch_bounds__(kT) void k_refit_level(WideNode *__restrict__ nodes, uint32_t lo, uint32_t hi,
                                                    const float *__restrict__ verts, const uint32_t *__restrict__ idx,
                                                    float4 *__restrict__ tri_v, const int32_t *__restrict__ vb)
{
    const uint32_t w = lo + blockIdx.x * kT + threadIdx.x;
    if (w >= hi) return;
    float mag = 0.0f;
    for (int a = 0; a < 6; ++a) {
        const float m = fabsf(ord2f(vb[a]));
        if (m <= FLT_MAX) mag = fmaxf(mag, m);
    }
    WideNode wn = nodes[w];
    for (int c = 0; c < 4; ++c) {
        const int32_t ref = wn.child[c];
        if (ref == DR_EMPTY_REF) continue;
        float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
        if (ref >= 0) {
            const WideNode &ch = nodes[ref];
            for (int cc = 0; cc < 4; ++cc) {
                if (ch.child[cc] == DR_EMPTY_REF) continue;
                l[0] = fminf(l[0], ch.lox[cc]); h[0] = fmaxf(h[0], ch.hix[cc]);
                l[1] = fminf(l[1], ch.loy[cc]); h[1] = fmaxf(h[1], ch.hiy[cc]);
                l[2] = fminf(l[2], ch.loz[cc]); h[2] = fmaxf(h[2], ch.hiz[cc]);
            }
        } else {
            const uint32_t first = leaf_first(ref), cnt = leaf_count(ref);
            for (uint32_t r = first; r < first + cnt; ++r) {
                float4 *tv = tri_v + 3 * (size_t)r;
                const float4 w0 = tv[0], w1 = tv[1], w2 = tv[2];
                const uint32_t id = (uint32_t)__float_as_int(w0.w);
                const uint32_t pw = (uint32_t)__float_as_int(w1.w), pieces = (uint32_t)__float_as_int(w2.w);
                double v[3][3];
                float tl[3] = {INFINITY, INFINITY, INFINITY}, th[3] = {-INFINITY, -INFINITY, -INFINITY};
                float f[3][3];
                for (int q = 0; q < 3; ++q) {
                    const uint32_t vi = idx[3 * (size_t)id + q];
                    for (int a = 0; a < 3; ++a) {
                        f[q][a] = verts[3 * (size_t)vi + a];
                        v[q][a] = f[q][a];
                        tl[a] = fminf(tl[a], f[q][a]);
                        th[a] = fmaxf(th[a], f[q][a]);
                    }
                }
                tv[0] = make_float4(f[0][0], f[0][1], f[0][2], w0.w);
                tv[1] = make_float4(f[1][0], f[1][1], f[1][2], w1.w);
                tv[2] = make_float4(f[2][0], f[2][1], f[2][2], w2.w);
                bool finite = true;
                for (int a = 0; a < 3; ++a) finite = finite && isfinite(tl[a]) && isfinite(th[a]);
                if (!finite) continue;
                float plo[3] = {tl[0], tl[1], tl[2]}, phi[3] = {th[0], th[1], th[2]};
                if (pieces > 1) {
                    float m = mag;
                    for (int a = 0; a < 3; ++a) m = fmaxf(m, fmaxf(fabsf(tl[a]), fabsf(th[a])));
                    piece_box(v, tl, th, (int)(pw >> 16), pw & 0xFFFFu, pieces, m * 0x1p-17f, plo, phi);
                }
                for (int a = 0; a < 3; ++a) {
                    l[a] = fminf(l[a], plo[a]);
                    h[a] = fmaxf(h[a], phi[a]);
                }
            }
        }
        wn.lox[c] = down(l[0]); wn.loy[c] = down(l[1]); wn.loz[c] = down(l[2]);
        wn.hix[c] = up(h[0]); wn.hiy[c] = up(h[1]); wn.hiz[c] = up(h[2]);
    }
    nodes[w] = wn;
}

}  // namespace

dr_status refit_bvh(Scene *s, hipStream_t stream)
{
    if (!s->built || s->level_ofs.size() < 2) return fail(DR_E_INVALID, "scene BVH not built");
    if (dr_status st = s->refit_tmp.ensure(8 * sizeof(int32_t))) return st;
    int32_t *vb = s->refit_tmp.as<int32_t>();
    k_init_bounds<<<1, 64, 0, stream>>>(vb);
    k_vertex_bounds<<<std::min(blocks(s->n_verts), kBoundsBlocks), kT, 0, stream>>>(s->verts.as<float>(), s->n_verts,
                                                                                   vb);
    DR_HIP_TRY(hipGetLastError());
    for (size_t d = s->level_ofs.size() - 1; d-- > 0;) {
        const uint32_t lo = s->level_ofs[d], hi = s->level_ofs[d + 1];
        if (hi <= lo) continue;
        k_refit_level<<<blocks(hi - lo), kT, 0, stream>>>(s->bnodes.as<WideNode>(), lo, hi, s->verts.as<float>(),
                                                          s->indices.as<uint32_t>(), s->tri_v.as<float4>(), vb);
        DR_HIP_TRY(hipGetLastError());
    }
    k_quantize<<<blocks(s->n_nodes), kT, 0, stream>>>(s->bnodes.as<WideNode>(), s->n_nodes, s->nodes.as<QNode>(),
                                                      reinterpret_cast<uint32_t *>(vb + 6));
    DR_HIP_TRY(hipGetLastError());
    int32_t b[7];
    DR_HIP_TRY(hipMemcpyAsync(b, vb, sizeof(b), hipMemcpyDeviceToHost, stream));
    DR_HIP_TRY(hipStreamSynchronize(stream));
    if (b[6]) return fail(DR_E_INVALID, "vertex coordinates must be below 2^126 in magnitude (rebuild or refit again)");
    for (int a = 0; a < 6; ++a) {
        const int32_t i = b[a] >= 0 ? b[a] : b[a] ^ 0x7FFFFFFF;
        float f;
        std::memcpy(&f, &i, sizeof(f));
        s->bounds[a] = f;
    }
    return DR_OK;
}

dr_status build_bvh(Scene *s, const dr_bvh_params *p, hipStream_t stream)
{
    const uint32_t nt = s->n_tris;
    uint32_t leaf_max = p ? p->leaf_max : 4u;
    if (leaf_max < 1 || leaf_max > (uint32_t)kLeafMaxCap) return fail(DR_E_INVALID, "leaf_max must be 1..8");
    if (nt == 0) return fail(DR_E_INVALID, "scene has no triangles");
    if (nt > kMaxLeafRefs) return fail(DR_E_UNSUPPORTED, "more than 2^28 - 2 triangles per scene");
    const uint32_t flags = p ? p->flags : 0u;
    if (flags & ~(DR_BVH_LBVH | DR_BVH_NO_SPLIT | DR_BVH_RADIUS_MASK | DR_BVH_SPLIT_MASK | DR_BVH_SPLIT_RATIO_MASK))
        return fail(DR_E_INVALID, "unknown BVH flags");
    const bool ploc = !(flags & DR_BVH_LBVH);
    const int radius = (flags & DR_BVH_RADIUS_MASK) ? (int)((flags & DR_BVH_RADIUS_MASK) >> 8) : 16;
    if (radius > kPlocRMax) return fail(DR_E_INVALID, "PLOC search radius must be 1..64");
    uint32_t split_cap_q = (flags & DR_BVH_SPLIT_MASK) >> DR_BVH_SPLIT_SHIFT;
    uint32_t split_ratio_q = (flags & DR_BVH_SPLIT_RATIO_MASK) >> DR_BVH_SPLIT_RATIO_SHIFT;
    if (split_cap_q == 0) split_cap_q = kSplitCapDefaultQuarters;
    if (split_ratio_q == 0) split_ratio_q = kSplitRatioDefaultQuarters;
    if (flags & DR_BVH_NO_SPLIT) split_cap_q = 0;
    Context *ctx = s->ctx;

    hipEvent_t e0, e1;
    DR_HIP_TRY(hipEventCreate(&e0));
    DR_HIP_TRY(hipEventCreate(&e1));
    DR_HIP_TRY(hipEventRecord(e0, stream));

    BuildTmp t;
    dr_status st = DR_OK;
    auto need = [&](DevBuf &b, size_t bytes) { if (st == DR_OK) st = b.ensure(bytes); };
    need(t.tri_lo, sizeof(float4) * nt);
    need(t.tri_hi, sizeof(float4) * nt);
    need(t.counters, 128);
    if (st) return st;
    uint32_t *counters = t.counters.as<uint32_t>();
    int32_t *cbounds = reinterpret_cast<int32_t *>(counters + 8);
    int32_t *rbounds = reinterpret_cast<int32_t *>(counters + 16);
    k_init_counters<<<1, 64, 0, stream>>>(counters, cbounds, rbounds);
    k_tri_bounds<<<std::min(blocks(nt), kBoundsBlocks), kT, 0, stream>>>(s->verts.as<float>(), s->indices.as<uint32_t>(), nt,
                                                t.tri_lo.as<float4>(), t.tri_hi.as<float4>(), cbounds);
    DR_HIP_TRY(hipGetLastError());

    // leaf references: the triangles, or their split pieces
    uint32_t n = nt;
    t.rlo = t.tri_lo.as<float4>();
    t.rhi = t.tri_hi.as<float4>();
    t.rtri = nullptr;
    const int32_t *ref_bounds = cbounds;
    if (split_cap_q > 0 && nt > 1) {
        need(t.offs, sizeof(uint32_t) * (nt + 1));
        if (st) return st;
        uint32_t *offs = t.offs.as<uint32_t>();
        const float r0 = 0.25f * (float)split_ratio_q;
        const double cap = fmin(0.25 * split_cap_q * (double)nt, (double)(kMaxRefs - nt));
        uint32_t total_refs = 0;
        for (int pass = 0; pass < 2; ++pass) {
            k_split_counts<<<blocks(nt + 1), kT, 0, stream>>>(s->verts.as<float>(), s->indices.as<uint32_t>(), nt,
                                                              t.tri_lo.as<float4>(), t.tri_hi.as<float4>(), r0, offs);
            if (pass == 1) {
                const double extra = (double)total_refs - (double)nt;
                k_split_rescale<<<blocks(nt), kT, 0, stream>>>(offs, nt, cap / extra);
            }
            DR_HIP_TRY(hipGetLastError());
            if ((st = exclusive_scan_u32(ctx, offs, nt + 1, stream))) return st;
            DR_HIP_TRY(hipMemcpyAsync(&total_refs, offs + nt, 4, hipMemcpyDeviceToHost, stream));
            DR_HIP_TRY(hipStreamSynchronize(stream));
            if ((double)total_refs - (double)nt <= cap) break;
        }
        if (total_refs < nt || total_refs > kMaxRefs) return fail(DR_E_DEVICE, "split reference count out of range");
        n = total_refs;
        need(t.ref_lo, sizeof(float4) * n);
        need(t.ref_hi, sizeof(float4) * n);
        need(t.ref_tri, sizeof(uint32_t) * n);
        if (st) return st;
        k_split<<<std::min(blocks(nt), kBoundsBlocks), kT, 0, stream>>>(s->verts.as<float>(), s->indices.as<uint32_t>(), nt, t.tri_lo.as<float4>(),
                                       t.tri_hi.as<float4>(), t.offs.as<uint32_t>(), cbounds, t.ref_lo.as<float4>(),
                                       t.ref_hi.as<float4>(), t.ref_tri.as<uint32_t>(), rbounds);
        DR_HIP_TRY(hipGetLastError());
        t.rlo = t.ref_lo.as<float4>();
        t.rhi = t.ref_hi.as<float4>();
        t.rtri = t.ref_tri.as<uint32_t>();
        ref_bounds = rbounds;
    }

    const uint32_t ni = n > 1 ? n - 1 : 1;
    need(t.keys, sizeof(uint32_t) * n);
    need(t.vals, sizeof(uint32_t) * n);
    need(t.leaf_lo, sizeof(float4) * n);
    need(t.leaf_hi, sizeof(float4) * n);
    need(t.child_l, sizeof(int32_t) * ni);
    need(t.child_r, sizeof(int32_t) * ni);
    need(t.parent_i, sizeof(int32_t) * ni);
    need(t.parent_l, sizeof(int32_t) * n);
    need(t.range_first, sizeof(uint32_t) * ni);
    need(t.count, sizeof(uint32_t) * ni);
    need(t.node_lo, sizeof(float4) * ni);
    need(t.node_hi, sizeof(float4) * ni);
    need(t.cost, sizeof(float) * ni);
    need(t.visit, sizeof(uint32_t) * ni);
    need(t.collapse, sizeof(uint32_t) * ni);
    need(t.queue_a, sizeof(int2) * ni);
    need(t.queue_b, sizeof(int2) * ni);
    need(s->tri_v, sizeof(float4) * 3 * (size_t)n);
    // Wide nodes with a wide child have 4 children, the others only leaf
    // children (>= 2), so W <= 2L/3 + 1 <= 2n/3 + 1.
    const uint32_t wide_cap = (uint32_t)(2ull * n / 3 + 4);
    need(s->bnodes, sizeof(WideNode) * (size_t)wide_cap);
    need(s->nodes, sizeof(QNode) * (size_t)wide_cap);
    if (st) return st;

    k_morton<<<blocks(n), kT, 0, stream>>>(t.rlo, t.rhi, n, ref_bounds, t.keys.as<uint32_t>(), t.vals.as<uint32_t>());
    DR_HIP_TRY(hipGetLastError());
    if ((st = radix_sort_pairs(ctx, t.keys.as<uint32_t>(), t.vals.as<uint32_t>(), n, 30, stream))) return st;
    k_gather<<<blocks(n), kT, 0, stream>>>(s->verts.as<float>(), s->indices.as<uint32_t>(), t.vals.as<uint32_t>(), n,
                                           t.rlo, t.rhi, t.rtri, t.rtri ? t.offs.as<uint32_t>() : nullptr,
                                           t.leaf_lo.as<float4>(), t.leaf_hi.as<float4>(), s->tri_v.as<float4>());
    DR_HIP_TRY(hipGetLastError());

    const float ct = 1.0f, ci = 1.0f;
    uint32_t n_nodes = 1, n_leaves = 1, depth = 1;
    if (n == 1) {
        k_single_leaf_root<<<1, 1, 0, stream>>>(t.leaf_lo.as<float4>(), t.leaf_hi.as<float4>(), 1, s->bnodes.as<WideNode>(),
                                                nullptr, nullptr);
        DR_HIP_TRY(hipGetLastError());
        float4 lo, hi;
        DR_HIP_TRY(hipMemcpyAsync(&lo, t.leaf_lo.as<float4>(), 16, hipMemcpyDeviceToHost, stream));
        DR_HIP_TRY(hipMemcpyAsync(&hi, t.leaf_hi.as<float4>(), 16, hipMemcpyDeviceToHost, stream));
        DR_HIP_TRY(hipStreamSynchronize(stream));
        const float b[6] = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};
        for (int a = 0; a < 6; ++a) s->bounds[a] = b[a];
    } else {
        if (ploc) {
            if ((st = build_ploc(s, t, n, radius, leaf_max, ct, ci, stream))) return st;
        } else {
            k_karras<<<blocks(n - 1), kT, 0, stream>>>(t.keys.as<uint32_t>(), (int)n, t.child_l.as<int32_t>(),
                                                       t.child_r.as<int32_t>(), t.parent_i.as<int32_t>(),
                                                       t.parent_l.as<int32_t>(), t.range_first.as<uint32_t>(),
                                                       t.count.as<uint32_t>());
            k_fill_u32<<<blocks(n - 1), kT, 0, stream>>>(t.visit.as<uint32_t>(), n - 1, 0u);
            k_fit<<<blocks(n), kT, 0, stream>>>((int)n, t.child_l.as<int32_t>(), t.child_r.as<int32_t>(),
                                                t.parent_i.as<int32_t>(), t.parent_l.as<int32_t>(),
                                                t.count.as<uint32_t>(), t.leaf_lo.as<float4>(), t.leaf_hi.as<float4>(),
                                                t.node_lo.as<float4>(), t.node_hi.as<float4>(), t.cost.as<float>(),
                                                t.visit.as<uint32_t>(), t.collapse.as<uint32_t>(), leaf_max, ct, ci);
        }
        DR_HIP_TRY(hipGetLastError());
        uint32_t root_collapse = 0;
        float root_cost = 0.0f;
        float4 root_lo, root_hi;
        DR_HIP_TRY(hipMemcpyAsync(&root_collapse, t.collapse.as<uint32_t>(), 4, hipMemcpyDeviceToHost, stream));
        DR_HIP_TRY(hipMemcpyAsync(&root_cost, t.cost.as<float>(), 4, hipMemcpyDeviceToHost, stream));
        DR_HIP_TRY(hipMemcpyAsync(&root_lo, t.node_lo.as<float4>(), 16, hipMemcpyDeviceToHost, stream));
        DR_HIP_TRY(hipMemcpyAsync(&root_hi, t.node_hi.as<float4>(), 16, hipMemcpyDeviceToHost, stream));
        DR_HIP_TRY(hipStreamSynchronize(stream));
        const float ra = (root_hi.x - root_lo.x) * (root_hi.y - root_lo.y) + (root_hi.y - root_lo.y) * (root_hi.z - root_lo.z) +
                         (root_hi.z - root_lo.z) * (root_hi.x - root_lo.x);
        s->sah_cost = ra > 0 ? (double)root_cost / ra : 0.0;
        const float b[6] = {root_lo.x, root_lo.y, root_lo.z, root_hi.x, root_hi.y, root_hi.z};
        for (int a = 0; a < 6; ++a) s->bounds[a] = b[a];
        if (root_collapse) {
            k_single_leaf_root<<<1, 1, 0, stream>>>(t.leaf_lo.as<float4>(), t.leaf_hi.as<float4>(), n,
                                                    s->bnodes.as<WideNode>(), t.node_lo.as<float4>(), t.node_hi.as<float4>());
            DR_HIP_TRY(hipGetLastError());
            s->level_ofs = {0u, 1u};
        } else {
            // level-synchronous collapse; counters: [0] wide nodes, [1] queue out, [2] leaves
            k_collapse_init<<<1, 1, 0, stream>>>(counters, t.queue_a.as<int2>());
            uint32_t nin = 1;
            int2 *qa = t.queue_a.as<int2>(), *qb = t.queue_b.as<int2>();
            depth = 0;
            s->level_ofs = {0u, 1u};  // the root; each level allocates the next one's slots
            while (nin > 0) {
                k_collapse<<<blocks(nin), kT, 0, stream>>>(qa, nin, qb, counters, t.child_l.as<int32_t>(),
                                                          t.child_r.as<int32_t>(), t.range_first.as<uint32_t>(),
                                                          t.count.as<uint32_t>(), t.node_lo.as<float4>(),
                                                          t.node_hi.as<float4>(), t.leaf_lo.as<float4>(),
                                                          t.leaf_hi.as<float4>(), t.collapse.as<uint32_t>(),
                                                          s->bnodes.as<WideNode>(), wide_cap);
                DR_HIP_TRY(hipGetLastError());
                uint32_t c[4];
                DR_HIP_TRY(hipMemcpyAsync(c, counters, sizeof(c), hipMemcpyDeviceToHost, stream));
                DR_HIP_TRY(hipStreamSynchronize(stream));
                k_reset_queue_count<<<1, 1, 0, stream>>>(counters);
                nin = c[1];
                n_nodes = c[0];
                n_leaves = c[2];
                if (nin) s->level_ofs.push_back(n_nodes);
                ++depth;
                if (c[3]) return fail(DR_E_DEVICE, "BVH collapse exceeded its node bound");
                int2 *tmp = qa; qa = qb; qb = tmp;
                if (depth > kMaxWideDepth)
                    return fail(DR_E_OVERFLOW, "BVH deeper than " + std::to_string(kMaxWideDepth) +
                                                   " wide levels: the traversal stack bound");
            }
        }
    }
    // the traversal's quantised copy; counters[4] flags planes out of range
    k_quantize<<<blocks(n_nodes), kT, 0, stream>>>(s->bnodes.as<WideNode>(), n_nodes, s->nodes.as<QNode>(),
                                                   counters + 4);
    DR_HIP_TRY(hipGetLastError());
    DR_HIP_TRY(hipEventRecord(e1, stream));
    DR_HIP_TRY(hipEventSynchronize(e1));
    uint32_t qerr = 0;
    DR_HIP_TRY(hipMemcpy(&qerr, counters + 4, 4, hipMemcpyDeviceToHost));
    if (qerr) return fail(DR_E_INVALID, "vertex coordinates must be below 2^126 in magnitude");
    float ms = 0.0f;
    DR_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    s->n_refs = n;
    s->n_nodes = n_nodes;
    s->n_leaves = n_leaves;
    s->depth = depth;
    s->build_ms = ms;
    s->built = true;
    for (DevBuf *b : {&t.tri_lo, &t.tri_hi, &t.keys, &t.vals, &t.leaf_lo, &t.leaf_hi, &t.child_l, &t.child_r,
                      &t.parent_i, &t.parent_l, &t.range_first, &t.count, &t.node_lo, &t.node_hi, &t.cost,
                      &t.visit, &t.collapse, &t.queue_a, &t.queue_b, &t.counters})
        b->release();
    return DR_OK;
}

}  // namespace dr
