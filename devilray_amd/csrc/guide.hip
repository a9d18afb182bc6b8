// Learned path guide on the matrix cores: BF16 MFMA inference for wavefront
// queues, a fused forward/backward training kernel with per-block partial
// gradients, a deterministic reduction + Adam kernel, training-sample
// completion, and the C-ABI entry points.
//
// Matrix layout. All products use v_mfma_f32_16x16x32_bf16 with one wave per
// 16 samples: A (16 x 32) holds weights, B (32 x 16) activations with
// samples as columns, C (16 x 16) FP32. Lane l holds A[l % 16][8 (l / 16) + j]
// and B[8 (l / 16) + j][l % 16] for j = 0..7, and C[4 (l / 16) + r][l % 16]
// for r = 0..3. The K order of a hidden layer is permuted so that the C
// tiles of one layer are, lane for lane, the B operand of the next: k-step
// ks, lane group g, element j reads input feature
//     phi(ks, g, j) = (2 ks + j / 4) * 16 + 4 g + j % 4,
// and the weight fragments are stored pre-permuted (1 KB per 16 x 32 tile,
// one 16-byte load per lane). The input layer's K order is the identity.
#include "dr_internal.hpp"
#include "guide_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace dr {

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int kNP = DR_GUIDE_N_PARAMS;
constexpr int kT = 256;
constexpr int kFragElems = 512;  // 64 lanes x 8 BF16
constexpr int kFragF = 34;       // forward fragments: 4 + 3 x 8 + 6
constexpr int kFragB = 32;       // transposed fragments of layers 1..4: 4 x 8
constexpr float kLobeShare = 0.5f;  // mixture share of the lobes' training pdf

__host__ __device__ constexpr int layer_in(int l) { return l == 0 ? 32 : 64; }
__host__ __device__ constexpr int layer_out(int l) { return l == 4 ? 48 : 64; }
__host__ __device__ constexpr int layer_ks(int l) { return l == 0 ? 1 : 2; }
__host__ __device__ constexpr int layer_ob(int l) { return l == 4 ? 3 : 4; }
__host__ __device__ constexpr int w_off(int l)
{
    return l == 0 ? 0 : w_off(l - 1) + layer_out(l - 1) * layer_in(l - 1) + layer_out(l - 1);
}
__host__ __device__ constexpr int b_off(int l) { return w_off(l) + layer_out(l) * layer_in(l); }
__host__ __device__ constexpr int f_base(int l) { return l == 0 ? 0 : f_base(l - 1) + layer_ob(l - 1) * layer_ks(l - 1); }
static_assert(b_off(4) + 48 == kNP, "parameter count");
static_assert(f_base(4) + 6 == kFragF, "forward fragment count");

// k position (ks, g, j) of input feature c of layer l
__host__ __device__ inline void k_slot(int l, int c, int &ks, int &g, int &j)
{
    if (l == 0) {
        ks = 0;
        g = c / 8;
        j = c % 8;
        return;
    }
    const int blk = c / 16, rem = c % 16;
    ks = blk / 2;
    g = rem / 4;
    j = (blk % 2) * 4 + rem % 4;
}

// Element offsets of weight W_l[o][c] in the forward fragments and (l >= 1)
// in the transposed fragments (-1 if none).
__host__ __device__ inline void frag_pos(int l, int o, int c, int &pf, int &pb)
{
    int ks, g, j;
    k_slot(l, c, ks, g, j);
    pf = ((f_base(l) + (o / 16) * layer_ks(l) + ks) * 64 + (o % 16) + 16 * g) * 8 + j;
    pb = -1;
    if (l >= 1) {
        k_slot(1, o, ks, g, j);  // K of the transposed product = output rows, hidden order
        pb = (((l - 1) * 8 + (c / 16) * 2 + ks) * 64 + (c % 16) + 16 * g) * 8 + j;
    }
}

// Parameter index -> (layer, row, column); column -1 for a bias.
__host__ __device__ inline int param_layer(int p, int &o, int &c)
{
    int l = 0;
    while (l < 4 && p >= w_off(l + 1)) ++l;
    const int q = p - w_off(l);
    if (q >= layer_out(l) * layer_in(l)) {
        o = q - layer_out(l) * layer_in(l);
        c = -1;
    } else {
        o = q / layer_in(l);
        c = q % layer_in(l);
    }
    return l;
}

struct Net {
    const bf16x8 *ff0;     // all forward fragments, from layer 0's first
    const bf16x8 *ff;      // forward fragments from layer 1's first (ff0 + kFrag0 * 64)
    const bf16x8 *fb;      // transposed fragments
    const float *bias[5];  // FP32 biases of each layer
};
constexpr int kFrag0 = 4;  // forward fragments of layer 0

struct Box {
    float lo[3], ext[3];
};

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x4 bias4(const Net &net, int l, int ob, int g)
{
    const float *b = net.bias[l] + ob * 16 + 4 * g;
    return f32x4{b[0], b[1], b[2], b[3]};
}

typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// two floats to a packed BF16 pair (one conversion instruction)
__device__ __forceinline__ unsigned int pk_bf16(float a, float b)
{
    return __builtin_bit_cast(unsigned int, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

// eight floats (lo then hi) as a BF16 operand
__device__ __forceinline__ bf16x8 to_bf16x8(f32x4 lo, f32x4 hi)
{
    const u32x4 w = {pk_bf16(lo[0], lo[1]), pk_bf16(lo[2], lo[3]), pk_bf16(hi[0], hi[1]), pk_bf16(hi[2], hi[3])};
    return __builtin_bit_cast(bf16x8, w);
}

// BF16 conversion, then ReLU on the bit patterns: as 16-bit integers the
// negative values and -0 are exactly the negative patterns, so a packed
// integer max with 0 rounds as relu-then-round does and yields +0 for them.
__device__ __forceinline__ bf16x8 pack_relu(f32x4 lo, f32x4 hi)
{
    const s16x8 v = __builtin_elementwise_max(__builtin_bit_cast(s16x8, to_bf16x8(lo, hi)), (s16x8)0);
    return __builtin_bit_cast(bf16x8, v);
}

// e^x as one exp2 instruction (relative error about 1e-7 (1 + |x|)): the
// loss head's exponentials, compared with their references at BF16 tolerance
__device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }

// Input features 8g .. 8g+7 of one sample: for g < 3, sin and cos of
// 2^(k+1) pi u (k = 0..3; periods 1 to 1/8 of the box) with u the position
// along axis g normalised to the box; for g = 3 the normal, the outgoing
// direction, 1 and 0.
// The box's lower corner and extent along the axis that lane group g
// encodes (g < 3).
struct BoxAxis {
    float lo, ext;
};
__device__ __forceinline__ BoxAxis box_axis(const Box &box, int g)
{
    // the six values are made opaque first, so the selections stay register
    // selects (left alone, the compiler turns them into a per-lane indexed
    // load from a copy of the argument)
    float l0 = box.lo[0], l1 = box.lo[1], l2 = box.lo[2], e0 = box.ext[0], e1 = box.ext[1], e2 = box.ext[2];
    asm volatile("" : "+s"(l0), "+s"(l1), "+s"(l2), "+s"(e0), "+s"(e1), "+s"(e2));
    return BoxAxis{g == 0 ? l0 : (g == 1 ? l1 : l2), g == 0 ? e0 : (g == 1 ? e1 : e2)};
}

__device__ __forceinline__ bf16x8 encode8(int g, V3 p, V3 n, V3 wo, BoxAxis ax)
{
    bf16x8 f;
    if (g < 3) {
        const float x = g == 0 ? p.x : (g == 1 ? p.y : p.z);
        const float lo = ax.lo, ext = ax.ext;
        float u = (x - lo) / ext;
        u = fminf(fmaxf(u, 0.0f), 1.0f);
        // sin and cos of 2 pi u, then of each doubled angle:
        // sin 2a = 2 sin a cos a, cos 2a = (cos a - sin a)(cos a + sin a)
        // (a few ulps after three doublings; the features are rounded to BF16)
        float s, c;
        sincos_2pi(u - floorf(u), s, c);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            f[2 * k] = (__bf16)s;
            f[2 * k + 1] = (__bf16)c;
            const float s2 = 2.0f * s * c;
            c = (c - s) * (c + s);
            s = s2;
        }
    } else {
        f[0] = (__bf16)n.x; f[1] = (__bf16)n.y; f[2] = (__bf16)n.z;
        f[3] = (__bf16)wo.x; f[4] = (__bf16)wo.y; f[5] = (__bf16)wo.z;
        f[6] = (__bf16)1.0f; f[7] = (__bf16)0.0f;
    }
    return f;
}

__device__ __forceinline__ bf16x8 encode8(int g, V3 p, V3 n, V3 wo, const Box &box)
{
    return encode8(g, p, n, wo, box_axis(box, g));
}

// Forward pass of one wave's 16 samples. xh[l] = hidden activation l+1 as
// B operands (kept for the backward masks); y = output tiles (rows ob*16 +
// 4g + r of the lane's sample).
// Forward pass of NG groups of 16 samples in one wave; each weight fragment
// is loaded once and applied to every group. xh[q][l] = hidden activation
// l+1 of group q as B operands (kept for the backward masks); y[q] = output
// tiles (rows ob*16 + 4g + r of the lane's sample).
template <int NG>
__device__ __forceinline__ void net_forward_g(const Net &net, int lane, const bf16x8 (&x0)[NG],
                                              bf16x8 (&xh)[NG][4][2], f32x4 (&y)[NG][3])
{
    const int g = lane >> 4;
    f32x4 acc[NG][4];
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
        const bf16x8 a = net.ff0[ob * 64 + lane];
        const f32x4 b0 = bias4(net, 0, ob, g);
#pragma unroll
        for (int q = 0; q < NG; ++q) acc[q][ob] = mfma(a, x0[q], b0);
    }
#pragma unroll
    for (int q = 0; q < NG; ++q) {
        xh[q][0][0] = pack_relu(acc[q][0], acc[q][1]);
        xh[q][0][1] = pack_relu(acc[q][2], acc[q][3]);
    }
#pragma unroll
    for (int l = 1; l < 4; ++l) {
#pragma unroll
        for (int ob = 0; ob < 4; ++ob) {
            const f32x4 b0 = bias4(net, l, ob, g);
#pragma unroll
            for (int q = 0; q < NG; ++q) acc[q][ob] = b0;
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                const bf16x8 a = net.ff[(f_base(l) - kFrag0 + ob * 2 + ks) * 64 + lane];
#pragma unroll
                for (int q = 0; q < NG; ++q) acc[q][ob] = mfma(a, xh[q][l - 1][ks], acc[q][ob]);
            }
        }
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            xh[q][l][0] = pack_relu(acc[q][0], acc[q][1]);
            xh[q][l][1] = pack_relu(acc[q][2], acc[q][3]);
        }
    }
#pragma unroll
    for (int ob = 0; ob < 3; ++ob) {
        const f32x4 b0 = bias4(net, 4, ob, g);
#pragma unroll
        for (int q = 0; q < NG; ++q) y[q][ob] = b0;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const bf16x8 a = net.ff[(f_base(4) - kFrag0 + ob * 2 + ks) * 64 + lane];
#pragma unroll
            for (int q = 0; q < NG; ++q) y[q][ob] = mfma(a, xh[q][3][ks], y[q][ob]);
        }
    }
}

// One group of 16 samples (inference and evaluation).
__device__ __forceinline__ void net_forward(const Net &net, int lane, bf16x8 x0, bf16x8 (&xh)[4][2], f32x4 (&y)[3])
{
    const bf16x8 xs[1] = {x0};
    bf16x8 xq[1][4][2];
    f32x4 yq[1][3];
    net_forward_g<1>(net, lane, xs, xq, yq);
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        xh[l][0] = xq[0][l][0];
        xh[l][1] = xq[0][l][1];
    }
#pragma unroll
    for (int ob = 0; ob < 3; ++ob) y[ob] = yq[0][ob];
}

// Decoded lobes 2g, 2g+1 of the lane's sample (local output slot q = 4 ob + r:
// lobe 2g uses q = 0..4, lobe 2g+1 q = 5..9, group 0 holds the selection
// logit at q = 10). Softmax and the selection logit are combined across the
// four lane groups of the sample.
struct HeadLane {
    float mu[2][3], r[2], kr[2], kap[2], pi[2], sig;
};

// Combining across the four 16-lane rows with the row-swap instructions
// (VALU, no LDS round trip): permlane16_swap(x, x) yields x of the even and
// of the odd row of each row pair, permlane32_swap(x, x) x of the lower and
// of the upper half wave (checked with an in-bounds probe on the GPU).
// Sums and maxima are symmetric, so every lane gets what a butterfly of
// shuffles gives it.
__device__ __forceinline__ void row_pair16(float x, float &a, float &b)
{
    const unsigned u = __builtin_bit_cast(unsigned, x);
    const auto p = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    a = __builtin_bit_cast(float, (unsigned)p[0]);
    b = __builtin_bit_cast(float, (unsigned)p[1]);
}
__device__ __forceinline__ void row_pair32(float x, float &a, float &b)
{
    const unsigned u = __builtin_bit_cast(unsigned, x);
    const auto p = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    a = __builtin_bit_cast(float, (unsigned)p[0]);
    b = __builtin_bit_cast(float, (unsigned)p[1]);
}
__device__ __forceinline__ float rows_sum(float x)
{
    float a, b;
    row_pair16(x, a, b);
    row_pair32(a + b, a, b);
    return a + b;
}
__device__ __forceinline__ float rows_max(float x)
{
    float a, b;
    row_pair16(x, a, b);
    row_pair32(fmaxf(a, b), a, b);
    return fmaxf(a, b);
}
// the value of lane (lane & 15), row 0's, in every row
__device__ __forceinline__ float row0_bcast(float x)
{
    float a, b;
    row_pair16(x, a, b);
    row_pair32(a, a, b);
    return a;
}

__device__ __forceinline__ void head_decode(const f32x4 (&y)[3], int lane, HeadLane &h)
{
    float yl[12];
#pragma unroll
    for (int ob = 0; ob < 3; ++ob)
#pragma unroll
        for (int r = 0; r < 4; ++r) yl[ob * 4 + r] = y[ob][r];
    float w[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int q0 = 5 * t;
        const float mx = yl[q0], my = yl[q0 + 1], mz = yl[q0 + 2];
        const float rr = fmaxf(sqrtf(mx * mx + my * my + mz * mz), kMuMinNorm);
        const float ir = __builtin_amdgcn_rcpf(rr);
        h.r[t] = rr;
        h.mu[t][0] = mx * ir;
        h.mu[t][1] = my * ir;
        h.mu[t][2] = mz * ir;
        h.kr[t] = yl[q0 + 3];
        h.kap[t] = fexp(fminf(fmaxf(h.kr[t], kLogKappaLo), kLogKappaHi));
        w[t] = yl[q0 + 4];
    }
    const float m = rows_max(fmaxf(w[0], w[1]));
    const float e0 = fexp(w[0] - m), e1 = fexp(w[1] - m);
    const float sum = rows_sum(e0 + e1);
    const float isum = __builtin_amdgcn_rcpf(sum);
    h.pi[0] = e0 * isum;
    h.pi[1] = e1 * isum;
    const float sl = __builtin_amdgcn_rcpf(1.0f + fexp(-yl[10]));
    h.sig = row0_bcast(sl);
}

struct PathIn {
    const float *verts;
    const uint32_t *idx;
    const uint32_t *tri_mat;
    const dr_material *mats;
    const dr_ray *rays;
    const dr_hit *hits;
    const uint32_t *list;
    const uint32_t *count;
    float *head;
};

__device__ __forceinline__ V3 vload(const float *v, uint32_t i)
{
    return V3{v[3 * (size_t)i], v[3 * (size_t)i + 1], v[3 * (size_t)i + 2]};
}

// One wave per 16 queue slots; every lane of a wave reaches the matrix
// instructions (invalid slots run on neutral inputs and store nothing).
// Grid of the path inference: a few blocks per CU, each copying the forward
// weights into LDS once and then striding over 64-path tiles (one block per
// tile read the 34 KB of weights from L2 for every 64 paths).
constexpr uint32_t kInferBlocks = 1024;

// A path's shading point for the guide: position, geometric normal facing the
// incoming ray, and the direction back along it; valid for a hit on a
// Lambert triangle.
struct PathGeo {
    uint32_t path;
    bool valid;
    V3 p, nn, wo;
};

__device__ __forceinline__ PathGeo path_geo(const PathIn &in, uint32_t slot, uint32_t n)
{
    PathGeo q{0u, slot < n, V3{0.0f, 0.0f, 0.0f}, V3{0.0f, 0.0f, 1.0f}, V3{0.0f, 0.0f, 1.0f}};
    if (q.valid) {
        q.path = in.list[slot];
        const dr_hit h = in.hits[q.path];
        q.valid = h.id >= 0;
        if (q.valid) {
            const uint32_t tri = (uint32_t)h.id;
            q.valid = in.mats[in.tri_mat[tri]].kind == DR_MAT_LAMBERT;
            if (q.valid) {
                const V3 A = vload(in.verts, in.idx[3 * (size_t)tri]);
                const V3 B = vload(in.verts, in.idx[3 * (size_t)tri + 1]);
                const V3 C = vload(in.verts, in.idx[3 * (size_t)tri + 2]);
                const dr_ray r = in.rays[q.path];
                const V3 d{r.dx, r.dy, r.dz};
                q.p = bary_point(A, B, C, h.u, h.v);
                const V3 ng = normalize(cross(sub(B, A), sub(C, A)));
                q.nn = dot(ng, d) > 0.0f ? mul(ng, -1.0f) : ng;
                q.wo = mul(d, -1.0f);
            }
        }
    }
    return q;
}

__global__ __launch_bounds__(kT) void k_guide_infer(Net gnet, Box box, PathIn in)
{
    __shared__ bf16x8 Wf[kFragF * 64];
    __shared__ float Bs[5][64];
    // Most launches give a block one tile, so its time is the chain of
    // memory round trips before the network runs. The weight loads are
    // issued first (they need no count), and the first tile's path data is
    // gathered before the weights are stored and the block meets, so those
    // latencies overlap.
    constexpr int kCF = (kFragF * 64 + kT - 1) / kT;
    constexpr int kCB = (5 * 64 + kT - 1) / kT;
    bf16x8 tf[kCF];
    float tb[kCB];
#pragma unroll
    for (int r = 0; r < kCF; ++r) {
        const int k = (int)threadIdx.x + r * kT;
        if (k < kFragF * 64) tf[r] = gnet.ff0[k];
    }
#pragma unroll
    for (int r = 0; r < kCB; ++r) {
        const int k = (int)threadIdx.x + r * kT;
        const int l = k / 64, o = k % 64;
        // the layer's bias array by selects: indexing the argument's pointer
        // array with a per-lane value would copy it to scratch memory
        const float *b = l == 0 ? gnet.bias[0] : l == 1 ? gnet.bias[1] : l == 2 ? gnet.bias[2]
                       : l == 3 ? gnet.bias[3] : gnet.bias[4];
        tb[r] = (k < 5 * 64 && o < layer_out(l)) ? b[o] : 0.0f;
    }
    const uint32_t n = *in.count;
    if ((uint64_t)blockIdx.x * 64u >= n) return;  // block-uniform: no tile for this block
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int s = lane & 15, g = lane >> 4;
    const PathGeo first = path_geo(in, blockIdx.x * 64u + (uint32_t)wave * 16u + (uint32_t)s, n);
#pragma unroll
    for (int r = 0; r < kCF; ++r) {
        const int k = (int)threadIdx.x + r * kT;
        if (k < kFragF * 64) Wf[k] = tf[r];
    }
#pragma unroll
    for (int r = 0; r < kCB; ++r) {
        const int k = (int)threadIdx.x + r * kT;
        if (k < 5 * 64) Bs[k / 64][k % 64] = tb[r];
    }
    __syncthreads();
    const Net net{Wf, Wf + kFrag0 * 64, nullptr, {Bs[0], Bs[1], Bs[2], Bs[3], Bs[4]}};
    // the lane's box axis, selected once (not a per-lane argument load per tile)
    BoxAxis ax = box_axis(box, g < 3 ? g : 0);
    asm volatile("" : "+v"(ax.lo), "+v"(ax.ext));
    for (uint32_t tile = blockIdx.x; (uint64_t)tile * 64u < n; tile += gridDim.x) {
        const uint32_t base = tile * 64u + (uint32_t)wave * 16u;
        if (base >= n) continue;  // wave-uniform
        const PathGeo q = tile == blockIdx.x ? first : path_geo(in, base + (uint32_t)s, n);
        const bool valid = q.valid;
        const uint32_t path = q.path;
        const V3 p = q.p, nn = q.nn, wo = q.wo;
        const bf16x8 x0 = encode8(g, p, nn, wo, ax);
        bf16x8 xh[4][2];
        f32x4 y[3];
        net_forward(net, lane, x0, xh, y);
        HeadLane h;
        head_decode(y, lane, h);
        // a non-finite head (diverged parameters) becomes a neutral one that
        // the shading kernel never selects (alpha = 0), so images stay finite
        float chk = h.sig;
#pragma unroll
        for (int t = 0; t < 2; ++t) chk += h.mu[t][0] + h.mu[t][1] + h.mu[t][2] + h.kap[t] + h.pi[t];
        int bad = isfinite(chk) ? 0 : 1;
        bad |= __shfl_xor(bad, 16);
        bad |= __shfl_xor(bad, 32);
        if (!valid) continue;
        if (bad) {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                h.mu[t][0] = 0.0f;
                h.mu[t][1] = 0.0f;
                h.mu[t][2] = 1.0f;
                h.kap[t] = 1.0f;
                h.pi[t] = 1.0f / kGuideLobes;
            }
        }
        float *o = in.head + (size_t)kHeadFloats * path;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            float *l = o + 5 * (2 * g + t);
            l[0] = h.mu[t][0];
            l[1] = h.mu[t][1];
            l[2] = h.mu[t][2];
            l[3] = h.kap[t];
            l[4] = h.pi[t];
        }
        if (g == 0) {
            o[40] = bad ? 0.0f : kAlphaLo + kAlphaSpan * h.sig;
            o[41] = 0.0f;
            o[42] = 0.0f;
            o[43] = 0.0f;
        }
    }
}

// ---- evaluation of caller samples (same network path as the render) -------
__global__ __launch_bounds__(kT) void k_guide_eval(Net net, Box box, const dr_guide_sample *__restrict__ smp,
                                                   uint32_t n, float *__restrict__ head)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t base = (blockIdx.x * (kT / 64) + wave) * 16u;
    if (base >= n) return;
    const int s = lane & 15, g = lane >> 4;
    const uint32_t i = base + (uint32_t)s;
    const bool valid = i < n;
    V3 p{0.0f, 0.0f, 0.0f}, nn{0.0f, 0.0f, 1.0f}, wo{0.0f, 0.0f, 1.0f};
    if (valid) {
        const dr_guide_sample q = smp[i];
        p = V3{q.pos[0], q.pos[1], q.pos[2]};
        nn = V3{q.normal[0], q.normal[1], q.normal[2]};
        wo = V3{q.wo[0], q.wo[1], q.wo[2]};
    }
    bf16x8 xh[4][2];
    f32x4 y[3];
    net_forward(net, lane, encode8(g, p, nn, wo, box), xh, y);
    HeadLane h;
    head_decode(y, lane, h);
    if (!valid) return;
    float *o = head + (size_t)kHeadFloats * i;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        float *l = o + 5 * (2 * g + t);
        l[0] = h.mu[t][0];
        l[1] = h.mu[t][1];
        l[2] = h.mu[t][2];
        l[3] = h.kap[t];
        l[4] = h.pi[t];
    }
    if (g == 0) {
        o[40] = kAlphaLo + kAlphaSpan * h.sig;
        o[41] = o[42] = o[43] = 0.0f;
    }
}

// ---- training: fused forward + backward, partial gradients per block -------
// A block of 8 waves (two per SIMD) trains on tiles of 256 samples. Each wave
// takes two groups of 16 samples and applies every weight fragment it loads
// to both: half the fragment reads per sample, and two independent chains of
// matrix products to cover each other's latency. LDS holds all weight
// fragments, the biases and one 256-sample x 64-feature BF16 tile each of a
// layer's input activations X and output deltas D (134,432 B). The
// activations stay in registers from the forward pass; the backward pass
// visits one layer per phase: the waves write X_L and D_L, meet at a barrier,
// accumulate their share of dW_L = D_L X_L^T on the matrix cores while they
// compute the next delta, and meet again before the next phase rewrites the
// tiles. Each wave keeps its dW tiles in registers across all tiles.
constexpr int kTrainWaves = 8;
constexpr int kTrainThreads = 64 * kTrainWaves;
constexpr int kGroups = 2;                             // 16-sample groups per wave
constexpr int kTileCols = 16 * kGroups * kTrainWaves;  // samples per tile
constexpr int kTileElems = kTileCols * 64;             // BF16 per X or D tile: one 128-byte row per sample
constexpr int kTileKs = kTileCols / 32;                // K steps of a dW product

// Tile layout. Row r holds sample r's 64 features as 16 chunks of 8 bytes
// (4 features); chunk c sits at position c ^ chunk_swz(r). A lane writes its
// own sample's values one chunk per store, and the dW products read them back
// with ds_read_b64_tr_b16, which transposes within each 16-lane group: lane k
// receives element k % 4 of the chunks addressed by lanes k / 4 + 4 j,
// j = 0..3 (checked on the GPU with an in-bounds probe). So when lane m
// addresses row 4 g + m / 4 at chunk C(m % 4), lane k of group g gets one
// feature of 4 consecutive samples, an operand half of a K = samples product.
// Features to chunks: a hidden layer's feature 16 b + 4 g + e (the lane
// layout of the B operands) is element e of chunk 4 g + b; the input's
// feature 8 g + 4 h + e is element e of chunk 4 g + h. The XOR patterns give
// each half wave 64 distinct banks both when it writes (16 rows x 2 chunks)
// and when it reads (8 rows x 4 chunks).
__device__ __forceinline__ int chunk_swz(int row) { return (int)((0xA381B290u >> (4 * ((row >> 1) & 7))) & 15u); }
__device__ __forceinline__ int chunk_off(int row, int c) { return row * 64 + ((c ^ chunk_swz(row)) << 2); }

typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) s16x4 lds_s16x4;

__device__ __forceinline__ void put_chunk(__bf16 *t, int row, int c, unsigned long long v)
{
    *reinterpret_cast<unsigned long long *>(t + chunk_off(row, c)) = v;
}

// The operand of one K step (32 samples) at the lane's address p: elements
// 0..3 are samples 4 g .. 4 g + 3 of the step, elements 4..7 samples
// 16 + 4 g .. 16 + 4 g + 3 (A and B operands use the same order, so each
// product sums over every sample of the step once).
__device__ __forceinline__ bf16x8 tr_operand(const __bf16 *p)
{
    const s16x4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4 *)p);
    const s16x4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4 *)(p + 16 * 64));
    return __builtin_bit_cast(bf16x8, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
}

// dW tiles of a wave: out block ob = wave % 4 (layer 4 has three, so waves 3
// and 7 take none there); in blocks 0 .. IB/2 - 1 plus the bias tile for the
// first four waves, IB/2 .. IB - 1 for the others. The wave reads its D rows
// once per K step for all its tiles.
__host__ __device__ constexpr int n_in_half(int l) { return layer_in(l) / 32; }       // in blocks per wave
__host__ __device__ constexpr int n_slots(int l) { return n_in_half(l) + 1; }         // with the bias tile
__host__ __device__ constexpr int acc_base(int l) { return l == 4 ? 0 : acc_base(l + 1) + n_slots(l + 1); }
constexpr int kAcc = acc_base(0) + n_slots(0);
static_assert(kAcc == 14, "accumulator tiles per wave");

// Row length of the per-block partial gradients: kNP + 1 floats padded to a
// multiple of 64, so a wave's 64 consecutive floats of one row are one
// aligned 256-byte segment when the Adam kernel sums them.
constexpr int kPartStride = (kNP + 1 + 63) / 64 * 64;

struct TrainIn {
    const dr_guide_sample *smp;
    const uint32_t *count;  // device count (NULL: n)
    uint32_t n, cap;
    uint32_t batch, window;  // batch > 0: samples (window * batch + i) mod count, i < batch
    float *partials;        // gridDim.x rows of kPartStride floats: kNP + 1 used
    uint32_t sel_variance;  // selection logit: 0 = KL objective, 1 = second-moment objective
};

// The training sample in slot i of the batch: sample (start + i) mod avail
// of the ring, or past the batch's end a zero-weight placeholder.
__device__ __forceinline__ dr_guide_sample train_sample(const dr_guide_sample *smp, uint32_t i, uint32_t n,
                                                        uint32_t start, uint32_t avail)
{
    if (i < n) return smp[start + i < avail ? start + i : start + i - avail];
    dr_guide_sample z = {};
    z.normal[2] = 1.0f;
    z.wo[2] = 1.0f;
    z.wi[2] = 1.0f;
    return z;
}

// The wave's share of dW_L += D_L X_L^T over the tile's samples (K).
template <int L>
__device__ __forceinline__ void dw_layer(f32x4 (&acc)[kAcc], const __bf16 *Dt, const __bf16 *Xt, int ob, int half,
                                         int lane, bf16x8 ones)
{
    constexpr int NI = n_in_half(L);
    if (ob >= layer_ob(L)) return;  // wave-uniform
    const int m = lane & 15, g = lane >> 4;
    const int lrow = 4 * g + (m >> 2), sw = chunk_swz(lrow);  // the swizzle of rows lrow + 16 k
    const __bf16 *dp = Dt + lrow * 64 + (((4 * (m & 3) + ob) ^ sw) << 2);
    const __bf16 *xp[NI];
#pragma unroll
    for (int u = 0; u < NI; ++u) {
        const int blk = half * NI + u;  // in features 16 blk .. 16 blk + 15
        const int c = L == 0 ? 8 * blk + 4 * ((m & 3) >> 1) + (m & 1) : 4 * (m & 3) + blk;
        xp[u] = Xt + lrow * 64 + ((c ^ sw) << 2);
    }
#pragma unroll
    for (int ks = 0; ks < kTileKs; ++ks) {
        const bf16x8 a = tr_operand(dp + 32 * 64 * ks);
#pragma unroll
        for (int u = 0; u < NI; ++u)
            acc[acc_base(L) + u] = mfma(a, tr_operand(xp[u] + 32 * 64 * ks), acc[acc_base(L) + u]);
        if (half == 0) acc[acc_base(L) + NI] = mfma(a, ones, acc[acc_base(L) + NI]);
    }
}

template <int L>
__device__ __forceinline__ void dw_store(const f32x4 (&acc)[kAcc], float *P, int ob, int half, int lane)
{
    constexpr int NI = n_in_half(L);
    constexpr int kW = w_off(L), kB = b_off(L);
    if (ob >= layer_ob(L)) return;
    // the empty asm makes the lane index opaque here, so the address
    // arithmetic happens at the stores and is not hoisted out of the tile
    // loop (where its registers would be live throughout)
    asm volatile("" : "+v"(lane));
    const int j = lane & 15, g = lane >> 4;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = ob * 16 + 4 * g + r;
#pragma unroll
        for (int u = 0; u < NI; ++u) P[kW + row * layer_in(L) + (half * NI + u) * 16 + j] = acc[acc_base(L) + u][r];
        if (half == 0 && j == 0) P[kB + row] = acc[acc_base(L) + NI][r];
    }
}

// A hidden layer's values in B-operand order (element j of operand ks is
// feature (2 ks + j / 4) * 16 + 4 g + j % 4 of the lane's sample) into the
// tile row of that sample: four chunk stores.
__device__ __forceinline__ void store_hidden(__bf16 *t, const bf16x8 (&v)[2], int g, int row)
{
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        const u64x2 w = __builtin_bit_cast(u64x2, v[ks]);
#pragma unroll
        for (int h = 0; h < 2; ++h) put_chunk(t, row, 4 * g + 2 * ks + h, w[h]);
    }
}

// Zero the BF16 elements of d whose ReLU activation x is not positive. x
// comes from pack_relu, so its patterns are 16-bit integers in [0, 0x7FFF]:
// 0 - x is negative exactly where x > 0, and its arithmetic shift by 15 is
// then all ones. Written as two packed 16-bit instructions per pair (the
// compiler otherwise turns the expression into per-element compares).
__device__ __forceinline__ bf16x8 relu_mask(bf16x8 d, bf16x8 x)
{
    u32x4 dw = __builtin_bit_cast(u32x4, d);
    const u32x4 xw = __builtin_bit_cast(u32x4, x);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        unsigned int m;
        asm("v_pk_sub_u16 %0, %1, %2\n\tv_pk_ashrrev_i16 %0, %3, %0"
            : "=&v"(m)
            : "v"(0u), "v"(xw[k]), "v"(0x000F000Fu));
        dw[k] &= m;
    }
    return __builtin_bit_cast(bf16x8, dw);
}

// deltas of hidden layer L (1..4) of both groups from the deltas of layer L+1
// (B operands in dB): W_L^T dB, masked by the ReLU of activation L; returned
// in dB as the next B operands. Each transposed fragment is loaded once.
template <int L>
__device__ __forceinline__ void delta_layer(const Net &net, int lane, bf16x8 (&dB)[kGroups][2],
                                            const bf16x8 (&xh)[kGroups][4][2])
{
    f32x4 c[kGroups][4];
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
#pragma unroll
        for (int q = 0; q < kGroups; ++q) c[q][ib] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const bf16x8 a = net.fb[((L - 1) * 8 + ib * 2 + ks) * 64 + lane];
#pragma unroll
            for (int q = 0; q < kGroups; ++q) c[q][ib] = mfma(a, dB[q][ks], c[q][ib]);
        }
    }
#pragma unroll
    for (int q = 0; q < kGroups; ++q)
#pragma unroll
        for (int h = 0; h < 2; ++h) dB[q][h] = relu_mask(to_bf16x8(c[q][2 * h], c[q][2 * h + 1]), xh[q][L - 1][h]);
}

// Fast reciprocal for the loss head (the gradient is compared with its
// reference at BF16 tolerance; ~1 ulp).
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Loss gradient at the network's outputs for the lane's sample of one group:
// dy (slot q = 4 ob + r of output block ob) and the output layer's deltas as
// B operands.
// loss w * [-log(q/2 + cos+/(2 pi)) - log(alpha q' + (1 - alpha) cos+/pi)]
__device__ __forceinline__ void train_head(const f32x4 (&y)[3], int lane, V3 nn, V3 wi, float w, float ps,
                                           uint32_t sel_variance, float (&dy)[12], bf16x8 (&dB)[2])
{
    const int g = lane >> 4;
    HeadLane h;
    head_decode(y, lane, h);
    const float alpha = kAlphaLo + kAlphaSpan * h.sig;
    float f[2], c[2], em[2];  // em: e^(2 kappa) - 1 per lobe
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const float k = h.kap[t];
        c[t] = h.mu[t][0] * wi.x + h.mu[t][1] * wi.y + h.mu[t][2] * wi.z;
        // vMF density k / (2 pi (1 - e^-2k)) e^(k (c - 1))
        em[t] = fexp(2.0f * k) - 1.0f;
        f[t] = fdiv(k * fexp(k * (c[t] - 1.0f)), kTwoPi * (1.0f - fexp(-2.0f * k)));
    }
    const float q = rows_sum(h.pi[0] * f[0] + h.pi[1] * f[1]);
    const float cb = dot(nn, wi);
    const float b = cb > 0.0f ? cb * (1.0f / kPi) : 0.0f;
    // lobes: KL of the fixed half-and-half proposal; selection: KL of the
    // learned one-sample MIS pdf with the mixture held fixed
    const float pm = alpha * q + (1.0f - alpha) * b;
    const float G = w != 0.0f ? -fdiv(w, pm) : 0.0f;
    const float dq = w != 0.0f ? -fdiv(w, kLobeShare * q + (1.0f - kLobeShare) * b) * kLobeShare : 0.0f;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int q0 = 5 * t;
        const float tt = h.pi[t] * f[t];
        const float k = h.kap[t];
        const float ir = __builtin_amdgcn_rcpf(h.r[t]);
        dy[q0 + 4] = dq * (tt - h.pi[t] * q);
        dy[q0 + 3] = (h.kr[t] > kLogKappaLo && h.kr[t] < kLogKappaHi)
                         ? dq * tt * (1.0f - fdiv(2.0f * k, em[t]) + k * (c[t] - 1.0f))
                         : 0.0f;
        const float gx = dq * tt * k * wi.x, gy = dq * tt * k * wi.y, gz = dq * tt * k * wi.z;
        const float mg = h.mu[t][0] * gx + h.mu[t][1] * gy + h.mu[t][2] * gz;
        dy[q0] = (gx - h.mu[t][0] * mg) * ir;
        dy[q0 + 1] = (gy - h.mu[t][1] * mg) * ir;
        dy[q0 + 2] = (gz - h.mu[t][2] * mg) * ir;
    }
    // second-moment objective sum_i w_i^2 p_i / pm (Rath et al. 2020,
    // "Variance-aware path guiding"): the KL gradient scaled by w p / pm, the
    // sample's weight under the current mixture; it trains the selection row
    // of the output layer only (no delta into the hidden layers, whose
    // features stay the lobes')
    const float rs = sel_variance ? (pm > 0.0f ? fdiv(w * ps, pm) : 0.0f) : 1.0f;
    dy[10] = g == 0 ? G * rs * (q - b) * kAlphaSpan * h.sig * (1.0f - h.sig) : 0.0f;
    dy[11] = 0.0f;
    const float d10 = sel_variance ? 0.0f : dy[10];
    dB[0] = __builtin_bit_cast(bf16x8, (u32x4){pk_bf16(dy[0], dy[1]), pk_bf16(dy[2], dy[3]), pk_bf16(dy[4], dy[5]),
                                               pk_bf16(dy[6], dy[7])});
    dB[1] = __builtin_bit_cast(bf16x8, (u32x4){pk_bf16(dy[8], dy[9]), pk_bf16(d10, dy[11]), 0u, 0u});
}

__global__ __launch_bounds__(kTrainThreads) void k_guide_train(Net gnet, Box box, TrainIn in)
{
    __shared__ bf16x8 Wf[kFragF * 64];
    __shared__ bf16x8 Wb[kFragB * 64];
    __shared__ __attribute__((aligned(16))) __bf16 Xt[kTileElems];
    __shared__ __attribute__((aligned(16))) __bf16 Dt[kTileElems];
    __shared__ float Bs[5][64];
    __shared__ float red[kTrainWaves];
    const int lane = threadIdx.x & 63;
    // wave-uniform, so the wave's dW tiles are chosen by scalar branches
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    // waves go to the SIMDs round-robin: out block = wave % 4, so each SIMD
    // runs one wave of each half (the first half also takes the bias tiles)
    const int ob = wave & 3, half = wave >> 2;
    const int s = lane & 15, g = lane >> 4;
    const uint32_t avail = in.count ? min(*in.count, in.cap) : in.n;
    const uint32_t n = in.batch ? min(avail, in.batch) : avail;
    const uint32_t start = (in.batch && avail) ? (uint32_t)(((uint64_t)in.window * in.batch) % avail) : 0u;
    const uint32_t ntiles = (n + (kTileCols - 1)) / kTileCols;
    int col[kGroups];
    // the samples of the block's next tile: loaded one tile ahead, so their
    // latency hides under the last backward phases of the tile before (the
    // first tile's under the weight copy)
    dr_guide_sample nxt[kGroups];
#pragma unroll
    for (int q = 0; q < kGroups; ++q) {
        col[q] = (wave * kGroups + q) * 16 + s;
        nxt[q] = train_sample(in.smp, blockIdx.x * (uint32_t)kTileCols + (uint32_t)col[q], n, start, avail);
    }
    {
        // every load is issued before the first store, so the copy costs one
        // memory round trip rather than one per fragment
        constexpr int kCF = (kFragF * 64 + kTrainThreads - 1) / kTrainThreads;
        constexpr int kCB = (kFragB * 64 + kTrainThreads - 1) / kTrainThreads;
        bf16x8 tf[kCF], tb[kCB];
#pragma unroll
        for (int r = 0; r < kCF; ++r) {
            const int k = (int)threadIdx.x + r * kTrainThreads;
            if (k < kFragF * 64) tf[r] = gnet.ff0[k];
        }
#pragma unroll
        for (int r = 0; r < kCB; ++r) {
            const int k = (int)threadIdx.x + r * kTrainThreads;
            if (k < kFragB * 64) tb[r] = gnet.fb[k];
        }
#pragma unroll
        for (int r = 0; r < kCF; ++r) {
            const int k = (int)threadIdx.x + r * kTrainThreads;
            if (k < kFragF * 64) Wf[k] = tf[r];
        }
#pragma unroll
        for (int r = 0; r < kCB; ++r) {
            const int k = (int)threadIdx.x + r * kTrainThreads;
            if (k < kFragB * 64) Wb[k] = tb[r];
        }
    }
    for (int k = threadIdx.x; k < 5 * 64; k += kTrainThreads) {
        const int l = k / 64, o = k % 64;
        Bs[l][o] = o < layer_out(l) ? gnet.bias[l][o] : 0.0f;
    }
    __syncthreads();
    const Net net{Wf, Wf + kFrag0 * 64, Wb, {Bs[0], Bs[1], Bs[2], Bs[3], Bs[4]}};
    f32x4 acc[kAcc];
#pragma unroll
    for (int k = 0; k < kAcc; ++k) acc[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // BF16 ones, the B operand of the bias-gradient products; the empty asm
    // hides the value, so the compiler keeps it in registers instead of
    // rematerialising the constant before every product
    u32x4 ones_w = {0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u};
    asm volatile("" : "+v"(ones_w));
    const bf16x8 ones = __builtin_bit_cast(bf16x8, ones_w);
    float wsum = 0.0f;
    // the lane's box axis, selected once: inside the loop the compiler turned
    // the selection into a per-lane load from the kernel arguments every tile
    BoxAxis ax = box_axis(box, g < 3 ? g : 0);
    asm volatile("" : "+v"(ax.lo), "+v"(ax.ext));
    // the block's partial gradient; a layer's share is stored as soon as the
    // block's last tile has added to it, so the stores overlap the remaining
    // phases instead of all leaving at the end
    float *P = in.partials + (size_t)blockIdx.x * kPartStride;
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const bool last = tile + gridDim.x >= ntiles;  // block-uniform
        bf16x8 x0[kGroups];
        V3 nn[kGroups], wi[kGroups];
        float w[kGroups], ps[kGroups];
#pragma unroll
        for (int q = 0; q < kGroups; ++q) {
            const dr_guide_sample &smp = nxt[q];
            const V3 p{smp.pos[0], smp.pos[1], smp.pos[2]}, wo{smp.wo[0], smp.wo[1], smp.wo[2]};
            nn[q] = V3{smp.normal[0], smp.normal[1], smp.normal[2]};
            wi[q] = V3{smp.wi[0], smp.wi[1], smp.wi[2]};
            w[q] = smp.weight;
            ps[q] = smp.pdf;
            if (g == 0) wsum += w[q];
            x0[q] = encode8(g, p, nn[q], wo, ax);
        }
        bf16x8 xh[kGroups][4][2];
        f32x4 y[kGroups][3];
        net_forward_g<kGroups>(net, lane, x0, xh, y);
        float dy[kGroups][12];
        bf16x8 dB[kGroups][2];
#pragma unroll
        for (int q = 0; q < kGroups; ++q) train_head(y[q], lane, nn[q], wi[q], w[q], ps[q], in.sel_variance, dy[q], dB[q]);

        // backward, one layer per phase through the X / D tiles (the first
        // barrier of a phase waits for every wave to finish reading the tiles)
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kGroups; ++q) {
            // output deltas: slot qq is feature (qq / 4) * 16 + 4 g + qq % 4
#pragma unroll
            for (int b = 0; b < 3; ++b)
                put_chunk(Dt, col[q], 4 * g + b,
                          (unsigned long long)pk_bf16(dy[q][4 * b], dy[q][4 * b + 1]) |
                              ((unsigned long long)pk_bf16(dy[q][4 * b + 2], dy[q][4 * b + 3]) << 32));
            store_hidden(Xt, xh[q][3], g, col[q]);
        }
        __syncthreads();
        dw_layer<4>(acc, Dt, Xt, ob, half, lane, ones);
        if (last) dw_store<4>(acc, P, ob, half, lane);
        delta_layer<4>(net, lane, dB, xh);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kGroups; ++q) {
            store_hidden(Dt, dB[q], g, col[q]);
            store_hidden(Xt, xh[q][2], g, col[q]);
        }
        __syncthreads();
        dw_layer<3>(acc, Dt, Xt, ob, half, lane, ones);
        if (last) dw_store<3>(acc, P, ob, half, lane);
        delta_layer<3>(net, lane, dB, xh);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kGroups; ++q) {
            store_hidden(Dt, dB[q], g, col[q]);
            store_hidden(Xt, xh[q][1], g, col[q]);
        }
        __syncthreads();
        dw_layer<2>(acc, Dt, Xt, ob, half, lane, ones);
        if (last) dw_store<2>(acc, P, ob, half, lane);
        delta_layer<2>(net, lane, dB, xh);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kGroups; ++q) {
            store_hidden(Dt, dB[q], g, col[q]);
            store_hidden(Xt, xh[q][0], g, col[q]);
        }
        const uint32_t next = tile + gridDim.x;
#pragma unroll
        for (int q = 0; q < kGroups; ++q)
            nxt[q] = train_sample(in.smp, next * (uint32_t)kTileCols + (uint32_t)col[q], n, start, avail);
        __syncthreads();
        dw_layer<1>(acc, Dt, Xt, ob, half, lane, ones);
        if (last) dw_store<1>(acc, P, ob, half, lane);
        delta_layer<1>(net, lane, dB, xh);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kGroups; ++q) {
            store_hidden(Dt, dB[q], g, col[q]);
            const u64x2 w = __builtin_bit_cast(u64x2, x0[q]);  // input features 8 g .. 8 g + 7
            put_chunk(Xt, col[q], 4 * g, w[0]);
            put_chunk(Xt, col[q], 4 * g + 1, w[1]);
        }
        __syncthreads();
        dw_layer<0>(acc, Dt, Xt, ob, half, lane, ones);
        if (last) dw_store<0>(acc, P, ob, half, lane);
    }
    if (blockIdx.x >= ntiles) {  // no tile for this block: a zero partial
        dw_store<4>(acc, P, ob, half, lane);
        dw_store<3>(acc, P, ob, half, lane);
        dw_store<2>(acc, P, ob, half, lane);
        dw_store<1>(acc, P, ob, half, lane);
        dw_store<0>(acc, P, ob, half, lane);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) wsum += __shfl_xor(wsum, o);
    if (lane == 0) red[wave] = wsum;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
#pragma unroll
        for (int k = 0; k < kTrainWaves; ++k) t += red[k];
        P[kNP] = t;
    }
}

__device__ __forceinline__ void pack_param(int p, float w, __bf16 *ff, __bf16 *fb)
{
    int o, c;
    const int l = param_layer(p, o, c);
    if (c < 0) return;  // biases are read from the FP32 parameters
    int pf, pb;
    frag_pos(l, o, c, pf, pb);
    const __bf16 v = (__bf16)w;
    ff[pf] = v;
    if (pb >= 0) fb[pb] = v;
}

__global__ __launch_bounds__(kT) void k_guide_pack(const float *__restrict__ master, __bf16 *ff, __bf16 *fb)
{
    const int p = blockIdx.x * kT + threadIdx.x;
    if (p < kNP) pack_param(p, master[p], ff, fb);
}

struct AdamIn {
    float lr, beta1, beta2, eps, c1, c2;  // c1 = 1/(1 - beta1^t), c2 = 1/(1 - beta2^t)
};

// Sum of the block partials in a fixed order (deterministic): a block owns 64
// parameters; each of its 16 waves sums every 16th partial of them (one
// coalesced 256-byte row per load) and wave 0 adds the 16 slices in order.
// With grad set, only the sums are written; otherwise one Adam step on the
// loss normalised by the total sample weight, and the BF16 fragments are
// refreshed.
constexpr int kAdamSlices = 16;
static_assert(kGuideTrainBlocks % kAdamSlices == 0 && kGuideTrainBlocks % 64 == 0, "partial slots per slice and lane");
__global__ __launch_bounds__(64 * kAdamSlices) void k_guide_adam(const float *__restrict__ partials, int nblk,
                                                                 float *master, float *m1, float *m2, __bf16 *ff,
                                                                 __bf16 *fb, AdamIn ad, float *grad)
{
    __shared__ float part[kAdamSlices][64];
    __shared__ float s_sw;
    const int lane = threadIdx.x & 63, slice = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    // all of the slice's loads are issued before the sum (in block order)
    constexpr int kPer = kGuideTrainBlocks / kAdamSlices;
    float pv[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const int b = slice + r * kAdamSlices;
        pv[r] = (p <= kNP && b < nblk) ? partials[(size_t)b * kPartStride + p] : 0.0f;
    }
    // the parameter's Adam state, loaded under the partial sums
    float m1p = 0.0f, m2p = 0.0f, wp = 0.0f;
    if (slice == 0 && !grad && p < kNP) {
        m1p = m1[p];
        m2p = m2[p];
        wp = master[p];
    }
    float acc = 0.0f;
#pragma unroll
    for (int r = 0; r < kPer; ++r)
        if (slice + r * kAdamSlices < nblk) acc += pv[r];
    part[slice][lane] = acc;
    if (slice == 1) {  // the total sample weight (column kNP), the same in every block
        constexpr int kPerLane = kGuideTrainBlocks / 64;
        float wv[kPerLane];
#pragma unroll
        for (int r = 0; r < kPerLane; ++r) {
            const int b = lane + 64 * r;
            wv[r] = b < nblk ? partials[(size_t)b * kPartStride + kNP] : 0.0f;
        }
        float w = 0.0f;
#pragma unroll
        for (int r = 0; r < kPerLane; ++r)
            if (lane + 64 * r < nblk) w += wv[r];
        for (int off = 32; off >= 1; off >>= 1) w += __shfl_xor(w, off, 64);
        if (lane == 0) s_sw = w;
    }
    __syncthreads();
    if (slice != 0 || p > kNP) return;
    float sum = 0.0f;
    for (int k = 0; k < kAdamSlices; ++k) sum += part[k][lane];
    if (grad) {
        grad[p] = sum;
        return;
    }
    if (p == kNP) return;
    const float sw = s_sw;
    if (!(sw > 0.0f) || !isfinite(sw)) return;
    const float gr = sum / sw;
    if (!isfinite(gr)) return;
    const float m = ad.beta1 * m1p + (1.0f - ad.beta1) * gr;
    const float v = ad.beta2 * m2p + (1.0f - ad.beta2) * gr * gr;
    m1[p] = m;
    m2[p] = v;
    const float w = wp - ad.lr * (m * ad.c1) / (sqrtf(v * ad.c2) + ad.eps);
    master[p] = w;
    pack_param(p, w, ff, fb);
}

// Starts a collection: the stage count only (the training ring persists).
__global__ void k_guide_zero_counters(uint32_t *c)
{
    if (threadIdx.x == 0) c[0] = 0u;
}

// weight = luminance((L_end - L_snap) / beta) * max(0, n . wi) / pdf
__global__ __launch_bounds__(kT) void k_guide_finalize(dr_guide_sample *__restrict__ smp, GuideAux *__restrict__ aux,
                                                       const uint32_t *__restrict__ counters, uint32_t cap,
                                                       const float4 *__restrict__ rad)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i >= min(counters[0], cap)) return;
    const GuideAux x = aux[i];
    if (x.path == kGuideNoSample) return;
    const float4 L = rad[x.path];
    const float s0 = L.x - x.lsnap[0], s1 = L.y - x.lsnap[1], s2 = L.z - x.lsnap[2];
    const float l0 = x.beta[0] > 0.0f ? s0 / x.beta[0] : 0.0f;
    const float l1 = x.beta[1] > 0.0f ? s1 / x.beta[1] : 0.0f;
    const float l2 = x.beta[2] > 0.0f ? s2 / x.beta[2] : 0.0f;
    const float lum = fmaxf(0.0f, 0.2126f * l0 + 0.7152f * l1 + 0.0722f * l2);
    const dr_guide_sample q = smp[i];
    const float cs = fmaxf(0.0f, q.normal[0] * q.wi[0] + q.normal[1] * q.wi[1] + q.normal[2] * q.wi[2]);
    float w = x.pdf > 0.0f ? lum * cs / x.pdf : 0.0f;
    if (!isfinite(w)) w = 0.0f;
    smp[i].weight = w;
    aux[i].path = kGuideNoSample;
}

// Sort input over the whole stage: perm = identity, and the slots no sample
// took get the sentinel key, the largest key of the key width. A sample's key
// may equal it, but the sort is stable and such a sample's slot is below
// every unused slot, so after the sort the first n positions of perm are the
// n staged samples in key order.
__global__ __launch_bounds__(kT) void k_guide_sort_init(uint32_t *__restrict__ key_lo, uint32_t *__restrict__ key_hi,
                                                        uint32_t *__restrict__ perm,
                                                        const uint32_t *__restrict__ counters, uint32_t stage_cap,
                                                        uint32_t sent_lo, uint32_t sent_hi)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i >= stage_cap) return;
    perm[i] = i;
    if (i >= min(counters[0], stage_cap)) {
        key_lo[i] = sent_lo;
        key_hi[i] = sent_hi;
    }
}

// Second digit group of a 64-bit key: the high words in the order of the
// first sort (a stable sort by them then orders by the whole key).
__global__ __launch_bounds__(kT) void k_guide_gather_hi(uint32_t *__restrict__ dst, const uint32_t *__restrict__ key_hi,
                                                        const uint32_t *__restrict__ perm, uint32_t stage_cap)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i < stage_cap) dst[i] = key_hi[perm[i]];
}

__global__ __launch_bounds__(kT) void k_guide_flag(const dr_guide_sample *__restrict__ smp,
                                                   const uint32_t *__restrict__ perm,
                                                   const uint32_t *__restrict__ counters, uint32_t stage_cap,
                                                   uint32_t *__restrict__ flags)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i > stage_cap) return;
    const uint32_t n = min(counters[0], stage_cap);
    flags[i] = (i < n && smp[perm[i]].weight > 0.0f) ? 1u : 0u;
}

// flags holds exclusive ranks over the key order (flags[stage_cap] = number
// of non-zero samples); the sample of rank r goes to ring slot (head + r) mod
// ring_cap. Only the first ring_cap ranks are written, so every slot has one
// writer.
__global__ __launch_bounds__(kT) void k_guide_scatter(const dr_guide_sample *__restrict__ smp,
                                                      const uint32_t *__restrict__ perm,
                                                      const uint32_t *__restrict__ counters, uint32_t stage_cap,
                                                      const uint32_t *__restrict__ flags, dr_guide_sample *ring,
                                                      uint32_t ring_cap)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    const uint32_t n = min(counters[0], stage_cap);
    if (i >= n) return;
    const uint32_t r = flags[i];
    if (flags[i + 1] == r || r >= ring_cap) return;  // zero weight, or beyond one ring's worth
    ring[(counters[1] + r) % ring_cap] = smp[perm[i]];
}

__global__ void k_guide_ring_advance(uint32_t *counters, const uint32_t *__restrict__ flags, uint32_t stage_cap,
                                     uint32_t ring_cap)
{
    if (threadIdx.x != 0) return;
    const uint32_t nnz = min(flags[stage_cap], ring_cap);
    counters[1] = (counters[1] + nnz) % ring_cap;
    counters[2] = min(counters[2] + nnz, ring_cap);
}

inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

Net net_of(const Guide *g)
{
    const float *m = g->master.as<const float>();
    const bf16x8 *ff = g->frag_f.as<const bf16x8>();
    return Net{ff, ff + kFrag0 * 64, g->frag_b.as<const bf16x8>(),
               {m + b_off(0), m + b_off(1), m + b_off(2), m + b_off(3), m + b_off(4)}};
}

Box box_of(const Guide *g)
{
    Box b;
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = g->prm.bounds[k];
        b.ext[k] = g->prm.bounds[3 + k] - g->prm.bounds[k];
    }
    return b;
}

// He-uniform hidden weights, output layer scaled by 0.1, zero biases; from
// Philox(seed, parameter index).
void init_params(std::vector<float> &P, uint64_t seed)
{
    P.assign(kNP, 0.0f);
    for (int p = 0; p < kNP; ++p) {
        int o, c;
        const int l = param_layer(p, o, c);
        if (c < 0) continue;
        const U4 r = philox4x32_10(U4{(uint32_t)p, 0x6D1Du, 0u, (uint32_t)(seed >> 32)}, 0x51DEu, (uint32_t)seed);
        const float lim = std::sqrt(6.0f / (float)layer_in(l)) * (l == 4 ? 0.1f : 1.0f);
        P[p] = (2.0f * u01(r.x) - 1.0f) * lim;
    }
}

dr_status upload_params(Guide *g, const float *host)
{
    DR_HIP_TRY(hipMemcpy(g->master.ptr, host, sizeof(float) * kNP, hipMemcpyHostToDevice));
    DR_HIP_TRY(hipMemset(g->m1.ptr, 0, sizeof(float) * kNP));
    DR_HIP_TRY(hipMemset(g->m2.ptr, 0, sizeof(float) * kNP));
    k_guide_pack<<<cdiv(kNP, kT), kT>>>(g->master.as<float>(), g->frag_f.as<__bf16>(), g->frag_b.as<__bf16>());
    DR_HIP_TRY(hipGetLastError());
    DR_HIP_TRY(hipDeviceSynchronize());
    g->step = 0;
    return DR_OK;
}

dr_status run_train(Guide *g, const dr_guide_sample *smp, const uint32_t *count, uint32_t n, uint32_t cap,
                    float *grad, hipStream_t stream, uint32_t batch = 0)
{
    const uint32_t n_max = batch ? std::min(batch, count ? cap : n) : (count ? cap : n);
    const int nblk = (int)std::max<uint32_t>(1u, std::min<uint32_t>(kGuideTrainBlocks, cdiv(n_max, kTileCols)));
    TrainIn in{smp, count, n, cap, batch, g->window++, g->partials.as<float>(), g->sel_variance};
    k_guide_train<<<nblk, kTrainThreads, 0, stream>>>(net_of(g), box_of(g), in);
    DR_HIP_TRY(hipGetLastError());
    AdamIn ad{};
    if (!grad) {
        g->step += 1;
        const double t = (double)g->step;
        ad.lr = g->prm.lr;
        ad.beta1 = g->prm.beta1;
        ad.beta2 = g->prm.beta2;
        ad.eps = g->prm.eps;
        ad.c1 = (float)(1.0 / (1.0 - std::pow((double)g->prm.beta1, t)));
        ad.c2 = (float)(1.0 / (1.0 - std::pow((double)g->prm.beta2, t)));
    }
    k_guide_adam<<<cdiv(kNP + 1, 64), 64 * kAdamSlices, 0, stream>>>(g->partials.as<float>(), nblk, g->master.as<float>(),
                                                       g->m1.as<float>(), g->m2.as<float>(), g->frag_f.as<__bf16>(),
                                                       g->frag_b.as<__bf16>(), ad, grad);
    DR_HIP_TRY(hipGetLastError());
    return DR_OK;
}

}  // namespace

dr_status guide_infer_paths(Guide *g, const Scene *s, const dr_ray *rays, const dr_hit *hits, const uint32_t *list,
                            const uint32_t *count, uint32_t n_max, float *head, hipStream_t stream)
{
    PathIn in{s->verts.as<float>(), s->indices.as<uint32_t>(), s->tri_mat.as<uint32_t>(),
              s->materials.as<dr_material>(), rays, hits, list, count, head};
    k_guide_infer<<<std::min<uint32_t>(cdiv(n_max, 64), kInferBlocks), kT, 0, stream>>>(net_of(g), box_of(g), in);
    DR_HIP_TRY(hipGetLastError());
    return DR_OK;
}

dr_status guide_compact(Guide *g, uint32_t key_bits, hipStream_t stream)
{
    const uint32_t S = g->stage_cap;
    uint32_t *lo = g->keys.as<uint32_t>(), *hi = lo + S, *perm = g->perm.as<uint32_t>();
    if (key_bits > 64) return fail(DR_E_INVALID, "guide sample keys wider than 64 bits");
    const int B = std::max(1, (int)key_bits);
    const uint64_t sent = B >= 64 ? ~0ull : (1ull << B) - 1ull;
    k_guide_sort_init<<<cdiv(S, kT), kT, 0, stream>>>(lo, hi, perm, g->counters.as<uint32_t>(), S, (uint32_t)sent,
                                                      (uint32_t)(sent >> 32));
    DR_HIP_TRY(hipGetLastError());
    if (dr_status st = radix_sort_pairs(g->ctx, lo, perm, S, std::min(B, 32), stream)) return st;
    if (B > 32) {
        k_guide_gather_hi<<<cdiv(S, kT), kT, 0, stream>>>(lo, hi, perm, S);
        DR_HIP_TRY(hipGetLastError());
        if (dr_status st = radix_sort_pairs(g->ctx, lo, perm, S, B - 32, stream)) return st;
    }
    k_guide_flag<<<cdiv(S + 1, kT), kT, 0, stream>>>(g->samples.as<dr_guide_sample>(), perm,
                                                    g->counters.as<uint32_t>(), S, g->flags.as<uint32_t>());
    DR_HIP_TRY(hipGetLastError());
    if (dr_status st = exclusive_scan_u32(g->ctx, g->flags.as<uint32_t>(), S + 1, stream)) return st;
    k_guide_scatter<<<cdiv(S, kT), kT, 0, stream>>>(g->samples.as<dr_guide_sample>(), perm,
                                                   g->counters.as<uint32_t>(), S, g->flags.as<uint32_t>(),
                                                   g->ring.as<dr_guide_sample>(), g->prm.capacity);
    DR_HIP_TRY(hipGetLastError());
    k_guide_ring_advance<<<1, 64, 0, stream>>>(g->counters.as<uint32_t>(), g->flags.as<uint32_t>(), S,
                                               g->prm.capacity);
    DR_HIP_TRY(hipGetLastError());
    return DR_OK;
}

dr_status guide_finalize(Guide *g, const float4 *radiance, hipStream_t stream)
{
    const uint32_t cap = g->stage_cap;
    k_guide_finalize<<<cdiv(cap, kT), kT, 0, stream>>>(g->samples.as<dr_guide_sample>(), g->aux.as<GuideAux>(),
                                                       g->counters.as<uint32_t>(), cap, radiance);
    DR_HIP_TRY(hipGetLastError());
    return DR_OK;
}

dr_status guide_reset_counters(Guide *g, hipStream_t stream)
{
    k_guide_zero_counters<<<1, 64, 0, stream>>>(g->counters.as<uint32_t>());
    DR_HIP_TRY(hipGetLastError());
    return DR_OK;
}

}  // namespace dr

using namespace dr;

extern "C" {

dr_status dr_guide_create(dr_context *cx, const dr_guide_params *params, dr_guide **out)
{
    if (!cx || !params || !out) return fail(DR_E_INVALID, "null argument");
    *out = nullptr;
    const dr_guide_params &p = *params;
    for (int k = 0; k < 3; ++k)
        if (!(p.bounds[3 + k] > p.bounds[k]) || !std::isfinite(p.bounds[k]) || !std::isfinite(p.bounds[3 + k]))
            return fail(DR_E_INVALID, "guide bounds must satisfy lo < hi on every axis");
    if (p.capacity == 0 || p.capacity > (1u << 22)) return fail(DR_E_INVALID, "capacity must be 1..2^22");
    if (!(p.lr > 0.0f) || !(p.beta1 >= 0.0f && p.beta1 < 1.0f) || !(p.beta2 >= 0.0f && p.beta2 < 1.0f) ||
        !(p.eps > 0.0f))
        return fail(DR_E_INVALID, "Adam parameters out of range");
    auto *ctx = reinterpret_cast<Context *>(cx);
    DR_HIP_TRY(hipSetDevice(ctx->device));
    auto *g = new (std::nothrow) Guide();
    if (!g) return fail(DR_E_OOM, "host allocation");
    g->ctx = ctx;
    g->prm = p;
    dr_status st = DR_OK;
    auto need = [&](DevBuf &b, size_t bytes) { if (st == DR_OK) st = b.ensure(bytes); };
    need(g->master, sizeof(float) * kNP);
    need(g->m1, sizeof(float) * kNP);
    need(g->m2, sizeof(float) * kNP);
    need(g->frag_f, 2ull * kFragElems * kFragF);
    need(g->frag_b, 2ull * kFragElems * kFragB);
    need(g->partials, sizeof(float) * (size_t)kGuideTrainBlocks * kPartStride);
    g->stage_cap = p.capacity * 4u;
    need(g->samples, sizeof(dr_guide_sample) * (size_t)g->stage_cap);
    need(g->aux, sizeof(GuideAux) * (size_t)g->stage_cap);
    need(g->ring, sizeof(dr_guide_sample) * (size_t)p.capacity);
    need(g->flags, 4ull * ((size_t)g->stage_cap + 1));
    need(g->keys, 8ull * g->stage_cap);
    need(g->perm, 4ull * g->stage_cap);
    need(g->counters, 16);
    if (st) {
        delete g;
        return st;
    }
    hipError_t e = hipMemset(g->frag_f.ptr, 0, g->frag_f.bytes);
    if (e == hipSuccess) e = hipMemset(g->frag_b.ptr, 0, g->frag_b.bytes);
    if (e == hipSuccess) e = hipMemset(g->counters.ptr, 0, 16);
    if (e != hipSuccess) {
        delete g;
        return fail(DR_E_DEVICE, std::string("hipMemset: ") + hipGetErrorString(e));
    }
    std::vector<float> P;
    init_params(P, p.seed);
    if ((st = upload_params(g, P.data()))) {
        delete g;
        return st;
    }
    *out = reinterpret_cast<dr_guide *>(g);
    return DR_OK;
}

dr_status dr_guide_destroy(dr_guide *gd)
{
    if (!gd) return DR_OK;
    auto *g = reinterpret_cast<Guide *>(gd);
    (void)hipSetDevice(g->ctx->device);
    (void)hipDeviceSynchronize();
    delete g;
    return DR_OK;
}

dr_status dr_guide_get_params(dr_guide *gd, float *params)
{
    if (!gd || !params) return fail(DR_E_INVALID, "null argument");
    auto *g = reinterpret_cast<Guide *>(gd);
    DR_HIP_TRY(hipSetDevice(g->ctx->device));
    DR_HIP_TRY(hipDeviceSynchronize());
    DR_HIP_TRY(hipMemcpy(params, g->master.ptr, sizeof(float) * kNP, hipMemcpyDeviceToHost));
    return DR_OK;
}

dr_status dr_guide_set_params(dr_guide *gd, const float *params)
{
    if (!gd || !params) return fail(DR_E_INVALID, "null argument");
    auto *g = reinterpret_cast<Guide *>(gd);
    for (int p = 0; p < kNP; ++p)
        if (!std::isfinite(params[p])) return fail(DR_E_INVALID, "non-finite parameter");
    DR_HIP_TRY(hipSetDevice(g->ctx->device));
    DR_HIP_TRY(hipDeviceSynchronize());
    return upload_params(g, params);
}

dr_status dr_guide_eval(dr_guide *gd, const dr_guide_sample *samples, uint32_t n, float *head, dr_stream stream)
{
    if (!gd || (n && (!samples || !head))) return fail(DR_E_INVALID, "null argument");
    if (n == 0) return DR_OK;
    auto *g = reinterpret_cast<Guide *>(gd);
    DR_HIP_TRY(hipSetDevice(g->ctx->device));
    k_guide_eval<<<cdiv(n, 64), kT, 0, reinterpret_cast<hipStream_t>(stream)>>>(net_of(g), box_of(g), samples, n,
                                                                                 head);
    DR_HIP_TRY(hipGetLastError());
    return DR_OK;
}

dr_status dr_guide_gradient(dr_guide *gd, const dr_guide_sample *samples, uint32_t n, float *grad, dr_stream stream)
{
    if (!gd || !grad || (n && !samples)) return fail(DR_E_INVALID, "null argument");
    auto *g = reinterpret_cast<Guide *>(gd);
    DR_HIP_TRY(hipSetDevice(g->ctx->device));
    return run_train(g, samples, nullptr, n, n, grad, reinterpret_cast<hipStream_t>(stream));
}

dr_status dr_guide_train_step(dr_guide *gd, const dr_guide_sample *samples, uint32_t n, dr_stream stream)
{
    if (!gd) return fail(DR_E_INVALID, "guide is null");
    auto *g = reinterpret_cast<Guide *>(gd);
    DR_HIP_TRY(hipSetDevice(g->ctx->device));
    const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (samples) return run_train(g, samples, nullptr, n, n, nullptr, st);
    return run_train(g, g->ring.as<dr_guide_sample>(), g->counters.as<uint32_t>() + 2, 0, g->prm.capacity, nullptr,
                     st, g->train_batch);
}

dr_status dr_guide_copy_samples(dr_guide *gd, dr_guide_sample *host, uint32_t max_n, uint32_t *n_kept,
                                uint32_t *n_total)
{
    if (!gd) return fail(DR_E_INVALID, "guide is null");
    auto *g = reinterpret_cast<Guide *>(gd);
    DR_HIP_TRY(hipSetDevice(g->ctx->device));
    DR_HIP_TRY(hipDeviceSynchronize());
    uint32_t c[4];
    DR_HIP_TRY(hipMemcpy(c, g->counters.ptr, sizeof c, hipMemcpyDeviceToHost));
    const uint32_t kept = std::min(c[0], g->stage_cap);
    if (n_kept) *n_kept = kept;
    if (n_total) *n_total = c[0];
    if (host && max_n && kept) {
        // in key order: the staged samples occupy slots [0, kept), and the
        // first kept entries of perm are those slots sorted by key
        std::vector<dr_guide_sample> stage(kept);
        std::vector<uint32_t> order(kept);
        DR_HIP_TRY(hipMemcpy(stage.data(), g->samples.ptr, sizeof(dr_guide_sample) * kept, hipMemcpyDeviceToHost));
        DR_HIP_TRY(hipMemcpy(order.data(), g->perm.ptr, 4ull * kept, hipMemcpyDeviceToHost));
        const uint32_t m = std::min(kept, max_n);
        for (uint32_t i = 0; i < m; ++i) {
            if (order[i] >= kept) return fail(DR_E_DEVICE, "stage order out of range");
            host[i] = stage[order[i]];
        }
    }
    return DR_OK;
}

dr_status dr_guide_set_options(dr_guide *gd, float collect_fraction, uint32_t guided_vertices, uint32_t train_batch)
{
    if (!gd) return fail(DR_E_INVALID, "guide is null");
    if (!(collect_fraction >= 0.0f && collect_fraction <= 1.0f)) return fail(DR_E_INVALID, "fraction must be in [0, 1]");
    auto *g = reinterpret_cast<Guide *>(gd);
    g->collect_fraction = collect_fraction;
    g->guided_vertices = guided_vertices;
    g->train_batch = train_batch;
    return DR_OK;
}

dr_status dr_guide_set_selection_objective(dr_guide *gd, int32_t objective)
{
    if (!gd) return fail(DR_E_INVALID, "guide is null");
    if (objective != DR_GUIDE_SELECT_KL && objective != DR_GUIDE_SELECT_VARIANCE)
        return fail(DR_E_INVALID, "selection objective must be DR_GUIDE_SELECT_KL or DR_GUIDE_SELECT_VARIANCE");
    reinterpret_cast<Guide *>(gd)->sel_variance = objective == DR_GUIDE_SELECT_VARIANCE ? 1u : 0u;
    return DR_OK;
}

dr_status dr_guide_set_learning_rate(dr_guide *gd, float lr)
{
    if (!gd) return fail(DR_E_INVALID, "guide is null");
    if (!(lr > 0.0f && lr <= 1.0f)) return fail(DR_E_INVALID, "learning rate must be in (0, 1]");
    reinterpret_cast<Guide *>(gd)->prm.lr = lr;
    return DR_OK;
}

}  // extern "C"
