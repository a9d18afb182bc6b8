// LSD radix sort of (u32 key, u32 value) pairs and an exclusive u32 scan.
//
// One pass per 8-bit digit: a per-tile digit histogram, an exclusive scan of
// the digit-major histogram table (which yields every tile's global offset
// per digit), then a stable scatter. The scatter ranks keys inside a wave
// with ballots (lanes whose digit matches form a peer mask; rank = peers below
// the lane), wave by wave and chunk by chunk in input order, so the sort is
// stable. Tiles are 2048 keys: 4 wave64s x 8 chunks of 64.
#include "dr_internal.hpp"

namespace dr {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kChunks = 8;
constexpr uint32_t kTile = kThreads * kChunks;  // 2048

__global__ __launch_bounds__(kThreads) void k_radix_hist(const uint32_t *__restrict__ keys, uint32_t n,
                                                         int shift, uint32_t *__restrict__ hist,
                                                         uint32_t nblocks)
{
    __shared__ uint32_t h[256];
    const uint32_t tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kTile;
    for (int c = 0; c < kChunks; ++c) {
        const uint32_t idx = base + c * kThreads + tid;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[tid * nblocks + blockIdx.x] = h[tid];
}

__global__ __launch_bounds__(kThreads) void k_radix_scatter(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin, uint32_t *__restrict__ kout,
    uint32_t *__restrict__ vout, uint32_t n, int shift, const uint32_t *__restrict__ offsets,
    uint32_t nblocks)
{
    __shared__ uint32_t wh[kWaves][256];
    const uint32_t tid = threadIdx.x;
    const uint32_t w = tid >> 6, lane = tid & 63u;
    for (uint32_t k = tid; k < kWaves * 256; k += kThreads) (&wh[0][0])[k] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kTile + w * (64u * kChunks);
    for (int c = 0; c < kChunks; ++c) {
        const uint32_t idx = base + c * 64u + lane;
        if (idx < n) atomicAdd(&wh[w][(kin[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    {
        const uint32_t d = tid;  // kThreads == 256 digits
        uint32_t run = offsets[d * nblocks + blockIdx.x];
        for (uint32_t ww = 0; ww < kWaves; ++ww) {
            const uint32_t cnt = wh[ww][d];
            wh[ww][d] = run;
            run += cnt;
        }
    }
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64u - lane));
    for (int c = 0; c < kChunks; ++c) {
        const uint32_t idx = base + c * 64u + lane;
        const bool valid = idx < n;
        const uint32_t key = valid ? kin[idx] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        const int leader = __ffsll((unsigned long long)peers) - 1;
        uint32_t old = 0;
        if (valid && (int)lane == leader) old = atomicAdd(&wh[w][d], (uint32_t)__popcll(peers));
        old = __shfl(old, leader < 0 ? 0 : leader, 64);
        if (valid) {
            const uint32_t pos = old + rank;
            kout[pos] = key;
            vout[pos] = vin[idx];
        }
    }
}

// ---- exclusive scan (u32 / u64, exact integer sums) --------------------------
template <class T>
__device__ __forceinline__ T wave_incl_scan(T x, uint32_t lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T y = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += y;
    }
    return x;
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_scan_block(T *__restrict__ data, uint32_t n, T *__restrict__ sums)
{
    __shared__ T wsum[kWaves];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint32_t base = blockIdx.x * kTile + tid * kChunks;
    T v[kChunks];
    T s = 0;
#pragma unroll
    for (int k = 0; k < kChunks; ++k) {
        const uint32_t idx = base + k;
        v[k] = idx < n ? data[idx] : T(0);
        s += v[k];
    }
    const T incl = wave_incl_scan<T>(s, lane);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    T wpre = 0;
    for (uint32_t ww = 0; ww < w; ++ww) wpre += wsum[ww];
    T run = wpre + incl - s;
#pragma unroll
    for (int k = 0; k < kChunks; ++k) {
        const uint32_t idx = base + k;
        if (idx < n) data[idx] = run;
        run += v[k];
    }
    if (tid == kThreads - 1 && sums) sums[blockIdx.x] = wpre + incl;
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_scan_add(T *__restrict__ data, uint32_t n, const T *__restrict__ sums)
{
    const T add = sums[blockIdx.x];
    const uint32_t base = blockIdx.x * kTile;
    for (int k = 0; k < kChunks; ++k) {
        const uint32_t idx = base + k * kThreads + threadIdx.x;
        if (idx < n) data[idx] += add;
    }
}

template <class T>
dr_status scan_rec(Context *ctx, T *data, uint32_t n, int level, hipStream_t stream)
{
    if (n == 0) return DR_OK;
    const uint32_t nb = (n + kTile - 1) / kTile;
    if (nb == 1) {
        k_scan_block<T><<<1, kThreads, 0, stream>>>(data, n, nullptr);
        DR_HIP_TRY(hipGetLastError());
        return DR_OK;
    }
    if (level >= 3) return fail(DR_E_INVALID, "scan: input too large");
    DevBuf &tmp = ctx->scratch_sort[3 + level];
    if (dr_status st = tmp.ensure(sizeof(T) * nb)) return st;
    k_scan_block<T><<<nb, kThreads, 0, stream>>>(data, n, tmp.as<T>());
    DR_HIP_TRY(hipGetLastError());
    if (dr_status st = scan_rec<T>(ctx, tmp.as<T>(), nb, level + 1, stream)) return st;
    k_scan_add<T><<<nb, kThreads, 0, stream>>>(data, n, tmp.as<T>());
    DR_HIP_TRY(hipGetLastError());
    return DR_OK;
}

}  // namespace

dr_status exclusive_scan_u32(Context *ctx, uint32_t *data, uint32_t n, hipStream_t stream)
{
    return scan_rec<uint32_t>(ctx, data, n, 0, stream);
}

dr_status exclusive_scan_u64(Context *ctx, uint64_t *data, uint32_t n, hipStream_t stream)
{
    return scan_rec<uint64_t>(ctx, data, n, 0, stream);
}

dr_status radix_sort_pairs(Context *ctx, uint32_t *keys, uint32_t *vals, uint32_t n, int key_bits,
                           hipStream_t stream)
{
    if (n <= 1) return DR_OK;
    const uint32_t nb = (n + kTile - 1) / kTile;
    DevBuf &kalt = ctx->scratch_sort[0];
    DevBuf &valt = ctx->scratch_sort[1];
    DevBuf &hist = ctx->scratch_sort[2];
    if (dr_status st = kalt.ensure(sizeof(uint32_t) * n)) return st;
    if (dr_status st = valt.ensure(sizeof(uint32_t) * n)) return st;
    if (dr_status st = hist.ensure(sizeof(uint32_t) * 256u * nb)) return st;
    uint32_t *ka = keys, *va = vals, *kb = kalt.as<uint32_t>(), *vb = valt.as<uint32_t>();
    const int passes = (key_bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        k_radix_hist<<<nb, kThreads, 0, stream>>>(ka, n, shift, hist.as<uint32_t>(), nb);
        DR_HIP_TRY(hipGetLastError());
        if (dr_status st = exclusive_scan_u32(ctx, hist.as<uint32_t>(), 256u * nb, stream)) return st;
        k_radix_scatter<<<nb, kThreads, 0, stream>>>(ka, va, kb, vb, n, shift, hist.as<uint32_t>(), nb);
        DR_HIP_TRY(hipGetLastError());
        uint32_t *t = ka; ka = kb; kb = t;
        t = va; va = vb; vb = t;
    }
    if (ka != keys) {
        DR_HIP_TRY(hipMemcpyAsync(keys, ka, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, stream));
        DR_HIP_TRY(hipMemcpyAsync(vals, va, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, stream));
    }
    return DR_OK;
}

}  // namespace dr
