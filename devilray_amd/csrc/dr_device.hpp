// Device-side geometry: ray set-up, watertight triangle test, robust slab test.
//
// This translation unit family is compiled with -ffp-contract=off: every
// product and sum rounds as written, so the triangle test reproduces a CPU
// evaluation of the same expression bit for bit. FMAs are spelled out with
// fmaf() where wanted.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace dr {

struct RayPre {
    float ox, oy, oz;
    float tmin, tmax;
    float ix, iy, iz;  // 1 / d, IEEE division
    float sx, sy, sz;  // shear constants of the watertight test
    int kz;            // dimension of largest |d|; kx = kz+1, ky = kz+2 (mod 3)
};

__device__ __forceinline__ float comp(float x, float y, float z, int k)
{
    return k == 0 ? x : (k == 1 ? y : z);
}

// Per-ray precomputation of the watertight test (Woop, Benthin, Wald 2013):
// kz = dimension of largest |d| (first on ties), kx, ky the next two cyclic.
// Woop et al. swap kx and ky when d[kz] < 0 to keep the winding for culling;
// for a two-sided test the swap only negates U, V, W, det and T exactly
// (IEEE rounding is symmetric), so t, u, v are identical without it.
__device__ __forceinline__ void ray_prepare(float ox, float oy, float oz, float tmin, float dx,
                                            float dy, float dz, float tmax, RayPre &p)
{
    p.ox = ox; p.oy = oy; p.oz = oz;
    p.tmin = tmin; p.tmax = tmax;
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    const int kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    const int kx = kz + 1 == 3 ? 0 : kz + 1;
    const int ky = kx + 1 == 3 ? 0 : kx + 1;
    const float dkz = comp(dx, dy, dz, kz);
    p.kz = kz;
    p.sx = comp(dx, dy, dz, kx) / dkz;
    p.sy = comp(dx, dy, dz, ky) / dkz;
    p.sz = 1.0f / dkz;
    p.ix = 1.0f / dx;
    p.iy = 1.0f / dy;
    p.iz = 1.0f / dz;
}

// Watertight two-sided test on vertices already expressed in the ray's
// (kx, ky, kz) order relative to the origin. Returns true with t, u (weight of
// the 2nd vertex), v (weight of the 3rd) when the ray's line crosses the
// triangle and det != 0. Must match the oracle's evaluation order exactly.
__device__ __forceinline__ bool tri_watertight_perm(const RayPre &p, float Akx, float Aky, float Akz, float Bkx,
                                                    float Bky, float Bkz, float Ckx, float Cky, float Ckz,
                                                    float &t, float &u, float &v)
{
    const float Ax = Akx - p.sx * Akz;
    const float Ay = Aky - p.sy * Akz;
    const float Bx = Bkx - p.sx * Bkz;
    const float By = Bky - p.sy * Bkz;
    const float Cx = Ckx - p.sx * Ckz;
    const float Cy = Cky - p.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        // products of two floats are exact in double
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f || det != det) return false;
    const float Az = p.sz * Akz;
    const float Bz = p.sz * Bkz;
    const float Cz = p.sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    const float rcp = 1.0f / det;
    t = T * rcp;
    u = V * rcp;
    v = W * rcp;
    return true;
}

// KZ >= 0: the ray's dominant axis is known at compile time (uniform across
// the wave); KZ < 0: chosen per lane with selects.
template <int KZ>
__device__ __forceinline__ bool tri_watertight(const RayPre &p, float4 a, float4 b, float4 c, float &t, float &u,
                                               float &v)
{
    const float A0 = a.x - p.ox, A1 = a.y - p.oy, A2 = a.z - p.oz;
    const float B0 = b.x - p.ox, B1 = b.y - p.oy, B2 = b.z - p.oz;
    const float C0 = c.x - p.ox, C1 = c.y - p.oy, C2 = c.z - p.oz;
    if (KZ == 0) return tri_watertight_perm(p, A1, A2, A0, B1, B2, B0, C1, C2, C0, t, u, v);
    if (KZ == 1) return tri_watertight_perm(p, A2, A0, A1, B2, B0, B1, C2, C0, C1, t, u, v);
    if (KZ == 2) return tri_watertight_perm(p, A0, A1, A2, B0, B1, B2, C0, C1, C2, t, u, v);
    const int kz = p.kz;
    const int kx = kz + 1 == 3 ? 0 : kz + 1;
    const int ky = kx + 1 == 3 ? 0 : kx + 1;
    return tri_watertight_perm(p, comp(A0, A1, A2, kx), comp(A0, A1, A2, ky), comp(A0, A1, A2, kz),
                               comp(B0, B1, B2, kx), comp(B0, B1, B2, ky), comp(B0, B1, B2, kz),
                               comp(C0, C1, C2, kx), comp(C0, C1, C2, ky), comp(C0, C1, C2, kz), t, u, v);
}

// Robust slab factor (Ize, JCGT 2013): far distances are scaled by
// 1 + 2*gamma(3) so rounding never drops a box the exact ray touches. The
// best-hit bound is relaxed by 2^-18 relative to cover the triangle test's
// own rounding when pruning.
constexpr float kFarScale = 1.0f + 2.0f * (3.0f * 0x1p-24f) / (1.0f - 3.0f * 0x1p-24f);
constexpr float kBestSlack = 1.0f + 0x1p-18f;

// Closest-hit acceptance: tmin <= t <= tbest, equal t resolved to the lower id.
__device__ __forceinline__ bool accept_hit(float t, uint32_t id, float tmin, float tbest, uint32_t best_id)
{
    if (!(t >= tmin)) return false;
    if (t < tbest) return true;
    return t == tbest && id < best_id;
}

}  // namespace dr
