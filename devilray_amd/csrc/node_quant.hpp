// Quantisation of a wide node's child boxes to one byte per plane (QNode).
//
// Per axis, the children's planes are put on a grid origin + q * step,
// q = 0..255, with step = m * 2^e (m an integer below 2^8) and origin =
// k * 2^e (k an integer). Every grid point is then (k + q m) * 2^e, an exact
// float while |k| + 255 m < 2^24, so fma(q, step, origin) computes it with no
// rounding. The step is the smallest of that form whose grid, from the
// multiple of 2^e at or below the lowest plane, reaches the highest plane in
// 255 steps, so the grid uses almost the whole byte range; e >= -100 keeps
// every nonzero plane a normal float. Lower planes round down and upper planes
// round up, so the quantised box of a child contains its box. The double
// arithmetic is exact (integers below 2^53 times powers of two, floor and
// ceil of floats and of their differences with grid points).
//
// Only standard headers: the CPU test of this file compiles it with the
// host compiler.
#pragma once

#include <cmath>
#include <cstdint>

namespace dr {

// Planes of magnitude at or above this are not quantised (the build fails).
constexpr double kQuantMaxPlane = 0x1p126;

// One axis: lo[c], hi[c] for the children with use[c]; writes the grid origin
// and step and the bytes of each child's lower / upper plane index (byte c).
// A child without a box gets q = 255 below and 0 above. Returns false when a
// plane's magnitude is 2^126 or more.
__host__ __device__ inline bool quantize_axis(const float lo[4], const float hi[4], const bool use[4], float &origin,
                                              float &step, uint32_t &wlo, uint32_t &whi)
{
    double plo = INFINITY, phi = -INFINITY;
    for (int c = 0; c < 4; ++c)
        if (use[c]) {
            plo = fmin(plo, (double)lo[c]);
            phi = fmax(phi, (double)hi[c]);
        }
    wlo = 0u;
    whi = 0u;
    if (!(plo <= phi)) {  // no child with a box
        origin = 0.0f;
        step = 1.0f;
        wlo = 0xFFFFFFFFu;
        return true;
    }
    if (!(fabs(plo) < kQuantMaxPlane && fabs(phi) < kQuantMaxPlane)) return false;
    // start where m has 8 bits: 2^e <= span / 255 / 128
    int e = -100;
    const double span = phi - plo;
    if (span > 0.0) {
        int x;
        frexp(span / 255.0, &x);  // span / 255 in [2^(x-1), 2^x)
        e = x - 1 - 7 > -100 ? x - 1 - 7 : -100;
    }
    double u = 0.0, k = 0.0, m = 1.0;
    for (;; ++e) {
        if (e > 126) return false;
        u = ldexp(1.0, e);
        k = floor(plo / u);
        // the smallest m with k + 255 m reaching the highest plane
        m = ceil((ceil(phi / u) - k) / 255.0);
        if (m < 1.0) m = 1.0;
        if (m < 256.0 && fabs(k) + 255.0 * m < 16777216.0) break;
    }
    origin = (float)(k * u);
    step = (float)(m * u);
    const double st = m * u;
    // grid point q: (k + q m) u, exact in double, so the comparisons below
    // are exact; the quotients only give a starting guess
    auto point = [&](double q) { return (k + q * m) * u; };
    for (int c = 0; c < 4; ++c) {
        uint32_t ql = 255u, qh = 0u;
        if (use[c]) {
            double a = floor(((double)lo[c] - k * u) / st);
            a = fmin(fmax(a, 0.0), 255.0);
            while (a > 0.0 && point(a) > (double)lo[c]) a -= 1.0;
            while (a < 255.0 && point(a + 1.0) <= (double)lo[c]) a += 1.0;
            double b = ceil(((double)hi[c] - k * u) / st);
            b = fmin(fmax(b, 0.0), 255.0);
            while (b < 255.0 && point(b) < (double)hi[c]) b += 1.0;
            while (b > 0.0 && point(b - 1.0) >= (double)hi[c]) b -= 1.0;
            ql = (uint32_t)a;
            qh = (uint32_t)b;
        }
        wlo |= ql << (8 * c);
        whi |= qh << (8 * c);
    }
    return true;
}

// The plane of byte c of word w: what the traversal computes.
__host__ __device__ inline float quant_plane(uint32_t w, int c, float step, float origin)
{
    return fmaf((float)((w >> (8 * c)) & 0xFFu), step, origin);
}

}  // namespace dr

namespace dr {

// ---- the box test of a QNode's four children --------------------------------
//
// Per child, the entry and exit distances nr <= fr of a ray o + t d against
// the quantised box (near planes picked by the ray's octant, inv = 1 / d per
// axis); the caller takes the child when nr <= fr. Each plane is decoded
// exactly (fma(q, step, origin)), then (plane - o) * inv, and the exit
// distance is scaled by 1 + 2 gamma(3), the robust slab test (Ize, JCGT
// 2013): inv, the subtraction and the product round once each. So a child
// whose box the exact ray meets within [tmin, lim] is never dropped
// (tmin >= 0).

constexpr float kQuantU = 0x1p-24f;
constexpr float kFarScaleExact = 1.0f + 2.0f * (3.0f * kQuantU) / (1.0f - 3.0f * kQuantU);

// A QNode as the traversal loads it: origin, step, and the lower and upper
// plane bytes of the three axes.
struct QNodeWords {
    float ox, oy, oz, sx, sy, sz;
    uint32_t lo[3], hi[3];
};

struct SlabRay {
    float ox, oy, oz, ix, iy, iz, tmin;
    bool negx, negy, negz;  // the upper plane of the axis is near
};

__host__ __device__ inline float qbyte(uint32_t w, int c) { return (float)((w >> (8 * c)) & 0xFFu); }

__host__ __device__ inline void qnode_slabs(const QNodeWords &n, const SlabRay &r, float lim, float (&nr)[4],
                                            float (&fr)[4])
{
    const uint32_t wnx = r.negx ? n.hi[0] : n.lo[0], wfx = r.negx ? n.lo[0] : n.hi[0];
    const uint32_t wny = r.negy ? n.hi[1] : n.lo[1], wfy = r.negy ? n.lo[1] : n.hi[1];
    const uint32_t wnz = r.negz ? n.hi[2] : n.lo[2], wfz = r.negz ? n.lo[2] : n.hi[2];
    for (int c = 0; c < 4; ++c) {
        const float t0x = (fmaf(qbyte(wnx, c), n.sx, n.ox) - r.ox) * r.ix;
        const float t1x = (fmaf(qbyte(wfx, c), n.sx, n.ox) - r.ox) * r.ix;
        const float t0y = (fmaf(qbyte(wny, c), n.sy, n.oy) - r.oy) * r.iy;
        const float t1y = (fmaf(qbyte(wfy, c), n.sy, n.oy) - r.oy) * r.iy;
        const float t0z = (fmaf(qbyte(wnz, c), n.sz, n.oz) - r.oz) * r.iz;
        const float t1z = (fmaf(qbyte(wfz, c), n.sz, n.oz) - r.oz) * r.iz;
        nr[c] = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, r.tmin));
        fr[c] = fminf(fminf(fminf(t1x, t1y), t1z) * kFarScaleExact, lim);
    }
}

}  // namespace dr
