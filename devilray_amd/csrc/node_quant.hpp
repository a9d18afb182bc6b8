// Quantisation of a wide node's child boxes to one byte per plane (QNode).
//
// Per axis, the children's planes are put on a grid origin + q * 2^e,
// q = 0..255. e is the smallest exponent (at least -100, so every nonzero
// plane is a normal float) for which the grid from origin = floor(lo / 2^e) *
// 2^e spans the highest plane in 255 steps with |origin| below 2^24 steps.
// Then every grid point is an exact float, and fma(q, 2^e, origin) computes it
// with no rounding. Lower planes round down and upper planes round up, so the
// quantised box of a child contains its box. The double arithmetic is exact
// (scaling by powers of two, floor and ceil of floats).
//
// Only standard headers: the CPU test of this file compiles it with the
// host compiler.
#pragma once

#include <cmath>
#include <cstdint>

namespace dr {

// Planes of magnitude at or above this are not quantised (the build fails).
constexpr double kQuantMaxPlane = 0x1p126;

// One axis: lo[c], hi[c] for the children with use[c]; writes the grid origin,
// the biased exponent (e + 127) and the bytes of each child's lower / upper
// plane index (byte c). A child without a box gets q = 255 below and 0 above.
// Returns false when a plane's magnitude is 2^126 or more.
__host__ __device__ inline bool quantize_axis(const float lo[4], const float hi[4], const bool use[4], float &origin,
                                              uint32_t &ebias, uint32_t &wlo, uint32_t &whi)
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
        ebias = 127u;
        wlo = 0xFFFFFFFFu;
        return true;
    }
    if (!(fabs(plo) < kQuantMaxPlane && fabs(phi) < kQuantMaxPlane)) return false;
    int e = -100;
    const double span = (phi - plo) / 255.0;
    if (span > 0.0) {
        frexp(span, &e);  // span < 2^e
        e = e - 1 > -100 ? e - 1 : -100;
    }
    double st = 0.0, k = 0.0;
    for (;; ++e) {
        if (e > 126) return false;
        st = ldexp(1.0, e);
        k = floor(plo / st);
        const double kh = ceil(phi / st);
        if (kh - k <= 255.0 && fabs(k) + 255.0 < 16777216.0) break;
    }
    origin = (float)(k * st);
    ebias = (uint32_t)(e + 127);
    for (int c = 0; c < 4; ++c) {
        const uint32_t ql = use[c] ? (uint32_t)(floor((double)lo[c] / st) - k) : 255u;
        const uint32_t qh = use[c] ? (uint32_t)(ceil((double)hi[c] / st) - k) : 0u;
        wlo |= ql << (8 * c);
        whi |= qh << (8 * c);
    }
    return true;
}

// The plane of byte c of word w: what the traversal computes.
__host__ __device__ inline float quant_plane(uint32_t w, int c, uint32_t ebias, float origin)
{
    float step;
    const uint32_t bits = ebias << 23;
    __builtin_memcpy(&step, &bits, 4);
    return fmaf((float)((w >> (8 * c)) & 0xFFu), step, origin);
}

}  // namespace dr
