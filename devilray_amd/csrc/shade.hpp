// Shading math shared by the wavefront kernels. Every expression is written
// in a fixed evaluation order (the translation units are compiled with
// -ffp-contract=off), and the only transcendental used is the project's own
// polynomial sin/cos, so that a CPU evaluation of the same steps reproduces
// the device results bit for bit given the same random numbers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "philox.hpp"

namespace dr {

constexpr float kPi = 3.14159274f;

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 divs(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 normalize(V3 a) { return divs(a, sqrtf(dot(a, a))); }

// sin and cos of 2*pi*u for u in [0, 1): quadrant reduction (exact in
// binary), then Taylor polynomials to x^11 / x^12 on [0, pi/2) evaluated by
// Horner's rule; absolute error below 4e-7.
__device__ __forceinline__ void sincos_2pi(float u, float &s, float &c)
{
    const float q4 = u * 4.0f;
    int q = (int)q4;
    if (q > 3) q = 3;
    const float f = q4 - (float)q;
    const float x = f * 1.57079637f;
    const float x2 = x * x;
    float ps = -2.50521084e-8f;
    ps = ps * x2 + 2.75573192e-6f;
    ps = ps * x2 + -1.98412698e-4f;
    ps = ps * x2 + 8.33333333e-3f;
    ps = ps * x2 + -1.66666667e-1f;
    ps = ps * x2 + 1.0f;
    const float sn = ps * x;
    float pc = 2.08767570e-9f;
    pc = pc * x2 + -2.75573192e-7f;
    pc = pc * x2 + 2.48015873e-5f;
    pc = pc * x2 + -1.38888889e-3f;
    pc = pc * x2 + 4.16666667e-2f;
    pc = pc * x2 + -0.5f;
    pc = pc * x2 + 1.0f;
    const float cs = pc;
    if (q == 0) { s = sn; c = cs; }
    else if (q == 1) { s = cs; c = -sn; }
    else if (q == 2) { s = -sn; c = -cs; }
    else { s = -cs; c = sn; }
}

// Uniform direction on the sphere (the reference's uniformSphereSample,
// include/tracing/DistributionSamplers.hpp:8-20, with sin(acos z) written as
// sqrt(1 - z^2)).
__device__ __forceinline__ V3 uniform_sphere(float r1, float r2)
{
    const float z = 1.0f - 2.0f * r2;
    const float st = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float s0, c0;
    sincos_2pi(r1, s0, c0);
    return V3{st * s0, st * c0, z};
}

// Cosine-weighted direction about n: normalize(uniform_sphere + n)
// (cosineWeightedHemisphereSample, DistributionSamplers.hpp:33-42).
__device__ __forceinline__ V3 cosine_dir(V3 n, float r1, float r2)
{
    const V3 w = add(uniform_sphere(r1, r2), n);
    const float l2 = dot(w, w);
    if (!(l2 > 1e-20f)) return n;
    return divs(w, sqrtf(l2));
}

__device__ __forceinline__ float max_abs(V3 a) { return fmaxf(fabsf(a.x), fmaxf(fabsf(a.y), fabsf(a.z))); }

// Offset scale for new ray origins on a triangle: the reference's 1e-5,
// scaled by the largest vertex coordinate so it stays above the rounding of
// the hit point and of the triangle test's distance (both grow with the
// triangle's coordinates, not with the hit point's).
__device__ __forceinline__ float origin_eps(V3 a, V3 b, V3 c)
{
    return 1e-5f * fmaxf(1.0f, fmaxf(max_abs(a), fmaxf(max_abs(b), max_abs(c))));
}

// Hit point from barycentrics (accurate to the triangle's plane, unlike
// o + t d for long rays): A + (B - A) u + (C - A) v.
__device__ __forceinline__ V3 bary_point(V3 a, V3 b, V3 c, float u, float v)
{
    return add(add(a, mul(sub(b, a), u)), mul(sub(c, a), v));
}

// power heuristic (beta = 2); 0 when both densities are 0, as the reference's
// misWeightedEmission gives no contribution then (test_lighting.cpp:53-59)
__device__ __forceinline__ float power_heuristic(float a, float b)
{
    const float a2 = a * a, b2 = b * b;
    const float s = a2 + b2;
    return s > 0.0f ? a2 / s : 0.0f;
}

// Fresnel reflectance for a ray arriving with cos_i from index n1 into n2.
// Schlick's form is evaluated with the reference's types (schlick,
// PathGeneration.hpp:75-86): float powers of 1 - cos_i, then a double
// r0 + (1 - r0) c^5 (its 1.0 literal is a double) rounded once to float.
__device__ __forceinline__ float fresnel(float cos_i, float n1, float n2, float cos_t, bool schlick)
{
    if (schlick) {
        const float r = (n1 - n2) / (n1 + n2);
        const float r0 = r * r;
        const float c = 1.0f - cos_i;
        const float c2 = c * c;
        const float c4 = c2 * c2;
        const float c5 = c4 * c;
        return (float)((double)r0 + (1.0 - (double)r0) * (double)c5);
    }
    const float rs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t);
    const float rp = (n2 * cos_i - n1 * cos_t) / (n2 * cos_i + n1 * cos_t);
    return 0.5f * (rs * rs + rp * rp);
}

// Random numbers: key (pixel, seed low word), counter (sample, vertex,
// stream, seed high word). Stream 0 drives BSDF sampling / roulette /
// Fresnel, stream 1 light sampling; the camera uses vertex 0xFFFFFFFF.
__device__ __forceinline__ U4 path_rng(uint32_t pixel, uint32_t sample, uint32_t vertex, uint32_t stream, uint64_t seed)
{
    return philox4x32_10(U4{sample, vertex, stream, (uint32_t)(seed >> 32)}, pixel, (uint32_t)seed);
}

// Pinhole camera ray of a pixel sample (the reference's cameraRay,
// include/tracing/CameraRay.hpp:6-34): the pixel centre, or a Philox jitter
// drawn from the (pixel, sample) stream, through the sensor at focal length f,
// rotated by the camera transform; origin at its translation.
__device__ __forceinline__ dr_ray camera_ray(const CamParams &cam, uint32_t pixel, uint32_t sample, uint64_t seed)
{
    const uint32_t x = pixel % cam.w, y = pixel / cam.w;
    float sx = 0.5f, sy = 0.5f;
    if (cam.sampling == DR_PIXEL_UNIFORM_RANDOM) {
        const U4 r = path_rng(pixel, sample, 0xFFFFFFFFu, 0u, seed);
        sx = u01(r.x);
        sy = u01(r.y);
    }
    const float pcx = ((float)x + sx) * cam.psx - cam.cx;
    const float pcy = ((float)y + sy) * cam.psy - cam.cy;
    const V3 v = normalize(V3{pcx / cam.f, pcy / cam.f, 1.0f});
    const float *m = cam.m;
    dr_ray r;
    r.ox = m[3]; r.oy = m[7]; r.oz = m[11];
    r.dx = m[0] * v.x + m[1] * v.y + m[2] * v.z;
    r.dy = m[4] * v.x + m[5] * v.y + m[6] * v.z;
    r.dz = m[8] * v.x + m[9] * v.y + m[10] * v.z;
    r.tmin = 0.0f;
    r.tmax = INFINITY;
    return r;
}

}  // namespace dr
