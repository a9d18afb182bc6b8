// Device helpers of the learned path guide shared by the network kernels
// (guide.hip) and the shading kernel (render.hip): the decoded mixture record
// a path carries, von Mises-Fisher density and sampling, and the one-sample
// MIS combination with the cosine BSDF.
#pragma once

#include "shade.hpp"

namespace dr {

constexpr int kGuideLobes = 8;
constexpr int kHeadFloats = 44;     // 8 x {mu x, mu y, mu z, kappa, weight}, alpha, 3 pad
constexpr float kLogKappaLo = -3.0f;  // log concentration clamp
constexpr float kLogKappaHi = 7.0f;
constexpr float kMuMinNorm = 1e-6f;
constexpr float kAlphaLo = 0.1f;    // selection probability 0.1 + 0.8 sigmoid(s)
constexpr float kAlphaSpan = 0.8f;
constexpr float kTwoPi = 6.28318548f;

// vMF density with concentration k about mu at cosine c = mu . w:
// k / (2 pi (1 - e^-2k)) e^(k (c - 1)), written with expm1 so small k keeps
// its precision and large k cannot overflow.
__device__ __forceinline__ float vmf_density(float k, float c)
{
    return k / (kTwoPi * -expm1f(-2.0f * k)) * expf(k * (c - 1.0f));
}

// Mixture density q(w) of a decoded head record.
__device__ __forceinline__ float head_mixture(const float *__restrict__ h, V3 w)
{
    float q = 0.0f;
#pragma unroll
    for (int j = 0; j < kGuideLobes; ++j) {
        const float *l = h + 5 * j;
        const float c = l[0] * w.x + l[1] * w.y + l[2] * w.z;
        q += l[4] * vmf_density(l[3], c);
    }
    return q;
}

// Orthonormal basis about a unit vector (Duff et al. 2017, branchless).
__device__ __forceinline__ void onb(V3 n, V3 &t, V3 &b)
{
    const float sg = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sg + n.z);
    const float bb = n.x * n.y * a;
    t = V3{1.0f + sg * n.x * n.x * a, sg * bb, -sg * n.x};
    b = V3{bb, sg + n.y * n.y * a, -n.y};
}

// W. Jakob's numerically stable vMF sampler (2012): the cosine about mu is
// 1 + log(x + (1 - x) e^-2k) / k with x in (0, 1], the azimuth uniform.
__device__ __forceinline__ V3 vmf_sample(V3 mu, float k, float x1, float x2)
{
    const float x = 1.0f - x1;  // (0, 1]
    float w = 1.0f + logf(x + (1.0f - x) * expf(-2.0f * k)) / k;
    w = fminf(1.0f, fmaxf(-1.0f, w));
    const float st = sqrtf(fmaxf(0.0f, 1.0f - w * w));
    float s, c;
    sincos_2pi(x2, s, c);
    V3 t, b;
    onb(mu, t, b);
    return V3{w * mu.x + st * (c * t.x + s * b.x), w * mu.y + st * (c * t.y + s * b.y),
              w * mu.z + st * (c * t.z + s * b.z)};
}

}  // namespace dr
