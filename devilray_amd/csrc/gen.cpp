// Deterministic workload generators (host): icospheres with seeded radial
// displacement, incoherent surface rays and pinhole camera rays. All random
// draws come from Philox keyed by the seed, so every run and every consumer
// sees identical inputs.
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../../include/dr_api.h"
#include "philox.hpp"

namespace {

struct D3 {
    double x, y, z;
};
inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 mul(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline D3 norm(D3 a)
{
    const double l = std::sqrt(dot(a, a));
    return {a.x / l, a.y / l, a.z / l};
}

dr::U4 draw(uint64_t seed, uint32_t a, uint32_t b)
{
    return dr::philox4x32_10(dr::U4{a, b, 0u, 0u}, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// (0, 1] and [0, 1) doubles with 32 random bits
inline double open01(uint32_t x) { return ((double)x + 1.0) * (1.0 / 4294967296.0); }
inline double unit01(uint32_t x) { return (double)x * (1.0 / 4294967296.0); }

void icosphere_dirs(uint32_t subdiv, std::vector<D3> &v, std::vector<uint32_t> &f)
{
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto &p : v) p = norm(p);
    f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
         3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (uint32_t s = 0; s < subdiv; ++s) {
        std::unordered_map<uint64_t, uint32_t> mid;
        mid.reserve(f.size());
        auto midpoint = [&](uint32_t a, uint32_t b) {
            const uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            const uint32_t idx = (uint32_t)v.size();
            v.push_back(norm(mul(add(v[a], v[b]), 0.5)));
            mid.emplace(key, idx);
            return idx;
        };
        std::vector<uint32_t> nf;
        nf.reserve(f.size() * 4);
        for (size_t k = 0; k < f.size(); k += 3) {
            const uint32_t a = f[k], b = f[k + 1], c = f[k + 2];
            const uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            const uint32_t tri[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), tri, tri + 12);
        }
        f.swap(nf);
    }
}

// Duff et al., "Building an Orthonormal Basis, Revisited", JCGT 6(1) 2017.
void onb(D3 n, D3 &t, D3 &b)
{
    const double s = n.z >= 0.0 ? 1.0 : -1.0;
    const double a = -1.0 / (s + n.z);
    const double c = n.x * n.y * a;
    t = {1.0 + s * n.x * n.x * a, s * c, -s * n.x};
    b = {c, s + n.y * n.y * a, -n.y};
}

}  // namespace

extern "C" {

dr_status dr_gen_icosphere(uint32_t subdiv, float sigma, uint64_t seed, float *vertices, uint32_t *n_vertices,
                           uint32_t *indices, uint32_t *n_triangles)
{
    if (subdiv > 10) return DR_E_INVALID;
    const uint64_t nt = 20ull << (2 * subdiv);
    const uint64_t nv = 10ull * (1ull << (2 * subdiv)) + 2;
    if (n_vertices) *n_vertices = (uint32_t)nv;
    if (n_triangles) *n_triangles = (uint32_t)nt;
    if (!vertices && !indices) return DR_OK;
    std::vector<D3> v;
    std::vector<uint32_t> f;
    icosphere_dirs(subdiv, v, f);
    if (v.size() != nv || f.size() != 3 * nt) return DR_E_DEVICE;
    if (vertices) {
        for (size_t i = 0; i < v.size(); ++i) {
            const dr::U4 r = draw(seed, (uint32_t)i, 0x1C05u);
            const double g = std::sqrt(-2.0 * std::log(open01(r.x))) * std::cos(2.0 * M_PI * unit01(r.y));
            const double rad = 1.0 + (double)sigma * g;
            vertices[3 * i] = (float)(v[i].x * rad);
            vertices[3 * i + 1] = (float)(v[i].y * rad);
            vertices[3 * i + 2] = (float)(v[i].z * rad);
        }
    }
    if (indices) std::memcpy(indices, f.data(), sizeof(uint32_t) * f.size());
    return DR_OK;
}

dr_status dr_gen_rays_surface(const float *vertices, const uint32_t *indices, uint32_t n_triangles, uint32_t n_rays,
                              uint64_t seed, int32_t inward, float offset, dr_ray *rays)
{
    if (!vertices || !indices || !rays || n_triangles == 0) return DR_E_INVALID;
    std::vector<double> cdf(n_triangles);
    double acc = 0.0;
    auto vert = [&](uint32_t k) {
        return D3{vertices[3 * (size_t)k], vertices[3 * (size_t)k + 1], vertices[3 * (size_t)k + 2]};
    };
    for (uint32_t t = 0; t < n_triangles; ++t) {
        const D3 a = vert(indices[3 * (size_t)t]), b = vert(indices[3 * (size_t)t + 1]), c = vert(indices[3 * (size_t)t + 2]);
        const D3 n = cross(sub(b, a), sub(c, a));
        acc += 0.5 * std::sqrt(dot(n, n));
        cdf[t] = acc;
    }
    for (uint32_t i = 0; i < n_rays; ++i) {
        const dr::U4 r0 = draw(seed, i, 0x5A1Du);
        const dr::U4 r1 = draw(seed, i, 0x5A1Eu);
        const double target = unit01(r0.x) * acc;
        uint32_t lo = 0, hi = n_triangles - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (cdf[mid] > target) hi = mid; else lo = mid + 1;
        }
        const D3 a = vert(indices[3 * (size_t)lo]), b = vert(indices[3 * (size_t)lo + 1]), c = vert(indices[3 * (size_t)lo + 2]);
        const double su = std::sqrt(unit01(r0.y));
        const double b0 = 1.0 - su, b1 = unit01(r0.z) * su, b2 = 1.0 - b0 - b1;
        const D3 p = add(add(mul(a, b0), mul(b, b1)), mul(c, b2));
        D3 n = norm(cross(sub(b, a), sub(c, a)));
        if (inward) n = mul(n, -1.0);
        D3 tt, bb;
        onb(n, tt, bb);
        const double rr = std::sqrt(unit01(r1.x));
        const double phi = 2.0 * M_PI * unit01(r1.y);
        const double z = std::sqrt(std::fmax(0.0, 1.0 - rr * rr));
        const D3 d = norm(add(add(mul(tt, rr * std::cos(phi)), mul(bb, rr * std::sin(phi))), mul(n, z)));
        const D3 o = add(p, mul(d, (double)offset));
        rays[i] = dr_ray{(float)o.x, (float)o.y, (float)o.z, 0.0f, (float)d.x, (float)d.y, (float)d.z, INFINITY};
    }
    return DR_OK;
}

dr_status dr_gen_rays_camera(const float eye[3], const float target[3], const float up[3], float fov_y_deg,
                             uint32_t width, uint32_t height, dr_ray *rays)
{
    if (!eye || !target || !up || !rays || width == 0 || height == 0) return DR_E_INVALID;
    const D3 e{eye[0], eye[1], eye[2]};
    const D3 fwd = norm(sub(D3{target[0], target[1], target[2]}, e));
    const D3 right = norm(cross(fwd, D3{up[0], up[1], up[2]}));
    const D3 u = cross(right, fwd);
    const double th = std::tan(0.5 * fov_y_deg * M_PI / 180.0);
    const double aspect = (double)width / (double)height;
    for (uint32_t y = 0; y < height; ++y)
        for (uint32_t x = 0; x < width; ++x) {
            const double sx = (2.0 * (x + 0.5) / width - 1.0) * th * aspect;
            const double sy = (1.0 - 2.0 * (y + 0.5) / height) * th;
            const D3 d = norm(add(add(fwd, mul(right, sx)), mul(u, sy)));
            rays[(size_t)y * width + x] = dr_ray{(float)e.x, (float)e.y, (float)e.z, 0.0f,
                                                 (float)d.x, (float)d.y, (float)d.z, INFINITY};
        }
    return DR_OK;
}

}  // extern "C"
