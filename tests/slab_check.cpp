// CPU check of the traversal's box test on quantised nodes (qnode_slabs in
// devilray_amd/csrc/node_quant.hpp): whenever a ray meets a child's quantised
// box within [tmin, lim] by an exact evaluation (long double: the plane
// differences are exact, the quotients good to 2^-64), the test must take the
// child. Rays: random ones, and grazing ones aimed at
// points on the edges and corners of the decoded boxes, where the exact entry
// and exit distances nearly coincide and only the far scale keeps the child;
// directions with zero and tiny components. A child that the ray clearly
// misses (its box grown by a thousandth is still missed) must be culled.
// Prints "ok <rays>".
#define __host__
#define __device__
#include "../devilray_amd/csrc/node_quant.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

using namespace dr;

static int failures = 0;

struct Box {
    float lo[3], hi[3];
};

// exact slab test on one decoded box
static bool exact_hit(const Box &b, const float o[3], const float d[3], float tmin, float lim)
{
    long double tn = tmin, tf = lim;
    for (int a = 0; a < 3; ++a) {
        if (d[a] == 0.0f) {
            if (!(o[a] >= b.lo[a] && o[a] <= b.hi[a])) return false;
            continue;
        }
        long double t0 = ((long double)b.lo[a] - o[a]) / d[a], t1 = ((long double)b.hi[a] - o[a]) / d[a];
        if (t0 > t1) std::swap(t0, t1);
        tn = std::fmax(tn, t0);
        tf = std::fmin(tf, t1);
    }
    return tn <= tf;
}

int main(int argc, char **argv)
{
    const long n_nodes = argc > 1 ? std::atol(argv[1]) : 20000;
    std::mt19937_64 rng(77);
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    long rays = 0;
    for (long it = 0; it < n_nodes; ++it) {
        // four child boxes around a random centre at a random scale
        const double scale = std::ldexp(1.0, (int)(u01(rng) * 40) - 20);
        const double cx = (u01(rng) * 2 - 1) * scale * std::ldexp(1.0, (int)(u01(rng) * 12));
        float lo[3][4], hi[3][4];
        bool use[4];
        for (int c = 0; c < 4; ++c) {
            use[c] = c == 0 || u01(rng) < 0.9;
            for (int a = 0; a < 3; ++a) {
                const double p = cx + (u01(rng) * 2 - 1) * scale;
                const double e = u01(rng) < 0.2 ? 0.0 : scale * u01(rng) * u01(rng);  // some flat boxes
                lo[a][c] = (float)p;
                hi[a][c] = (float)(p + e);
                if (hi[a][c] < lo[a][c]) hi[a][c] = lo[a][c];
            }
        }
        QNodeWords w;
        float *org[3] = {&w.ox, &w.oy, &w.oz}, *stp[3] = {&w.sx, &w.sy, &w.sz};
        bool ok = true;
        for (int a = 0; a < 3; ++a) ok = quantize_axis(lo[a], hi[a], use, *org[a], *stp[a], w.lo[a], w.hi[a]) && ok;
        if (!ok) continue;
        Box qb[4];
        for (int c = 0; c < 4; ++c)
            for (int a = 0; a < 3; ++a) {
                qb[c].lo[a] = quant_plane(w.lo[a], c, *stp[a], *org[a]);
                qb[c].hi[a] = quant_plane(w.hi[a], c, *stp[a], *org[a]);
            }
        for (int k = 0; k < 24; ++k) {
            float o[3], d[3];
            // origin: near the node or far from it
            const double reach = scale * (u01(rng) < 0.5 ? 1.5 : std::ldexp(1.0, (int)(u01(rng) * 16)));
            for (int a = 0; a < 3; ++a) o[a] = (float)(cx + (u01(rng) * 2 - 1) * reach);
            const int mode = k % 4;
            double tgt[3];
            const int c = (int)(u01(rng) * 4) & 3;
            if (mode == 0) {  // random direction
                for (int a = 0; a < 3; ++a) tgt[a] = o[a] + (u01(rng) * 2 - 1);
            } else {  // a point on an edge or corner of child c's box
                for (int a = 0; a < 3; ++a) {
                    const double l = qb[c].lo[a], h = qb[c].hi[a];
                    tgt[a] = (mode == 3 || a != k % 3) ? (u01(rng) < 0.5 ? l : h) : l + (h - l) * u01(rng);
                }
            }
            for (int a = 0; a < 3; ++a) d[a] = (float)(tgt[a] - o[a]);
            if (k % 7 == 5) d[k % 3] = 0.0f;                        // a zero component
            if (k % 11 == 7) d[(k + 1) % 3] = (float)(d[(k + 1) % 3] * 1e-30);  // a tiny one
            if (d[0] == 0.0f && d[1] == 0.0f && d[2] == 0.0f) continue;
            SlabRay r;
            r.ox = o[0]; r.oy = o[1]; r.oz = o[2];
            r.ix = 1.0f / d[0]; r.iy = 1.0f / d[1]; r.iz = 1.0f / d[2];
            r.negx = std::signbit(r.ix); r.negy = std::signbit(r.iy); r.negz = std::signbit(r.iz);
            r.tmin = u01(rng) < 0.5 ? 0.0f : (float)(u01(rng) * 0.5);
            const float lim = u01(rng) < 0.5 ? INFINITY : (float)(0.5 + u01(rng) * 2.0);
            float nre[4], fre[4];
            qnode_slabs(w, r, lim, nre, fre);
            ++rays;
            for (int cc = 0; cc < 4; ++cc) {
                if (!use[cc]) continue;
                if (!exact_hit(qb[cc], o, d, r.tmin, lim)) {
                    // a clear miss is culled: the box grown by a thousandth
                    // of its size and position and by 2^-16 of its distance
                    // from the origin (the relative slack of the far scale is
                    // 2^-20) is still missed
                    Box g = qb[cc];
                    double dist = 0.0;
                    for (int a = 0; a < 3; ++a)
                        dist = std::fmax(dist, std::fmax(std::fabs((double)g.lo[a] - o[a]), std::fabs((double)g.hi[a] - o[a])));
                    for (int a = 0; a < 3; ++a) {
                        const float m = 1e-3f * (g.hi[a] - g.lo[a]) + 1e-3f * std::fmax(std::fabs(g.lo[a]), std::fabs(g.hi[a])) +
                                        (float)(dist * 0x1p-16) + 1e-30f;
                        g.lo[a] -= m;
                        g.hi[a] += m;
                    }
                    if (!exact_hit(g, o, d, r.tmin, lim) && nre[cc] <= fre[cc]) {
                        if (++failures <= 8)
                            std::printf("FAIL clear miss taken: o (%a %a %a) d (%a %a %a) box (%a %a %a)-(%a %a %a) tmin %a lim %a nr %a fr %a\n",
                                        o[0], o[1], o[2], d[0], d[1], d[2], qb[cc].lo[0], qb[cc].lo[1], qb[cc].lo[2],
                                        qb[cc].hi[0], qb[cc].hi[1], qb[cc].hi[2], r.tmin, lim, nre[cc], fre[cc]);
                    }
                    continue;
                }
                if (!(nre[cc] <= fre[cc])) {
                    if (++failures <= 8)
                        std::printf("FAIL hit dropped: o (%a %a %a) d (%a %a %a) box (%a %a %a)-(%a %a %a) tmin %a lim %a\n",
                                    o[0], o[1], o[2], d[0], d[1], d[2], qb[cc].lo[0], qb[cc].lo[1], qb[cc].lo[2],
                                    qb[cc].hi[0], qb[cc].hi[1], qb[cc].hi[2], r.tmin, lim);
                }
            }
        }
    }
    if (failures) {
        std::printf("%d failures in %ld rays\n", failures, rays);
        return 1;
    }
    std::printf("ok %ld\n", rays);
    return 0;
}
