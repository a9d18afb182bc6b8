// CPU check of the node-box quantisation (devilray_amd/csrc/node_quant.hpp)
// that the traversal relies on for exact hits: every quantised lower plane is
// at or below the child's plane and every upper plane at or above it, each
// plane the float fma computes is exactly origin + q * step, each is within
// one step of the plane it bounds and the tightest grid point that bounds it,
// and the step exceeds span / 255 by at most 1/64 of it plus the grid unit
// that 2^24 units from zero to the box need. Random boxes over magnitudes
// 2^-140 .. 2^125 and spans down to a few ulps of their position, plus the
// edge cases. Prints "ok <cases>".
#define __host__
#define __device__
#include "../devilray_amd/csrc/node_quant.hpp"

#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <random>

using dr::quant_plane;
using dr::quantize_axis;

static int failures = 0;

static void fail(const char *what, const float lo[4], const float hi[4], const bool use[4])
{
    if (++failures <= 10) {
        std::printf("FAIL %s:", what);
        for (int c = 0; c < 4; ++c)
            if (use[c]) std::printf(" [%a, %a]", lo[c], hi[c]);
        std::printf("\n");
    }
}

static void check(const float lo[4], const float hi[4], const bool use[4])
{
    double plo = INFINITY, phi = -INFINITY;
    for (int c = 0; c < 4; ++c)
        if (use[c]) {
            plo = std::fmin(plo, lo[c]);
            phi = std::fmax(phi, hi[c]);
        }
    float origin, step;
    uint32_t wlo, whi;
    const bool ok = quantize_axis(lo, hi, use, origin, step, wlo, whi);
    const bool in_range = !(plo <= phi) || (std::fabs(plo) < 0x1p126 && std::fabs(phi) < 0x1p126);
    if (ok != in_range) return fail(ok ? "accepted out of range" : "rejected in range", lo, hi, use);
    if (!ok) return;
    if (!(plo <= phi)) {
        if (wlo != 0xFFFFFFFFu || whi != 0u) fail("empty node bytes", lo, hi, use);
        return;
    }
    if (!(step > 0.0f) || !std::isfinite(step) || !std::isfinite(origin)) return fail("grid not finite", lo, hi, use);
    const double span = phi - plo, st = step;
    // the grid unit is 2^-7 of span / 255, or as large as 2^24 units from
    // zero to the box need (about 2^-23 of the box's position)
    if (!(st <= span / 255.0 * (1.0 + 1.0 / 64.0) + std::fabs(plo) * 0x1p-21 + 0x1p-92))
        fail("step not tight", lo, hi, use);
    for (int c = 0; c < 4; ++c) {
        const uint32_t ql = (wlo >> (8 * c)) & 0xFFu, qh = (whi >> (8 * c)) & 0xFFu;
        const float pl = quant_plane(wlo, c, step, origin), ph = quant_plane(whi, c, step, origin);
        if (!use[c]) {
            if (ql != 255u || qh != 0u) fail("unused child bytes", lo, hi, use);
            continue;
        }
        // origin + q step in double is exact (both are multiples of the
        // grid unit, below 2^53 units), so this compares with the real value
        if ((double)pl != (double)origin + ql * st || (double)ph != (double)origin + qh * st)
            fail("plane not exact", lo, hi, use);
        if (!(pl <= lo[c]) || !(ph >= hi[c])) fail("box not contained", lo, hi, use);
        if (ql < 255u && (double)origin + (ql + 1) * st <= (double)lo[c]) fail("lower plane not the tightest", lo, hi, use);
        if (qh > 0u && (double)origin + (qh - 1) * st >= (double)hi[c]) fail("upper plane not the tightest", lo, hi, use);
        if (!((double)lo[c] - pl <= st) || !((double)ph - hi[c] <= st)) fail("plane more than a step out", lo, hi, use);
    }
}

int main(int argc, char **argv)
{
    const long n = argc > 1 ? std::atol(argv[1]) : 200000;
    std::mt19937_64 rng(20240611);
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    std::uniform_int_distribution<int> mag(-140, 125), rel(0, 40), nuse(1, 4);
    long cases = 0;
    for (long i = 0; i < n; ++i) {
        const double R = std::ldexp(1.0, mag(rng));
        const double centre = (2.0 * u01(rng) - 1.0) * R;
        const double span = R * std::ldexp(1.0, -rel(rng)) * u01(rng);
        float lo[4], hi[4];
        bool use[4] = {false, false, false, false};
        const int k = nuse(rng);
        for (int c = 0; c < 4; ++c) {
            use[c] = c < k;
            double a = centre + span * u01(rng), b = a + span * u01(rng) * (u01(rng) < 0.1 ? 0.0 : 1.0);
            lo[c] = (float)a;
            hi[c] = (float)b;
            if (hi[c] < lo[c]) hi[c] = lo[c];
            if (!std::isfinite(lo[c]) || !std::isfinite(hi[c])) use[c] = false, lo[c] = INFINITY, hi[c] = -INFINITY;
        }
        check(lo, hi, use);
        ++cases;
    }
    // edge cases: planes on the grid, zero and negative zero, denormals,
    // the largest accepted and the first rejected magnitudes, one point,
    // all four children identical, no child in use
    const float edge[][2] = {{0.0f, 0.0f}, {-0.0f, 0.0f}, {1e-40f, 2e-40f}, {-1e-45f, 1e-45f},
                             {1.0f, 1.0f}, {-1.0f, 255.0f}, {0.0f, 256.0f}, {-0x1.fffffep125f, 0x1.fffffep125f},
                             {0x1p126f, 0x1p126f}, {-0x1p126f, 0.0f}, {3e25f, 3.0000002e25f},
                             {16777215.0f, 16777216.0f}, {-FLT_MIN, FLT_MIN}};
    for (const auto &b : edge) {
        const float lo[4] = {b[0], b[0], b[0], b[0]}, hi[4] = {b[1], b[1], b[1], b[1]};
        for (int k = 1; k <= 4; ++k) {
            bool use[4];
            for (int c = 0; c < 4; ++c) use[c] = c < k;
            check(lo, hi, use);
            ++cases;
        }
    }
    // planes far below the grid unit next to planes of order one: the
    // difference with the grid origin rounds in double, and only the exact
    // comparisons with grid points keep such a plane inside its box
    const float tiny[] = {1e-30f, -1e-30f, 1e-38f, 3e-20f, -7e-12f};
    // (in [-1, 2.9745] the grid is -8192 + 128 q units of 2^-13, so grid
    // point 64 is exactly zero, and a plane at +-1e-30 loses its offset from
    // it in the double difference with the origin; likewise the others)
    const float big[][2] = {{-1.0f, 2.9745002f}, {-2.25f, 5.6936255f}, {-6.0f, 5.9136553f}, {-0.078125f, 0.91507417f},
                            {-1.0f, 1.0f}, {-3.0f, 0.5f}, {-1e3f, 2e3f}, {-0.7f, 0.3f}};
    for (float t : tiny)
        for (const auto &b : big) {
            const float lo[4] = {b[0], t, t, b[0]}, hi[4] = {t, t, b[1], b[1]};
            const bool use[4] = {true, true, true, true};
            check(lo, hi, use);
            ++cases;
        }
    {
        const float lo[4] = {INFINITY, INFINITY, INFINITY, INFINITY}, hi[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
        const bool use[4] = {false, false, false, false};
        check(lo, hi, use);
        ++cases;
    }
    if (failures) {
        std::printf("%d failures in %ld cases\n", failures, cases);
        return 1;
    }
    std::printf("ok %ld\n", cases);
    return 0;
}
