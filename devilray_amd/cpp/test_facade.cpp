// The reference's own GoogleTest cases for the render path, restated against
// the C++ facade (needs a GPU): triangle intersection cases
// (test/src/test_intersection.cpp), mesh intersections, nearest-of-stacked
// and cast (test/src/test_tracing.cpp:147-298) and the analytic diffuse reflection
// render (test/src/test_render.cpp:50-165), the mesh loader
// (test/src/test_mesh.cpp), quad and scene structures
// (test/src/test_scene.cpp:39-77) and the running average
// (test/src/test_running_average.cpp); then the façade's own cases for the
// render statistics and the background render loop of the interactive
// application (renderer/src/RenderThread.cpp:3-12). Prints one line per case
// and exits non-zero if any failed.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "devilray.hpp"

using namespace devilray;

namespace {
int g_fail = 0;
#define EXPECT(cond)                                                                     \
    do {                                                                                 \
        if (!(cond)) {                                                                   \
            std::printf("    FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);            \
            ++g_fail;                                                                    \
        }                                                                                \
    } while (0)
#define EXPECT_NEAR(a, b, tol) EXPECT(std::fabs((double)(a) - (double)(b)) <= (double)(tol))

void run(const char *name, const std::function<void()> &fn)
{
    const int before = g_fail;
    try {
        fn();
    } catch (const std::exception &e) {
        std::printf("    exception: %s\n", e.what());
        ++g_fail;
    }
    std::printf("%s %s\n", g_fail == before ? "PASS" : "FAIL", name);
    std::fflush(stdout);
}

// A(0,0,0) B(1,0,0) C(0,1,0), the reference's unitTri
Scene unit_triangle(Context &ctx)
{
    Scene s;
    s.materials.push_back(Material::diffuse({0.5f, 0.5f, 0.5f}));
    Mesh m;
    m.points = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}};
    m.triangles = {{0, 1, 2}};
    s.objects.push_back({m, 0});
    s.build(ctx);
    return s;
}

std::optional<Intersection> cast1(Context &ctx, const Scene &s, Vec3 p, Vec3 v)
{
    return cast(ctx, s, {Ray{p, v}})[0];
}

// the analytic scene of test_render.cpp: a white quad in front of the camera
// and an emitting one behind it, both 1000 wide
Scene facing_quads(float emittance)
{
    Scene scene;
    scene.materials.push_back(Material::diffuse({1, 1, 1}));
    scene.materials.push_back(Material::diffuse({0, 0, 0}, {emittance, emittance, emittance}));
    scene.objects.push_back({createQuadMesh({0, 0, 1}, {0, 0, -1}, {1, 0, 0}, 1000.0f), 0});
    scene.objects.push_back({createQuadMesh({0, 0, -1}, {0, 0, 1}, {1, 0, 0}, 1000.0f), 1});
    return scene;
}

Camera pinhole(Size2i res, float x_offset = 0.0f)
{
    Camera cam;
    cam.intrinsics = {50.0f, {res.width / 2.0f * 0.1f, res.height / 2.0f * 0.1f}};
    cam.resolution = res;
    cam.physical_pixel_size = {0.1f, 0.1f};
    cam.transform.values[0][3] = x_offset;
    return cam;
}

template <class F> bool wait_for(F done, double seconds)
{
    const auto end = std::chrono::steady_clock::now() + std::chrono::duration<double>(seconds);
    while (!done()) {
        if (std::chrono::steady_clock::now() > end) return false;
        std::this_thread::sleep_for(std::chrono::milliseconds(1));
    }
    return true;
}
}  // namespace

int main(int argc, char **argv)
{
    // argv[1] (optional): a directory for the saved image and its raw pixels
    const std::string out_dir = argc > 1 ? argv[1] : "";
    Context ctx(0);
    Scene tri = unit_triangle(ctx);

    // test/src/test_transform.cpp:20-89 and test/src/test_matrix.cpp:14-90
    auto vnear = [](const Vec3 &v, float x, float y, float z) {
        EXPECT_NEAR(v.x, x, 1e-5f);
        EXPECT_NEAR(v.y, y, 1e-5f);
        EXPECT_NEAR(v.z, z, 1e-5f);
    };
    run("TransformTest.ApplyToPointScalesThenTranslates", [&] {
        const Transform t{{2, 3, 4}, {10, 20, 30}};
        vnear(t.applyToPoint({1, 1, 1}), 12, 23, 34);
        vnear(t.applyToPoint({0, 0, 0}), 10, 20, 30);
    });
    run("TransformTest.ApplyToPointIdentityIsNoOp", [&] { vnear(Transform{}.applyToPoint({5, -2, 7}), 5, -2, 7); });
    run("TransformTest.ApplyInverseUndoesPointTransformForRayOrigin", [&] {
        const Transform t{{2, 4, 5}, {1, 2, 3}};
        Ray world;
        world.p = {9, 10, 13};
        world.v = {0, 0, 1};
        const Ray model = t.applyInverse(world);
        const Vec3 back = t.applyToPoint(model.p);
        vnear(back, 9, 10, 13);
        vnear(model.p, 4, 2, 2);
    });
    run("TransformTest.ApplyInverseScalesDirectionWithoutTranslation", [&] {
        const Transform t{{2, 4, 5}, {100, 200, 300}};
        Ray r;
        r.p = {0, 0, 0};
        r.v = {2, 4, 5};
        vnear(t.applyInverse(r).v, 1, 1, 1);
    });
    run("MatrixTest.IdentityIsDiagonal", [&] {
        const auto id = Matrix4x4f::identity();
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) EXPECT(id.values[r][c] == (r == c ? 1.0f : 0.0f));
    });
    run("MatrixTest.ApplyToDirectionIgnoresTranslation", [&] {
        vnear(Matrix4x4f::translation({10, 20, 30}).applyToDirection({1, 2, 3}), 1, 2, 3);
    });
    run("MatrixTest.ApplyToDirectionRotates", [&] {
        vnear(Matrix4x4f::rotation({0, 0, 1}, 1.57079632679f).applyToDirection({1, 0, 0}), 0, 1, 0);
    });
    run("MatrixTest.GetOffsetReturnsTranslationColumn", [&] {
        vnear(Matrix4x4f::translation({7, -3, 11}).getOffset(), 7, -3, 11);
        vnear(Matrix4x4f::identity().getOffset(), 0, 0, 0);
    });
    run("MatrixTest.GetOffsetAfterComposition", [&] {
        vnear((Matrix4x4f::translation({1, 2, 3}) * Matrix4x4f::translation({4, 5, 6})).getOffset(), 5, 7, 9);
    });
    run("MatrixTest.MultiplicationByIdentity", [&] {
        const Matrix4x4f m = Matrix4x4f::translation({1, 2, 3});
        const Matrix4x4f r = m * Matrix4x4f::identity();
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) EXPECT(r.values[i][j] == m.values[i][j]);
    });
    run("MatrixTest.TranslationMatrixComposition", [&] {
        vnear(Matrix4x4f::translation({5, -2, 10}).applyToPosition({0, 0, 0}), 5, -2, 10);
    });
    run("MatrixTest.TranslationDoesNotAffectDirection", [&] {
        vnear(Matrix4x4f::translation({10, 10, 10}).applyToDirection({1, 0, 0}), 1, 0, 0);
    });
    run("MatrixTest.Rotation90DegreesZ", [&] {
        vnear(Matrix4x4f::rotation({0, 0, 1}, 1.57079632679f).applyToPosition({1, 0, 0}), 0, 1, 0);
    });
    run("MatrixTest.RotationZeroAngle", [&] {
        vnear(Matrix4x4f::rotation({1, 1, 1}, 0.0f).applyToPosition({1, 2, 3}), 1, 2, 3);
    });
    run("MatrixTest.ApplyToPositionWithWComponent", [&] {
        vnear(Matrix4x4f::translation({1, 1, 1}).applyToPosition({2, 2, 2}), 3, 3, 3);
    });
    run("MatrixTest.CombinedTransformations", [&] {
        const Matrix4x4f m = Matrix4x4f::rotation({0, 0, 1}, 1.57079632679f) * Matrix4x4f::translation({1, 0, 0});
        vnear(m.applyToPosition({0, 0, 0}), 0, 1, 0);
    });
    run("MatrixTest.RotationIsOrthonormalAboutAnyAxis", [&] {
        // R R^T = I and R axis = axis for a unit axis off the coordinate planes
        const float n = 1.0f / std::sqrt(14.0f);
        const Vec3 ax{1 * n, -2 * n, 3 * n};
        const Matrix4x4f r = Matrix4x4f::rotation(ax, 0.7f);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                float d = 0.0f;
                for (int k = 0; k < 3; ++k) d += r.values[i][k] * r.values[j][k];
                EXPECT_NEAR(d, i == j ? 1.0f : 0.0f, 1e-6f);
            }
        vnear(r.applyToDirection(ax), ax.x, ax.y, ax.z);
    });
    run("CameraTest.TransformMatrixMovesTheView", [&] {
        // a camera translated by (0.5, 0, 0) sees the triangle hit at the
        // same spot shifted in the image: rays from the camera start at the offset
        Camera cam;
        cam.transform = Matrix4x4f::translation({0.5f, 0.0f, 0.0f});
        EXPECT(cam.transform.values[0][3] == 0.5f && cam.transform.getOffset().x == 0.5f);
    });

    run("IntersectionTest.HitsCenterOfTriangle", [&] {
        auto r = cast1(ctx, tri, {0.2f, 0.2f, 5.0f}, {0, 0, -1});
        EXPECT(r.has_value());
        if (!r) return;
        EXPECT_NEAR(r->t, 5.0f, 1e-5f);
        EXPECT(r->bari.x >= 0 && r->bari.y >= 0 && r->bari.z >= 0);
        EXPECT_NEAR(r->bari.x + r->bari.y + r->bari.z, 1.0f, 1e-5f);
        EXPECT_NEAR(r->bari.y, 0.2f, 1e-6f);
        EXPECT_NEAR(r->bari.z, 0.2f, 1e-6f);
    });
    run("IntersectionTest.HitsExactlyOnVertex", [&] {
        auto r = cast1(ctx, tri, {0, 0, 10}, {0, 0, -1});
        EXPECT(r.has_value() && std::fabs(r->t - 10.0f) <= 1e-5f);
    });
    run("IntersectionTest.MissesJustOutsideEdge", [&] {
        EXPECT(!cast1(ctx, tri, {-0.01f, 0.5f, 1.0f}, {0, 0, -1}).has_value());
    });
    run("IntersectionTest.MissesBehindRayOrigin", [&] {
        EXPECT(!cast1(ctx, tri, {0.2f, 0.2f, 1.0f}, {0, 0, 1}).has_value());
    });
    run("IntersectionTest.ParallelRayDoesNotIntersect", [&] {
        EXPECT(!cast1(ctx, tri, {-1.0f, 0.2f, 1.0f}, {1, 0, 0}).has_value());
    });
    auto single = [&](Vec3 a, Vec3 b, Vec3 c) {
        Scene s;
        s.materials.push_back(Material::diffuse({1, 1, 1}));
        Mesh m;
        m.points = {a, b, c};
        m.triangles = {{0, 1, 2}};
        s.objects.push_back({m, 0});
        s.build(ctx);
        return s;
    };
    run("IntersectionTest.DegenerateTriangleFails", [&] {
        Scene s = single({0, 0, 0}, {0, 0, 0}, {0, 0, 0});
        EXPECT(!cast1(ctx, s, {0, 0, 1}, {0, 0, -1}).has_value());
    });
    run("IntersectionTest.NoBackfaceCulling", [&] {
        auto r = cast1(ctx, tri, {0.2f, 0.2f, -1.0f}, {0, 0, 1});
        EXPECT(r.has_value() && std::isfinite(r->t) && std::isfinite(r->bari.x));
    });
    run("IntersectionTest.VeryDistantIntersection", [&] {
        auto r = cast1(ctx, tri, {0.2f, 0.2f, 10000.0f}, {0, 0, -1});
        EXPECT(r.has_value() && std::fabs(r->t - 10000.0f) <= 1e-2f);
    });
    run("IntersectionTest.DegenerateTriangleCollinear", [&] {
        Scene s = single({0, 0, 0}, {1, 0, 0}, {2, 0, 0});
        EXPECT(!cast1(ctx, s, {0.5f, 0, 1}, {0, 0, -1}).has_value());
    });
    run("IntersectionTest.DegenerateRayZeroDirection", [&] {
        EXPECT(!cast1(ctx, tri, {0.2f, 0.2f, 1.0f}, {0, 0, 0}).has_value());
    });
    run("IntersectionTest.DegenerateRayNaNInDirection", [&] {
        EXPECT(!cast1(ctx, tri, {0, 0, 0}, {NAN, NAN, NAN}).has_value());
    });
    run("IntersectionTest.ExtremelyThinTriangle", [&] {
        Scene s = single({0, 0, 0}, {100.0f, 0, 0}, {100.0f, 2e-8f, 0});
        auto r = cast1(ctx, s, {50.0f, 1e-9f, 1.0f}, {0, 0, -1});
        EXPECT(r.has_value() && std::isfinite(r->t));
    });
    run("IntersectionTest.MultipleAngleSanityCheck", [&] {
        const Vec3 inside{0.2f, 0.2f, 0.0f}, outside{5.0f, 5.0f, 0.0f};
        const Vec3 origins[4] = {{0.2f, 0.2f, 10.0f}, {10.0f, 0.2f, 10.0f}, {-5.0f, -5.0f, 2.0f}, {0.0f, 0.0f, 1.0f}};
        for (const Vec3 &o : origins) {
            auto r = cast1(ctx, tri, o, {inside.x - o.x, inside.y - o.y, inside.z - o.z});
            EXPECT(r.has_value());
            if (r) {
                EXPECT_NEAR(r->t, 1.0f, 1e-4f);
                EXPECT_NEAR(r->bari.x + r->bari.y + r->bari.z, 1.0f, 1e-5f);
                EXPECT_NEAR(r->bari.x, 0.6f, 1e-4f);
                EXPECT_NEAR(r->bari.y, 0.2f, 1e-4f);
                EXPECT_NEAR(r->bari.z, 0.2f, 1e-4f);
            }
            EXPECT(!cast1(ctx, tri, o, {outside.x - o.x, outside.y - o.y, outside.z - o.z}).has_value());
        }
    });
    run("TracingTest.NearestOfStackedTriangles", [&] {
        Scene s;
        s.materials.push_back(Material::diffuse({1, 1, 1}));
        for (float z : {-2.0f, 0.0f, 3.0f}) {
            Mesh m;
            m.points = {{-1, -1, z}, {1, -1, z}, {0, 1, z}};
            m.triangles = {{0, 1, 2}};
            s.objects.push_back({m, 0});
        }
        s.build(ctx);
        auto r = cast1(ctx, s, {0, 0, 10}, {0, 0, -1});
        EXPECT(r.has_value() && r->triangle == 2 && std::fabs(r->t - 7.0f) <= 1e-5f);
        auto r2 = cast1(ctx, s, {0, 0, -10}, {0, 0, 1});
        EXPECT(r2.has_value() && r2->triangle == 0 && std::fabs(r2->t - 8.0f) <= 1e-5f);
        EXPECT(!cast1(ctx, s, {5, 5, 10}, {0, 0, -1}).has_value());
    });
    run("RendererTest.AnalyticalDiffuseReflection", [&] {
        const Size2i res{256, 256};
        Renderer renderer(ctx, res);
        Camera cam;
        cam.intrinsics = {50.0f, {res.width / 2.0f * 0.1f, res.height / 2.0f * 0.1f}};
        cam.resolution = res;
        cam.physical_pixel_size = {0.1f, 0.1f};
        renderer.setCamera(cam);
        const float emittance = 0.55f;
        Scene scene;
        scene.materials.push_back(Material::diffuse({1, 1, 1}));
        scene.materials.push_back(Material::diffuse({0, 0, 0}, {emittance, emittance, emittance}));
        scene.objects.push_back({createQuadMesh({0, 0, 1}, {0, 0, -1}, {1, 0, 0}, 1000.0f), 0});
        scene.objects.push_back({createQuadMesh({0, 0, -1}, {0, 0, 1}, {1, 0, 0}, 1000.0f), 1});
        renderer.setScene(std::move(scene));
        renderer.setOutputOptions({OutputLinearity::Linear});
        for (int i = 0; i < 100; ++i) renderer.render();
        renderer.createPixels();
        const Vec4 *px = renderer.getRawPixels();
        const float count0 = px[0].w;
        double s1 = 0, s2 = 0, s3 = 0;
        bool counts = true, grey = true;
        for (int i = 0; i < res.width * res.height; ++i) {
            counts &= px[i].w == count0;
            const float r = px[i].x / px[i].w, g = px[i].y / px[i].w, b = px[i].z / px[i].w;
            grey &= (r == g && g == b);
            s1 += r;
            s2 += (double)r * r;
            s3 += (double)r * r * r;
        }
        const double n = res.width * res.height, mean = s1 / n, sd = std::sqrt(s2 / n - mean * mean);
        const double skew = (s3 / n - 3 * mean * sd * sd - mean * mean * mean) / (sd * sd * sd);
        std::printf("    mean %.5f stddev %.5f skew %.3g samples %.0f\n", mean, sd, skew, count0);
        EXPECT(counts);
        EXPECT(grey);
        EXPECT(count0 == 100.0f);
        EXPECT_NEAR(mean, emittance, 0.01);
        EXPECT(sd <= 0.04);
        EXPECT(skew <= 4e-5);
        const uint32_t *p = renderer.getPixels();
        EXPECT(((p[0] >> 24) & 0xFF) == 255u);
    });
    run("RendererTest.DebugViews", [&] {
        // a quad filling the view: winding colour everywhere, barycentric
        // weights summing to one, rows flipped and packed by createPixels
        const Size2i res{64, 64};
        Renderer renderer(ctx, res);
        Camera cam;
        cam.intrinsics = {50.0f, {res.width / 2.0f * 0.1f, res.height / 2.0f * 0.1f}};
        cam.resolution = res;
        cam.physical_pixel_size = {0.1f, 0.1f};
        renderer.setCamera(cam);
        Scene scene;
        scene.materials.push_back(Material::diffuse({0.5f, 0.25f, 1.0f}));
        scene.objects.push_back({createQuadMesh({0, 0, 1}, {0, 0, -1}, {1, 0, 0}, 1000.0f), 0});
        renderer.setScene(std::move(scene));
        renderer.setOutputOptions({OutputLinearity::Linear});
        renderer.setDebug(DebugOptions::WindingOrder);
        renderer.render(2);
        renderer.createPixels();
        const uint32_t *p = renderer.getPixels();
        const uint32_t ccw = 255u | (186u << 8) | (120u << 16) | (255u << 24);
        const uint32_t cw = 135u | (209u << 8) | (255u << 16) | (255u << 24);
        bool same = true;
        for (int i = 0; i < res.width * res.height; ++i) same &= p[i] == p[0];
        EXPECT(same && (p[0] == ccw || p[0] == cw));
        if (!out_dir.empty()) {
            EXPECT(renderer.saveImage(out_dir + "/winding.png"));
            FILE *f = std::fopen((out_dir + "/winding.rgba").c_str(), "wb");
            EXPECT(f && std::fwrite(p, 4, (size_t)res.width * res.height, f) == (size_t)res.width * res.height);
            if (f) std::fclose(f);
        }
        renderer.setDebug(DebugOptions::BariCoords);
        renderer.render(3);
        const Vec4 *px = renderer.getRawPixels();
        bool sums = true;
        for (int i = 0; i < res.width * res.height; ++i)
            sums &= px[i].w == 3.0f && std::fabs(px[i].x + px[i].y + px[i].z - 3.0f) <= 1e-5f;
        EXPECT(sums);
        renderer.setDebug(DebugOptions::UVChecker);
        renderer.render(1);
        px = renderer.getRawPixels();
        EXPECT(px[0].w == 1.0f && px[0].x == 0.25f && px[0].y == 0.125f && px[0].z == 0.5f);
    });
    // ---- the reference's mesh tests (test/src/test_mesh.cpp), restated ----
    const std::string tmp = out_dir.empty() ? std::string("/tmp") : out_dir;
    auto write_obj = [&](const std::string &name, const std::string &text) {
        const std::string path = tmp + "/" + name;
        FILE *f = std::fopen(path.c_str(), "wb");
        if (f) { std::fwrite(text.data(), 1, text.size(), f); std::fclose(f); }
        return path;
    };
    auto throws = [](const std::string &path) {
        try { loadMesh(path); } catch (const std::exception &) { return true; }
        return false;
    };
    run("MeshLoaderTest.HandlesFullVertexDefinition", [&] {
        Mesh m = loadMesh(write_obj("full_vertex.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0.5 0.5\nvn 0 0 1\nf 1/1/1 2/1/1 3/1/1\n"));
        EXPECT(m.triangles.size() == 1);
        if (m.triangles.size() != 1) return;
        EXPECT(m.triangles[0][0] == 0 && m.triangles[0][1] == 1 && m.triangles[0][2] == 2);
        EXPECT(m.normal_indices[0][0] == 0 && m.normal_indices[0][1] == 0 && m.normal_indices[0][2] == 0);
    });
    run("MeshLoaderTest.QuadsThrowError", [&] {
        EXPECT(throws(write_obj("quad.obj", "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n")));
    });
    run("MeshLoaderTest.NGonsThrow", [&] {
        EXPECT(throws(write_obj("ngon.obj", "v 0 0 0\nv 1 0 0\nv 2 1 0\nv 1 2 0\nv 0 1 0\nf 1 2 3 4 5\n")));
    });
    run("MeshLoaderTest.DiscardsDegenerateFaces", [&] {
        Mesh m = loadMesh(write_obj("degenerate.obj", "v 0 0 0\nv 1 1 1\nf 1//1 2//1\nf 1//1\n"));
        EXPECT(m.triangles.empty());
    });
    run("MeshLoaderTest.HandlesAllFaceIndexFormats", [&] {
        Mesh m = loadMesh(write_obj("index_formats.obj",
                                    "v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0.5 0.5\nvn 0 0 1\nf 1 2 3\nf 1/1 2/1 3/1\n"
                                    "f 1//1 2//1 3//1\nf 1/1/1 2/1/1 3/1/1\n"));
        EXPECT(m.triangles.size() == 4);
        if (m.triangles.size() != 4) return;
        EXPECT(m.triangles[0][0] == 0);
        EXPECT(m.triangles[1][1] == 1);
        EXPECT(m.triangles[2][2] == 2 && m.normal_indices[2][2] == 0);
        EXPECT(m.triangles[3][0] == 0 && m.normal_indices[3][0] == 0);
    });
    run("MeshLoaderTest.ThrowsOnInvalidContent", [&] {
        const std::vector<std::pair<std::string, std::string>> bad = {
            {"oob_pos.obj", "v 0 0 0\nf 1 2 3\n"},
            {"oob_norm.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//5 2//5 3//5\n"},
            {"bad_coords.obj", "v 0.0 apple 1.0\n"},
            {"bad_face_syntax.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1///1 2///1 3///1\n"},
            {"zero_index.obj", "v 0 0 0\nf 0 0 0\n"},
            {"negative_oob.obj", "v 0 0 0\nf -2 -2 -2\n"},
            {"incomplete_vertex.obj", "v 0.1 0.2\nf 1 1 1\n"},
            {"overflow_index.obj", "v 0 0 0\nf 9999999999999999999999 1 1\n"},
        };
        for (const auto &[name, text] : bad) EXPECT(throws(write_obj(name, text)));
        EXPECT(throws(tmp + "/no_such_file.obj"));
    });
    auto tri_mesh = [](std::vector<Vec3> pts) {
        Mesh m;
        m.points = std::move(pts);
        m.triangles = {{0, 1, 2}};
        return m;
    };
    run("MeshGeometryTest.CalculateBoundsFindsCentroidAndExtent", [&] {
        const MeshBounds b = calculateMeshBounds(tri_mesh({{0, 0, 0}, {2, 0, 0}, {0, 2, 0}}));
        EXPECT_NEAR(b.center.x, 2.0f / 3.0f, 1e-5f);
        EXPECT_NEAR(b.center.y, 2.0f / 3.0f, 1e-5f);
        EXPECT_NEAR(b.center.z, 0.0f, 1e-5f);
        const float dx = 2.0f - b.center.x, dy = -b.center.y;
        EXPECT_NEAR(b.extent, std::sqrt(dx * dx + dy * dy), 1e-5f);
    });
    run("MeshGeometryTest.NormalizeRecentersAndRescales", [&] {
        Mesh m = tri_mesh({{0, 0, 0}, {2, 0, 0}, {0, 2, 0}});
        normalizeMeshSize(m);
        const MeshBounds b = calculateMeshBounds(m);
        EXPECT_NEAR(b.center.x, 0.0f, 1e-5f);
        EXPECT_NEAR(b.center.y, 0.0f, 1e-5f);
        EXPECT_NEAR(b.extent, 1.0f, 1e-5f);
    });
    run("MeshGeometryTest.CoarseNormalsAreFaceNormals", [&] {
        Mesh m = tri_mesh({{0, 0, 0}, {1, 0, 0}, {0, 1, 0}});
        generateCoarseNormals(m);
        EXPECT(m.normals.size() == 1);
        if (m.normals.size() != 1) return;
        EXPECT_NEAR(m.normals[0].x, 0.0f, 1e-6f);
        EXPECT_NEAR(m.normals[0].y, 0.0f, 1e-6f);
        EXPECT_NEAR(m.normals[0].z, 1.0f, 1e-6f);
        EXPECT(m.normal_indices[0][0] == 0 && m.normal_indices[0][1] == 0 && m.normal_indices[0][2] == 0);
    });
    run("MeshGeometryTest.CoarseNormalsAreGeneratedPerTriangle", [&] {
        Mesh m;
        m.points = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        m.triangles = {{0, 1, 2}, {0, 1, 3}};
        generateCoarseNormals(m);
        EXPECT(m.normals.size() == 2);
        EXPECT(m.normal_indices.size() == 2 && m.normal_indices[0][0] == 0 && m.normal_indices[1][0] == 1);
    });
    // ---- the reference's models, and placed objects ----
    const std::string models = argc > 2 ? argv[2] : "";
    run("Models.LoadReferenceModels", [&] {
        if (models.empty()) return;
        struct Expect { const char *file; size_t points, normals, tris; };
        for (const Expect &e : {Expect{"cube.obj", 8, 6, 12}, Expect{"ico.obj", 12, 20, 20},
                                Expect{"suzanne.obj", 507, 942, 967}}) {
            Mesh m = loadMesh(models + "/" + e.file);
            EXPECT(m.points.size() == e.points && m.normals.size() == e.normals && m.triangles.size() == e.tris);
        }
    });
    run("Models.TransformedObjectsAreHitWherePlaced", [&] {
        if (models.empty()) return;
        Mesh cube = loadMesh(models + "/cube.obj");
        normalizeMeshSize(cube);  // centred, radius 1
        Scene s;
        s.materials.push_back(Material::diffuse({0.5f, 0.5f, 0.5f}));
        s.objects.push_back({cube, 0});
        s.objects.push_back({cube, 0});
        s.objects[1].setScale({2.0f, 2.0f, 2.0f});
        s.objects[1].setPosition({10.0f, 0.0f, 0.0f});
        s.build(ctx);
        auto a = cast1(ctx, s, {0, 0, 5}, {0, 0, -1});
        auto b = cast1(ctx, s, {10, 0, 5}, {0, 0, -1});
        const float half = 1.0f / std::sqrt(3.0f);  // the normalised cube's half side
        EXPECT(a.has_value() && a->triangle < 12 && std::fabs(a->t - (5.0f - half)) <= 1e-5f);
        EXPECT(b.has_value() && b->triangle >= 12 && std::fabs(b->t - (5.0f - 2.0f * half)) <= 1e-5f);
        EXPECT(!cast1(ctx, s, {5, 0, 5}, {0, 0, -1}).has_value());
    });
    run("Models.RefitFollowsMovedObjects", [&] {
        // two unit quads; moving the second by refit (no rebuild) is seen by
        // cast, and by the renderer's moveObject
        Scene s;
        s.materials.push_back(Material::diffuse({0.5f, 0.5f, 0.5f}));
        s.objects.push_back({createQuadMesh({0, 0, 0}, {0, 0, 1}, {1, 0, 0}, 1.0f), 0});
        s.objects.push_back({createQuadMesh({0, 0, 0}, {0, 0, 1}, {1, 0, 0}, 1.0f), 0});
        s.objects[1].setPosition({3, 0, 0});
        s.build(ctx);
        const auto before = cast1(ctx, s, {3, 0, 5}, {0, 0, -1});
        EXPECT(before.has_value() && before->triangle >= 2);
        s.objects[1].setPosition({0, 0, -2});  // now 4 wide, 2 behind the first quad
        s.objects[1].setScale({4, 4, 4});
        s.refit(ctx);
        EXPECT(!cast1(ctx, s, {3, 0, 5}, {0, 0, -1}).has_value());
        const auto behind = cast1(ctx, s, {1.5f, 0, 5}, {0, 0, -1});
        EXPECT(behind.has_value() && behind->triangle >= 2 && std::fabs(behind->t - 7.0f) <= 1e-5f);
        const auto front = cast1(ctx, s, {0.2f, 0.2f, 5}, {0, 0, -1});
        EXPECT(front.has_value() && front->triangle < 2 && std::fabs(front->t - 5.0f) <= 1e-5f);
        s.objects.pop_back();
        bool threw = false;
        try { s.refit(ctx); } catch (const Error &) { threw = true; }
        EXPECT(threw);
        // the renderer: move the emitter of the analytic scene out of view,
        // the next frame is black
        const Size2i res{32, 32};
        Renderer renderer(ctx, res);
        renderer.setCamera(pinhole(res));
        renderer.setScene(facing_quads(0.55f));
        renderer.setPathOptions(4, 4);
        renderer.render(2);
        const Vec4 *px = renderer.getRawPixels();
        EXPECT(px[0].x > 0.0f);
        renderer.moveObject(1, {0, 0, 5000}, {1, 1, 1});
        renderer.render(2);
        px = renderer.getRawPixels();
        double sum = 0;
        for (int i = 0; i < res.width * res.height; ++i) sum += px[i].x;
        EXPECT(sum == 0.0 && px[0].w == 2.0f && renderer.getSampleCount() == 2);
    });
    // ---- test/src/test_tracing.cpp:147-219 (MeshIntersectionTest), restated ----
    auto unit_object = [&](Transform tf) {
        Scene s;
        s.materials.push_back(Material::diffuse({0.5f, 0.5f, 0.5f}));
        s.materials.push_back(Material::diffuse({0.2f, 0.2f, 0.2f}));
        Mesh m;
        m.points = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}};
        m.triangles = {{0, 1, 2}};
        Scene::Object o{m, 1};
        o.transform = tf;
        s.objects.push_back(o);
        s.build(ctx);
        return s;
    };
    auto near3 = [](Vec3 a, Vec3 b, float tol) {
        return std::fabs(a.x - b.x) <= tol && std::fabs(a.y - b.y) <= tol && std::fabs(a.z - b.z) <= tol;
    };
    run("MeshIntersectionTest.HitsSingleTriangle", [&] {
        Scene s = unit_object({});
        const auto hit = cast1(ctx, s, {0.2f, 0.2f, 5}, {0, 0, -1});
        EXPECT(hit.has_value() && std::fabs(hit->t - 5.0f) <= 1e-5f && near3(hit->p, {0.2f, 0.2f, 0}, 1e-5f));
        EXPECT(hit && hit->area == 0.5f && hit->mat == 1u);
        EXPECT(!cast1(ctx, s, {5, 5, 5}, {0, 0, -1}).has_value());
    });
    run("MeshIntersectionTest.AppliesModelToWorldTranslation", [&] {
        Scene s = unit_object(Transform{{1, 1, 1}, {10, 0, 0}});
        const auto hit = cast1(ctx, s, {10.2f, 0.2f, 5}, {0, 0, -1});
        EXPECT(hit.has_value() && near3(hit->p, {10.2f, 0.2f, 0}, 1e-5f));
    });
    run("MeshIntersectionTest.WorldSpaceAreaScalesWithTransform", [&] {
        Scene s = unit_object(Transform{{2, 3, 1}, {0, 0, 0}});
        const auto hit = cast1(ctx, s, {0.2f, 0.2f, 5}, {0, 0, -1});
        EXPECT(hit.has_value() && std::fabs(hit->area - 3.0f) <= 1e-5f);
    });
    run("MeshIntersectionTest.NormalFlipsTowardIncomingRayAndWindingFollowsIt", [&] {
        Scene s = unit_object({});
        const auto above = cast1(ctx, s, {0.2f, 0.2f, 5}, {0, 0, -1});
        const auto below = cast1(ctx, s, {0.2f, 0.2f, -5}, {0, 0, 1});
        EXPECT(above.has_value() && near3(above->n, {0, 0, 1}, 1e-6f) && !above->ccw);
        EXPECT(below.has_value() && near3(below->n, {0, 0, -1}, 1e-6f) && below->ccw);
    });
    run("CastTest.ReturnsNearestObjectAndEmptySceneMisses", [&] {  // test_tracing.cpp:276-298
        Scene s;
        for (int k = 0; k < 23; ++k) s.materials.push_back(Material::diffuse({0.5f, 0.5f, 0.5f}));
        Mesh m;
        m.points = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}};
        m.triangles = {{0, 1, 2}};
        s.objects.push_back({m, 11});
        s.objects.push_back({m, 22});
        s.objects[1].setPosition({0, 0, -3});
        s.build(ctx);
        const auto hit = cast1(ctx, s, {0.2f, 0.2f, 5}, {0, 0, -1});
        EXPECT(hit.has_value() && hit->mat == 11u && std::fabs(hit->t - 5.0f) <= 1e-5f);
        Scene empty;
        empty.build(ctx);
        EXPECT(!cast1(ctx, empty, {0, 0, 0}, {0, 0, -1}).has_value());
    });
    run("RendererTest.DebugRenderSquare", [&] {  // test_render.cpp:168-237
        const Size2i res{512, 512};
        Renderer renderer(ctx, res);
        renderer.setCamera(pinhole(res));
        Scene scene;
        Material white = Material::diffuse({1, 1, 1});
        white.debug_color = Vec4{0.9f, 0.7f, 0.5f, 0.0f};
        scene.materials.push_back(white);
        scene.objects.push_back({createQuadMesh({0, 0, 1}, {0, 0, -1}, {1, 0, 0}, 1.0f), 0});
        renderer.setScene(std::move(scene));
        renderer.setOutputOptions({OutputLinearity::GammaCorrected});
        renderer.setDebug(DebugOptions::UVChecker);
        renderer.render();
        if (!out_dir.empty()) EXPECT(renderer.saveImage(out_dir + "/debug_render_1_ray.png"));
        const uint32_t *px = renderer.getPixels();
        bool ok = true;
        for (int i = 0; i < res.height; ++i)
            for (int j = 0; j < res.width; ++j) {
                const uint32_t p = px[i * res.width + j];
                const float r = (p & 0xFF) / 255.0f, g = ((p >> 8) & 0xFF) / 255.0f, b = ((p >> 16) & 0xFF) / 255.0f;
                const bool rim = i < 6 || j < 6 || i >= res.height - 6 || j >= res.width - 6;
                ok &= rim ? (r == 0.0f && g == 0.0f && b == 0.0f) : (r >= 0.65f && g >= 0.6f && b >= 0.5f);
            }
        EXPECT(ok);
    });
    // ---- test/src/test_scene.cpp, restated ----
    run("SceneTest.CreateQuadMeshProducesAxisAlignedSquare", [&] {
        const Mesh quad = createQuadMesh({0, 0, 0}, {0, 0, 1}, {1, 0, 0}, 2.0f);
        EXPECT(quad.points.size() == 4 && quad.triangles.size() == 2);
        double area = 0;
        for (const Vec3 &p : quad.points)
            EXPECT(std::fabs(std::fabs(p.x) - 1.0f) <= 1e-5f && std::fabs(std::fabs(p.y) - 1.0f) <= 1e-5f &&
                   std::fabs(p.z) <= 1e-5f);
        for (const auto &t : quad.triangles) {
            const Vec3 a = quad.points[t[0]], b = quad.points[t[1]], c = quad.points[t[2]];
            const double ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z, vx = c.x - a.x, vy = c.y - a.y, vz = c.z - a.z;
            const double cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
            area += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
        }
        EXPECT_NEAR(area, 4.0, 1e-5);
    });
    run("SceneTest.BuildCopiesGeometryAndBuildsStructures", [&] {
        // a tetrahedron: the device scene holds its triangles, a BVH, and a light
        // table with one entry per triangle when it emits, none when it does not
        Mesh tet;
        tet.points = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        tet.triangles = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};
        for (int emits = 1; emits >= 0; --emits) {
            Scene s;
            s.materials.push_back(emits ? Material::diffuse({0.5f, 0.5f, 0.5f}, {1, 1, 1})
                                        : Material::diffuse({0.5f, 0.5f, 0.5f}));
            s.objects.push_back({tet, 0});
            s.build(ctx);
            dr_scene_info info{};
            check(dr_scene_get_info(s.handle(), &info));
            uint32_t n_emit = 0;
            check(dr_scene_light_count(s.handle(), &n_emit));
            EXPECT(s.triangleCount() == 4 && info.n_triangles == 4 && info.n_nodes > 0);
            EXPECT(n_emit == (emits ? 4u : 0u));
        }
    });
    // ---- test/src/test_running_average.cpp, restated ----
    run("RunningAverageTest.EmptyAverageIsZero", [&] { EXPECT(RunningAverage(5).mean() == 0.0); });
    run("RunningAverageTest.AveragesAddedValues", [&] {
        RunningAverage a(5);
        a.add(2.0);
        a.add(4.0);
        a.add(6.0);
        EXPECT(a.mean() == 4.0);
    });
    run("RunningAverageTest.DropsOldestBeyondWindow", [&] {
        RunningAverage a(3);
        for (double v : {1.0, 2.0, 3.0, 4.0}) a.add(v);
        EXPECT(a.mean() == 3.0);
    });
    run("RunningAverageTest.KeepsOnlyWindowSizeSamples", [&] {
        RunningAverage a(2);
        for (int i = 0; i < 100; ++i) a.add((double)i);
        EXPECT(a.mean() == 98.5);
    });
    run("RunningAverageTest.ResetClearsHistory", [&] {
        RunningAverage a(5);
        a.add(10.0);
        a.add(20.0);
        a.reset();
        EXPECT(a.mean() == 0.0);
    });
    // ---- render statistics and the render loop ----
    run("RendererStats.CastsAndTimesSinceClear", [&] {
        const Size2i res{48, 32};
        const uint64_t npix = (uint64_t)res.width * res.height;
        Renderer renderer(ctx, res);
        renderer.setCamera(pinhole(res));
        renderer.setScene(facing_quads(0.55f));
        EXPECT(renderer.getTotalCasts() == 0 && renderer.getMeanRenderTimes() == 0.0f);
        renderer.setPathOptions(1, 1);  // camera rays only
        renderer.render(3);
        EXPECT(renderer.getTotalCasts() == 3 * npix);
        EXPECT(renderer.getMeanRenderTimes() > 0.0f);
        renderer.setPathOptions(10, 10);  // clears: counts start again
        renderer.render(2);
        const uint64_t c2 = renderer.getTotalCasts();
        std::printf("    casts %llu for 2 spp of %llu pixels\n", (unsigned long long)c2, (unsigned long long)npix);
        EXPECT(c2 > 2 * npix && c2 <= 20 * npix);
        renderer.render(1);
        EXPECT(renderer.getTotalCasts() > c2);
        renderer.setDebug(DebugOptions::BariCoords);  // one cast per debug sample
        renderer.render(4);
        EXPECT(renderer.getTotalCasts() == 4 * npix);
        EXPECT(renderer.getSampleCount() == 4);
    });
    run("RenderThread.ProgressiveUnderConcurrentChanges", [&] {
        // a render loop runs while this thread moves the camera and reads the
        // display pixels; what it accumulates after the move is the image of
        // the same number of samples rendered in one call
        const Size2i res{96, 64};
        const Camera cam_a = pinhole(res), cam_b = pinhole(res, 0.25f);
        Renderer renderer(ctx, res);
        renderer.setCamera(cam_a);
        renderer.setScene(facing_quads(0.55f));
        renderer.setPathOptions(6, 3, DR_FRESNEL_SCHLICK, 5);
        uint64_t reads = 0;
        {
            RenderLoop loop(renderer);
            EXPECT(wait_for([&] { return loop.passes() >= 3 || loop.error(); }, 60.0));
            renderer.setCamera(cam_b);
            const uint64_t p0 = loop.passes();
            EXPECT(wait_for([&] {
                reads += (renderer.getPixels()[0] >> 24) == 255u;
                return loop.passes() >= p0 + 6 || loop.error();
            }, 60.0));
            loop.stop();
            EXPECT(!loop.error());
            EXPECT(loop.averageRenderTime() > 0.0f);
            EXPECT(loop.totalRayCasts() == renderer.getTotalCasts());
        }
        const uint32_t n = renderer.getSampleCount();
        std::printf("    %u samples after the camera moved, %llu pixel reads\n", n, (unsigned long long)reads);
        EXPECT(n >= 4 && reads > 0);
        const size_t count = (size_t)res.width * res.height * 4;
        std::vector<double> progressive(renderer.getRawPixelsFP64(), renderer.getRawPixelsFP64() + count);
        Renderer once(ctx, res);
        once.setCamera(cam_b);
        once.setScene(facing_quads(0.55f));
        once.setPathOptions(6, 3, DR_FRESNEL_SCHLICK, 5);
        once.render(n);
        EXPECT(std::memcmp(progressive.data(), once.getRawPixelsFP64(), count * sizeof(double)) == 0);
        EXPECT(once.getTotalCasts() == renderer.getTotalCasts());
        EXPECT(std::memcmp(renderer.getPixels(), once.getPixels(), (size_t)res.width * res.height * 4) == 0);
    });
    run("RenderThread.StopsAndKeepsTheError", [&] {
        Renderer renderer(ctx, {8, 8});  // no scene: render() throws
        RenderLoop loop(renderer);
        EXPECT(wait_for([&] { return (bool)loop.error(); }, 30.0));
        loop.stop();
        bool invalid = false;
        try {
            std::rethrow_exception(loop.error());
        } catch (const Error &e) {
            invalid = e.status == DR_E_INVALID;
        } catch (...) {
        }
        EXPECT(invalid && loop.passes() == 0);
    });
    std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "OK", g_fail);
    return g_fail ? 1 : 0;
}
