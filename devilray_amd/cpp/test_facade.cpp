// The reference's own GoogleTest cases for the render path, restated against
// the C++ facade with the reference's own API (needs a GPU; every query runs
// on the device): triangle intersection (test/src/test_intersection.cpp),
// mesh intersection, cast and the benchmark counters
// (test/src/test_tracing.cpp:147-357), the analytic diffuse reflection and
// debug renders (test/src/test_render.cpp:50-237), the mesh loader
// (test/src/test_mesh.cpp), GpuTris / quad / view structures
// (test/src/test_scene.cpp), transforms and matrices
// (test/src/test_transform.cpp, test_matrix.cpp) and the running average
// (test/src/test_running_average.cpp); then the facade's own cases for the
// batched cast, refit, render statistics, the renderer's buffers and the
// background render loop of the interactive application
// (renderer/src/RenderThread.cpp:3-12). Prints one line per case and exits
// non-zero if any failed.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <random>
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

bool finite(const TriangleIntersection &r)
{
    return std::isfinite(r.t) && std::isfinite(r.bari.x) && std::isfinite(r.bari.y) && std::isfinite(r.bari.z);
}

bool near3(Vec3 a, Vec3 b, float tol = 1e-5f)
{
    return std::fabs(a.x - b.x) <= tol && std::fabs(a.y - b.y) <= tol && std::fabs(a.z - b.z) <= tol;
}

// test::DeterministicRng (test/src/TracingTestHelpers.hpp:35-43)
struct DeterministicRng {
    std::mt19937 generator;
    std::uniform_real_distribution<float> distribution{0.0f, 1.0f};
    explicit DeterministicRng(uint32_t seed = 0xC0FFEEu) : generator(seed) {}
    float rnd() { return distribution(generator); }
};

// test::HostObject (TracingTestHelpers.hpp:60-85): a mesh with a placement
// and a material, viewed as one object
struct HostObject {
    GpuTris tris;
    Transform transform{};
    int material = 0;
    TriangleMesh view()
    {
        TriangleMesh object = viewGpuTris(tris);
        object.model_to_world = transform;
        object.material = material;
        return object;
    }
};
HostObject makeHostObject(Mesh mesh) { return HostObject{convertMeshToTris(mesh)}; }

Mesh unitTriangleMesh()
{
    Mesh mesh;
    mesh.name = "unit-triangle";
    mesh.points = {Vec3{0, 0, 0}, Vec3{1, 0, 0}, Vec3{0, 1, 0}};
    mesh.normals = {Vec3{0, 0, 1}};
    mesh.triangles = {Triangle{.a = {0, 0}, .b = {1, 0}, .c = {2, 0}}};
    return mesh;
}

// test_tracing.cpp:53-70
Mesh scatteredTriangleMesh(DeterministicRng &rng, int count)
{
    Mesh mesh;
    mesh.name = "scattered";
    mesh.normals = {Vec3{0, 0, 1}};
    for (int i = 0; i < count; ++i) {
        const float cx = (rng.rnd() - 0.5f) * 4.0f;
        const float cy = (rng.rnd() - 0.5f) * 4.0f;
        const float cz = rng.rnd() * -5.0f;
        const uint32_t base = (uint32_t)mesh.points.size();
        mesh.points.push_back(Vec3{cx, cy, cz});
        mesh.points.push_back(Vec3{cx + 0.6f, cy, cz});
        mesh.points.push_back(Vec3{cx, cy + 0.6f, cz});
        mesh.triangles.push_back(Triangle{.a = {base, 0}, .b = {base + 1, 0}, .c = {base + 2, 0}});
    }
    return mesh;
}

float bruteForceClosestT(const Mesh &mesh, const Ray &ray)
{
    float best = std::numeric_limits<float>::infinity();
    for (const auto &tri : mesh.triangles) {
        const TriangleVertices verts{mesh.points[tri.a.pi], mesh.points[tri.b.pi], mesh.points[tri.c.pi]};
        const auto hit = testTriangleIntersection(ray, verts);
        if (hit.has_value()) best = std::min(best, hit->t);
    }
    return best;
}

// test_render.cpp:20-37
Camera createCamera(Size2i resolution, float focal_length, float physical_pixel_size, float x_offset = 0.0f)
{
    Camera cam{
        .transform = Matrix4x4f::translation(Vec3{x_offset, 0, 0}),
        .intrinsics = Intrinsics{.focal_length = focal_length,
                                 .center = Vec2{resolution.width / 2.f * physical_pixel_size,
                                                resolution.height / 2.f * physical_pixel_size}},
        .resolution = resolution,
        .physical_pixel_size = Size2f{physical_pixel_size, physical_pixel_size},
    };
    return cam;
}

int addObject(Scene &scene, GpuTris storage, int material)
{
    scene.mesh_storage.push_back(std::move(storage));
    auto view = viewGpuTris(scene.mesh_storage.back());
    view.material = material;
    scene.objects.push_back(view);
    return (int)scene.objects.size() - 1;
}

DiffuseMaterial diffuse(Vec4 reflectance, Vec4 emission = Vec4{0, 0, 0, 0})
{
    return DiffuseMaterial{.emission = emission, .diffuse_reflectance = reflectance};
}

// the analytic scene of test_render.cpp:58-106: a white quad in front of the
// camera and an emitting one behind it, both 1000 wide
Scene facing_quads(float emittance)
{
    Scene scene;
    auto white = diffuse(Vec4{1, 1, 1, 0});
    white.debug_color = Vec4{0.9f, 0.7f, 0.5f, 0.0f};
    scene.materials.push_back(white);
    scene.materials.push_back(diffuse(Vec4{0, 0, 0, 0}, Vec4{emittance, emittance, emittance, 0}));
    addObject(scene, createQuadMesh(Vec3{0, 0, 1.0}, Vec3{0, 0, -1}, Vec3{1, 0, 0}, 1000.0), 0);
    addObject(scene, createQuadMesh(Vec3{0, 0, -1.0}, Vec3{0, 0, 1}, Vec3{1, 0, 0}, 1000.0), 1);
    return scene;
}

std::optional<Intersection> cast1(Context &ctx, const Scene &s, Vec3 p, Vec3 v)
{
    return cast(ctx, s, {Ray{p, v}})[0];
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
    // argv[1] (optional): a directory for the saved images and raw pixels;
    // argv[2] (optional): the directory of the reference's model files
    const std::string out_dir = argc > 1 ? argv[1] : "";
    Context &ctx = defaultContext();

    // ---- test/src/test_transform.cpp:20-89 and test/src/test_matrix.cpp:14-90 ----
    auto vnear = [](const Vec3 &v, float x, float y, float z) {
        EXPECT_NEAR(v.x, x, 1e-5f);
        EXPECT_NEAR(v.y, y, 1e-5f);
        EXPECT_NEAR(v.z, z, 1e-5f);
    };
    run("TransformTest.ApplyToPointScalesThenTranslates", [&] {
        const Transform t{.s = {2, 3, 4}, .p = {10, 20, 30}};
        vnear(t.applyToPoint({1, 1, 1}), 12, 23, 34);
        vnear(t.applyToPoint({0, 0, 0}), 10, 20, 30);
    });
    run("TransformTest.ApplyToPointIdentityIsNoOp", [&] { vnear(Transform{}.applyToPoint({5, -2, 7}), 5, -2, 7); });
    run("TransformTest.ApplyInverseUndoesPointTransformForRayOrigin", [&] {
        const Transform t{.s = {2, 4, 5}, .p = {1, 2, 3}};
        const Ray model = t.applyInverse(Ray{.p = {9, 10, 13}, .v = {0, 0, 1}});
        vnear(t.applyToPoint(model.p), 9, 10, 13);
        vnear(model.p, 4, 2, 2);
    });
    run("TransformTest.ApplyInverseScalesDirectionWithoutTranslation", [&] {
        const Transform t{.s = {2, 4, 5}, .p = {100, 200, 300}};
        vnear(t.applyInverse(Ray{.p = {0, 0, 0}, .v = {2, 4, 5}}).v, 1, 1, 1);
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
        const Vec3 ax = Vec3{1, -2, 3}.normalized();
        const Matrix4x4f r = Matrix4x4f::rotation(ax, 0.7f);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                float d = 0.0f;
                for (int k = 0; k < 3; ++k) d += r.values[i][k] * r.values[j][k];
                EXPECT_NEAR(d, i == j ? 1.0f : 0.0f, 1e-6f);
            }
        vnear(r.applyToDirection(ax), ax.x, ax.y, ax.z);
    });

    // ---- test/src/test_intersection.cpp, on the device ----
    const TriangleVertices unitTri = {{0.0f, 0.0f, 0.0f}, {1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}};
    run("IntersectionTest.HitsCenterOfTriangle", [&] {
        auto result = testTriangleIntersection(Ray{{0.2f, 0.2f, 5.0f}, {0.0f, 0.0f, -1.0f}}, unitTri);
        EXPECT(result.has_value());
        if (!result) return;
        EXPECT(finite(*result));
        EXPECT_NEAR(result->t, 5.0f, 1e-5f);
        EXPECT(result->bari.x >= 0 && result->bari.y >= 0 && result->bari.z >= 0);
        EXPECT_NEAR(result->bari.x + result->bari.y + result->bari.z, 1.0f, 1e-5f);
    });
    run("IntersectionTest.HitsExactlyOnVertex", [&] {
        auto result = testTriangleIntersection(Ray{{0.0f, 0.0f, 10.0f}, {0.0f, 0.0f, -1.0f}}, unitTri);
        EXPECT(result.has_value() && finite(*result) && std::fabs(result->t - 10.0f) <= 1e-5f);
    });
    run("IntersectionTest.MissesJustOutsideEdge", [&] {
        EXPECT(!testTriangleIntersection(Ray{{-0.01f, 0.5f, 1.0f}, {0.0f, 0.0f, -1.0f}}, unitTri).has_value());
    });
    run("IntersectionTest.MissesBehindRayOrigin", [&] {
        EXPECT(!testTriangleIntersection(Ray{{0.2f, 0.2f, 1.0f}, {0.0f, 0.0f, 1.0f}}, unitTri).has_value());
    });
    run("IntersectionTest.ParallelRayDoesNotIntersect", [&] {
        EXPECT(!testTriangleIntersection(Ray{{-1.0f, 0.2f, 1.0f}, {1.0f, 0.0f, 0.0f}}, unitTri).has_value());
    });
    run("IntersectionTest.DegenerateTriangleFails", [&] {
        const TriangleVertices flatTri = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        EXPECT(!testTriangleIntersection(Ray{{0, 0, 1}, {0, 0, -1}}, flatTri).has_value());
    });
    run("IntersectionTest.NoBackfaceCulling", [&] {
        auto result = testTriangleIntersection(Ray{{0.2f, 0.2f, -1.0f}, {0.0f, 0.0f, 1.0f}}, unitTri);
        EXPECT(result.has_value() && finite(*result));
    });
    run("IntersectionTest.VeryDistantIntersection", [&] {
        auto result = testTriangleIntersection(Ray{{0.2f, 0.2f, 10000.0f}, {0.0f, 0.0f, -1.0f}}, unitTri);
        EXPECT(result.has_value() && std::fabs(result->t - 10000.0f) <= 1e-2f);
    });
    run("IntersectionTest.DegenerateTriangleCollinear", [&] {
        const TriangleVertices lineTri = {{0, 0, 0}, {1, 0, 0}, {2, 0, 0}};
        EXPECT(!testTriangleIntersection(Ray{{0.5f, 0, 1}, {0, 0, -1}}, lineTri).has_value());
    });
    run("IntersectionTest.DegenerateRayZeroDirection", [&] {
        EXPECT(!testTriangleIntersection(Ray{{0.2f, 0.2f, 1.0f}, {0, 0, 0}}, unitTri).has_value());
    });
    run("IntersectionTest.DegenerateRayNaNInDirection", [&] {
        const float nan = std::numeric_limits<float>::quiet_NaN();
        EXPECT(!testTriangleIntersection(Ray{{0, 0, 0}, {nan, nan, nan}}, unitTri).has_value());
    });
    run("IntersectionTest.ExtremelyThinTriangle", [&] {
        const TriangleVertices sliverTri = {{0, 0, 0}, {100.0f, 0, 0}, {100.0f, 2e-8f, 0}};
        auto result = testTriangleIntersection(Ray{{50.0f, 1e-9f, 1.0f}, {0, 0, -1}}, sliverTri);
        EXPECT(result.has_value() && finite(*result));
    });
    run("IntersectionTest.MultipleAngleSanityCheck", [&] {
        const Vec3 targetInside{0.2f, 0.2f, 0.0f}, targetOutside{5.0f, 5.0f, 0.0f};
        const std::vector<Vec3> origins = {{0.2f, 0.2f, 10.0f}, {10.0f, 0.2f, 10.0f}, {-5.0f, -5.0f, 2.0f},
                                           {0.0f, 0.0f, 1.0f}};
        for (const auto &origin : origins) {
            const auto resultHit = testTriangleIntersection(Ray{origin, targetInside - origin}, unitTri);
            EXPECT(resultHit.has_value());
            if (resultHit.has_value()) {
                EXPECT(finite(*resultHit));
                EXPECT_NEAR(resultHit->t, 1.0f, 1e-4f);
                EXPECT_NEAR(resultHit->bari.x + resultHit->bari.y + resultHit->bari.z, 1.0f, 1e-5f);
                EXPECT_NEAR(resultHit->bari.x, 0.6f, 1e-4f);
                EXPECT_NEAR(resultHit->bari.y, 0.2f, 1e-4f);
                EXPECT_NEAR(resultHit->bari.z, 0.2f, 1e-4f);
            }
            EXPECT(!testTriangleIntersection(Ray{origin, targetOutside - origin}, unitTri).has_value());
        }
    });

    // ---- test/src/test_tracing.cpp:147-298: getIntersection and cast ----
    run("MeshIntersectionTest.HitsSingleTriangle", [&] {
        HostObject object = makeHostObject(unitTriangleMesh());
        const TriangleMesh mesh = object.view();
        const auto hit = getIntersection(Ray{.p = {0.2f, 0.2f, 5}, .v = {0, 0, -1}}, mesh);
        EXPECT(hit.has_value());
        if (!hit) return;
        EXPECT_NEAR(hit->t, 5.0f, 1e-5f);
        EXPECT(near3(hit->p, {0.2f, 0.2f, 0}));
        EXPECT(hit->triangle.area == 0.5f);
        EXPECT(hit->object == &mesh && hit->triangle_id == 0);
    });
    run("MeshIntersectionTest.MissReturnsNullopt", [&] {
        HostObject object = makeHostObject(unitTriangleMesh());
        EXPECT(!getIntersection(Ray{.p = {5, 5, 5}, .v = {0, 0, -1}}, object.view()).has_value());
    });
    run("MeshIntersectionTest.AppliesModelToWorldTranslation", [&] {
        HostObject object = makeHostObject(unitTriangleMesh());
        object.transform = Transform{.s = {1, 1, 1}, .p = {10, 0, 0}};
        const auto hit = getIntersection(Ray{.p = {10.2f, 0.2f, 5}, .v = {0, 0, -1}}, object.view());
        EXPECT(hit.has_value() && near3(hit->p, {10.2f, 0.2f, 0}));
    });
    run("MeshIntersectionTest.WorldSpaceAreaScalesWithTransform", [&] {
        HostObject object = makeHostObject(unitTriangleMesh());
        object.transform = Transform{.s = {2, 3, 1}, .p = {0, 0, 0}};
        const auto hit = getIntersection(Ray{.p = {0.2f, 0.2f, 5}, .v = {0, 0, -1}}, object.view());
        EXPECT(hit.has_value() && std::fabs(hit->triangle.area - 0.5f * 2.0f * 3.0f) <= 1e-5f);
    });
    run("MeshIntersectionTest.NormalFlipsTowardIncomingRay", [&] {
        HostObject object = makeHostObject(unitTriangleMesh());
        const TriangleMesh mesh = object.view();
        const auto fromAbove = getIntersection(Ray{.p = {0.2f, 0.2f, 5}, .v = {0, 0, -1}}, mesh);
        EXPECT(fromAbove.has_value() && near3(fromAbove->n, {0, 0, 1}) && fromAbove->n.dot(Vec3{0, 0, -1}) <= 0.0f);
        const auto fromBelow = getIntersection(Ray{.p = {0.2f, 0.2f, -5}, .v = {0, 0, 1}}, mesh);
        EXPECT(fromBelow.has_value() && near3(fromBelow->n, {0, 0, -1}));
    });
    run("MeshIntersectionTest.WindingFlagFollowsRayDirection", [&] {
        HostObject object = makeHostObject(unitTriangleMesh());
        const TriangleMesh mesh = object.view();
        const auto fromAbove = getIntersection(Ray{.p = {0.2f, 0.2f, 5}, .v = {0, 0, -1}}, mesh);
        EXPECT(fromAbove.has_value() && !fromAbove->triangle.ccw);
        const auto fromBelow = getIntersection(Ray{.p = {0.2f, 0.2f, -5}, .v = {0, 0, 1}}, mesh);
        EXPECT(fromBelow.has_value() && fromBelow->triangle.ccw);
    });
    run("MeshIntersectionTest.ReturnsClosestOfStackedTriangles", [&] {
        Mesh mesh;
        mesh.name = "stack";
        mesh.normals = {Vec3{0, 0, 1}};
        for (float z : {-2.0f, 0.0f, -4.0f}) {
            const uint32_t base = (uint32_t)mesh.points.size();
            mesh.points.push_back(Vec3{0, 0, z});
            mesh.points.push_back(Vec3{1, 0, z});
            mesh.points.push_back(Vec3{0, 1, z});
            mesh.triangles.push_back(Triangle{.a = {base, 0}, .b = {base + 1, 0}, .c = {base + 2, 0}});
        }
        HostObject object = makeHostObject(std::move(mesh));
        const auto hit = getIntersection(Ray{.p = {0.2f, 0.2f, 5}, .v = {0, 0, -1}}, object.view());
        EXPECT(hit.has_value() && std::fabs(hit->t - 5.0f) <= 1e-5f && std::fabs(hit->p.z) <= 1e-5f);
        EXPECT(hit && hit->triangle_id == 1);
    });
    run("MeshIntersectionTest.TraversalMatchesBruteForce", [&] {
        DeterministicRng rng;
        Mesh mesh = scatteredTriangleMesh(rng, 40);
        const Mesh meshCopy = mesh;
        HostObject object = makeHostObject(std::move(mesh));
        const TriangleMesh view = object.view();
        DeterministicRng rayRng{7};
        for (int i = 0; i < 25; ++i) {
            const Ray ray{.p = {(rayRng.rnd() - 0.5f) * 4.0f, (rayRng.rnd() - 0.5f) * 4.0f, 5.0f}, .v = {0, 0, -1}};
            const auto hit = getIntersection(ray, view);
            const float bruteT = bruteForceClosestT(meshCopy, ray);
            if (std::isinf(bruteT)) {
                EXPECT(!hit.has_value());
            } else {
                EXPECT(hit.has_value());
                if (hit) EXPECT_NEAR(hit->t, bruteT, 1e-4f);
            }
        }
    });
    run("CastTest.ReturnsNearestObject", [&] {
        HostObject near = makeHostObject(unitTriangleMesh());
        near.material = 11;
        HostObject far = makeHostObject(unitTriangleMesh());
        far.material = 22;
        far.transform = Transform{.s = {1, 1, 1}, .p = {0, 0, -3}};
        const std::array<TriangleMesh, 2> objects{near.view(), far.view()};
        const ObjectsInfo info{.total_radiant_power = 1.0f};
        const auto hit = cast(Ray{.p = {0.2f, 0.2f, 5}, .v = {0, 0, -1}}, objects, info);
        EXPECT(hit.has_value() && hit->mat == 11 && std::fabs(hit->t - 5.0f) <= 1e-5f);
        EXPECT(hit && hit->object == &objects[0]);
        const auto behind = cast(Ray{.p = {0.2f, 0.2f, -5}, .v = {0, 0, 1}}, objects, info);
        EXPECT(behind.has_value() && behind->mat == 22 && behind->object == &objects[1]);
    });
    run("CastTest.EmptySceneMisses", [&] {
        const ObjectsInfo info{.total_radiant_power = 1.0f};
        EXPECT(!cast(Ray{.p = {0, 0, 0}, .v = {0, 0, -1}}, std::span<const TriangleMesh>{}, info).has_value());
    });

    // ---- test/src/test_tracing.cpp:302-357: the benchmark counters, on the
    // device's 4-wide BVH ----
    run("BenchmarkCountsTest.SingleTriangleHitCountsOneOfEach", [&] {
        HostObject object = makeHostObject(unitTriangleMesh());
        benchmark::HitTests counts;
        const auto hit = getIntersectionBenchmark(Ray{.p = {0.2f, 0.2f, 5}, .v = {0, 0, -1}}, object.view(), counts);
        EXPECT(hit.has_value());
        EXPECT(counts.bbox_tests == 1 && counts.triangle_tests == 1);
    });
    run("BenchmarkCountsTest.MissingBoxSkipsTriangleTests", [&] {
        HostObject object = makeHostObject(unitTriangleMesh());
        benchmark::HitTests counts;
        const auto hit = getIntersectionBenchmark(Ray{.p = {5, 5, 5}, .v = {0, 0, -1}}, object.view(), counts);
        EXPECT(!hit.has_value());
        EXPECT(counts.bbox_tests == 1 && counts.triangle_tests == 0);
    });
    run("BenchmarkCountsTest.HierarchyPrunesTriangleTests", [&] {
        DeterministicRng rng;
        HostObject object = makeHostObject(scatteredTriangleMesh(rng, 64));
        const TriangleMesh mesh = object.view();
        benchmark::HitTests counts;
        getIntersectionBenchmark(Ray{.p = {0, 0, 5}, .v = {0, 0, -1}}, mesh, counts);
        EXPECT(counts.bbox_tests > 0);
        EXPECT(counts.triangle_tests < mesh.triangle_count);
    });
    run("BenchmarkCountsTest.SkipAndCountingTraversalAgree", [&] {
        DeterministicRng rng;
        HostObject object = makeHostObject(scatteredTriangleMesh(rng, 32));
        const TriangleMesh mesh = object.view();
        const Ray ray{.p = {0.3f, -0.2f, 5}, .v = {0, 0, -1}};
        const auto plain = getIntersection(ray, mesh);
        benchmark::HitTests counts;
        const auto counted = getIntersectionBenchmark(ray, mesh, counts);
        EXPECT(plain.has_value() == counted.has_value());
        if (plain && counted) EXPECT(plain->t == counted->t && plain->triangle_id == counted->triangle_id);
    });

    // ---- test/src/test_scene.cpp ----
    auto tetrahedronMesh = [] {
        Mesh mesh;
        mesh.name = "tetra";
        mesh.points = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        mesh.triangles = {Triangle{.a = {0, 0}, .b = {1, 0}, .c = {2, 0}}, Triangle{.a = {0, 0}, .b = {1, 0}, .c = {3, 0}},
                          Triangle{.a = {0, 0}, .b = {2, 0}, .c = {3, 0}}, Triangle{.a = {1, 0}, .b = {2, 0}, .c = {3, 0}}};
        return mesh;
    };
    auto meshSurfaceArea = [](const GpuTris &tris) {
        float total = 0;
        for (const Triangle &t : tris.triangles)
            total += (tris.points[t.a.pi] - tris.points[t.b.pi]).cross(tris.points[t.a.pi] - tris.points[t.c.pi]).length() / 2;
        return total;
    };
    run("ConvertMeshToTrisTest.CopiesGeometry", [&] {
        Mesh mesh = tetrahedronMesh();
        GpuTris tris = convertMeshToTris(mesh);
        EXPECT(tris.points.size() == mesh.points.size() && tris.triangles.size() == mesh.triangles.size());
        EXPECT(tris.name == "tetra");
    });
    run("CreateQuadMeshTest.ProducesAxisAlignedSquare", [&] {
        GpuTris quad = createQuadMesh(Vec3{0, 0, 0}, Vec3{0, 0, 1}, Vec3{1, 0, 0}, 2.0f);
        EXPECT(quad.points.size() == 4u && quad.triangles.size() == 2u);
        for (const Vec3 &p : quad.points)
            EXPECT(std::fabs(std::fabs(p.x) - 1.0f) <= 1e-5f && std::fabs(std::fabs(p.y) - 1.0f) <= 1e-5f &&
                   std::fabs(p.z) <= 1e-5f);
        EXPECT_NEAR(meshSurfaceArea(quad), 4.0f, 1e-5f);
    });
    run("ViewGpuTrisTest.WiresUpPointersAndSurfaceArea", [&] {
        GpuTris quad = createQuadMesh(Vec3{0, 0, 0}, Vec3{0, 0, 1}, Vec3{1, 0, 0}, 2.0f);
        TriangleMesh view = viewGpuTris(quad);
        EXPECT(view.triangle_count == 2 && view.points != nullptr && view.normals != nullptr &&
               view.triangles != nullptr);
        EXPECT_NEAR(view.base_surface_area, 4.0f, 1e-5f);
        EXPECT(view.surface_area == view.base_surface_area);
        view.setScale(Vec3{2, 2, 2});  // src/tracing/TriangleMesh.cpp:8-15: area x (s^3)^(2/3)
        EXPECT_NEAR(view.surface_area, 16.0f, 1e-4f);
    });
    run("SceneTest.BuildCopiesGeometryAndBuildsStructures", [&] {
        // the device scene holds the tetrahedron's triangles, a BVH, and a light
        // table with one entry per triangle when it emits, none when it does not
        for (int emits = 1; emits >= 0; --emits) {
            Scene s;
            s.materials.push_back(diffuse(Vec4{0.5f, 0.5f, 0.5f, 0}, emits ? Vec4{1, 1, 1, 0} : Vec4{0, 0, 0, 0}));
            Mesh tet = tetrahedronMesh();
            addObject(s, convertMeshToTris(tet), 0);
            s.build(ctx);
            dr_scene_info info{};
            check(dr_scene_get_info(s.handle(), &info));
            uint32_t n_emit = 0;
            check(dr_scene_light_count(s.handle(), &n_emit));
            EXPECT(s.triangleCount() == 4 && info.n_triangles == 4 && info.n_nodes > 0);
            EXPECT(n_emit == (emits ? 4u : 0u));
            // calculateLightWeights: area x luminance(pi E)
            const float area = s.objects[0].surface_area;
            EXPECT_NEAR(s.info.total_radiant_power, emits ? area * 3.14159265f : 0.0f, 1e-5f);
        }
    });
    run("SceneTest.ViewsSurviveMovingTheScene", [&] {
        // objects point into mesh_storage (a std::list): moving the scene keeps them valid
        Scene a = facing_quads(0.55f);
        Scene b = std::move(a);
        EXPECT(b.objects.size() == 2 && b.objects[0].points == b.mesh_storage.front().points.data());
        b.build(ctx);
        const auto hit = cast1(ctx, b, {0, 0, 0}, {0, 0, 1});
        EXPECT(hit.has_value() && std::fabs(hit->t - 1.0f) <= 1e-5f && hit->mat == 0);
    });

    // ---- batched cast over a built scene (the fast path) ----
    run("BatchedCast.NearestOfStackedObjects", [&] {
        Scene s;
        s.materials.push_back(diffuse(Vec4{1, 1, 1, 0}));
        for (float z : {-2.0f, 0.0f, 3.0f}) {
            Mesh m;
            m.points = {{-1, -1, z}, {1, -1, z}, {0, 1, z}};
            m.triangles = {Triangle{{0, 0}, {1, 0}, {2, 0}}};
            addObject(s, convertMeshToTris(m), 0);
        }
        s.build(ctx);
        auto r = cast1(ctx, s, {0, 0, 10}, {0, 0, -1});
        EXPECT(r.has_value() && r->triangle_id == 2 && std::fabs(r->t - 7.0f) <= 1e-5f);
        auto r2 = cast1(ctx, s, {0, 0, -10}, {0, 0, 1});
        EXPECT(r2.has_value() && r2->triangle_id == 0 && std::fabs(r2->t - 8.0f) <= 1e-5f);
        EXPECT(!cast1(ctx, s, {5, 5, 10}, {0, 0, -1}).has_value());
        Scene empty;
        empty.build(ctx);
        EXPECT(!cast1(ctx, empty, {0, 0, 0}, {0, 0, -1}).has_value());
    });
    run("BatchedCast.AgreesWithGetIntersection", [&] {
        DeterministicRng rng(3);
        Scene s;
        s.materials.push_back(diffuse(Vec4{1, 1, 1, 0}));
        Mesh scattered = scatteredTriangleMesh(rng, 48);
        addObject(s, convertMeshToTris(scattered), 0);
        s.objects[0].setPosition(Vec3{0.5f, -0.25f, 1.0f});
        s.objects[0].setScale(Vec3{1.5f, 0.75f, 1.25f});
        s.build(ctx);
        std::vector<Ray> rays;
        for (int i = 0; i < 64; ++i)
            rays.push_back(Ray{.p = {(rng.rnd() - 0.5f) * 6.0f, (rng.rnd() - 0.5f) * 4.0f, 5.0f}, .v = {0, 0, -1}});
        const auto batched = cast(ctx, s, rays);
        int hits = 0;
        for (size_t i = 0; i < rays.size(); ++i) {
            const auto one = getIntersection(rays[i], s.objects[0]);
            EXPECT(one.has_value() == batched[i].has_value());
            if (one && batched[i]) {
                EXPECT(one->t == batched[i]->t && one->triangle_id == batched[i]->triangle_id);
                EXPECT(near3(one->p, batched[i]->p, 0.0f) && one->triangle.area == batched[i]->triangle.area);
                ++hits;
            }
        }
        EXPECT(hits > 8);
    });

    // ---- test/src/test_render.cpp:50-237, with the reference's own setup ----
    run("RendererTest.AnalyticalDiffuseReflection", [&] {
        const Size2i resolution{256, 256};
        Renderer renderer(resolution);
        renderer.setCamera(createCamera(resolution, 50.0f, 0.1f));
        const double emittance = 0.55;
        Scene scene;
        const int white_mat_idx = (int)scene.materials.size();
        {
            auto material = DiffuseMaterial{
                .emission = Vec4{0, 0, 0, 0},
                .diffuse_reflectance = Vec4{1.0, 1.0, 1.0, 0.0},
            };
            material.debug_color = Vec4{0.9f, 0.7f, 0.5f, 0.0f};
            scene.materials.push_back(material);
        }
        const int light_mat_idx = (int)scene.materials.size();
        {
            auto material = DiffuseMaterial{
                .emission = Vec4{(float)emittance, (float)emittance, (float)emittance, 0.0},
                .diffuse_reflectance = Vec4{0.0, 0.0, 0.0, 0.0},
            };
            scene.materials.push_back(material);
        }
        {
            scene.mesh_storage.push_back(createQuadMesh(Vec3{0, 0, 1.0}, Vec3{0, 0, -1}, Vec3{1, 0, 0}, 1000.0));
            auto receiver = viewGpuTris(scene.mesh_storage.back());
            receiver.material = white_mat_idx;
            scene.objects.push_back(std::move(receiver));
        }
        {
            scene.mesh_storage.push_back(createQuadMesh(Vec3{0, 0, -1.0}, Vec3{0, 0, 1}, Vec3{1, 0, 0}, 1000.0));
            auto source = viewGpuTris(scene.mesh_storage.back());
            source.material = light_mat_idx;
            scene.objects.push_back(std::move(source));
        }
        renderer.setScene(std::move(scene));
        renderer.setOutputOptions(OutputOptions{.linearity = OutputLinearity::Linear});
        renderer.setDebug(DebugOptions::Off);
        for (int i = 0; i < 100; ++i) renderer.render();
        renderer.createPixels();
        if (!out_dir.empty()) EXPECT(renderer.saveImage(std::filesystem::path(out_dir) / "analytic_diffuse_planes.png"));
        const Vec4 *pixels = renderer.getRawPixels();
        const float count0 = pixels[0].w;
        double s1 = 0, s2 = 0, s3 = 0;
        bool counts = true, grey = true;
        for (int i = 0; i < resolution.area(); ++i) {
            counts &= pixels[i].w == count0;
            const float r = pixels[i].x / pixels[i].w, g = pixels[i].y / pixels[i].w, b = pixels[i].z / pixels[i].w;
            grey &= (r == g && g == b);
            s1 += r;
            s2 += (double)r * r;
            s3 += (double)r * r * r;
        }
        const double n = resolution.area(), mean = s1 / n, sd = std::sqrt(s2 / n - mean * mean);
        const double skew = (s3 / n - 3 * mean * sd * sd - mean * mean * mean) / (sd * sd * sd);
        std::printf("    mean %.5f stddev %.5f skew %.3g samples %.0f\n", mean, sd, skew, count0);
        EXPECT(counts && grey && count0 == 100.0f);
        EXPECT_NEAR(mean, emittance, 0.01);
        EXPECT(sd <= 0.04);
        EXPECT(skew <= 4e-5);
        EXPECT(((renderer.getPixels()[0] >> 24) & 0xFF) == 255u);
    });
    run("RendererTest.DebugRenderSquare", [&] {
        const Size2i resolution{512, 512};
        Renderer renderer(resolution);
        renderer.setCamera(createCamera(resolution, 50.0f, 0.1f));
        Scene scene;
        const int white_mat_idx = (int)scene.materials.size();
        {
            auto material = DiffuseMaterial{
                .emission = Vec4{0, 0, 0, 0},
                .diffuse_reflectance = Vec4{1.0, 1.0, 1.0, 0.0},
            };
            material.debug_color = Vec4{0.9f, 0.7f, 0.5f, 0.0f};
            scene.materials.push_back(material);
        }
        {
            scene.mesh_storage.push_back(createQuadMesh(Vec3{0, 0, 1.0}, Vec3{0, 0, -1}, Vec3{1, 0, 0}, 1.0));
            auto obj = viewGpuTris(scene.mesh_storage.back());
            obj.material = white_mat_idx;
            scene.objects.push_back(std::move(obj));
        }
        renderer.setScene(std::move(scene));
        renderer.setOutputOptions(OutputOptions{.linearity = OutputLinearity::GammaCorrected});
        renderer.setDebug(DebugOptions::UVChecker);
        renderer.render();
        renderer.createPixels();
        if (!out_dir.empty()) EXPECT(renderer.saveImage(std::filesystem::path(out_dir) / "debug_render_1_ray.png"));
        const uint32_t *pixels = renderer.getPixels();
        bool ok = true;
        for (int i = 0; i < resolution.height; ++i)
            for (int j = 0; j < resolution.width; ++j) {
                const auto px = pixels[i * resolution.width + j];
                const float r = (float)((px >> 0) & 0xff) / 255.0f;
                const float g = (float)((px >> 8) & 0xff) / 255.0f;
                const float b = (float)((px >> 16) & 0xff) / 255.0f;
                const bool is_rim = i < 6 || j < 6 || i >= resolution.height - 6 || j >= resolution.width - 6;
                ok &= is_rim ? (r == 0.0f && g == 0.0f && b == 0.0f) : (r >= 0.65f && g >= 0.6f && b >= 0.5f);
            }
        EXPECT(ok);
    });
    run("RendererTest.DebugViews", [&] {
        // a quad filling the view: winding colour everywhere, barycentric
        // weights summing to one, rows flipped and packed by createPixels
        const Size2i res{64, 64};
        Renderer renderer(res);
        renderer.setCamera(createCamera(res, 50.0f, 0.1f));
        Scene scene;
        auto m = diffuse(Vec4{0.5f, 0.25f, 1.0f, 0});
        m.debug_color = m.diffuse_reflectance;
        scene.materials.push_back(m);
        addObject(scene, createQuadMesh({0, 0, 1}, {0, 0, -1}, {1, 0, 0}, 1000.0f), 0);
        renderer.setScene(std::move(scene));
        renderer.setOutputOptions({OutputLinearity::Linear});
        renderer.setDebug(DebugOptions::WindingOrder);
        renderer.render(2);
        renderer.createPixels();
        const uint32_t *p = renderer.getPixels();
        const uint32_t ccw = 255u | (186u << 8) | (120u << 16) | (255u << 24);
        const uint32_t cw = 135u | (209u << 8) | (255u << 16) | (255u << 24);
        bool same = true;
        for (int i = 0; i < res.area(); ++i) same &= p[i] == p[0];
        EXPECT(same && (p[0] == ccw || p[0] == cw));
        if (!out_dir.empty()) {
            EXPECT(renderer.saveImage(std::filesystem::path(out_dir) / "winding.png"));
            FILE *f = std::fopen((out_dir + "/winding.rgba").c_str(), "wb");
            EXPECT(f && std::fwrite(p, 4, (size_t)res.area(), f) == (size_t)res.area());
            if (f) std::fclose(f);
        }
        renderer.setDebug(DebugOptions::BariCoords);
        renderer.render(3);
        const Vec4 *px = renderer.getRawPixels();
        bool sums = true;
        for (int i = 0; i < res.area(); ++i)
            sums &= px[i].w == 3.0f && std::fabs(px[i].x + px[i].y + px[i].z - 3.0f) <= 1e-5f;
        EXPECT(sums);
        renderer.setDebug(DebugOptions::UVChecker);
        renderer.render(1);
        px = renderer.getRawPixels();
        EXPECT(px[0].w == 1.0f && px[0].x == 0.25f && px[0].y == 0.125f && px[0].z == 0.5f);
    });
    run("RendererTest.BuffersHoldTheAccumulationAndCasts", [&] {
        // getBuffers (include/Renderer.hpp:48): the device image and the casts
        const Size2i res{16, 8};
        Renderer renderer(res);
        renderer.setCamera(createCamera(res, 50.0f, 0.1f));
        renderer.setScene(facing_quads(0.55f));
        renderer.setPathOptions(1, 1);  // camera rays only
        renderer.render(2);
        const Buffers &buffers = renderer.getBuffers();
        EXPECT(buffers.resolution.width == 16 && buffers.resolution.height == 8 && buffers.color != nullptr);
        EXPECT(buffers.totalCasts() == renderer.getTotalCasts() && buffers.totalCasts() == 2u * 16u * 8u);
        std::vector<double> host(4 * 16 * 8);
        check(dr_memcpy_dtoh(defaultContext().handle(), host.data(), buffers.color, host.size() * sizeof(double)));
        EXPECT(std::memcmp(host.data(), renderer.getRawPixelsFP64(), host.size() * sizeof(double)) == 0);
        EXPECT(host[3] == 2.0 && Buffers::maxPathLength == 10);
    });

    // ---- test/src/test_mesh.cpp ----
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
        EXPECT(m.triangles[0].a.pi == 0 && m.triangles[0].b.pi == 1 && m.triangles[0].c.pi == 2);
        EXPECT(m.triangles[0].a.ni == 0 && m.triangles[0].b.ni == 0 && m.triangles[0].c.ni == 0);
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
        EXPECT(m.triangles[0].a.pi == 0);
        EXPECT(m.triangles[1].b.pi == 1);
        EXPECT(m.triangles[2].c.pi == 2 && m.triangles[2].c.ni == 0);
        EXPECT(m.triangles[3].a.pi == 0 && m.triangles[3].a.ni == 0);
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
        m.triangles = {Triangle{{0, 0}, {1, 0}, {2, 0}}};
        return m;
    };
    run("MeshGeometryTest.CalculateBoundsFindsCentroidAndExtent", [&] {
        const MeshBounds b = calculateMeshBounds(tri_mesh({{0, 0, 0}, {2, 0, 0}, {0, 2, 0}}));
        EXPECT(near3(b.center, {2.0f / 3.0f, 2.0f / 3.0f, 0.0f}));
        EXPECT_NEAR(b.extent, (Vec3{2, 0, 0} - b.center).length(), 1e-5f);
    });
    run("MeshGeometryTest.NormalizeRecentersAndRescales", [&] {
        Mesh m = tri_mesh({{0, 0, 0}, {2, 0, 0}, {0, 2, 0}});
        normalizeMeshSize(m);
        const MeshBounds b = calculateMeshBounds(m);
        EXPECT(near3(b.center, {0, 0, 0}) && std::fabs(b.extent - 1.0f) <= 1e-5f);
    });
    run("MeshGeometryTest.CoarseNormalsAreFaceNormals", [&] {
        Mesh m = tri_mesh({{0, 0, 0}, {1, 0, 0}, {0, 1, 0}});
        generateCoarseNormals(m);
        EXPECT(m.normals.size() == 1);
        if (m.normals.size() != 1) return;
        EXPECT(near3(m.normals[0], {0, 0, 1}, 1e-6f));
        EXPECT(m.triangles[0].a.ni == 0 && m.triangles[0].b.ni == 0 && m.triangles[0].c.ni == 0);
    });
    run("MeshGeometryTest.CoarseNormalsAreGeneratedPerTriangle", [&] {
        Mesh m;
        m.points = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        m.triangles = {Triangle{{0, 0}, {1, 0}, {2, 0}}, Triangle{{0, 0}, {1, 0}, {3, 0}}};
        generateCoarseNormals(m);
        EXPECT(m.normals.size() == 2 && m.triangles[0].a.ni == 0 && m.triangles[1].a.ni == 1);
    });

    // ---- the reference's models, placed objects and refit ----
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
        s.materials.push_back(diffuse(Vec4{0.5f, 0.5f, 0.5f, 0}));
        s.mesh_storage.push_back(convertMeshToTris(cube));
        // two views of one storage, as the reference's instancing
        for (int k = 0; k < 2; ++k) s.objects.push_back(viewGpuTris(s.mesh_storage.back()));
        s.objects[1].setScale({2.0f, 2.0f, 2.0f});
        s.objects[1].setPosition({10.0f, 0.0f, 0.0f});
        s.build(ctx);
        auto a = cast1(ctx, s, {0, 0, 5}, {0, 0, -1});
        auto b = cast1(ctx, s, {10, 0, 5}, {0, 0, -1});
        const float half = 1.0f / std::sqrt(3.0f);  // the normalised cube's half side
        EXPECT(a.has_value() && a->triangle_id < 12 && std::fabs(a->t - (5.0f - half)) <= 1e-5f);
        EXPECT(b.has_value() && b->triangle_id >= 12 && std::fabs(b->t - (5.0f - 2.0f * half)) <= 1e-5f);
        EXPECT(!cast1(ctx, s, {5, 0, 5}, {0, 0, -1}).has_value());
    });
    run("Models.RefitFollowsMovedObjects", [&] {
        // two unit quads; moving the second by refit (no rebuild) is seen by
        // cast, and by the renderer's moveObject
        Scene s;
        s.materials.push_back(diffuse(Vec4{0.5f, 0.5f, 0.5f, 0}));
        addObject(s, createQuadMesh({0, 0, 0}, {0, 0, 1}, {1, 0, 0}, 1.0f), 0);
        addObject(s, createQuadMesh({0, 0, 0}, {0, 0, 1}, {1, 0, 0}, 1.0f), 0);
        s.objects[1].setPosition({3, 0, 0});
        s.build(ctx);
        const auto before = cast1(ctx, s, {3, 0, 5}, {0, 0, -1});
        EXPECT(before.has_value() && before->triangle_id >= 2);
        s.objects[1].setPosition({0, 0, -2});  // now 4 wide, 2 behind the first quad
        s.objects[1].setScale({4, 4, 4});
        s.refit(ctx);
        EXPECT(!cast1(ctx, s, {3, 0, 5}, {0, 0, -1}).has_value());
        const auto behind = cast1(ctx, s, {1.5f, 0, 5}, {0, 0, -1});
        EXPECT(behind.has_value() && behind->triangle_id >= 2 && std::fabs(behind->t - 7.0f) <= 1e-5f);
        const auto front = cast1(ctx, s, {0.2f, 0.2f, 5}, {0, 0, -1});
        EXPECT(front.has_value() && front->triangle_id < 2 && std::fabs(front->t - 5.0f) <= 1e-5f);
        s.objects.pop_back();
        bool threw = false;
        try { s.refit(ctx); } catch (const Error &) { threw = true; }
        EXPECT(threw);
        // the renderer: move the emitter of the analytic scene out of view,
        // the next frame is black
        const Size2i res{32, 32};
        Renderer renderer(res);
        renderer.setCamera(createCamera(res, 50.0f, 0.1f));
        renderer.setScene(facing_quads(0.55f));
        renderer.setPathOptions(4, 4);
        renderer.render(2);
        const Vec4 *px = renderer.getRawPixels();
        EXPECT(px[0].x > 0.0f);
        renderer.moveObject(1, {0, 0, 5000}, {1, 1, 1});
        renderer.render(2);
        px = renderer.getRawPixels();
        double sum = 0;
        for (int i = 0; i < res.area(); ++i) sum += px[i].x;
        EXPECT(sum == 0.0 && px[0].w == 2.0f && renderer.getSampleCount() == 2);
    });

    // ---- test/src/test_running_average.cpp ----
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
        const uint64_t npix = (uint64_t)res.area();
        Renderer renderer(res);
        renderer.setCamera(createCamera(res, 50.0f, 0.1f));
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
        const Camera cam_a = createCamera(res, 50.0f, 0.1f), cam_b = createCamera(res, 50.0f, 0.1f, 0.25f);
        Renderer renderer(res);
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
        const size_t count = (size_t)res.area() * 4;
        std::vector<double> progressive(renderer.getRawPixelsFP64(), renderer.getRawPixelsFP64() + count);
        Renderer once(res);
        once.setCamera(cam_b);
        once.setScene(facing_quads(0.55f));
        once.setPathOptions(6, 3, DR_FRESNEL_SCHLICK, 5);
        once.render(n);
        EXPECT(std::memcmp(progressive.data(), once.getRawPixelsFP64(), count * sizeof(double)) == 0);
        EXPECT(once.getTotalCasts() == renderer.getTotalCasts());
        EXPECT(std::memcmp(renderer.getPixels(), once.getPixels(), (size_t)res.area() * 4) == 0);
    });
    run("RenderThread.StopsAndKeepsTheError", [&] {
        Renderer renderer(Size2i{8, 8});  // no scene: render() throws
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
