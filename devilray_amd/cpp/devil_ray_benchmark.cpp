// Headless traversal benchmark with the reference's command line and report
// (benchmark/src/main.cpp:31-109, BenchmarkGenerator.cpp:3-46,
// benchmark.cu:16-55): load an OBJ mesh, cast random chords through it and
// print the HitTests sums (include/tracing/Benchmark.hpp:5-14): triangle hits,
// triangle tests and box tests, per ray.
//
// Rays: two directions uniform on the sphere, d0 and d1; with the mesh's
// centre c and radius r' = 1.1 x its extent, the ray starts at d0 r' + c and
// points at (d1 - d0) r' + c, as the reference's kernel writes it. The
// reference draws them from cuRAND; here a SplitMix64 stream (seed 0) does.
// They are traced on the device by dr_trace_closest_counted in batches; box
// tests count the child boxes of the 4-wide nodes visited.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <string>
#include <vector>

#include "devilray.hpp"

using namespace devilray;

namespace {

constexpr const char *kDefaultMesh = "models/bunny.obj";
constexpr int kDefaultRays = 10'000'000;
constexpr uint32_t kBatch = 1u << 22;

void print_usage(const std::string &program)
{
    std::printf("Usage: %s [mesh.obj] [ray_count]\n\n"
                "Casts random rays through a mesh's bounding-box hierarchy and reports\n"
                "triangle/bbox test counts.\n\n"
                "Arguments (both optional):\n"
                "  mesh.obj    Path to the mesh to benchmark (default: %s)\n"
                "  ray_count   Number of rays to cast (default: %d)\n\n"
                "Options:\n"
                "  -h, --help  Show this help and exit\n",
                program.c_str(), kDefaultMesh, kDefaultRays);
}

struct SplitMix {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float u01() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
};

Vec3 uniform_sphere(SplitMix &rng)
{
    const float z = 1.0f - 2.0f * rng.u01();
    const float r = std::sqrt(std::max(0.0f, 1.0f - z * z));
    const float phi = 6.2831853f * rng.u01();
    return Vec3{r * std::cos(phi), r * std::sin(phi), z};
}

}  // namespace

int main(int argc, char **argv)
{
    std::vector<std::string> args(argv, argv + argc);
    if (args.empty()) args.push_back("devil_ray_benchmark");
    for (const auto &a : args)
        if (a == "-h" || a == "--help") {
            print_usage(args[0]);
            return 0;
        }
    if (args.size() > 3) {
        std::fprintf(stderr, "Error: too many arguments.\n\n");
        print_usage(args[0]);
        return 1;
    }
    const std::string mesh_file = args.size() > 1 ? args[1] : kDefaultMesh;
    int ray_count = kDefaultRays;
    if (args.size() > 2) {
        try {
            size_t used = 0;
            ray_count = std::stoi(args[2], &used);
            if (used != args[2].size()) throw std::invalid_argument("trailing characters");
        } catch (const std::exception &) {
            std::fprintf(stderr, "Error: ray_count must be an integer, got '%s'.\n\n", args[2].c_str());
            print_usage(args[0]);
            return 1;
        }
        if (ray_count <= 0) {
            std::fprintf(stderr, "Error: ray_count must be positive, got %d.\n\n", ray_count);
            print_usage(args[0]);
            return 1;
        }
    }
    try {
        Mesh mesh = loadMesh(mesh_file);
        generateCoarseNormals(mesh);
        std::printf("Loaded mesh '%s':\n\t%zu points, \n\t%zu normals, \n\t%zu tris\n\n", mesh.name.c_str(),
                    mesh.points.size(), mesh.normals.size(), mesh.triangles.size());
        std::printf("== Benchmark Started ==\n");
        std::fflush(stdout);
        const auto t0 = std::chrono::steady_clock::now();

        Context ctx(0);
        const MeshBounds bounds = calculateMeshBounds(mesh);
        Scene scene;
        scene.materials.push_back(DiffuseMaterial{{}, Vec4{0, 0, 0, 0}, Vec4{0.5f, 0.5f, 0.5f, 0.0f}});
        scene.mesh_storage.push_back(convertMeshToTris(mesh, false));
        scene.objects.push_back(viewGpuTris(scene.mesh_storage.back()));
        scene.build(ctx);

        SplitMix rng{0};
        const float r = bounds.extent * 1.1f;
        const uint32_t batch = std::min<uint32_t>(kBatch, (uint32_t)ray_count);
        std::vector<dr_ray> rays(batch);
        std::vector<dr_hit> hits(batch);
        std::vector<uint32_t> boxes(batch), tris(batch);
        void *d_rays = nullptr, *d_hits = nullptr, *d_box = nullptr, *d_tri = nullptr;
        check(dr_device_alloc(ctx.handle(), sizeof(dr_ray) * batch, &d_rays));
        check(dr_device_alloc(ctx.handle(), sizeof(dr_hit) * batch, &d_hits));
        check(dr_device_alloc(ctx.handle(), 4ull * batch, &d_box));
        check(dr_device_alloc(ctx.handle(), 4ull * batch, &d_tri));
        int64_t triangle_hits = 0, triangle_tests = 0, bbox_tests = 0;
        for (int done = 0; done < ray_count;) {
            const uint32_t n = (uint32_t)std::min<int64_t>(batch, (int64_t)ray_count - done);
            for (uint32_t i = 0; i < n; ++i) {
                const Vec3 d0 = uniform_sphere(rng), d1 = uniform_sphere(rng);
                const Vec3 p0{d0.x * r + bounds.center.x, d0.y * r + bounds.center.y, d0.z * r + bounds.center.z};
                const Vec3 p1{(d1.x - d0.x) * r + bounds.center.x, (d1.y - d0.y) * r + bounds.center.y,
                              (d1.z - d0.z) * r + bounds.center.z};
                Vec3 v{p1.x - p0.x, p1.y - p0.y, p1.z - p0.z};
                const float l = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
                v = Vec3{v.x / l, v.y / l, v.z / l};
                rays[i] = dr_ray{p0.x, p0.y, p0.z, 0.0f, v.x, v.y, v.z, INFINITY};
            }
            check(dr_memcpy_htod(ctx.handle(), d_rays, rays.data(), sizeof(dr_ray) * n));
            check(dr_trace_closest_counted(scene.handle(), (const dr_ray *)d_rays, n, (dr_hit *)d_hits,
                                           (uint32_t *)d_box, (uint32_t *)d_tri, nullptr));
            check(dr_memcpy_dtoh(ctx.handle(), hits.data(), d_hits, sizeof(dr_hit) * n));
            check(dr_memcpy_dtoh(ctx.handle(), boxes.data(), d_box, 4ull * n));
            check(dr_memcpy_dtoh(ctx.handle(), tris.data(), d_tri, 4ull * n));
            for (uint32_t i = 0; i < n; ++i) {
                triangle_hits += hits[i].id >= 0;
                triangle_tests += tris[i];
                bbox_tests += boxes[i];
            }
            done += (int)n;
        }
        for (void *p : {d_rays, d_hits, d_box, d_tri}) dr_device_free(ctx.handle(), p);
        const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const double nr = (double)ray_count;
        std::printf("Number of rays: %d\n", ray_count);
        std::printf("Triangle hits: %lld (%g%%)\n", (long long)triangle_hits, triangle_hits / nr * 100.0);
        std::printf("Triangle tests: %lld (%g/ray)\n", (long long)triangle_tests, triangle_tests / nr);
        std::printf("BBox tests: %lld (%g/ray)\n", (long long)bbox_tests, bbox_tests / nr);
        std::printf("== Benchmark Finished in %gs ==\n", elapsed);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "Benchmark failed: %s\n", e.what());
        return 1;
    }
    return 0;
}
