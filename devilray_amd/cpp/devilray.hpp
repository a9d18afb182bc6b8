// C++ host facade over the C ABI (include/dr_api.h). It keeps the reference's
// names, types and contracts for the render path, so that the reference's
// callers (its render thread, its tests, its benchmark) compile against it:
//   math            include/Utils.hpp:56-190 (Vec2/3/4, Size2, Ray)
//   models          include/models/Mesh.hpp (Vertex, Triangle, Mesh, loadMesh),
//                   include/Transform.hpp, include/models/Matrix.hpp
//   materials       include/tracing/Material.hpp, Medium.hpp
//   scene           include/tracing/GpuTris.hpp (GpuTris, convertMeshToTris,
//                   viewGpuTris, createQuadMesh), TriangleMesh.hpp, Scene.hpp
//   tracing         include/tracing/Intersection.hpp (testTriangleIntersection,
//                   getIntersection, getIntersectionBenchmark), Benchmark.hpp,
//                   cast (IntersectionTestsImpl.hpp:234-251)
//   Renderer        include/Renderer.hpp:23-74, Buffers.hpp, Camera.hpp,
//                   OutputOptions.hpp, PixelSampling.hpp, DebugOptions.hpp,
//                   RunningAverage.hpp
// Every query runs on the device through the C ABI. Where the reference's
// structures hold device views (TriangleMesh points into GpuTris storage), the
// views here point into host storage and the scene is copied to the device as
// one flattened scene with one BVH when it is built (Scene::build, or
// Renderer::setScene). The single-ray tracing functions build a device scene
// for their call, so batched cast() over a built Scene is the fast path.
// Errors: a non-OK dr_status becomes devilray::Error (a std::runtime_error);
// no hit is std::nullopt. Threading: as the reference, setters take one mutex
// and the display copy another (include/Renderer.hpp:43-44); a third
// serialises the frame work of render(), createPixels() and the raw-pixel
// reads, so a setter called from another thread waits at most for a snapshot
// of the settings, not for a render (setScene alone waits for the render in
// flight, whose scene it replaces).
#pragma once

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <deque>
#include <exception>
#include <filesystem>
#include <list>
#include <memory>
#include <mutex>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <thread>
#include <variant>
#include <vector>

#include "../../include/dr_api.h"

namespace devilray {

struct Error : std::runtime_error {
    dr_status status;
    Error(dr_status s, const std::string &msg) : std::runtime_error(msg), status(s) {}
};
void check(dr_status s);

// ---- math (include/Utils.hpp:56-190) ----------------------------------------
struct Vec4 {
    float x, y, z, w;
    constexpr Vec4 operator+(const Vec4 &v) const { return Vec4{x + v.x, y + v.y, z + v.z, w + v.w}; }
    constexpr Vec4 operator*(const Vec4 &v) const { return Vec4{x * v.x, y * v.y, z * v.z, w * v.w}; }
    constexpr Vec4 operator*(float f) const { return Vec4{x * f, y * f, z * f, w * f}; }
    constexpr Vec4 operator/(float f) const { return *this * (1.0f / f); }
    constexpr float max() const { return std::max(std::max(x, y), z); }
};

struct Vec3 {
    float x, y, z;
    constexpr Vec3 operator-(const Vec3 &v) const { return Vec3{x - v.x, y - v.y, z - v.z}; }
    constexpr Vec3 operator+(const Vec3 &v) const { return Vec3{x + v.x, y + v.y, z + v.z}; }
    constexpr Vec3 operator*(const Vec3 &v) const { return Vec3{x * v.x, y * v.y, z * v.z}; }
    constexpr Vec3 operator*(float f) const { return Vec3{x * f, y * f, z * f}; }
    constexpr Vec3 operator/(float f) const { return *this * (1.0f / f); }
    constexpr Vec3 &operator+=(const Vec3 &v) { return *this = *this + v; }
    float length() const { return std::sqrt(x * x + y * y + z * z); }
    Vec3 normalized() const
    {
        const float l = length();
        return Vec3{x / l, y / l, z / l};
    }
    constexpr Vec3 cross(const Vec3 &v) const { return Vec3{y * v.z - z * v.y, z * v.x - x * v.z, x * v.y - y * v.x}; }
    constexpr float dot(const Vec3 &v) const { return x * v.x + y * v.y + z * v.z; }
    bool anyNan() const { return std::isnan(x) || std::isnan(y) || std::isnan(z); }
    constexpr Vec3 inv() const { return Vec3{1 / x, 1 / y, 1 / z}; }
};

template <typename T>
struct Vec2 {
    T x, y;
    constexpr Vec2 operator+(const Vec2 &v) const { return Vec2{x + v.x, y + v.y}; }
    constexpr Vec2 operator-(const Vec2 &v) const { return Vec2{x - v.x, y - v.y}; }
    constexpr Vec2 operator*(const Vec2 &v) const { return Vec2{x * v.x, y * v.y}; }
    constexpr Vec2 operator/(const Vec2 &v) const { return Vec2{x / v.x, y / v.y}; }
    constexpr Vec2 operator*(const T &f) const { return Vec2{x * f, y * f}; }
    constexpr Vec2 operator/(const T &f) const { return Vec2{x / f, y / f}; }
};
using Vec2i = Vec2<int>;
using Vec2f = Vec2<float>;

template <typename T>
struct Size2 {
    T width, height;
    constexpr Size2 operator/(const T &v) const { return Size2{width / v, height / v}; }
    constexpr T area() const { return width * height; }
    constexpr Vec2<T> toVec() const { return Vec2<T>{width, height}; }
};
using Size2i = Size2<int>;
using Size2f = Size2<float>;

// The reference's Ray {p, v}; tmin / tmax bound the accepted hit distances
// (the reference accepts any t >= 0).
struct Ray {
    Vec3 p, v;
    float tmin = 0.0f, tmax = 3.4e38f;
};

// ---- models (include/models/Mesh.hpp, include/Transform.hpp, Matrix.hpp) ---
struct Vertex {
    uint32_t pi;  // position index
    uint32_t ni;  // normal index
};
struct Triangle {
    Vertex a, b, c;
};
struct Mesh {
    std::vector<Vec3> points;
    std::vector<Vec3> normals;
    std::vector<Triangle> triangles;
    std::string name;
};
struct MeshBounds {
    Vec3 center;
    float extent;
};
// OBJ reading (loadMesh, src/models/Mesh.cpp:117-186), with the reference's
// rules: v / vn / o / f are read; #, mtllib, s and vt lines are skipped; a
// face is v, v/vt, v//vn or v/vt/vn with 1-based (or negative, relative)
// indices; faces of fewer than three vertices are dropped; polygons of more
// than three, unknown tags, bad numbers or out-of-range indices throw
// std::runtime_error (std::out_of_range for indices beyond int). Without vn
// lines every vertex gets normal index 0 of a (0, 0, 0) normal.
Mesh loadMesh(const std::string &fileName);
// One normal per triangle, normalize((A - B) x (A - C)) (src/models/Mesh.cpp:197-218).
void generateCoarseNormals(Mesh &mesh);
// Mean of the points and the largest distance from it (src/models/Mesh.cpp:220-245).
MeshBounds calculateMeshBounds(const Mesh &mesh);
// Points moved to (p - center) / extent (src/models/Mesh.cpp:188-195).
void normalizeMeshSize(Mesh &mesh);

// Scale-and-translate placement (include/Transform.hpp:5-26): world point =
// p * s + position; a world ray in model space has origin (o - p) / s and
// direction v / s (not renormalised, so t stays the world distance).
struct Transform {
    Vec3 s = Vec3{1, 1, 1};
    Vec3 p = Vec3{0, 0, 0};
    Ray applyInverse(const Ray &ray_in_world) const
    {
        const Vec3 inv_s = s.inv();
        Ray r = ray_in_world;
        r.p = (ray_in_world.p - p) * inv_s;
        r.v = ray_in_world.v * inv_s;
        return r;
    }
    Vec3 applyToPoint(const Vec3 &point) const { return point * s + p; }
};

// Row-major 4x4 affine transform acting on column vectors
// (include/models/Matrix.hpp:8-126); values[r][3] is the translation.
template <typename T>
struct Matrix4x4 {
    T values[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};

    static Matrix4x4 identity() { return {}; }
    static Matrix4x4 translation(const Vec3 &offset)
    {
        Matrix4x4 m;
        m.values[0][3] = offset.x;
        m.values[1][3] = offset.y;
        m.values[2][3] = offset.z;
        return m;
    }
    // Right-handed rotation by `angle` radians about the unit vector `axis`:
    // R = cos(a) I + sin(a) [axis]_x + (1 - cos(a)) axis axis^T.
    static Matrix4x4 rotation(const Vec3 &axis, float angle)
    {
        const T c = std::cos(angle), s = std::sin(angle), t = 1 - c;
        const T x = axis.x, y = axis.y, z = axis.z;
        Matrix4x4 m;
        m.values[0][0] = c + t * x * x;
        m.values[0][1] = t * x * y - s * z;
        m.values[0][2] = t * x * z + s * y;
        m.values[1][0] = t * x * y + s * z;
        m.values[1][1] = c + t * y * y;
        m.values[1][2] = t * y * z - s * x;
        m.values[2][0] = t * x * z - s * y;
        m.values[2][1] = t * y * z + s * x;
        m.values[2][2] = c + t * z * z;
        return m;
    }
    Matrix4x4 operator*(const Matrix4x4 &o) const
    {
        Matrix4x4 r;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                T acc = 0;
                for (int k = 0; k < 4; ++k) acc += values[i][k] * o.values[k][j];
                r.values[i][j] = acc;
            }
        return r;
    }
    Vec3 applyToPosition(const Vec3 &q) const  // w = 1: rotation and translation
    {
        return Vec3{values[0][0] * q.x + values[0][1] * q.y + values[0][2] * q.z + values[0][3],
                    values[1][0] * q.x + values[1][1] * q.y + values[1][2] * q.z + values[1][3],
                    values[2][0] * q.x + values[2][1] * q.y + values[2][2] * q.z + values[2][3]};
    }
    Vec3 applyToDirection(const Vec3 &d) const  // w = 0: the 3x3 part only
    {
        return Vec3{values[0][0] * d.x + values[0][1] * d.y + values[0][2] * d.z,
                    values[1][0] * d.x + values[1][1] * d.y + values[1][2] * d.z,
                    values[2][0] * d.x + values[2][1] * d.y + values[2][2] * d.z};
    }
    Vec3 getOffset() const { return Vec3{values[0][3], values[1][3], values[2][3]}; }
};
using Matrix4x4f = Matrix4x4<float>;

// ---- materials (include/tracing/Material.hpp:9-25, Medium.hpp) --------------
struct Medium {
    float ior;  ///< index of refraction
    static const Medium air;
};
inline const Medium Medium::air{1.0003f};

struct BaseMaterial {
    Vec4 debug_color;  // shown by the UV-checker debug view
};
struct TransparentMaterial : BaseMaterial {
    Medium inside_medium;
};
struct DiffuseMaterial : BaseMaterial {
    Vec4 emission;
    Vec4 diffuse_reflectance;
};
using Material = std::variant<TransparentMaterial, DiffuseMaterial>;
Vec4 getDebugColor(const Material &mat);

// ---- camera and output (include/tracing/Camera.hpp, ...) --------------------
struct Intrinsics {
    float focal_length;
    Vec2f center;
};
// the camera-to-world transform (its rows 0..2 are used), pinhole
// intrinsics, physical pixel size
struct Camera {
    mutable Matrix4x4f transform;
    Intrinsics intrinsics;
    Size2i resolution;
    Size2f physical_pixel_size;
};
enum class PixelSampling : int { Center = 0, UniformRandom = 1 };
enum class OutputLinearity { Linear, GammaCorrected };
struct OutputOptions {
    OutputLinearity linearity;
};
enum class DebugOptions : int { Off = 0, UVChecker = 1, BariCoords = 2, WindingOrder = 3 };

// Mean of the last `sampleCount` values added, 0 when there are none
// (include/RunningAverage.hpp:5-16).
struct RunningAverage {
    explicit RunningAverage(int sampleCount) : m_sampleCount(sampleCount) {}
    void add(double value);
    double mean() const;
    void reset() { m_samples.clear(); }

private:
    std::deque<double> m_samples;
    int m_sampleCount;
};

// ---- tracing types (include/tracing/Intersection.hpp, Benchmark.hpp) --------
struct TriangleIntersection {
    float t;
    Vec3 bari;  // weights of the vertices A, B, C
};
struct TriangleVertices {
    Vec3 a, b, c;
};
struct TriangleHitData {
    Vec3 bari;
    float area;  // world triangle area
    bool ccw;    // ((A - B) x (A - C)) . v > 0
};
struct TriangleMesh;
struct Intersection {
    float t;
    Vec3 p;                      // world hit point
    Vec2f uv;                    // (0, 0), as in the reference
    Vec3 n;                      // unit face normal turned towards the incoming ray
    int mat;
    const TriangleMesh *object;  // the object hit (nullptr from a batched cast)
    TriangleHitData triangle;
    int32_t triangle_id = -1;    // index among the scene's (or object's) triangles, in input order
};
namespace benchmark {
struct HitTests {
    int64_t triangle_tests = 0;
    int64_t triangle_hits = 0;
    int64_t bbox_tests = 0;
    constexpr void registerTriangleTest() { ++triangle_tests; }
    constexpr void registerTriangleHit() { ++triangle_hits; }
    constexpr void registerBBoxTest() { ++bbox_tests; }
};
}  // namespace benchmark

// ---- scene (include/tracing/GpuTris.hpp, TriangleMesh.hpp, Scene.hpp) -------
// The storage of one mesh (host copies of its points, normals and triangles;
// the device copies live in the built scene).
struct GpuTris {
    std::vector<Vec3> points;
    std::vector<Vec3> normals;
    std::vector<Triangle> triangles;
    std::string name;
};
// A non-owning view of a GpuTris placed in the world by model_to_world, with
// one material; the storage must outlive it (src/GpuTris.cpp:61-85).
struct TriangleMesh {
    const Vec3 *points = nullptr;
    const Vec3 *normals = nullptr;
    const Triangle *triangles = nullptr;
    int triangle_count = 0;
    int point_count = 0;
    Transform model_to_world{};
    int material = 0;
    float surface_area = 0.0f;       // world area: base area x (sx sy sz)^(2/3)
    float base_surface_area = 0.0f;  // area of the mesh as stored
    // src/tracing/TriangleMesh.cpp:3-15
    void setPosition(const Vec3 &pos);
    void setScale(const Vec3 &scale);
};
// The reference also builds a per-mesh triangle alias table and BBH here;
// the built scene's flattened light table and device BVH replace them.
GpuTris convertMeshToTris(Mesh &mesh, bool generateTriangleSampler = true);
// Two-triangle square of side `size` (src/GpuTris.cpp:87-107): up = right x
// normal, corners (-,-), (+,-), (+,+), (-,+), triangles (0,2,1), (0,3,2).
GpuTris createQuadMesh(Vec3 center, Vec3 normal, Vec3 right, float size);
TriangleMesh viewGpuTris(GpuTris &tris);

struct ObjectsInfo {
    float total_radiant_power;
};

// A device (dr_context). defaultContext() is the one the reference-style
// constructors use: device 0, created on first use, kept for the process.
class Context {
public:
    explicit Context(int device = 0);
    ~Context();
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    dr_context *handle() const { return ctx_; }

private:
    dr_context *ctx_ = nullptr;
};
Context &defaultContext();

// include/tracing/Scene.hpp:15-26: objects (views), their materials and the
// storage the views point into. build() copies the placed objects into one
// device scene with one BVH (Renderer::setScene builds if needed) and sets
// info.total_radiant_power = sum of object area x luminance(pi E).
struct Scene {
    ObjectsInfo info{0.0f};
    std::vector<TriangleMesh> objects;
    std::vector<Material> materials;
    std::list<GpuTris> mesh_storage;

    Scene() = default;
    Scene(Scene &&) noexcept;
    Scene &operator=(Scene &&) noexcept;
    ~Scene();

    void build(Context &ctx = defaultContext(), uint32_t leaf_max = 4);
    // After objects moved (setPosition / setScale) or their points changed,
    // with the same objects and triangles as at build(): refits the built BVH
    // to the new world positions instead of rebuilding it (dr_scene_refit).
    void refit(Context &ctx = defaultContext());
    dr_scene *handle() const { return scene_; }
    uint32_t triangleCount() const { return n_tris_; }
    bool built() const { return built_; }
    // the placed triangles as built (or last refitted): xyz per point,
    // point indices and material per triangle, in object order
    const std::vector<float> &worldVertices() const { return world_; }
    const std::vector<uint32_t> &worldIndices() const { return idx_; }
    const std::vector<uint32_t> &triangleMaterials() const { return tri_mat_; }

private:
    dr_scene *scene_ = nullptr;
    uint32_t n_tris_ = 0, n_points_ = 0;
    bool built_ = false;  // (a scene built without triangles has no device scene)
    std::vector<float> world_;
    std::vector<uint32_t> idx_, tri_mat_;
};

// ---- tracing (include/tracing/Intersection.hpp:53-56, cast) ------------------
// The reference's semantics: two-sided, any t >= 0, no hit for a NaN ray, a
// degenerate triangle or a ray parallel to the plane. The device evaluates the
// watertight test (Woop et al. 2013), so t and the weights can differ from the
// reference's plane-based formula in the last bits.
std::optional<TriangleIntersection> testTriangleIntersection(const Ray &ray, const TriangleVertices &triangle);
std::optional<Intersection> getIntersection(const Ray &ray, const TriangleMesh &tris);
// With the device traversal's counters: box tests are the child boxes tested
// at the visited nodes of the 4-wide BVH, triangle tests the watertight tests
// (the reference counts them on its binary median-split BBH).
std::optional<Intersection> getIntersectionBenchmark(const Ray &ray, const TriangleMesh &tris,
                                                     benchmark::HitTests &benchmark);
std::optional<Intersection> cast(const Ray &ray, std::span<const TriangleMesh> objects, const ObjectsInfo &info);
// Batched closest hits over a built scene (std::nullopt on a miss); triangle
// ids index the scene's triangles in object order.
std::vector<std::optional<Intersection>> cast(Context &ctx, const Scene &scene, const std::vector<Ray> &rays);

// include/Buffers.hpp:9-22: the renderer's device accumulation image and its
// cast count. color is FP64 {sum r, g, b, sample count} per pixel (the
// reference keeps FP32 Vec4); the host copies are Renderer::getRawPixels
// and getRawPixelsFP64.
struct Buffers {
    static constexpr int maxPathLength = 10;
    Size2i resolution{0, 0};
    const double *color = nullptr;  // device memory
    uint64_t totalCasts() const { return casts ? casts->load() : 0; }
    const std::atomic<uint64_t> *casts = nullptr;
};

class Renderer {
public:
    explicit Renderer(Size2i resolution);  // on defaultContext()
    Renderer(Context &ctx, Size2i resolution);
    ~Renderer();
    Renderer(const Renderer &) = delete;
    Renderer &operator=(const Renderer &) = delete;

    const Buffers &getBuffers() const { return buffers_; }
    // Mean wall time in ms of the last 20 render() calls since the last clear
    // (include/Renderer.hpp:50), and the camera and bounce rays cast since
    // the last clear, shadow rays excluded (include/Renderer.hpp:51).
    float getMeanRenderTimes() const;
    uint64_t getTotalCasts() const { return total_casts_.load(); }

    // Debug views of the first hit instead of the path tracer; clears the
    // image, as the reference's setDebug (src/Renderer.cpp:155-161).
    void setDebug(DebugOptions dbg);
    void setCamera(Camera cam);
    // Takes the scene, builds it on the device if needed and rebuilds the
    // light table (calculateLightWeights).
    void setScene(Scene scene);
    void setPixelSampling(PixelSampling sampling);
    void setOutputOptions(OutputOptions options);

    // Adds exactly one sample per pixel, updates the display pixels and waits
    // for the device (as the reference's render(), src/Renderer.cpp:198-233);
    // render(n) adds n in one launch sequence.
    void render(uint32_t spp = 1);
    void clear();

    void createPixels();
    // The display pixels as an 8-bit RGBA PNG (Renderer::saveImage,
    // src/Renderer.cpp:92-97); returns false, with a warning, if the file
    // cannot be written.
    bool saveImage(const std::filesystem::path &path);
    const uint32_t *getPixels();
    const Vec4 *getRawPixels();  // FP32 view {sum r, g, b, count}
    // Emissive-triangle light table of the scene (src/Renderer.cpp:113-152).
    void calculateLightWeights();

    // Beyond the reference: the FP64 accumulation itself, the sample count,
    // path tracer parameters (max_depth counts path vertices; the reference's
    // fixed behaviour is max_depth 10, no Russian roulette, Schlick Fresnel),
    // and moving object `index` of the scene (TriangleMesh::setPosition /
    // setScale on the renderer's scene; refits the BVH and clears the image).
    const double *getRawPixelsFP64();
    uint32_t getSampleCount() const { return samples_.load(); }
    void setPathOptions(uint32_t max_depth, uint32_t rr_depth, uint32_t fresnel = DR_FRESNEL_EXACT,
                        uint64_t seed = 0);
    void moveObject(size_t index, Vec3 position, Vec3 scale);

private:
    void createPixelsLocked(uint32_t encoding);
    void calculateLightWeightsLocked();
    Context &ctx_;
    Size2i res_;
    Scene scene_;
    Camera cam_{};
    PixelSampling sampling_ = PixelSampling::UniformRandom;
    OutputOptions output_{OutputLinearity::GammaCorrected};
    DebugOptions debug_ = DebugOptions::Off;
    void *rgba_ = nullptr;  // device RGBA8 display pixels
    uint32_t max_depth_ = 10, rr_depth_ = 10, fresnel_ = DR_FRESNEL_SCHLICK;
    uint64_t seed_ = 0;
    void *accum_ = nullptr;  // device FP64 RGBW
    Buffers buffers_;
    std::atomic<uint32_t> samples_{0};
    std::atomic<uint64_t> total_casts_{0};
    RunningAverage render_times_{20};
    bool needs_clear_ = true;
    std::vector<double> host_fp64_;
    std::vector<Vec4> host_vec4_;
    std::vector<uint32_t> pixels_, display_;
    std::mutex frame_mutex_, render_mutex_, display_mutex_;
    mutable std::mutex stats_mutex_;
};

// Application::renderWorker (renderer/src/RenderThread.cpp:3-12): a
// background thread that calls render() until stopped and publishes the mean
// render time and the ray casts after every pass, while other threads change
// the camera or settings (each change clears the image at the start of the
// next pass) and read getPixels(). An exception from render() stops the loop
// and is kept for error(). The loop must be stopped (or destroyed) before the
// renderer it drives.
class RenderLoop {
public:
    explicit RenderLoop(Renderer &renderer);
    ~RenderLoop();  // stop()
    RenderLoop(const RenderLoop &) = delete;
    RenderLoop &operator=(const RenderLoop &) = delete;

    void stop();  // asks the loop to end after the pass in flight and joins it
    float averageRenderTime() const { return average_render_time_.load(std::memory_order_relaxed); }
    uint64_t totalRayCasts() const { return total_ray_casts_.load(std::memory_order_relaxed); }
    uint64_t passes() const { return passes_.load(); }
    std::exception_ptr error() const;

private:
    Renderer &renderer_;
    std::atomic<bool> stop_{false};
    std::atomic<float> average_render_time_{0.0f};
    std::atomic<uint64_t> total_ray_casts_{0};
    std::atomic<uint64_t> passes_{0};
    mutable std::mutex error_mutex_;
    std::exception_ptr error_;
    std::thread thread_;
};

}  // namespace devilray
