// C++ host facade over the C ABI (include/dr_api.h), keeping the reference's
// names and contracts for the render path so that its callers switch over
// unchanged:
//   Renderer      include/Renderer.hpp:23-74 (setScene / setCamera /
//                 setPixelSampling / setOutputOptions / render / clear /
//                 createPixels / getPixels / getRawPixels)
//   Scene         include/tracing/Scene.hpp:15-26 (materials + objects),
//                 built with createQuadMesh (include/tracing/GpuTris.hpp:22-24)
//   cast          include/tracing/IntersectionTestsImpl.hpp:234-251, here
//                 over a batch of rays (the reference casts one ray per call)
// Errors: a non-OK dr_status becomes devilray::Error (the reference throws
// std::runtime_error on misuse, DeviceMemory.cu:51); no-hit is std::nullopt.
// Threading: as the reference, setters take one mutex and the display copy
// another (include/Renderer.hpp:43-44); a third serialises the frame work of
// render(), createPixels() and the raw-pixel reads, so a setter called from
// another thread waits at most for a snapshot of the settings, not for a
// render (setScene alone waits for the render in flight, whose scene it
// replaces).
#pragma once

#include <array>
#include <atomic>
#include <cstdint>
#include <deque>
#include <exception>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dr_api.h"

namespace devilray {

struct Error : std::runtime_error {
    dr_status status;
    Error(dr_status s, const std::string &msg) : std::runtime_error(msg), status(s) {}
};
void check(dr_status s);

struct Vec2 { float x, y; };
struct Vec3 { float x, y, z; };
struct Vec4 { float x, y, z, w; };
struct Size2i { int width, height; };

struct Ray {
    Vec3 p, v;  // origin, direction
    float tmin = 0.0f, tmax = 3.4e38f;
};

// Closest hit: t and the barycentric weights (w_A, w_B, w_C) of the
// triangle's vertices, as the reference's TriangleIntersection
// (include/tracing/Intersection.hpp:12-20), plus the triangle id; and the
// reference's Intersection fields (Intersection.hpp:31-40): the world hit
// point, the unit face normal turned towards the incoming ray, the world
// triangle area, ccw = ((A - B) x (A - C)) . v > 0, and the object's material.
// (The reference scales a model-space normal by 1/S without renormalising;
// here the normal is that of the placed triangle, unit length.)
struct Intersection {
    float t;
    Vec3 bari;
    int32_t triangle;
    Vec3 p{0, 0, 0};
    Vec3 n{0, 0, 0};
    float area = 0.0f;
    bool ccw = false;
    uint32_t mat = 0;
};

// DiffuseMaterial / TransparentMaterial (include/tracing/Material.hpp:9-25),
// with the debug colour the UV-checker view shows (the albedo when unset).
struct Material {
    static Material diffuse(Vec3 reflectance, Vec3 emission = {0, 0, 0});
    static Material transparent(float ior);
    dr_material m{};
    std::optional<Vec4> debug_color;
};

// Mesh (include/models/Mesh.hpp:7-27): positions, normals, and per triangle
// the position indices and (when loaded or generated) the normal indices.
struct Mesh {
    std::string name;
    std::vector<Vec3> points;
    std::vector<std::array<uint32_t, 3>> triangles;
    std::vector<Vec3> normals;
    std::vector<std::array<uint32_t, 3>> normal_indices;
};
// Two-triangle square of side `size` (createQuadMesh, src/GpuTris.cpp:87-107).
Mesh createQuadMesh(Vec3 center, Vec3 normal, Vec3 right, float size);

// OBJ reading (loadMesh, src/models/Mesh.cpp:117-186), with the reference's
// rules: v / vn / o / f are read; #, mtllib, s and vt lines are skipped; a
// face is v, v/vt, v//vn or v/vt/vn with 1-based (or negative, relative)
// indices; faces of fewer than three vertices are dropped; polygons of more
// than three, unknown tags, bad numbers or out-of-range indices throw
// std::runtime_error (std::out_of_range for indices beyond int). Without vn
// lines every triangle gets a (0, 0, 0) normal at index 0.
Mesh loadMesh(const std::string &path);
struct MeshBounds { Vec3 center; float extent; };
// Mean of the points and the largest distance from it (src/models/Mesh.cpp:220-245).
MeshBounds calculateMeshBounds(const Mesh &mesh);
// Points moved to (p - center) / extent (src/models/Mesh.cpp:188-195).
void normalizeMeshSize(Mesh &mesh);
// One normal per triangle, normalize((A - B) x (A - C)) (src/models/Mesh.cpp:197-218).
void generateCoarseNormals(Mesh &mesh);

// Scale-and-translate placement of an object (include/Transform.hpp:5-26):
// world point = p * s + position, applied when the scene is built.
struct Transform {
    Vec3 s{1.0f, 1.0f, 1.0f};
    Vec3 p{0.0f, 0.0f, 0.0f};

    Vec3 applyToPoint(const Vec3 &q) const { return {q.x * s.x + p.x, q.y * s.y + p.y, q.z * s.z + p.z}; }
    // A world ray in the object's model space: origin (o - p) / s, direction
    // v / s (not renormalised), the interval kept.
    Ray applyInverse(const Ray &r) const
    {
        Ray m = r;
        m.p = {(r.p.x - p.x) / s.x, (r.p.y - p.y) / s.y, (r.p.z - p.z) / s.z};
        m.v = {r.v.x / s.x, r.v.y / s.y, r.v.z / s.z};
        return m;
    }
};

// Row-major 4x4 affine transform acting on column vectors
// (include/models/Matrix.hpp:8-126); values[r][3] is the translation.
struct Matrix4x4f {
    float values[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};

    static Matrix4x4f identity() { return {}; }
    static Matrix4x4f translation(const Vec3 &offset)
    {
        Matrix4x4f m;
        m.values[0][3] = offset.x;
        m.values[1][3] = offset.y;
        m.values[2][3] = offset.z;
        return m;
    }
    // Right-handed rotation by `angle` radians about the unit vector `axis`:
    // R = cos(a) I + sin(a) [axis]_x + (1 - cos(a)) axis axis^T.
    static Matrix4x4f rotation(const Vec3 &axis, float angle);
    Matrix4x4f operator*(const Matrix4x4f &o) const;
    Vec3 applyToPosition(const Vec3 &q) const;   // w = 1: rotation and translation
    Vec3 applyToDirection(const Vec3 &d) const;  // w = 0: the 3x3 part only
    Vec3 getOffset() const { return {values[0][3], values[1][3], values[2][3]}; }
};

enum class PixelSampling : int { Center = 0, UniformRandom = 1 };
enum class OutputLinearity { Linear, GammaCorrected };
// include/DebugOptions.hpp:3-8
enum class DebugOptions : int { Off = 0, UVChecker = 1, BariCoords = 2, WindingOrder = 3 };
struct OutputOptions { OutputLinearity linearity = OutputLinearity::GammaCorrected; };

struct Intrinsics {
    float focal_length;
    Vec2 center;
};
// Camera (include/tracing/Camera.hpp:6-19): the camera-to-world transform
// (its rows 0..2 are used), pinhole intrinsics, physical pixel size.
struct Camera {
    Matrix4x4f transform;
    Intrinsics intrinsics{1.0f, {0.0f, 0.0f}};
    Size2i resolution{0, 0};
    Vec2 physical_pixel_size{1.0f, 1.0f};
};

// Mean of the last `window` values added, 0 when there are none
// (include/RunningAverage.hpp:5-16).
class RunningAverage {
public:
    explicit RunningAverage(int window) : window_(window) {}
    void add(double value);
    double mean() const;
    void reset() { values_.clear(); }

private:
    std::deque<double> values_;
    size_t window_;
};

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

// Scene: materials plus objects (a mesh and a material index each), copied
// into one device scene and BVH at build() (setScene builds if needed).
class Scene {
public:
    std::vector<Material> materials;
    struct Object {
        Mesh mesh;
        uint32_t material;
        Transform transform{};
        // TriangleMesh::setPosition / setScale (src/tracing/TriangleMesh.cpp:3-15);
        // the flattened BVH is rebuilt at the next build()
        void setPosition(Vec3 p) { transform.p = p; }
        void setScale(Vec3 s) { transform.s = s; }
    };
    std::vector<Object> objects;

    Scene() = default;
    Scene(Scene &&) noexcept;
    Scene &operator=(Scene &&) noexcept;
    ~Scene();

    void build(Context &ctx, uint32_t leaf_max = 4);
    // After objects moved (setPosition / setScale) or their points changed,
    // with the same objects and triangles as at build(): refits the built BVH
    // to the new world positions instead of rebuilding it (dr_scene_refit).
    void refit(Context &ctx);
    dr_scene *handle() const { return scene_; }
    uint32_t triangleCount() const { return n_tris_; }
    bool built() const { return built_; }
    // the placed triangles as built (or last refitted): xyz per point,
    // point indices and material per triangle
    const std::vector<float> &worldVertices() const { return world_; }
    const std::vector<uint32_t> &worldIndices() const { return idx_; }
    const std::vector<uint32_t> &triangleMaterials() const { return tri_mat_; }

private:
    std::vector<float> worldPoints() const;
    dr_scene *scene_ = nullptr;
    uint32_t n_tris_ = 0, n_points_ = 0;
    bool built_ = false;  // (a scene built without triangles has no device scene)
    std::vector<float> world_;
    std::vector<uint32_t> idx_, tri_mat_;
};

// Closest hit for each ray (std::nullopt on a miss); triangle ids index the
// scene's triangles in object order.
std::vector<std::optional<Intersection>> cast(Context &ctx, const Scene &scene, const std::vector<Ray> &rays);

class Renderer {
public:
    Renderer(Context &ctx, Size2i resolution);
    ~Renderer();
    Renderer(const Renderer &) = delete;
    Renderer &operator=(const Renderer &) = delete;

    void setScene(Scene scene);
    void setCamera(Camera cam);
    void setPixelSampling(PixelSampling sampling);
    void setOutputOptions(OutputOptions options);
    // Debug views of the first hit instead of the path tracer; clears the
    // image, as the reference's setDebug (src/Renderer.cpp:155-161).
    void setDebug(DebugOptions dbg);
    // Path tracer parameters beyond the reference (max_depth counts path
    // vertices; the reference's fixed behaviour is max_depth 10, no Russian
    // roulette, Schlick Fresnel).
    void setPathOptions(uint32_t max_depth, uint32_t rr_depth, uint32_t fresnel = DR_FRESNEL_EXACT,
                        uint64_t seed = 0);
    // Moves object `index` of the scene set by setScene (TriangleMesh::
    // setPosition / setScale on the renderer's scene), refits the BVH and
    // clears the image.
    void moveObject(size_t index, Vec3 position, Vec3 scale);

    // Adds exactly one sample per pixel, updates the display pixels and waits
    // for the device (as the reference's render(), src/Renderer.cpp:198-233);
    // render(n) adds n in one launch sequence.
    void render(uint32_t spp = 1);
    void clear();

    // Mean wall time in ms of the last 20 render() calls since the last clear
    // (include/Renderer.hpp:50), and the camera and bounce rays cast since
    // the last clear, shadow rays excluded (include/Renderer.hpp:51).
    float getMeanRenderTimes() const;
    uint64_t getTotalCasts() const { return total_casts_.load(); }

    void createPixels();
    // The display pixels as an 8-bit RGBA PNG (Renderer::saveImage,
    // src/Renderer.cpp:92-97); returns false, with a warning, if the file
    // cannot be written.
    bool saveImage(const std::string &path);
    const uint32_t *getPixels();
    const Vec4 *getRawPixels();      // FP32 view {sum r, g, b, count}
    const double *getRawPixelsFP64();  // the FP64 accumulation itself
    uint32_t getSampleCount() const { return samples_.load(); }

private:
    void createPixelsLocked(uint32_t encoding);
    Context &ctx_;
    Size2i res_;
    Scene scene_;
    Camera cam_;
    PixelSampling sampling_ = PixelSampling::UniformRandom;
    OutputOptions output_;
    DebugOptions debug_ = DebugOptions::Off;
    void *rgba_ = nullptr;  // device RGBA8 display pixels
    uint32_t max_depth_ = 10, rr_depth_ = 10, fresnel_ = DR_FRESNEL_SCHLICK;
    uint64_t seed_ = 0;
    void *accum_ = nullptr;  // device FP64 RGBW
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
