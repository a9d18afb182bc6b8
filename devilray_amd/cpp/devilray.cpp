// C++ host facade over the C ABI: see devilray.hpp.
#include "devilray.hpp"

#include <cstdio>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>

namespace devilray {

void check(dr_status s)
{
    if (s != DR_OK) throw Error(s, std::string(dr_status_string(s)) + ": " + dr_last_error());
}

// ---- materials ----------------------------------------------------------------
Vec4 getDebugColor(const Material &mat)
{
    return std::visit([](const auto &m) { return m.debug_color; }, mat);
}

namespace {

dr_material to_dr(const Material &mat)
{
    dr_material d{};
    if (const auto *diffuse = std::get_if<DiffuseMaterial>(&mat)) {
        d.kind = DR_MAT_LAMBERT;
        d.albedo[0] = diffuse->diffuse_reflectance.x;
        d.albedo[1] = diffuse->diffuse_reflectance.y;
        d.albedo[2] = diffuse->diffuse_reflectance.z;
        d.emission[0] = diffuse->emission.x;
        d.emission[1] = diffuse->emission.y;
        d.emission[2] = diffuse->emission.z;
        d.ior = 1.0f;
    } else {
        d.kind = DR_MAT_DIELECTRIC;
        d.albedo[0] = d.albedo[1] = d.albedo[2] = 1.0f;
        d.ior = std::get<TransparentMaterial>(mat).inside_medium.ior;
    }
    return d;
}

// luminance(radiantExitance(material)): pi E for a diffuse material, 0 for a
// transparent one (include/tracing/LightSampling.hpp:7-13, Utils.hpp:226-229)
float exitance_luminance(const Material &mat)
{
    const auto *diffuse = std::get_if<DiffuseMaterial>(&mat);
    if (!diffuse) return 0.0f;
    const float pi = 3.14159265358979f;
    return 0.2126f * diffuse->emission.x * pi + 0.7152f * diffuse->emission.y * pi +
           0.0722f * diffuse->emission.z * pi;
}

float triangle_area(const Vec3 &A, const Vec3 &B, const Vec3 &C) { return (A - B).cross(A - C).length() / 2; }

void trim_right(std::string &line)
{
    size_t i = line.size();
    while (i > 0 && (line[i - 1] == ' ' || line[i - 1] == '\t' || line[i - 1] == '\n' || line[i - 1] == '\r')) --i;
    line.resize(i);
}

// "pi", "pi/ti", "pi//ni" or "pi/ti/ni"; missing indices stay 1 (the first)
std::array<int, 2> parse_face_vertex(const std::string &str)
{
    for (const char c : str)
        if (std::string("0123456789/").find(c) == std::string::npos)
            throw std::runtime_error(std::string("Found unknown character in face descriptor: '") + c + "'");
    std::array<int, 2> v{1, 1};  // position, normal
    int kind = 0;
    size_t start = 0;
    for (;;) {
        const size_t slash = str.find('/', start);
        const std::string part = str.substr(start, slash == std::string::npos ? std::string::npos : slash - start);
        if (!part.empty() && kind < 3) {
            const int index = std::stoi(part);  // throws std::out_of_range beyond int
            if (kind == 0) v[0] = index;
            if (kind == 2) v[1] = index;
        }
        ++kind;
        if (slash == std::string::npos) break;
        start = slash + 1;
    }
    if (kind > 3) throw std::runtime_error("Unexpected number of indices in face vertex: " + std::to_string(kind));
    return v;
}

Vec3 read_vec3(std::istringstream &ss)
{
    Vec3 p{};
    ss >> p.x >> p.y >> p.z;
    if (ss.fail()) throw std::runtime_error("Failed to parse: " + ss.str());
    return p;
}

}  // namespace

Mesh loadMesh(const std::string &fileName)
{
    std::ifstream in(fileName);
    if (!in) throw std::runtime_error("File " + fileName + " does not exist");
    Mesh mesh;
    mesh.name = "Mesh";
    std::vector<std::array<std::array<int, 2>, 3>> faces;
    std::string line;
    while (std::getline(in, line)) {
        trim_right(line);
        std::istringstream ss(line);
        std::string tag;
        ss >> tag;
        if (tag.empty() || tag[0] == '#' || tag == "mtllib" || tag == "s" || tag == "vt") continue;
        if (tag == "v") { mesh.points.push_back(read_vec3(ss)); continue; }
        if (tag == "vn") { mesh.normals.push_back(read_vec3(ss)); continue; }
        if (tag == "o") { ss >> mesh.name; continue; }
        if (tag == "f") {
            std::vector<std::array<int, 2>> vs;
            std::string token;
            while (ss >> token) vs.push_back(parse_face_vertex(token));
            if (vs.size() < 3) continue;
            if (vs.size() > 3)
                throw std::runtime_error("Triangulation not supported. Found face with " + std::to_string(vs.size()) +
                                         " vertices");
            faces.push_back({vs[0], vs[1], vs[2]});
            continue;
        }
        throw std::runtime_error("Unrecognized tag: " + tag);
    }
    if (mesh.normals.empty()) mesh.normals.push_back(Vec3{0.0f, 0.0f, 0.0f});
    const int np = (int)mesh.points.size(), nn = (int)mesh.normals.size();
    for (size_t i = 0; i < faces.size(); ++i) {
        Vertex v[3];
        for (int k = 0; k < 3; ++k) {
            int pi = faces[i][k][0], ni = faces[i][k][1];
            pi = pi < 0 ? np + pi : pi - 1;  // 1-based, or relative to the end
            ni = ni < 0 ? nn + ni : ni - 1;
            if (pi < 0 || pi >= np)
                throw std::runtime_error("Invalid position index on triangle " + std::to_string(i) + ": " +
                                         std::to_string(pi) + ". Number of points is " + std::to_string(np) + ".");
            if (ni < 0 || ni >= nn)
                throw std::runtime_error("Invalid normal index on triangle " + std::to_string(i) + ": " +
                                         std::to_string(ni) + ". Number of normals is " + std::to_string(nn) + ".");
            v[k] = Vertex{(uint32_t)pi, (uint32_t)ni};
        }
        mesh.triangles.push_back(Triangle{v[0], v[1], v[2]});
    }
    return mesh;
}

MeshBounds calculateMeshBounds(const Mesh &mesh)
{
    Vec3 sum{0.0f, 0.0f, 0.0f};
    for (const Vec3 &p : mesh.points) sum += p;
    const Vec3 c = sum / (float)mesh.points.size();
    float extent = 0.0f;
    for (const Vec3 &p : mesh.points) extent = std::max(extent, (p - c).length());
    return MeshBounds{c, extent};
}

void normalizeMeshSize(Mesh &mesh)
{
    const MeshBounds b = calculateMeshBounds(mesh);
    for (Vec3 &p : mesh.points) p = (p - b.center) / b.extent;
}

void generateCoarseNormals(Mesh &mesh)
{
    mesh.normals.clear();
    for (Triangle &t : mesh.triangles) {
        const Vec3 A = mesh.points[t.a.pi], B = mesh.points[t.b.pi], C = mesh.points[t.c.pi];
        const uint32_t k = (uint32_t)mesh.normals.size();
        mesh.normals.push_back((A - B).cross(A - C).normalized());
        t.a.ni = t.b.ni = t.c.ni = k;
    }
}

void RunningAverage::add(double value)
{
    m_samples.push_back(value);
    while ((int)m_samples.size() > m_sampleCount) m_samples.pop_front();
}

double RunningAverage::mean() const
{
    if (m_samples.empty()) return 0.0;
    return std::accumulate(m_samples.begin(), m_samples.end(), 0.0) / (double)m_samples.size();
}

// ---- scene storage and views ------------------------------------------------------
GpuTris convertMeshToTris(Mesh &mesh, bool)
{
    return GpuTris{mesh.points, mesh.normals, mesh.triangles, mesh.name};
}

GpuTris createQuadMesh(Vec3 center, Vec3 normal, Vec3 right, float size)
{
    const Vec3 up = right.cross(normal);
    const float half = size / 2.0f;
    Mesh mesh;
    mesh.name = "quad";
    mesh.points = {center + right * (-half) + up * (-half), center + right * half + up * (-half),
                   center + right * half + up * half, center + right * (-half) + up * half};
    mesh.normals = {normal};
    mesh.triangles = {Triangle{{0, 0}, {2, 0}, {1, 0}}, Triangle{{0, 0}, {3, 0}, {2, 0}}};
    return convertMeshToTris(mesh);
}

TriangleMesh viewGpuTris(GpuTris &tris)
{
    TriangleMesh obj;
    obj.points = tris.points.data();
    obj.normals = tris.normals.data();
    obj.triangles = tris.triangles.data();
    obj.triangle_count = (int)tris.triangles.size();
    obj.point_count = (int)tris.points.size();
    double total = 0;  // FP64 sum of the triangle areas (src/GpuTris.cpp:39-56)
    for (const Triangle &t : tris.triangles)
        total += triangle_area(tris.points[t.a.pi], tris.points[t.b.pi], tris.points[t.c.pi]);
    obj.base_surface_area = (float)total;
    obj.surface_area = obj.base_surface_area;
    return obj;
}

void TriangleMesh::setPosition(const Vec3 &pos) { model_to_world.p = pos; }

void TriangleMesh::setScale(const Vec3 &scale)
{
    model_to_world.s = scale;
    const float vol = scale.x * scale.y * scale.z;
    surface_area = base_surface_area * std::cbrt(vol * vol);
}

Context::Context(int device) { check(dr_context_create(device, &ctx_)); }
Context::~Context()
{
    if (ctx_) dr_context_destroy(ctx_);
}

Context &defaultContext()
{
    // created on first use and kept until the process ends, so that no
    // renderer or scene can outlive it in static destruction
    static Context *ctx = new Context(0);
    return *ctx;
}

namespace {

// number of points a view's triangles refer to
int view_points(const TriangleMesh &o)
{
    if (o.point_count > 0) return o.point_count;
    uint32_t m = 0;
    for (int i = 0; i < o.triangle_count; ++i)
        m = std::max({m, o.triangles[i].a.pi + 1, o.triangles[i].b.pi + 1, o.triangles[i].c.pi + 1});
    return (int)m;
}

// The placed objects flattened for the device: world points (p * s +
// position), point indices rebased per object, and each triangle's material.
struct Flat {
    std::vector<float> verts;
    std::vector<uint32_t> idx, tri_mat;
    std::vector<uint32_t> first_tri;  // per object
};

Flat flatten(std::span<const TriangleMesh> objects, size_t n_materials)
{
    Flat f;
    uint32_t base = 0;
    for (const TriangleMesh &o : objects) {
        if (o.material < 0 || (size_t)o.material >= std::max<size_t>(n_materials, 1))
            throw Error(DR_E_INVALID, "object material index out of range");
        const int np = view_points(o);
        f.first_tri.push_back((uint32_t)f.tri_mat.size());
        for (int k = 0; k < np; ++k) {
            const Vec3 w = o.model_to_world.applyToPoint(o.points[k]);
            f.verts.insert(f.verts.end(), {w.x, w.y, w.z});
        }
        for (int i = 0; i < o.triangle_count; ++i) {
            const Triangle &t = o.triangles[i];
            for (uint32_t k : {t.a.pi, t.b.pi, t.c.pi}) {
                if (k >= (uint32_t)np) throw Error(DR_E_INVALID, "triangle vertex index out of range");
                f.idx.push_back(base + k);
            }
            f.tri_mat.push_back((uint32_t)o.material);
        }
        base += (uint32_t)np;
    }
    return f;
}

dr_scene *create_device_scene(Context &ctx, const Flat &f, const std::vector<dr_material> &mats, uint32_t leaf_max)
{
    if (f.tri_mat.empty()) return nullptr;
    dr_scene_desc d{};
    d.vertices = f.verts.data();
    d.n_vertices = (uint32_t)(f.verts.size() / 3);
    d.indices = f.idx.data();
    d.n_triangles = (uint32_t)f.tri_mat.size();
    d.tri_material = f.tri_mat.data();
    d.materials = mats.data();
    d.n_materials = (uint32_t)mats.size();
    dr_scene *s = nullptr;
    check(dr_scene_create(ctx.handle(), &d, &s));
    const dr_bvh_params bp{leaf_max, 0};
    const dr_status st = dr_scene_build_bvh(s, &bp, nullptr);
    if (st != DR_OK) {
        dr_scene_destroy(s);
        check(st);
    }
    return s;
}

struct DeviceBuffer {
    Context &ctx;
    void *ptr = nullptr;
    DeviceBuffer(Context &c, size_t bytes) : ctx(c) { check(dr_device_alloc(c.handle(), bytes ? bytes : 16, &ptr)); }
    ~DeviceBuffer() { dr_device_free(ctx.handle(), ptr); }
};

struct SceneHandle {
    dr_scene *s = nullptr;
    ~SceneHandle()
    {
        if (s) dr_scene_destroy(s);
    }
};

// Closest hits of a ray batch on a device scene (counters when given).
std::vector<dr_hit> trace(Context &ctx, dr_scene *s, const std::vector<Ray> &rays, std::vector<uint32_t> *box = nullptr,
                          std::vector<uint32_t> *tri = nullptr)
{
    const size_t n = rays.size();
    std::vector<dr_hit> hits(n, dr_hit{-1, INFINITY, 0.0f, 0.0f});
    if (!s || n == 0) return hits;
    std::vector<dr_ray> in(n);
    for (size_t i = 0; i < n; ++i) {
        const Ray &r = rays[i];
        in[i] = dr_ray{r.p.x, r.p.y, r.p.z, r.tmin, r.v.x, r.v.y, r.v.z, r.tmax};
    }
    DeviceBuffer dr(ctx, sizeof(dr_ray) * n), dh(ctx, sizeof(dr_hit) * n);
    check(dr_memcpy_htod(ctx.handle(), dr.ptr, in.data(), sizeof(dr_ray) * n));
    if (box && tri) {
        DeviceBuffer db(ctx, 4 * n), dt(ctx, 4 * n);
        check(dr_trace_closest_counted(s, (const dr_ray *)dr.ptr, (uint32_t)n, (dr_hit *)dh.ptr, (uint32_t *)db.ptr,
                                       (uint32_t *)dt.ptr, nullptr));
        box->resize(n);
        tri->resize(n);
        check(dr_memcpy_dtoh(ctx.handle(), box->data(), db.ptr, 4 * n));
        check(dr_memcpy_dtoh(ctx.handle(), tri->data(), dt.ptr, 4 * n));
    } else {
        check(dr_trace_closest(s, (const dr_ray *)dr.ptr, (uint32_t)n, (dr_hit *)dh.ptr, nullptr));
    }
    check(dr_memcpy_dtoh(ctx.handle(), hits.data(), dh.ptr, sizeof(dr_hit) * n));
    return hits;
}

// The reference's Intersection of a device hit on the flattened triangles:
// hit point from the weights, unit face normal turned against the ray,
// world area, winding flag, material.
Intersection describe(const dr_hit &h, const Ray &ray, const std::vector<float> &w, const std::vector<uint32_t> &ix,
                      const std::vector<uint32_t> &tri_mat)
{
    const size_t tri = (size_t)h.id;
    const Vec3 A{w[3 * ix[3 * tri]], w[3 * ix[3 * tri] + 1], w[3 * ix[3 * tri] + 2]};
    const Vec3 B{w[3 * ix[3 * tri + 1]], w[3 * ix[3 * tri + 1] + 1], w[3 * ix[3 * tri + 1] + 2]};
    const Vec3 C{w[3 * ix[3 * tri + 2]], w[3 * ix[3 * tri + 2] + 1], w[3 * ix[3 * tri + 2] + 2]};
    const Vec3 nf = (A - B).cross(A - C);
    const float len = nf.length();
    const bool ccw = nf.dot(ray.v) > 0.0f;
    Intersection x{};
    x.t = h.t;
    x.p = A + (B - A) * h.u + (C - A) * h.v;
    x.uv = Vec2f{0.0f, 0.0f};
    x.n = nf * ((ccw ? -1.0f : 1.0f) / (len > 0.0f ? len : 1.0f));
    x.mat = (int)tri_mat[tri];
    x.object = nullptr;
    x.triangle = TriangleHitData{Vec3{1.0f - h.u - h.v, h.u, h.v}, 0.5f * len, ccw};
    x.triangle_id = h.id;
    return x;
}

std::vector<dr_material> device_materials(size_t n_needed)
{
    // objects traced without a scene's material list: one grey diffuse
    // material stands for every index the objects use
    dr_material m{};
    m.kind = DR_MAT_LAMBERT;
    m.albedo[0] = m.albedo[1] = m.albedo[2] = 0.5f;
    m.ior = 1.0f;
    return std::vector<dr_material>(std::max<size_t>(n_needed, 1), m);
}

// One ray against a set of objects on the device: the hit, which object it
// belongs to, and (optionally) the traversal counters.
std::optional<Intersection> trace_objects(const Ray &ray, std::span<const TriangleMesh> objects,
                                          benchmark::HitTests *counts)
{
    if (objects.empty()) return std::nullopt;
    int max_mat = 0;
    for (const TriangleMesh &o : objects) max_mat = std::max(max_mat, o.material);
    const Flat f = flatten(objects, (size_t)max_mat + 1);
    Context &ctx = defaultContext();
    SceneHandle sh;
    sh.s = create_device_scene(ctx, f, device_materials((size_t)max_mat + 1), 4);
    std::vector<uint32_t> box, tri;
    const std::vector<dr_hit> h = trace(ctx, sh.s, {ray}, counts ? &box : nullptr, counts ? &tri : nullptr);
    if (counts && !box.empty()) {
        counts->bbox_tests += box[0];
        counts->triangle_tests += tri[0];
    }
    if (h[0].id < 0) return std::nullopt;
    Intersection x = describe(h[0], ray, f.verts, f.idx, f.tri_mat);
    const size_t k = (size_t)(std::upper_bound(f.first_tri.begin(), f.first_tri.end(), (uint32_t)h[0].id) -
                              f.first_tri.begin()) - 1;
    x.object = &objects[k];
    x.triangle_id = h[0].id - (int32_t)f.first_tri[k];
    return x;
}

}  // namespace

// ---- the scene on the device ----------------------------------------------------------
Scene::Scene(Scene &&o) noexcept
    : info(o.info), objects(std::move(o.objects)), materials(std::move(o.materials)),
      mesh_storage(std::move(o.mesh_storage)), scene_(o.scene_), n_tris_(o.n_tris_), n_points_(o.n_points_),
      built_(o.built_), world_(std::move(o.world_)), idx_(std::move(o.idx_)), tri_mat_(std::move(o.tri_mat_))
{
    o.scene_ = nullptr;
}

Scene &Scene::operator=(Scene &&o) noexcept
{
    if (this != &o) {
        if (scene_) dr_scene_destroy(scene_);
        info = o.info;
        objects = std::move(o.objects);
        materials = std::move(o.materials);
        mesh_storage = std::move(o.mesh_storage);
        scene_ = o.scene_;
        n_tris_ = o.n_tris_;
        n_points_ = o.n_points_;
        built_ = o.built_;
        world_ = std::move(o.world_);
        idx_ = std::move(o.idx_);
        tri_mat_ = std::move(o.tri_mat_);
        o.scene_ = nullptr;
    }
    return *this;
}

Scene::~Scene()
{
    if (scene_) dr_scene_destroy(scene_);
}

void Scene::refit(Context &ctx)
{
    if (!scene_) {
        build(ctx);
        return;
    }
    Flat f = flatten(objects, materials.size());
    if (f.verts.size() != 3ull * n_points_ || f.tri_mat.size() != n_tris_)
        throw Error(DR_E_INVALID, "refit needs the objects and triangles of the last build");
    check(dr_scene_refit(scene_, f.verts.data(), nullptr));
    world_ = std::move(f.verts);
}

void Scene::build(Context &ctx, uint32_t leaf_max)
{
    Flat f = flatten(objects, materials.size());
    std::vector<dr_material> mats;
    for (const Material &m : materials) mats.push_back(to_dr(m));
    if (scene_) {
        dr_scene_destroy(scene_);
        scene_ = nullptr;
    }
    // calculateLightWeights (src/Renderer.cpp:113-152): object weight =
    // surface area x luminance of the radiant exitance
    double power = 0.0;
    for (const TriangleMesh &o : objects) power += (double)o.surface_area * exitance_luminance(materials[o.material]);
    info.total_radiant_power = (float)power;
    n_tris_ = (uint32_t)f.tri_mat.size();
    n_points_ = (uint32_t)(f.verts.size() / 3);
    built_ = true;
    scene_ = create_device_scene(ctx, f, mats, leaf_max);  // none without triangles: cast misses
    world_ = std::move(f.verts);
    idx_ = std::move(f.idx);
    tri_mat_ = std::move(f.tri_mat);
    if (scene_ && !materials.empty()) {
        std::vector<float> rgba;
        for (const Material &m : materials) {
            const Vec4 c = getDebugColor(m);
            rgba.insert(rgba.end(), {c.x, c.y, c.z, c.w});
        }
        check(dr_scene_set_debug_colors(scene_, rgba.data(), (uint32_t)materials.size()));
    }
}

// ---- tracing -------------------------------------------------------------------------------
std::optional<TriangleIntersection> testTriangleIntersection(const Ray &ray, const TriangleVertices &triangle)
{
    GpuTris tris;
    tris.points = {triangle.a, triangle.b, triangle.c};
    tris.triangles = {Triangle{{0, 0}, {1, 0}, {2, 0}}};
    const TriangleMesh view = viewGpuTris(tris);
    const auto hit = trace_objects(ray, std::span<const TriangleMesh>(&view, 1), nullptr);
    if (!hit) return std::nullopt;
    return TriangleIntersection{hit->t, hit->triangle.bari};
}

std::optional<Intersection> getIntersection(const Ray &ray, const TriangleMesh &tris)
{
    return trace_objects(ray, std::span<const TriangleMesh>(&tris, 1), nullptr);
}

std::optional<Intersection> getIntersectionBenchmark(const Ray &ray, const TriangleMesh &tris,
                                                     benchmark::HitTests &benchmark)
{
    return trace_objects(ray, std::span<const TriangleMesh>(&tris, 1), &benchmark);
}

std::optional<Intersection> cast(const Ray &ray, std::span<const TriangleMesh> objects, const ObjectsInfo &)
{
    return trace_objects(ray, objects, nullptr);
}

std::vector<std::optional<Intersection>> cast(Context &ctx, const Scene &scene, const std::vector<Ray> &rays)
{
    if (!scene.built()) throw Error(DR_E_INVALID, "scene not built");
    const size_t n = rays.size();
    std::vector<std::optional<Intersection>> out(n);
    if (!scene.handle() || n == 0) return out;  // no triangles
    const std::vector<dr_hit> hits = trace(ctx, scene.handle(), rays);
    for (size_t i = 0; i < n; ++i)
        if (hits[i].id >= 0)
            out[i] = describe(hits[i], rays[i], scene.worldVertices(), scene.worldIndices(), scene.triangleMaterials());
    return out;
}

// ---- the renderer --------------------------------------------------------------------------
Renderer::Renderer(Size2i resolution) : Renderer(defaultContext(), resolution) {}

Renderer::Renderer(Context &ctx, Size2i resolution) : ctx_(ctx), res_(resolution)
{
    if (resolution.width <= 0 || resolution.height <= 0) throw Error(DR_E_INVALID, "empty resolution");
    const size_t npix = (size_t)resolution.width * resolution.height;
    check(dr_device_alloc(ctx.handle(), npix * 4 * sizeof(double), &accum_));
    check(dr_fill_zero(ctx.handle(), accum_, npix * 4 * sizeof(double), nullptr));
    pixels_.assign(npix, 0u);
    display_.assign(npix, 0u);
    buffers_.resolution = resolution;
    buffers_.color = static_cast<const double *>(accum_);
    buffers_.casts = &total_casts_;
}

Renderer::~Renderer()
{
    if (accum_) dr_device_free(ctx_.handle(), accum_);
    if (rgba_) dr_device_free(ctx_.handle(), rgba_);
}

void Renderer::setScene(Scene scene)
{
    std::scoped_lock frame{frame_mutex_};
    std::scoped_lock guard{render_mutex_};
    scene_ = std::move(scene);
    if (!scene_.built()) scene_.build(ctx_);
    calculateLightWeightsLocked();
    needs_clear_ = true;
}

void Renderer::calculateLightWeights()
{
    std::scoped_lock frame{frame_mutex_};
    std::scoped_lock guard{render_mutex_};
    calculateLightWeightsLocked();
}

void Renderer::calculateLightWeightsLocked()
{
    if (scene_.handle()) check(dr_scene_build_lights(scene_.handle(), nullptr));
}

void Renderer::setCamera(Camera cam)
{
    std::scoped_lock guard{render_mutex_};
    cam_ = cam;
    needs_clear_ = true;
}

void Renderer::setPixelSampling(PixelSampling sampling)
{
    std::scoped_lock guard{render_mutex_};
    sampling_ = sampling;
    needs_clear_ = true;
}

void Renderer::setOutputOptions(OutputOptions options)
{
    std::scoped_lock guard{render_mutex_};
    output_ = options;
}

void Renderer::setDebug(DebugOptions dbg)
{
    std::scoped_lock guard{render_mutex_};
    debug_ = dbg;
    needs_clear_ = true;
}

void Renderer::setPathOptions(uint32_t max_depth, uint32_t rr_depth, uint32_t fresnel, uint64_t seed)
{
    std::scoped_lock guard{render_mutex_};
    max_depth_ = max_depth;
    rr_depth_ = rr_depth;
    fresnel_ = fresnel;
    seed_ = seed;
    needs_clear_ = true;
}

void Renderer::moveObject(size_t index, Vec3 position, Vec3 scale)
{
    std::scoped_lock frame{frame_mutex_};
    std::scoped_lock guard{render_mutex_};
    if (index >= scene_.objects.size()) throw Error(DR_E_INVALID, "object index out of range");
    scene_.objects[index].setPosition(position);
    scene_.objects[index].setScale(scale);
    scene_.refit(ctx_);
    needs_clear_ = true;
}

void Renderer::clear()
{
    std::scoped_lock guard{render_mutex_};
    needs_clear_ = true;
}

float Renderer::getMeanRenderTimes() const
{
    std::scoped_lock guard{stats_mutex_};
    return (float)render_times_.mean();
}

void Renderer::render(uint32_t spp)
{
    std::scoped_lock frame{frame_mutex_};
    dr_camera c{};
    dr_render_params p{};
    DebugOptions debug;
    uint32_t encoding;
    {
        std::scoped_lock guard{render_mutex_};
        if (!scene_.built()) throw Error(DR_E_INVALID, "no scene set");
        if (needs_clear_) {
            check(dr_fill_zero(ctx_.handle(), accum_, (size_t)res_.width * res_.height * 4 * sizeof(double), nullptr));
            uint64_t unused;
            if (scene_.handle()) check(dr_scene_ray_casts(scene_.handle(), &unused, 1, nullptr));
            samples_ = 0;
            total_casts_ = 0;
            std::scoped_lock stats{stats_mutex_};
            render_times_.reset();
            needs_clear_ = false;
        }
        std::memcpy(c.transform, cam_.transform.values, sizeof c.transform);  // rows 0..2
        c.focal_length = cam_.intrinsics.focal_length;
        c.center[0] = cam_.intrinsics.center.x;
        c.center[1] = cam_.intrinsics.center.y;
        c.pixel_size[0] = cam_.physical_pixel_size.width;
        c.pixel_size[1] = cam_.physical_pixel_size.height;
        c.width = (uint32_t)res_.width;
        c.height = (uint32_t)res_.height;
        c.pixel_sampling = (uint32_t)sampling_;
        p.spp = spp;
        p.sample_offset = samples_;
        p.max_depth = max_depth_;
        p.rr_depth = rr_depth_;
        p.seed = seed_;
        p.fresnel = fresnel_;
        debug = debug_;
        encoding = output_.linearity == OutputLinearity::GammaCorrected ? DR_OUTPUT_SRGB : DR_OUTPUT_LINEAR;
    }
    const auto t0 = std::chrono::steady_clock::now();
    if (!scene_.handle()) {
        // a scene without triangles: every path misses, the image stays black
    } else if (debug == DebugOptions::Off) {
        check(dr_render(scene_.handle(), &c, &p, (double *)accum_, nullptr));
    } else {
        check(dr_render_debug(scene_.handle(), &c, &p, (uint32_t)debug, (double *)accum_, nullptr));
    }
    createPixelsLocked(encoding);  // synchronises
    samples_ += spp;
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    {
        std::scoped_lock stats{stats_mutex_};
        render_times_.add(ms);
    }
    uint64_t casts = 0;
    if (scene_.handle()) check(dr_scene_ray_casts(scene_.handle(), &casts, 0, nullptr));
    total_casts_ = casts;
}

const double *Renderer::getRawPixelsFP64()
{
    std::scoped_lock frame{frame_mutex_};
    const size_t n = (size_t)res_.width * res_.height * 4;
    host_fp64_.resize(n);
    check(dr_memcpy_dtoh(ctx_.handle(), host_fp64_.data(), accum_, n * sizeof(double)));
    return host_fp64_.data();
}

const Vec4 *Renderer::getRawPixels()
{
    const double *a = getRawPixelsFP64();
    const size_t npix = (size_t)res_.width * res_.height;
    host_vec4_.resize(npix);
    for (size_t i = 0; i < npix; ++i)
        host_vec4_[i] = Vec4{(float)a[4 * i], (float)a[4 * i + 1], (float)a[4 * i + 2], (float)a[4 * i + 3]};
    return host_vec4_.data();
}

// sRGB transfer (clamped) when gamma-corrected, x255, round, pack
// r | g<<8 | b<<16 | a<<24, rows flipped (src/Renderer.cpp:51-90).
void Renderer::createPixels()
{
    std::scoped_lock frame{frame_mutex_};
    uint32_t enc;
    {
        std::scoped_lock guard{render_mutex_};
        enc = output_.linearity == OutputLinearity::GammaCorrected ? DR_OUTPUT_SRGB : DR_OUTPUT_LINEAR;
    }
    createPixelsLocked(enc);
}

void Renderer::createPixelsLocked(uint32_t enc)
{
    // the reference's host loop (src/Renderer.cpp:51-90) as one device pass;
    // only the 8-bit pixels cross to the host
    const size_t n = (size_t)res_.width * res_.height;
    if (!rgba_) check(dr_device_alloc(ctx_.handle(), n * sizeof(uint32_t), &rgba_));
    check(dr_resolve_rgba8(ctx_.handle(), (const double *)accum_, (uint32_t)res_.width, (uint32_t)res_.height, enc,
                           (uint32_t *)rgba_, nullptr));
    pixels_.resize(n);
    check(dr_memcpy_dtoh(ctx_.handle(), pixels_.data(), rgba_, n * sizeof(uint32_t)));
    std::scoped_lock guard{display_mutex_};
    display_ = pixels_;
}

namespace {

uint32_t crc32_png(const uint8_t *d, size_t n, uint32_t c = 0xFFFFFFFFu)
{
    static uint32_t table[256];
    static bool ready = false;
    if (!ready) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t v = i;
            for (int k = 0; k < 8; ++k) v = (v & 1u) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            table[i] = v;
        }
        ready = true;
    }
    for (size_t i = 0; i < n; ++i) c = table[(c ^ d[i]) & 0xFFu] ^ (c >> 8);
    return c;
}

void put_be32(std::vector<uint8_t> &o, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8) o.push_back((uint8_t)(v >> s));
}

void png_chunk(std::vector<uint8_t> &o, const char *type, const std::vector<uint8_t> &body)
{
    put_be32(o, (uint32_t)body.size());
    const size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), body.begin(), body.end());
    put_be32(o, crc32_png(o.data() + start, o.size() - start) ^ 0xFFFFFFFFu);
}

}  // namespace

bool Renderer::saveImage(const std::filesystem::path &path)
{
    // PNG with the image data in stored (uncompressed) deflate blocks: a
    // valid zlib stream that needs no compressor
    std::vector<uint32_t> px;
    {
        std::scoped_lock guard{display_mutex_};
        px = display_;
    }
    const uint32_t w = (uint32_t)res_.width, h = (uint32_t)res_.height;
    if (px.size() != (size_t)w * h) px.assign((size_t)w * h, 0xFF000000u);
    std::vector<uint8_t> raw;
    raw.reserve((size_t)h * (4 * w + 1));
    for (uint32_t y = 0; y < h; ++y) {
        raw.push_back(0);  // filter: none
        const uint8_t *row = reinterpret_cast<const uint8_t *>(px.data() + (size_t)y * w);
        raw.insert(raw.end(), row, row + 4 * (size_t)w);  // R, G, B, A (little-endian u32)
    }
    std::vector<uint8_t> z{0x78, 0x01};
    for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
        const size_t len = std::min<size_t>(65535, raw.size() - off);
        z.push_back(off + len >= raw.size() ? 1 : 0);
        z.push_back((uint8_t)len); z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)~len); z.push_back((uint8_t)(~len >> 8));
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + len);
        if (raw.empty()) break;
    }
    uint32_t a = 1, b = 0;  // Adler-32
    for (uint8_t v : raw) { a = (a + v) % 65521u; b = (b + a) % 65521u; }
    put_be32(z, (b << 16) | a);
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, w);
    put_be32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8 bits, RGBA, deflate, adaptive filters, no interlace
    std::vector<uint8_t> out{0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    png_chunk(out, "IHDR", ihdr);
    png_chunk(out, "IDAT", z);
    png_chunk(out, "IEND", {});
    FILE *f = std::fopen(path.string().c_str(), "wb");
    const bool ok = f && std::fwrite(out.data(), 1, out.size(), f) == out.size();
    if (f) std::fclose(f);
    if (!ok) std::fprintf(stderr, "WARN: Failed to save %s\n", path.string().c_str());
    return ok;
}

const uint32_t *Renderer::getPixels()
{
    std::scoped_lock guard{display_mutex_};
    return display_.data();
}

RenderLoop::RenderLoop(Renderer &renderer) : renderer_(renderer)
{
    thread_ = std::thread([this] {
        try {
            while (!stop_.load(std::memory_order_relaxed)) {
                renderer_.render();
                average_render_time_.store(renderer_.getMeanRenderTimes(), std::memory_order_relaxed);
                total_ray_casts_.store(renderer_.getTotalCasts(), std::memory_order_relaxed);
                passes_.fetch_add(1);
            }
        } catch (...) {
            std::scoped_lock guard{error_mutex_};
            error_ = std::current_exception();
        }
    });
}

RenderLoop::~RenderLoop() { stop(); }

void RenderLoop::stop()
{
    stop_.store(true);
    if (thread_.joinable()) thread_.join();
}

std::exception_ptr RenderLoop::error() const
{
    std::scoped_lock guard{error_mutex_};
    return error_;
}

}  // namespace devilray
