// C++ host facade over the C ABI: see devilray.hpp.
#include "devilray.hpp"

#include <cstdio>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <sstream>
#include <fstream>

namespace devilray {

void check(dr_status s)
{
    if (s != DR_OK) throw Error(s, std::string(dr_status_string(s)) + ": " + dr_last_error());
}

Material Material::diffuse(Vec3 reflectance, Vec3 emission)
{
    Material r;
    r.m.kind = DR_MAT_LAMBERT;
    r.m.albedo[0] = reflectance.x;
    r.m.albedo[1] = reflectance.y;
    r.m.albedo[2] = reflectance.z;
    r.m.emission[0] = emission.x;
    r.m.emission[1] = emission.y;
    r.m.emission[2] = emission.z;
    r.m.ior = 1.0f;
    return r;
}

Material Material::transparent(float ior)
{
    Material r;
    r.m.kind = DR_MAT_DIELECTRIC;
    r.m.albedo[0] = r.m.albedo[1] = r.m.albedo[2] = 1.0f;
    r.m.ior = ior;
    return r;
}

Mesh createQuadMesh(Vec3 center, Vec3 normal, Vec3 right, float size)
{
    // up = right x normal; corners (-,-), (+,-), (+,+), (-,+); triangles (0,2,1), (0,3,2)
    const Vec3 up{right.y * normal.z - right.z * normal.y, right.z * normal.x - right.x * normal.z,
                  right.x * normal.y - right.y * normal.x};
    const float h = size / 2.0f;
    auto corner = [&](float a, float b) {
        return Vec3{center.x + right.x * a + up.x * b, center.y + right.y * a + up.y * b,
                    center.z + right.z * a + up.z * b};
    };
    Mesh m;
    m.points = {corner(-h, -h), corner(h, -h), corner(h, h), corner(-h, h)};
    m.triangles = {{0, 2, 1}, {0, 3, 2}};
    return m;
}

namespace {

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

Mesh loadMesh(const std::string &path)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("File " + path + " does not exist");
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
        std::array<uint32_t, 3> t{}, tn{};
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
            t[k] = (uint32_t)pi;
            tn[k] = (uint32_t)ni;
        }
        mesh.triangles.push_back(t);
        mesh.normal_indices.push_back(tn);
    }
    return mesh;
}

MeshBounds calculateMeshBounds(const Mesh &mesh)
{
    Vec3 sum{0.0f, 0.0f, 0.0f};
    for (const Vec3 &p : mesh.points) { sum.x += p.x; sum.y += p.y; sum.z += p.z; }
    const float n = (float)mesh.points.size();
    const Vec3 c{sum.x / n, sum.y / n, sum.z / n};
    float extent = 0.0f;
    for (const Vec3 &p : mesh.points) {
        const float dx = p.x - c.x, dy = p.y - c.y, dz = p.z - c.z;
        extent = std::max(extent, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    return MeshBounds{c, extent};
}

void normalizeMeshSize(Mesh &mesh)
{
    const MeshBounds b = calculateMeshBounds(mesh);
    for (Vec3 &p : mesh.points)
        p = Vec3{(p.x - b.center.x) / b.extent, (p.y - b.center.y) / b.extent, (p.z - b.center.z) / b.extent};
}

void generateCoarseNormals(Mesh &mesh)
{
    mesh.normals.clear();
    mesh.normal_indices.clear();
    for (const auto &t : mesh.triangles) {
        const Vec3 A = mesh.points[t[0]], B = mesh.points[t[1]], C = mesh.points[t[2]];
        const Vec3 e1{A.x - B.x, A.y - B.y, A.z - B.z}, e2{A.x - C.x, A.y - C.y, A.z - C.z};
        Vec3 n{e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
        const float l = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
        n = Vec3{n.x / l, n.y / l, n.z / l};
        const uint32_t k = (uint32_t)mesh.normals.size();
        mesh.normals.push_back(n);
        mesh.normal_indices.push_back({k, k, k});
    }
}

Context::Context(int device) { check(dr_context_create(device, &ctx_)); }
Context::~Context()
{
    if (ctx_) dr_context_destroy(ctx_);
}

Scene::Scene(Scene &&o) noexcept
    : materials(std::move(o.materials)), objects(std::move(o.objects)), scene_(o.scene_), n_tris_(o.n_tris_),
      n_points_(o.n_points_), built_(o.built_), world_(std::move(o.world_)), idx_(std::move(o.idx_)),
      tri_mat_(std::move(o.tri_mat_))
{
    o.scene_ = nullptr;
}

Scene &Scene::operator=(Scene &&o) noexcept
{
    if (this != &o) {
        if (scene_) dr_scene_destroy(scene_);
        materials = std::move(o.materials);
        objects = std::move(o.objects);
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

// every object's points placed by its transform, p * s + position, in object order
std::vector<float> Scene::worldPoints() const
{
    std::vector<float> verts;
    for (const Object &o : objects) {
        const Transform &tf = o.transform;
        for (const Vec3 &p : o.mesh.points) {
            const Vec3 w = tf.applyToPoint(p);
            verts.push_back(w.x);
            verts.push_back(w.y);
            verts.push_back(w.z);
        }
    }
    return verts;
}

void Scene::refit(Context &ctx)
{
    if (!scene_) {
        build(ctx);
        return;
    }
    const std::vector<float> verts = worldPoints();
    uint32_t tris = 0;
    for (const Object &o : objects) tris += (uint32_t)o.mesh.triangles.size();
    if (verts.size() != 3ull * n_points_ || tris != n_tris_)
        throw Error(DR_E_INVALID, "refit needs the objects and triangles of the last build");
    check(dr_scene_refit(scene_, verts.data(), nullptr));
    world_ = verts;
}

void Scene::build(Context &ctx, uint32_t leaf_max)
{
    const std::vector<float> verts = worldPoints();
    std::vector<uint32_t> idx, tri_mat;
    uint32_t base = 0;
    for (const Object &o : objects) {
        if (o.material >= materials.size()) throw Error(DR_E_INVALID, "object material index out of range");
        for (const auto &t : o.mesh.triangles) {
            for (uint32_t k : t) {
                if (k >= o.mesh.points.size()) throw Error(DR_E_INVALID, "triangle vertex index out of range");
                idx.push_back(base + k);
            }
            tri_mat.push_back(o.material);
        }
        base += (uint32_t)o.mesh.points.size();
    }
    std::vector<dr_material> mats;
    for (const Material &m : materials) mats.push_back(m.m);
    dr_scene_desc d{};
    d.vertices = verts.data();  // (the scene copies them in)
    d.n_vertices = (uint32_t)(verts.size() / 3);
    d.indices = idx.data();
    d.n_triangles = (uint32_t)tri_mat.size();
    d.tri_material = tri_mat.data();
    d.materials = mats.data();
    d.n_materials = (uint32_t)mats.size();
    if (scene_) {
        dr_scene_destroy(scene_);
        scene_ = nullptr;
    }
    n_tris_ = d.n_triangles;
    n_points_ = d.n_vertices;
    world_ = verts;
    built_ = true;
    if (d.n_triangles == 0) return;  // nothing to trace: cast misses (CastTest.EmptySceneMisses)
    check(dr_scene_create(ctx.handle(), &d, &scene_));
    dr_bvh_params bp{leaf_max, 0};
    check(dr_scene_build_bvh(scene_, &bp, nullptr));
    bool any_debug = false;
    for (const Material &m : materials) any_debug = any_debug || m.debug_color.has_value();
    if (any_debug) {
        std::vector<float> rgba;
        for (const Material &m : materials) {
            const Vec4 c = m.debug_color.value_or(Vec4{m.m.albedo[0], m.m.albedo[1], m.m.albedo[2], 0.0f});
            rgba.insert(rgba.end(), {c.x, c.y, c.z, c.w});
        }
        check(dr_scene_set_debug_colors(scene_, rgba.data(), (uint32_t)materials.size()));
    }
    idx_ = std::move(idx);
    tri_mat_ = std::move(tri_mat);
}

namespace {
struct DeviceBuffer {
    Context &ctx;
    void *ptr = nullptr;
    DeviceBuffer(Context &c, size_t bytes) : ctx(c) { check(dr_device_alloc(c.handle(), bytes ? bytes : 16, &ptr)); }
    ~DeviceBuffer() { dr_device_free(ctx.handle(), ptr); }
};
}  // namespace

std::vector<std::optional<Intersection>> cast(Context &ctx, const Scene &scene, const std::vector<Ray> &rays)
{
    if (!scene.built()) throw Error(DR_E_INVALID, "scene not built");
    const size_t n = rays.size();
    if (!scene.handle() || n == 0) return std::vector<std::optional<Intersection>>(n);  // no triangles
    std::vector<dr_ray> in(n);
    for (size_t i = 0; i < n; ++i) {
        const Ray &r = rays[i];
        in[i] = dr_ray{r.p.x, r.p.y, r.p.z, r.tmin, r.v.x, r.v.y, r.v.z, r.tmax};
    }
    DeviceBuffer dr(ctx, sizeof(dr_ray) * n), dh(ctx, sizeof(dr_hit) * n);
    check(dr_memcpy_htod(ctx.handle(), dr.ptr, in.data(), sizeof(dr_ray) * n));
    check(dr_trace_closest(scene.handle(), (const dr_ray *)dr.ptr, (uint32_t)n, (dr_hit *)dh.ptr, nullptr));
    std::vector<dr_hit> hits(n);
    check(dr_memcpy_dtoh(ctx.handle(), hits.data(), dh.ptr, sizeof(dr_hit) * n));
    std::vector<std::optional<Intersection>> out(n);
    const std::vector<float> &w = scene.worldVertices();
    const std::vector<uint32_t> &ix = scene.worldIndices();
    for (size_t i = 0; i < n; ++i) {
        if (hits[i].id < 0) continue;
        const dr_hit &h = hits[i];
        Intersection x{h.t, Vec3{1.0f - h.u - h.v, h.u, h.v}, h.id};
        const size_t tri = (size_t)h.id;
        const float *A = &w[3 * (size_t)ix[3 * tri]], *B = &w[3 * (size_t)ix[3 * tri + 1]],
                    *C = &w[3 * (size_t)ix[3 * tri + 2]];
        const float ab[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]}, ac[3] = {A[0] - C[0], A[1] - C[1], A[2] - C[2]};
        const float nf[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
        const float len = std::sqrt(nf[0] * nf[0] + nf[1] * nf[1] + nf[2] * nf[2]);
        const Vec3 &v = rays[i].v;
        const float side = nf[0] * v.x + nf[1] * v.y + nf[2] * v.z;
        x.ccw = side > 0.0f;
        const float s = (side > 0.0f ? -1.0f : 1.0f) / (len > 0.0f ? len : 1.0f);
        x.n = Vec3{nf[0] * s, nf[1] * s, nf[2] * s};
        x.area = 0.5f * len;
        x.p = Vec3{A[0] + (B[0] - A[0]) * h.u + (C[0] - A[0]) * h.v, A[1] + (B[1] - A[1]) * h.u + (C[1] - A[1]) * h.v,
                   A[2] + (B[2] - A[2]) * h.u + (C[2] - A[2]) * h.v};
        x.mat = scene.triangleMaterials()[tri];
        out[i] = x;
    }
    return out;
}

Renderer::Renderer(Context &ctx, Size2i resolution) : ctx_(ctx), res_(resolution)
{
    if (resolution.width <= 0 || resolution.height <= 0) throw Error(DR_E_INVALID, "empty resolution");
    const size_t npix = (size_t)resolution.width * resolution.height;
    check(dr_device_alloc(ctx.handle(), npix * 4 * sizeof(double), &accum_));
    pixels_.assign(npix, 0u);
    display_.assign(npix, 0u);
}

Renderer::~Renderer()
{
    if (accum_) dr_device_free(ctx_.handle(), accum_);
    if (rgba_) dr_device_free(ctx_.handle(), rgba_);
}

void RunningAverage::add(double value)
{
    values_.push_back(value);
    while (values_.size() > window_) values_.pop_front();
}

double RunningAverage::mean() const
{
    if (values_.empty()) return 0.0;
    return std::accumulate(values_.begin(), values_.end(), 0.0) / (double)values_.size();
}

void Renderer::setScene(Scene scene)
{
    std::scoped_lock frame{frame_mutex_};
    std::scoped_lock guard{render_mutex_};
    scene_ = std::move(scene);
    if (!scene_.built()) scene_.build(ctx_);
    if (scene_.handle()) check(dr_scene_build_lights(scene_.handle(), nullptr));
    needs_clear_ = true;
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
        c.pixel_size[0] = cam_.physical_pixel_size.x;
        c.pixel_size[1] = cam_.physical_pixel_size.y;
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

bool Renderer::saveImage(const std::string &path)
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
    FILE *f = std::fopen(path.c_str(), "wb");
    const bool ok = f && std::fwrite(out.data(), 1, out.size(), f) == out.size();
    if (f) std::fclose(f);
    if (!ok) std::fprintf(stderr, "WARN: Failed to save %s\n", path.c_str());
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

Matrix4x4f Matrix4x4f::rotation(const Vec3 &a, float angle)
{
    const float c = std::cos(angle), sn = std::sin(angle), k = 1.0f - c;
    const float ax[3] = {a.x, a.y, a.z};
    // [axis]_x: the cross-product matrix, [axis]_x v = axis x v
    const float cross[3][3] = {{0.0f, -a.z, a.y}, {a.z, 0.0f, -a.x}, {-a.y, a.x, 0.0f}};
    Matrix4x4f m;
    for (int r = 0; r < 3; ++r)
        for (int col = 0; col < 3; ++col)
            m.values[r][col] = (r == col ? c : 0.0f) + sn * cross[r][col] + k * ax[r] * ax[col];
    return m;
}

Matrix4x4f Matrix4x4f::operator*(const Matrix4x4f &o) const
{
    Matrix4x4f m;
    for (int r = 0; r < 4; ++r)
        for (int col = 0; col < 4; ++col) {
            float acc = 0.0f;
            for (int k = 0; k < 4; ++k) acc += values[r][k] * o.values[k][col];
            m.values[r][col] = acc;
        }
    return m;
}

Vec3 Matrix4x4f::applyToPosition(const Vec3 &q) const
{
    const Vec3 d = applyToDirection(q);
    return {d.x + values[0][3], d.y + values[1][3], d.z + values[2][3]};
}

Vec3 Matrix4x4f::applyToDirection(const Vec3 &d) const
{
    return {values[0][0] * d.x + values[0][1] * d.y + values[0][2] * d.z,
            values[1][0] * d.x + values[1][1] * d.y + values[1][2] * d.z,
            values[2][0] * d.x + values[2][1] * d.y + values[2][2] * d.z};
}

}  // namespace devilray
