"""Deterministic scene builders for the render workloads.

Each returns a SceneData: world-space vertices, triangle indices, a material
index per triangle and the material list (dr_material tuples
(kind, albedo, emission, ior)). Geometry is baked to world space (the
reference's per-object scale/translate transforms are applied here).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import DR_MAT_DIELECTRIC, DR_MAT_LAMBERT, DrMaterial, gen_icosphere


@dataclass
class SceneData:
    vertices: np.ndarray
    indices: np.ndarray
    tri_material: np.ndarray
    materials: list = field(default_factory=list)
    camera: dict = field(default_factory=dict)
    # the reference's objects (TriangleMesh: one material, own BBH) as
    # contiguous triangle ranges (first, count); None = one object
    objects: np.ndarray | None = None

    def object_ranges(self) -> np.ndarray:
        if self.objects is None:
            return np.array([[0, self.indices.shape[0]]], np.int32)
        return np.asarray(self.objects, np.int32).reshape(-1, 2)

    def dr_materials(self):
        out = []
        for kind, alb, em, ior in self.materials:
            m = DrMaterial()
            m.kind = kind
            m.albedo[:] = list(alb)
            m.emission[:] = list(em)
            m.ior = ior
            out.append(m)
        return out


class _Builder:
    def __init__(self):
        self.v, self.f, self.m, self.mats = [], [], [], []
        self.nv = 0
        self.nt = 0
        self.objects = []  # (first triangle, count) per mesh

    def material(self, kind, albedo=(0, 0, 0), emission=(0, 0, 0), ior=1.0) -> int:
        self.mats.append((kind, tuple(float(x) for x in albedo), tuple(float(x) for x in emission), float(ior)))
        return len(self.mats) - 1

    def mesh(self, verts, tris, mat):
        verts = np.asarray(verts, np.float32).reshape(-1, 3)
        tris = np.asarray(tris, np.uint32).reshape(-1, 3)
        self.v.append(verts)
        self.f.append(tris + self.nv)
        self.m.append(np.full(tris.shape[0], mat, np.uint32))
        self.objects.append((self.nt, tris.shape[0]))
        self.nv += verts.shape[0]
        self.nt += tris.shape[0]

    def quad(self, center, normal, right, size_u, size_v, mat):
        """Quad like createQuadMesh (src/GpuTris.cpp:87-107): up = right x normal,
        triangles (0,2,1), (0,3,2)."""
        c, n, r = (np.asarray(x, np.float64) for x in (center, normal, right))
        up = np.cross(r, n)
        hu, hv = size_u / 2.0, size_v / 2.0
        pts = [c - r * hu - up * hv, c + r * hu - up * hv, c + r * hu + up * hv, c - r * hu + up * hv]
        self.mesh(pts, [[0, 2, 1], [0, 3, 2]], mat)

    def sphere(self, subdiv, center, radius, mat, sigma=0.0, seed=0):
        v, f = gen_icosphere(subdiv, sigma, seed)
        self.mesh(v.astype(np.float64) * radius + np.asarray(center, np.float64), f, mat)

    def build(self, camera=None) -> SceneData:
        return SceneData(np.concatenate(self.v).astype(np.float32), np.concatenate(self.f).astype(np.uint32),
                         np.concatenate(self.m).astype(np.uint32), self.mats, camera or {},
                         np.array(self.objects, np.int32).reshape(-1, 2))


def cornell_c0() -> SceneData:
    """Cornell-style box [-1,1]^3 open at +z: floor, ceiling, back (white 0.75),
    left red / right green (0.65); an emissive quad of radiance 15 covering 4%
    of the ceiling just below it; a Lambert (0.75) and a dielectric (IOR 1.5)
    icosphere at subdivision 3. 2,572 triangles. Camera at (0,0,3.4) looking at
    the origin, 40 degree vertical FOV."""
    b = _Builder()
    white = b.material(DR_MAT_LAMBERT, (0.75, 0.75, 0.75))
    red = b.material(DR_MAT_LAMBERT, (0.65, 0.05, 0.05))
    green = b.material(DR_MAT_LAMBERT, (0.05, 0.65, 0.05))
    light = b.material(DR_MAT_LAMBERT, (0.0, 0.0, 0.0), (15.0, 15.0, 15.0))
    glass = b.material(DR_MAT_DIELECTRIC, (1, 1, 1), (0, 0, 0), 1.5)
    b.quad((0, -1, 0), (0, 1, 0), (1, 0, 0), 2, 2, white)     # floor
    b.quad((0, 1, 0), (0, -1, 0), (1, 0, 0), 2, 2, white)     # ceiling
    b.quad((0, 0, -1), (0, 0, 1), (1, 0, 0), 2, 2, white)     # back
    b.quad((-1, 0, 0), (1, 0, 0), (0, 0, 1), 2, 2, red)       # left
    b.quad((1, 0, 0), (-1, 0, 0), (0, 0, 1), 2, 2, green)     # right
    b.quad((0, 0.999, 0), (0, -1, 0), (1, 0, 0), 0.4, 0.4, light)
    b.sphere(3, (-0.42, -0.65, -0.3), 0.35, white)
    b.sphere(3, (0.45, -0.65, 0.25), 0.35, glass)
    return b.build(dict(eye=(0.0, 0.0, 3.4), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=40.0,
                        width=256, height=256))


def furnace_enclosure(albedo: float = 0.5, emission: float = 1.0, glass_ior: float | None = None,
                      glass_subdiv: int = 3) -> SceneData:
    """A closed box [-1,1]^3 whose six walls are all Lambert with albedo a and
    emission Le, optionally with a dielectric icosphere (radius 0.45) at the
    centre; the camera is inside. Every point of the box then sees radiance
    Le (1 + a + a^2 + ...) = Le / (1 - a) in every direction, and a lossless
    dielectric in it changes nothing (an isothermal enclosure): a path of at
    most D Lambert vertices without roulette gives Le (1 - a^D) / (1 - a)."""
    b = _Builder()
    wall = b.material(DR_MAT_LAMBERT, (albedo,) * 3, (emission,) * 3)
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            c = [0.0, 0.0, 0.0]
            c[axis] = sgn
            n = [0.0, 0.0, 0.0]
            n[axis] = -sgn
            r = [0.0, 0.0, 0.0]
            r[(axis + 1) % 3] = 1.0
            b.quad(c, n, r, 2, 2, wall)
    if glass_ior is not None:
        glass = b.material(DR_MAT_DIELECTRIC, (1, 1, 1), (0, 0, 0), glass_ior)
        b.sphere(glass_subdiv, (0.0, 0.0, 0.0), 0.45, glass)
    return b.build(dict(eye=(0.35, 0.25, 0.9), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=60.0,
                        width=32, height=32))


def _random_rotations(rng, n):
    """Uniform random rotation matrices (Shoemake's quaternion method)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], 1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], 1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1)], axis=1)


def many_lights(n_spheres: int = 400, subdiv: int = 4, n_emissive: int = 100_000, seed: int = 0,
                extent: float = 20.0, partition: bool = False) -> SceneData:
    """The many-lights scene: n_spheres displaced icospheres (radius
    1 + 0.05 N(0,1) per vertex; 20*4^subdiv triangles each, 2,048,000 in
    total at the defaults) with random positions in [-extent, extent]^3,
    log-uniform scales in [0.5, 2.5] and uniform random rotations. 30% are
    dielectric with IOR ~ U[1.3, 1.7], the rest Lambert with albedo ~
    U[0.2, 0.9]. n_emissive triangles of Lambert spheres emit radiance
    LogNormal(0, 2) (grey). With partition=True a quarter of the spheres move
    behind an opaque partition at z = -extent - 4, all emitters are on those,
    and a closed room encloses the scene with a gap between the partition and
    the ceiling; the camera looks down at (0, -0.6 extent, 0), so that what it
    sees is lit indirectly, off the ceiling (the guided-sampling scene).
    Camera: 1920x1080, 45 degree FOV, looking at the origin from +z."""
    rng = np.random.default_rng(seed)
    base_v, base_f = gen_icosphere(subdiv, 0.0, 0)
    unit = base_v.astype(np.float64) / np.linalg.norm(base_v.astype(np.float64), axis=1, keepdims=True)
    nv, nt = unit.shape[0], base_f.shape[0]
    rot = _random_rotations(rng, n_spheres)
    pos = rng.uniform(-extent, extent, size=(n_spheres, 3))
    scale = np.exp(rng.uniform(np.log(0.5), np.log(2.5), n_spheres))
    dielectric = rng.random(n_spheres) < 0.3
    ior = rng.uniform(1.3, 1.7, n_spheres)
    albedo = rng.uniform(0.2, 0.9, n_spheres)
    if partition:
        # emitters live in a slab behind the wall
        pos[:, 2] = np.where(np.arange(n_spheres) % 4 == 0, rng.uniform(-extent - 14, -extent - 8, n_spheres),
                             pos[:, 2])
    verts = np.empty((n_spheres * nv, 3), np.float32)
    faces = np.empty((n_spheres * nt, 3), np.uint32)
    tri_mat = np.empty(n_spheres * nt, np.uint32)
    mats = []
    for s in range(n_spheres):
        r = 1.0 + 0.05 * rng.standard_normal(nv)
        v = (unit * r[:, None]) @ rot[s].T * scale[s] + pos[s]
        verts[s * nv:(s + 1) * nv] = v
        faces[s * nt:(s + 1) * nt] = base_f + s * nv
        if dielectric[s]:
            mats.append((DR_MAT_DIELECTRIC, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), float(ior[s])))
        else:
            a = float(albedo[s])
            mats.append((DR_MAT_LAMBERT, (a, a, a), (0.0, 0.0, 0.0), 1.0))
        tri_mat[s * nt:(s + 1) * nt] = s
    # emissive triangles: drawn from Lambert spheres (behind the wall when partitioned)
    eligible = np.flatnonzero(~dielectric[tri_mat])
    if partition:
        behind = pos[tri_mat[eligible], 2] < -extent - 4
        eligible = eligible[behind]
    chosen = np.sort(rng.choice(eligible, size=min(n_emissive, eligible.size), replace=False))
    radiance = rng.lognormal(0.0, 2.0, chosen.size)
    first = len(mats)
    for k, t in enumerate(chosen):
        a = mats[tri_mat[t]][1]
        mats.append((DR_MAT_LAMBERT, a, (float(radiance[k]),) * 3, 1.0))
        tri_mat[t] = first + k
    b_v, b_f, b_m = [verts], [faces], [tri_mat]
    if partition:
        # closed grey room around everything (the camera included) and an
        # opaque two-sided partition at z = -extent - 4 that leaves a gap of
        # `gap` units under the ceiling: light crossing the gap rises onto the
        # ceiling, which the camera (tilted down) does not see, so everything
        # in view is lit indirectly
        room = _Builder()
        room.nv = verts.shape[0]
        wm = len(mats)
        mats.append((DR_MAT_LAMBERT, (0.7, 0.7, 0.7), (0.0, 0.0, 0.0), 1.0))
        R = extent + 12.0
        z_lo, z_hi, z_p, gap = -extent - 18.0, 3.2 * extent + 6.0, -extent - 4.0, 6.0
        zc, lz = (z_lo + z_hi) / 2, z_hi - z_lo
        room.quad((0, -R, zc), (0, 1, 0), (1, 0, 0), 2 * R, lz, wm)
        room.quad((0, R, zc), (0, -1, 0), (1, 0, 0), 2 * R, lz, wm)
        room.quad((-R, 0, zc), (1, 0, 0), (0, 0, 1), lz, 2 * R, wm)
        room.quad((R, 0, zc), (-1, 0, 0), (0, 0, 1), lz, 2 * R, wm)
        room.quad((0, 0, z_lo), (0, 0, 1), (1, 0, 0), 2 * R, 2 * R, wm)
        room.quad((0, 0, z_hi), (0, 0, -1), (1, 0, 0), 2 * R, 2 * R, wm)
        room.quad((0, -gap / 2, z_p), (0, 0, 1), (1, 0, 0), 2 * R, 2 * R - gap, wm)
        room.quad((0, -gap / 2, z_p - 0.1), (0, 0, -1), (1, 0, 0), 2 * R, 2 * R - gap, wm)
        for wv, wf, wmat in zip(room.v, room.f, room.m):
            b_v.append(wv)
            b_f.append(wf)
            b_m.append(wmat)
    cam = dict(eye=(0.0, 0.0, 3.2 * extent), target=(0.0, -0.6 * extent if partition else 0.0, 0.0),
               up=(0.0, 1.0, 0.0), fov=45.0, width=1920, height=1080)
    # objects: one per sphere, then one per room quad
    objects = [(k * nt, nt) for k in range(n_spheres)]
    if partition:
        objects += [(n_spheres * nt + first, count) for first, count in room.objects]
    return SceneData(np.concatenate(b_v).astype(np.float32), np.concatenate(b_f).astype(np.uint32),
                     np.concatenate(b_m).astype(np.uint32), mats, cam, np.array(objects, np.int32).reshape(-1, 2))


def hidden_light_reflector() -> SceneData:
    """Guiding test scene: a grey floor (20 x 20, y = 0) lit only by a 3 x 3
    white reflector facing down at y = 8. The emitter (1 x 1, radiance 50,
    y = 7.5) sits under the reflector above a black blocker (1.6 x 1.6,
    y = 7.4), so no floor point sees it and next-event estimation from the
    floor always fails: the floor's light arrives from the reflector's small
    solid angle. Camera above the floor looking down at it."""
    b = _Builder()
    floor = b.material(DR_MAT_LAMBERT, (0.5, 0.5, 0.5))
    refl = b.material(DR_MAT_LAMBERT, (0.9, 0.9, 0.9))
    black = b.material(DR_MAT_LAMBERT, (0.0, 0.0, 0.0))
    light = b.material(DR_MAT_LAMBERT, (0.0, 0.0, 0.0), (50.0, 50.0, 50.0))
    b.quad((0, 0, 0), (0, 1, 0), (1, 0, 0), 20.0, 20.0, floor)
    b.quad((0, 8.0, 0), (0, -1, 0), (1, 0, 0), 3.0, 3.0, refl)
    b.quad((0, 7.5, 0), (0, 1, 0), (1, 0, 0), 1.0, 1.0, light)
    b.quad((0, 7.4, 0), (0, -1, 0), (1, 0, 0), 1.6, 1.6, black)
    cam = dict(eye=(0.0, 14.0, 16.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=50.0, width=256, height=256)
    return b.build(cam)


def furnace_planes(emittance: float = 0.55) -> SceneData:
    """The reference's analytic diffuse-reflection scene (test/src/test_render.cpp
    :50-165): a white (albedo 1) receiver quad at z = 1 facing the camera and an
    emissive quad at z = -1, both 1000 wide; seen from the origin along +z."""
    b = _Builder()
    white = b.material(DR_MAT_LAMBERT, (1.0, 1.0, 1.0))
    light = b.material(DR_MAT_LAMBERT, (0, 0, 0), (emittance,) * 3)
    b.quad((0, 0, 1.0), (0, 0, -1), (1, 0, 0), 1000.0, 1000.0, white)
    b.quad((0, 0, -1.0), (0, 0, 1), (1, 0, 0), 1000.0, 1000.0, light)
    return b.build()


def triangle_umbrella() -> SceneData:
    """The reference's 'triangle umbrella' test scene (test/src/test_render.cpp
    :239-336): one warm emissive triangle behind the camera and a 6x5
    half-sphere of single-triangle objects in red / green / blue."""
    b = _Builder()
    emissive = b.material(DR_MAT_LAMBERT, (0, 0, 0), (0.9, 0.8, 0.6))
    red = b.material(DR_MAT_LAMBERT, (0.9, 0.15, 0.15))
    green = b.material(DR_MAT_LAMBERT, (0.15, 0.9, 0.15))
    blue = b.material(DR_MAT_LAMBERT, (0.15, 0.15, 0.9))
    b.mesh([(-50.0, -50.0, -1.0), (0.0, 0.0, -1.0), (50.0, -50.0, -1.0)], [[0, 1, 2]], emissive)
    n_lon, n_lat, radius = 6, 5, np.float32(0.5)
    center = np.array([0.0, 0.0, 1.2], np.float32)
    pi = np.float32(np.pi)
    colors = [red, green, blue]

    def pt(theta, phi):
        th, ph = np.float32(theta), np.float32(phi)
        return (center[0] + radius * np.sin(th) * np.cos(ph),
                center[1] + radius * np.sin(th) * np.sin(ph),
                center[2] - radius * np.cos(th))

    for lat in range(n_lat):
        t0 = np.float32(lat) / np.float32(n_lat) * (pi / np.float32(2.0))
        t1 = np.float32(lat + 1) / np.float32(n_lat) * (pi / np.float32(2.0))
        for lon in range(n_lon):
            p0 = np.float32(lon) / np.float32(n_lon) * np.float32(2.0) * pi
            p1 = np.float32(lon + 1) / np.float32(n_lon) * np.float32(2.0) * pi
            mat = colors[lon % 3]
            if lat == 0:
                b.mesh([pt(t0, p0), pt(t1, p1), pt(t1, p0)], [[0, 1, 2]], mat)
            else:
                p00, p10, p01, p11 = pt(t0, p0), pt(t1, p0), pt(t0, p1), pt(t1, p1)
                b.mesh([p00, p01, p10], [[0, 1, 2]], mat)
                b.mesh([p01, p11, p10], [[0, 1, 2]], mat)
    return b.build()


def batch_c3(n_scenes: int = 64, seed: int = 0, tris_range=(1_000, 2_000_000), pixels_range=(128 ** 2, 1024 ** 2),
             spp_choices=(4, 16, 64)):
    """The batched workload: n_scenes scenes built like many_lights (displaced
    icospheres, 30% dielectric, 5% of the triangles emissive with LogNormal(0,
    2) radiance) with triangle counts log-uniform in tris_range, 16:9 images
    with pixel counts log-uniform in pixels_range and spp uniform over
    spp_choices. Returns [(SceneData, width, height, spp)]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_scenes):
        t = float(np.exp(rng.uniform(np.log(tris_range[0]), np.log(tris_range[1]))))
        subdiv = 4
        while subdiv > 1 and 20 * 4 ** subdiv * 8 > t:
            subdiv -= 1
        n_spheres = max(1, int(round(t / (20 * 4 ** subdiv))))
        n_tris = n_spheres * 20 * 4 ** subdiv
        extent = 20.0 * (n_spheres / 400.0) ** (1.0 / 3.0) + 2.0
        sd = many_lights(n_spheres=n_spheres, subdiv=subdiv, n_emissive=max(1, int(round(0.05 * n_tris))),
                         seed=seed * 1000 + k, extent=extent)
        px = float(np.exp(rng.uniform(np.log(pixels_range[0]), np.log(pixels_range[1]))))
        w = max(16, int(round(np.sqrt(px * 16.0 / 9.0))))
        h = max(9, int(round(w * 9.0 / 16.0)))
        spp = int(rng.choice(spp_choices))
        out.append((sd, w, h, spp))
    return out


def instanced_spheres(n_spheres: int = 400, n_lights: int = 40, subdiv: int = 5, seed: int = 0,
                      extent: float = 20.0):
    """An instanced workload: one Lambert and one glass icosphere mesh
    (subdivision `subdiv`, 20 * 4^subdiv triangles each) placed n_spheres
    times (30% glass) with random positions in [-extent, extent]^3 and
    log-uniform scales in [0.5, 2.5], n_lights placements of a small emissive
    icosphere, and a grey room around them. Returns (meshes, instances,
    camera): meshes as SceneData (model space, one material list each),
    instances as INSTANCE_DTYPE; the camera as for many_lights."""
    from . import make_instances
    rng = np.random.default_rng(seed)
    meshes = []
    for kind, alb, em, ior, sd in ((DR_MAT_LAMBERT, (0.6, 0.6, 0.6), (0, 0, 0), 1.0, subdiv),
                                   (DR_MAT_DIELECTRIC, (1, 1, 1), (0, 0, 0), 1.5, subdiv),
                                   (DR_MAT_LAMBERT, (0.5, 0.5, 0.5), (8.0, 7.0, 6.0), 1.0, 2)):
        b = _Builder()
        b.sphere(sd, (0.0, 0.0, 0.0), 1.0, b.material(kind, alb, em, ior))
        meshes.append(b.build())
    room = _Builder()
    grey = room.material(DR_MAT_LAMBERT, (0.7, 0.7, 0.7))
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            c = [0.0, 0.0, 0.0]
            c[axis] = sgn
            n = [0.0, 0.0, 0.0]
            n[axis] = -sgn
            r = [0.0, 0.0, 0.0]
            r[(axis + 1) % 3] = 1.0
            room.quad(c, n, r, 2, 2, grey)
    meshes.append(room.build())
    kind = np.where(rng.random(n_spheres) < 0.3, 1, 0)
    pos = rng.uniform(-extent, extent, (n_spheres, 3))
    scale = np.exp(rng.uniform(np.log(0.5), np.log(2.5), n_spheres))
    lpos = rng.uniform(-extent, extent, (n_lights, 3))
    lscale = rng.uniform(0.3, 0.6, n_lights)
    R = 3.2 * extent + 8.0
    inst = np.concatenate([
        make_instances(kind, np.repeat(scale[:, None], 3, 1), pos),
        make_instances(np.full(n_lights, 2), np.repeat(lscale[:, None], 3, 1), lpos),
        make_instances([3], (R, R, R), (0.0, 0.0, 0.0))])
    cam = dict(eye=(0.0, 0.0, 3.2 * extent), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=45.0,
               width=1920, height=1080)
    return meshes, inst, cam


def flatten_instances(meshes, instances) -> SceneData:
    """The world copy of an instanced set as one scene: vertices v * s + t in
    instance order (float32, one rounding per operation), materials of every
    mesh concatenated."""
    vs, fs, ms, off = [], [], [], 0
    moff = np.cumsum([0] + [len(m.materials) for m in meshes])
    for rec in instances:
        k = int(rec["scene"])
        m = meshes[k]
        w = (m.vertices.astype(np.float32) * rec["scale"].astype(np.float32)
             + rec["translate"].astype(np.float32)).astype(np.float32)
        vs.append(w)
        fs.append(m.indices.astype(np.uint32) + off)
        ms.append(m.tri_material.astype(np.uint32) + moff[k])
        off += w.shape[0]
    mats = [x for m in meshes for x in m.materials]
    return SceneData(np.concatenate(vs), np.concatenate(fs), np.concatenate(ms), mats)
