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

    def material(self, kind, albedo=(0, 0, 0), emission=(0, 0, 0), ior=1.0) -> int:
        self.mats.append((kind, tuple(float(x) for x in albedo), tuple(float(x) for x in emission), float(ior)))
        return len(self.mats) - 1

    def mesh(self, verts, tris, mat):
        verts = np.asarray(verts, np.float32).reshape(-1, 3)
        tris = np.asarray(tris, np.uint32).reshape(-1, 3)
        self.v.append(verts)
        self.f.append(tris + self.nv)
        self.m.append(np.full(tris.shape[0], mat, np.uint32))
        self.nv += verts.shape[0]

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
                         np.concatenate(self.m).astype(np.uint32), self.mats, camera or {})


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
