"""Instanced workloads shared by the oracle and device instancing tests.

Exact cases put every model vertex on a 2^-8 grid (|x| < 4), every scale at
+-2^k per axis with one magnitude per instance, every translation on a 2^-4
grid and every ray origin on a 2^-9 grid: then v * s + t, o - t and
(o - t) / s are all exact in float32, and a ray tested in model space sees
the same triangle as the flattened world copy up to a power-of-two factor, so
t, u and v agree bit for bit. General cases use arbitrary scales and
translations.
"""
from __future__ import annotations

import numpy as np

import devilray_amd as dr

GRID = 256.0  # model vertices on multiples of 1/256


def quantize(v, grid=GRID):
    return (np.round(np.asarray(v, np.float64) * grid) / grid).astype(np.float32)


def cube():
    v = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float32)
    f = np.array([[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5], [0, 4, 5], [0, 5, 1],
                  [2, 3, 7], [2, 7, 6], [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]], np.uint32)
    return v, f


def wavy_grid(n=6, seed=3):
    rng = np.random.default_rng(seed)
    xs = np.linspace(-1.5, 1.5, n + 1)
    v = np.array([[x, y, 0.0] for y in xs for x in xs], np.float64)
    v[:, 2] = rng.uniform(-0.3, 0.3, v.shape[0])
    f = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            f += [[a, a + 1, a + n + 2], [a, a + n + 2, a + n + 1]]
    return quantize(v), np.array(f, np.uint32)


def scenes_small():
    """Four small meshes on the grid: an icosphere (320 triangles), a cube, a
    wavy 6 x 6 grid, one triangle."""
    v, f = dr.gen_icosphere(2, 0.05, 1)
    tri = (quantize([[0, 0, 0], [1.5, 0, 0], [0, 1.25, 0.5]]), np.array([[0, 1, 2]], np.uint32))
    return [(quantize(v * 1.5), f), cube(), wavy_grid(), tri]


def exact_instances(n, n_scenes, seed=0, spread=10.0):
    rng = np.random.default_rng(seed)
    out = np.zeros(n, dr.INSTANCE_DTYPE)
    out["scene"] = rng.integers(0, n_scenes, n)
    mag = 2.0 ** rng.integers(-1, 2, n)
    sign = rng.choice([-1.0, 1.0], (n, 3), p=[0.25, 0.75])
    out["scale"] = (sign * mag[:, None]).astype(np.float32)
    out["translate"] = (np.round(rng.uniform(-spread, spread, (n, 3)) * 16) / 16).astype(np.float32)
    return out


def general_instances(n, n_scenes, seed=0, spread=10.0):
    rng = np.random.default_rng(seed)
    out = np.zeros(n, dr.INSTANCE_DTYPE)
    out["scene"] = rng.integers(0, n_scenes, n)
    sign = rng.choice([-1.0, 1.0], (n, 3), p=[0.2, 0.8])
    out["scale"] = (sign * rng.uniform(0.3, 2.5, (n, 3))).astype(np.float32)
    out["translate"] = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    return out


def flatten(scenes, instances):
    """World copies v * s + t of each instance's scene, in instance order."""
    vs, fs, off = [], [], 0
    for rec in instances:
        v, f = scenes[int(rec["scene"])]
        w = (np.asarray(v, np.float32) * rec["scale"].astype(np.float32)
             + rec["translate"].astype(np.float32)).astype(np.float32)
        vs.append(w)
        fs.append(np.asarray(f, np.uint32) + off)
        off += w.shape[0]
    return np.concatenate(vs), np.concatenate(fs)


def rays_in_box(lo, hi, n, seed=0, grid=512.0, tmax=np.inf):
    """Origins on a 1/grid lattice inside [lo, hi], uniform directions."""
    rng = np.random.default_rng(seed)
    r = np.zeros(n, dr.RAY_DTYPE)
    o = np.round(rng.uniform(lo, hi, (n, 3)) * grid) / grid
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r["ox"], r["oy"], r["oz"] = o.T.astype(np.float32)
    r["dx"], r["dy"], r["dz"] = d.T.astype(np.float32)
    r["tmin"] = 0.0
    r["tmax"] = tmax
    return r


def rays_at_targets(world_v, world_f, lo, hi, n, seed=0, grid=512.0):
    """Rays from lattice origins in [lo, hi] aimed at random triangle
    centroids of the flattened scene, so most of them hit something."""
    rng = np.random.default_rng(seed)
    r = rays_in_box(lo, hi, n, seed, grid)
    tri = world_f[rng.integers(0, world_f.shape[0], n)]
    c = world_v[tri].astype(np.float64).mean(axis=1)
    o = np.stack([r["ox"], r["oy"], r["oz"]], 1).astype(np.float64)
    d = c - o
    d /= np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-30)
    r["dx"], r["dy"], r["dz"] = d.T.astype(np.float32)
    return r


# ---- render workloads --------------------------------------------------------

LAMBERT, DIELECTRIC = dr.DR_MAT_LAMBERT, dr.DR_MAT_DIELECTRIC


def _box(half):
    v, f = cube()
    return quantize(v * half), f


def _quad(center, size):
    c = np.asarray(center, np.float64)
    h = size / 2.0
    v = np.array([c + [-h, 0, -h], c + [h, 0, -h], c + [h, 0, h], c + [-h, 0, h]], np.float64)
    return quantize(v), np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def _sphere(center, radius, subdiv=2):
    v, f = dr.gen_icosphere(subdiv, 0.0, 0)
    v = v.astype(np.float64) / np.linalg.norm(v.astype(np.float64), axis=1, keepdims=True)
    return quantize(v * radius + np.asarray(center, np.float64)), f


def render_scenes():
    """Model scenes on the 1/256 grid, each (vertices, triangles, per-triangle
    material, materials as (kind, albedo, emission, ior)): a grey box (the
    room), an emissive quad, a Lambert sphere off the origin, a glass sphere
    off the origin and a two-material pair of spheres (one half emits)."""
    out = []
    v, f = _box(2.0)
    out.append((v, f, np.zeros(f.shape[0], np.uint32), [(LAMBERT, (0.7, 0.7, 0.7), (0, 0, 0), 1.0)]))
    v, f = _quad((0.5, 1.75, 0.5), 0.5)
    out.append((v, f, np.zeros(2, np.uint32), [(LAMBERT, (0.5, 0.5, 0.5), (6.0, 5.0, 4.0), 1.0)]))
    v, f = _sphere((2.0, -1.0, 0.0), 0.5)
    out.append((v, f, np.zeros(f.shape[0], np.uint32), [(LAMBERT, (0.8, 0.3, 0.2), (0, 0, 0), 1.0)]))
    v, f = _sphere((-2.0, -2.0, 1.0), 0.625)
    out.append((v, f, np.zeros(f.shape[0], np.uint32), [(DIELECTRIC, (1, 1, 1), (0, 0, 0), 1.5)]))
    v, f = _sphere((1.0, 1.0, -2.0), 0.375, subdiv=1)
    m = (np.arange(f.shape[0]) % 2).astype(np.uint32)
    out.append((v, f, m, [(LAMBERT, (0.3, 0.8, 0.3), (0, 0, 0), 1.0),
                          (LAMBERT, (0.6, 0.6, 0.6), (1.5, 2.5, 3.0), 1.0)]))
    return out


def render_instances_exact():
    """Power-of-two scales of one magnitude per instance (signs mirror) and
    no translation: every vertex and every ray origin moves by an exact
    power of two, so the instanced traversal and shading see the flattened
    copy's numbers up to that factor."""
    rows = [(0, 4.0, (1, 1, 1)), (1, 2.0, (1, 1, 1)), (1, 2.0, (-1, 1, 1)), (2, 1.0, (1, 1, 1)),
            (2, 2.0, (1, 1, 1)), (2, 2.0, (-1, 1, -1)), (2, 2.0, (1, -1, 1)), (3, 1.0, (1, 1, 1)),
            (3, 2.0, (1, 1, 1)), (4, 2.0, (1, 1, 1)), (4, 2.0, (-1, -1, 1)), (4, 1.0, (1, 1, 1))]
    out = np.zeros(len(rows), dr.INSTANCE_DTYPE)
    for i, (sc, mag, sign) in enumerate(rows):
        out[i]["scene"] = sc
        out[i]["scale"] = (np.asarray(sign, np.float64) * mag).astype(np.float32)
    return out


def render_instances_general(seed=5):
    """The same placements with arbitrary scales and translations."""
    rng = np.random.default_rng(seed)
    out = render_instances_exact()
    for i in range(1, out.shape[0]):
        out[i]["scale"] = (out[i]["scale"] * rng.uniform(0.8, 1.2, 3)).astype(np.float32)
        out[i]["translate"] = rng.uniform(-0.4, 0.4, 3).astype(np.float32)
    return out


def flatten_render(scenes, instances):
    """The flattened copy of an instanced render set as a SceneData: world
    vertices v * s + t in instance order, each scene's materials once."""
    from devilray_amd.scenes import SceneData
    moff = np.cumsum([0] + [len(s[3]) for s in scenes])
    v, f = flatten([(s[0], s[1]) for s in scenes], instances)
    tm = np.concatenate([scenes[int(r["scene"])][2] + moff[int(r["scene"])] for r in instances]).astype(np.uint32)
    mats = [m for s in scenes for m in s[3]]
    return SceneData(v, f, tm, mats)


def build_render_scenes(ctx, scenes):
    out = []
    for v, f, tm, mats in scenes:
        from devilray_amd.scenes import SceneData
        sd = SceneData(v, f, tm, mats)
        s = dr.Scene(ctx, v, f, tm, sd.dr_materials())
        s.build_bvh(leaf_max=4)
        out.append(s)
    return out
