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
