"""C-ABI library checks that need no GPU: the library loads, every symbol
declared in include/*.h is exported and bound, host-side generators are
deterministic and well-formed, and misuse is reported without a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import devilray_amd as dr
from devilray_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(dr_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported_and_bound():
    names = declared_symbols()
    assert "dr_trace_closest" in names and "dr_scene_build_bvh" in names
    lib = dr.lib()
    for n in sorted(names):
        assert hasattr(lib, n), f"{n} not exported"
        assert n in _lib.PROTOTYPES, f"{n} not bound in devilray_amd._lib"


def test_status_strings_and_errors_without_device():
    lib = dr.lib()
    assert lib.dr_status_string(0) == b"ok"
    assert lib.dr_status_string(1) == b"invalid argument"
    assert lib.dr_scene_build_bvh(None, None, None) == _lib.DR_E_INVALID
    assert b"null" in lib.dr_last_error()
    assert lib.dr_trace_closest(None, None, 0, None, None) == _lib.DR_E_INVALID


def test_icosphere_counts_winding_and_determinism():
    for s in range(0, 5):
        v, f = dr.gen_icosphere(s, 0.0, 0)
        assert v.shape == (10 * 4 ** s + 2, 3) and f.shape == (20 * 4 ** s, 3)
        a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
        n = np.cross(b - a, c - a)
        assert np.all((n * (a + b + c)).sum(1) > 0), "faces wind outward"
        assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-6)
        # closed, edge-manifold: every undirected edge is used exactly twice
        e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        assert np.all(cnt == 2)
    v1, _ = dr.gen_icosphere(5, 0.05, 0)
    v2, _ = dr.gen_icosphere(5, 0.05, 0)
    v3, _ = dr.gen_icosphere(5, 0.05, 1)
    assert np.array_equal(v1, v2) and not np.array_equal(v1, v3)
    r = np.linalg.norm(v1.astype(np.float64), axis=1)
    assert abs(r.mean() - 1.0) < 0.005 and abs(r.std() - 0.05) < 0.005


def test_surface_rays_properties():
    v, f = dr.gen_icosphere(4, 0.0, 0)
    rays = dr.gen_rays_surface(v, f, 20000, seed=0, inward=True, offset=1e-5)
    o = np.stack([rays["ox"], rays["oy"], rays["oz"]], 1).astype(np.float64)
    d = np.stack([rays["dx"], rays["dy"], rays["dz"]], 1).astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert np.all(np.linalg.norm(o, axis=1) < 1.0 + 1e-6)
    # cosine-distributed about the inward normal: E[cos] = 2/3 (test_sampling.cpp:18-102)
    cos = -(d * o / np.linalg.norm(o, axis=1, keepdims=True)).sum(1)
    assert np.all(cos > -1e-3)
    assert cos.mean() == pytest.approx(2 / 3, abs=0.01)
    assert np.all(rays["tmin"] == 0) and np.all(np.isinf(rays["tmax"]))
    again = dr.gen_rays_surface(v, f, 20000, seed=0, inward=True, offset=1e-5)
    assert np.array_equal(rays.view(np.uint8), again.view(np.uint8))


def test_camera_rays():
    rays = dr.gen_rays_camera((0, 0, 3.5), (0, 0, 0), (0, 1, 0), 40.0, 64, 48)
    d = np.stack([rays["dx"], rays["dy"], rays["dz"]], 1).reshape(48, 64, 3)
    assert np.allclose(np.linalg.norm(d, axis=2), 1.0, atol=1e-6)
    assert np.all(d[..., 2] < 0)
    # top row looks up, left column looks left; vertical FOV 40 degrees
    assert d[0, 32, 1] > 0 and d[24, 0, 0] < 0
    half = np.degrees(np.arctan2(d[0, 32, 1], -d[0, 32, 2]))
    assert half == pytest.approx(20.0 * (1 - 1 / 48), abs=0.5)
