"""Pin the oracle's light sampler and path tracer.

Known answers: the reference's analytic diffuse-reflection test
(test/src/test_render.cpp:50-165: mean 0.55 +- 0.01, sigma <= 0.04, r = g = b),
its alias-table fidelity test (test/src/test_alias.cpp: implied distribution
matches the weights), closed forms for sin/cos, and invariants of the
estimator (determinism, progressive rendering equals one call).
"""
import numpy as np
import pytest

import devilray_amd as dr
import oracle
from devilray_amd import image, scenes


def test_sincos_polynomial():
    us = np.linspace(0.0, 1.0, 4001, endpoint=False)
    sc = np.array([oracle.sincos_2pi(float(u)) for u in us])
    assert np.max(np.abs(sc[:, 0] - np.sin(2 * np.pi * us))) < 4e-7
    assert np.max(np.abs(sc[:, 1] - np.cos(2 * np.pi * us))) < 4e-7
    assert oracle.sincos_2pi(0.0) == (0.0, 1.0)


def implied_distribution(ent):
    n = len(ent)
    index = {int(t): i for i, t in enumerate(ent["a"])}
    p = np.zeros(n)
    for s in range(n):
        p[index[int(ent["a"][s])]] += ent["p_a"][s] / n
        if ent["b"][s] >= 0:
            p[index[int(ent["b"][s])]] += (1.0 - ent["p_a"][s]) / n
    return p


def weighted_triangles(w):
    """Right triangles of area w_i (one emitter material), stacked in z."""
    n = len(w)
    a = np.sqrt(2.0 * w)
    v = np.zeros((3 * n, 3), np.float32)
    v[1::3, 0] = a
    v[2::3, 1] = a
    v[:, 2] = np.repeat(np.arange(n, dtype=np.float32), 3)
    f = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    return v, f


@pytest.mark.parametrize("n,sigma", [(1, 0.0), (7, 1.0), (1000, 1.0), (20000, 2.0)])
def test_alias_table_implied_distribution(n, sigma):
    rng = np.random.default_rng(42)
    w = rng.lognormal(0.0, sigma, n)
    v, f = weighted_triangles(w)
    mats = oracle.materials_array([(0, (0, 0, 0), (1, 1, 1), 1.0)])
    ent, pdf = oracle.lights_build(v, f, np.zeros(n, np.int32), mats)
    assert len(ent) == n
    # weights as the builder sees them (areas of the float triangles, FP64)
    vv = v.astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(vv[1::3] - vv[0::3], vv[2::3] - vv[0::3]), axis=1)
    target = area / area.sum()
    implied = implied_distribution(ent)
    assert np.max(np.abs(implied - target)) < 1e-6 + 1e-4 / n
    assert np.allclose(ent["pdf_a"], target, rtol=1e-5, atol=1e-9)
    assert np.all(ent["p_a"] <= 1.0) and np.all(ent["p_a"] >= 0.0)
    assert np.allclose(pdf * area, target, rtol=1e-5)


# test/src/test_alias.cpp:64-229: single element, uniform, simple biased,
# alternating high/low and the 200,000-entry log-normal stress case (seed 42);
# the empty and zero-sum cases are scenes without emitters, tested below
@pytest.mark.parametrize("case", ["single", "uniform", "biased", "alternating", "stress"])
def test_alias_table_reference_cases(case):
    if case == "single":
        w = np.float64([3.0])
    elif case == "uniform":
        w = np.ones(10)
    elif case == "biased":
        w = np.float64([1.0, 3.0])
    elif case == "alternating":
        w = np.tile([100.0, 1.0], 50)
    else:
        w = np.random.default_rng(42).lognormal(0.0, 1.0, 200_000)
    v, f = weighted_triangles(w)
    mats = oracle.materials_array([(0, (0, 0, 0), (1, 1, 1), 1.0)])
    ent, _ = oracle.lights_build(v, f, np.zeros(len(w), np.int32), mats)
    vv = v.astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(vv[1::3] - vv[0::3], vv[2::3] - vv[0::3]), axis=1)
    implied = implied_distribution(ent)
    tol = 1e-3 if case == "stress" else 1e-4     # the reference's bounds
    assert len(ent) == len(w) and np.max(np.abs(implied - area / area.sum())) < tol / len(w) + 1e-7
    if case == "single":
        assert ent["p_a"][0] == 1.0
    if case == "biased":
        assert np.allclose(implied, [0.25, 0.75], atol=1e-6)


def test_lights_skip_non_emitters_and_dielectrics():
    sc = scenes.cornell_c0()
    ent, pdf = oracle.lights_build(sc.vertices, sc.indices, sc.tri_material, oracle.materials_array(sc.materials))
    light_tris = np.nonzero(sc.tri_material == 3)[0]
    assert sorted(ent["a"]) == list(light_tris)
    assert np.all(pdf[light_tris] > 0) and np.count_nonzero(pdf) == len(light_tris)
    # two triangles of a 0.4 x 0.4 quad: uniform over the area
    assert np.allclose(pdf[light_tris], 1.0 / 0.16, rtol=1e-6)


def furnace_render(spp, size=48):
    fs = scenes.furnace_planes(0.55)
    bvh = oracle.BVH(fs.vertices, fs.indices)
    cam = oracle.Camera.from_buffer_copy(bytes(dr.camera_reference(size, size, 50.0, 0.1)))
    acc, _ = oracle.render(bvh, fs.tri_material, oracle.materials_array(fs.materials), cam, spp=spp,
                           max_depth=10, rr_depth=10)
    return acc


def test_oracle_analytic_furnace():
    """White plane lit by a facing emitter plane of radiance 0.55 reflects 0.55."""
    acc = furnace_render(64)
    img = acc[..., :3] / acc[..., 3:]
    assert np.all(acc[..., 3] == 64)
    assert np.array_equal(img[..., 0], img[..., 1]) and np.array_equal(img[..., 1], img[..., 2])
    assert img.mean() == pytest.approx(0.55, abs=0.01)
    assert img[..., 0].std() <= 0.04


def cornell(spp, size=32, sample_offset=0, accum=None, seed=0, hist=False):
    sc = scenes.cornell_c0()
    bvh = oracle.BVH(sc.vertices, sc.indices)
    c = sc.camera
    cam = oracle.Camera.from_buffer_copy(bytes(dr.camera_lookat(c["eye"], c["target"], c["up"], c["fov"], size, size)))
    return oracle.render(bvh, sc.tri_material, oracle.materials_array(sc.materials), cam, spp=spp, seed=seed,
                         sample_offset=sample_offset, accum=accum, vertex_hist=hist)


def test_oracle_progressive_equals_single_call():
    a, _ = cornell(4)
    b, _ = cornell(2)
    b, _ = cornell(2, sample_offset=2, accum=b)
    assert np.array_equal(a, b)
    c, _ = cornell(4, seed=1)
    assert not np.array_equal(a, c)


def test_oracle_roulette_and_depth():
    _, hist = cornell(8, hist=True)
    assert hist[0] == 32 * 32 * 8 and np.all(np.diff(hist) <= 0) and hist[-1] > 0


def test_create_pixels_and_png_roundtrip(tmp_path):
    acc = np.zeros((2, 3, 4))
    acc[..., 3] = 2
    acc[0, 0, :3] = [2.0, 1.0, 0.0]       # mean (1, .5, 0)
    acc[1, 2, :3] = [0.5, 0.5, 0.5]       # mean .25
    px = image.create_pixels(acc, gamma_corrected=False)
    rgba = image.unpack_rgba(px)
    assert tuple(rgba[1, 0]) == (255, 128, 0, 255)    # row 0 is flipped to the bottom
    assert tuple(rgba[0, 2]) == (64, 64, 64, 255)
    srgb = image.unpack_rgba(image.create_pixels(acc, gamma_corrected=True))
    assert srgb[0, 2, 0] == round((1.055 * 0.25 ** (1 / 2.4) - 0.055) * 255)
    p = str(tmp_path / "x.png")
    image.write_png(p, rgba)
    assert np.array_equal(image.read_png(p), rgba)


def test_umbrella_fixture_decodes():
    img = image.read_png("tests/golden/triangle_umbrella.png")
    assert img.shape == (256, 256, 4) and np.all(img[..., 3] == 255)
    assert img[..., :3].mean() > 5
