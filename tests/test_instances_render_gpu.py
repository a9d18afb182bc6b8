"""Rendering an instance set (dr_instances_render) on the device.

Workload: five small model scenes on a 1/256 grid (a grey room, an emissive
quad, a Lambert sphere, a glass sphere, a sphere with emitting and
non-emitting halves) placed by twelve instances. The bar:
- exact transforms (power-of-two scales, mirrors, no translation): the
  instanced render equals dr_render of the flattened copy bit for bit, its
  light table is the flattened scene's with emitters named by index, and it
  matches the oracle's render of the flattened copy (>= 99% of pixels
  bit-identical, means within 3 standard errors);
- arbitrary scales and translations: the instanced image agrees with the
  flattened copy's within rounding (>= 95% of pixels within 1e-4 relative,
  means within 1e-3), since the hits come from model-space ray tests;
- the render's cast count equals the flattened scene's.
"""
import numpy as np
import pytest

import devilray_amd as dr
import oracle
import instance_cases as ic

pytestmark = pytest.mark.gpu

W, H = 96, 72


def _camera(w=W, h=H):
    return dr.camera_lookat((0.25, 0.5, 7.5), (0.0, -0.5, 0.0), (0.0, 1.0, 0.0), 60.0, w, h)


@pytest.fixture(scope="module")
def rscenes(gpu_ctx):
    sc = ic.render_scenes()
    built = ic.build_render_scenes(gpu_ctx, sc)
    yield sc, built
    for s in built:
        s.close()


def _flat_scene(ctx, sd):
    s = dr.Scene(ctx, sd.vertices, sd.indices, sd.tri_material, sd.dr_materials())
    s.build_bvh(leaf_max=4)
    return s


@pytest.mark.parametrize("seed,fresnel", [(0, 0), (7, 1)])  # 1: Schlick
def test_exact_instances_render_equals_flattened(gpu_ctx, rscenes, seed, fresnel):
    sc, built = rscenes
    inst = ic.render_instances_exact()
    print(f"[case] instanced render, exact transforms, seed {seed} fresnel {fresnel}", flush=True)
    I = dr.Instances(gpu_ctx, built, inst)
    sd = ic.flatten_render(sc, inst)
    flat = _flat_scene(gpu_ctx, sd)
    cam = _camera()
    got = I.render(cam, spp=8, max_depth=8, rr_depth=3, seed=seed, fresnel=fresnel)
    ref = flat.render(cam, spp=8, max_depth=8, rr_depth=3, seed=seed, fresnel=fresnel)
    lit = ref[..., :3].sum()
    print(f"  image sum {lit:.4f}, equal pixels "
          f"{np.mean(np.all(got.view(np.uint64) == ref.view(np.uint64), axis=2)):.5f}", flush=True)
    assert lit > 0
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    # the light tables agree: the instance set names emitters by index
    ent, tri, pdf = I.lights()
    fent, fpdf = flat.lights()
    assert ent.shape[0] == fent.shape[0] > 0
    assert np.all(np.diff(tri.astype(np.int64)) > 0)
    for k in ("p_a", "pdf_a", "pdf_b"):
        assert np.array_equal(ent[k], fent[k])
    assert np.array_equal(tri[ent["a"]], fent["a"])
    hb = ent["b"] >= 0
    assert np.array_equal(hb, fent["b"] >= 0)
    assert np.array_equal(tri[ent["b"][hb]], fent["b"][hb])
    assert np.array_equal(pdf, fpdf[tri])
    flat.close()
    I.close()


def test_exact_instances_render_matches_oracle(gpu_ctx, rscenes):
    sc, built = rscenes
    inst = ic.render_instances_exact()
    print("[case] instanced render vs oracle on the flattened copy", flush=True)
    I = dr.Instances(gpu_ctx, built, inst)
    sd = ic.flatten_render(sc, inst)
    cam = _camera()
    got = I.render(cam, spp=16, max_depth=8, rr_depth=3, seed=3)
    ref, _ = oracle.render(oracle.BVH(sd.vertices, sd.indices), sd.tri_material, oracle.materials_array(sd.materials),
                           oracle.Camera.from_buffer_copy(bytes(cam)), spp=16, max_depth=8, rr_depth=3, seed=3)
    assert np.array_equal(got[..., 3], ref[..., 3])
    same = np.all(got[..., :3].view(np.uint64) == ref[..., :3].view(np.uint64), axis=2)
    g, r = got[..., :3] / got[..., 3:], ref[..., :3] / ref[..., 3:]
    se = r.std() / np.sqrt(r.size)
    print(f"  bit-identical pixels {same.mean():.5f}, means {g.mean():.6f} / {r.mean():.6f}", flush=True)
    assert same.mean() >= 0.99
    assert abs(g.mean() - r.mean()) < 3 * se + 1e-12
    I.close()


def test_general_instances_render_agrees_with_flattened(gpu_ctx, rscenes):
    sc, built = rscenes
    inst = ic.render_instances_general()
    print("[case] instanced render, general transforms", flush=True)
    I = dr.Instances(gpu_ctx, built, inst)
    sd = ic.flatten_render(sc, inst)
    flat = _flat_scene(gpu_ctx, sd)
    cam = _camera()
    got = I.render(cam, spp=8, max_depth=8, rr_depth=3, seed=11)
    ref = flat.render(cam, spp=8, max_depth=8, rr_depth=3, seed=11)
    g, r = got[..., :3] / got[..., 3:], ref[..., :3] / ref[..., 3:]
    close = np.all(np.abs(g - r) <= 1e-4 * np.maximum(np.abs(r), 1e-3), axis=2)
    print(f"  pixels within 1e-4: {close.mean():.5f}, means {g.mean():.6f} / {r.mean():.6f}", flush=True)
    assert close.mean() >= 0.95
    assert abs(g.mean() - r.mean()) <= 1e-3 * r.mean()
    # both lit the same emitters
    ent, tri, _ = I.lights()
    fent, _ = flat.lights()
    assert ent.shape[0] == fent.shape[0]
    flat.close()
    I.close()


def test_instances_render_arguments(gpu_ctx, rscenes):
    _, built = rscenes
    I = dr.Instances(gpu_ctx, built[:1], ic.render_instances_exact()[:1])
    acc = gpu_ctx.alloc(16 * 16 * 32)
    cam = _camera(16, 16)
    with pytest.raises(dr.DevilRayError):
        I.render_dev(cam, dr.DrRenderParams(1, 0, 0, 3, 0, 0, 0), acc.ptr)  # max_depth 0
    with pytest.raises(dr.DevilRayError):
        dr.check(dr.lib().dr_instances_render(None, dr.C.byref(cam), dr.C.byref(dr.DrRenderParams(1, 0, 8, 3, 0, 0, 0)),
                                              dr.C.c_void_p(acc.ptr), None))
    # a room without emitters renders black
    img = I.render(cam, spp=2)
    assert np.all(img[..., :3] == 0) and np.all(img[..., 3] == 2)
    assert I.lights()[0].shape[0] == 0
    I.close()


def test_exact_instances_render_in_lanes_equals_flattened(gpu_ctx, rscenes):
    """512 x 384 x 2 spp (393,216 paths, dealt to four render lanes on four
    streams): the instanced image still equals the flattened copy's bit for
    bit, and the flattened copy's equals its own one-lane render."""
    sc, built = rscenes
    inst = ic.render_instances_exact()
    print("[case] instanced render in four lanes, 512x384x2", flush=True)
    I = dr.Instances(gpu_ctx, built, inst)
    sd = ic.flatten_render(sc, inst)
    flat = _flat_scene(gpu_ctx, sd)
    cam = _camera(512, 384)
    got = I.render(cam, spp=2, max_depth=8, rr_depth=3, seed=5)
    ref = flat.render(cam, spp=2, max_depth=8, rr_depth=3, seed=5)
    flat.set_render_lanes(1)
    one = flat.render(cam, spp=2, max_depth=8, rr_depth=3, seed=5)
    assert got[..., :3].sum() > 0
    assert np.array_equal(ref.view(np.uint64), one.view(np.uint64))
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    flat.close()
    I.close()
