"""The render workloads beyond Cornell, on the device against the oracle.

- Many-lights scene at full geometry: 400 displaced icospheres, 2,048,000
  triangles, 100,000 log-normal emitters, 30% dielectric; rendered at a reduced
  resolution (192 x 108, 16 spp, 12 vertices, roulette from vertex 3) so the
  oracle finishes in seconds. Bar: >= 99% of pixels bit-identical, image
  means within 3 standard errors.
- Indirect-lit partition scene (the guided-sampling workload's geometry:
  the many-lights spheres with every emitter behind an opaque partition): same
  bar at 160 x 90, 16 spp.
- The batch of 64 uneven scenes (triangles log-uniform in [1e3, 2e6], pixels
  log-uniform in [128^2, 1024^2], spp in {4, 16, 64}): one batched call over
  all views equals one dr_render per scene bit for bit, and the view with the
  fewest samples equals the oracle at its full resolution and spp.
"""
import numpy as np
import pytest

import devilray_amd as dr
import oracle
from devilray_amd import scenes

pytestmark = pytest.mark.gpu


def assert_matches_oracle(got, ref, label):
    assert np.array_equal(got[..., 3], ref[..., 3])
    same = np.all(got[..., :3].view(np.uint64) == ref[..., :3].view(np.uint64), axis=2)
    g = got[..., :3] / got[..., 3:]
    r = ref[..., :3] / ref[..., 3:]
    se = r.std() / np.sqrt(r.size)
    print(f"{label}: bit-identical pixels {same.mean():.5f}, means {g.mean():.6f} / {r.mean():.6f} (se {se:.2e})",
          flush=True)
    assert same.mean() >= 0.99
    assert abs(g.mean() - r.mean()) < 3 * se + 1e-12


def render_both(ctx, sd, w, h, spp, max_depth=12, rr_depth=3, seed=0, fresnel=0):
    s = dr.Scene(ctx, sd.vertices, sd.indices, sd.tri_material, sd.dr_materials())
    s.build_bvh(leaf_max=4)
    c = sd.camera
    cam = dr.camera_lookat(c["eye"], c["target"], c["up"], c["fov"], w, h)
    got = s.render(cam, spp=spp, max_depth=max_depth, rr_depth=rr_depth, seed=seed, fresnel=fresnel)
    s.close()
    ref, _ = oracle.render(oracle.BVH(sd.vertices, sd.indices), sd.tri_material, oracle.materials_array(sd.materials),
                           oracle.Camera.from_buffer_copy(bytes(cam)), spp=spp, max_depth=max_depth,
                           rr_depth=rr_depth, seed=seed, fresnel=fresnel)
    return got, ref


def test_many_lights_full_geometry_matches_oracle(gpu_ctx):
    print("[case] many-lights 2,048,000 tris vs oracle, 192x108x16", flush=True)
    sd = scenes.many_lights()
    assert sd.indices.shape[0] == 2_048_000
    assert int(np.sum([np.any(np.asarray(m[2]) > 0) for m in sd.materials])) == 100_000
    got, ref = render_both(gpu_ctx, sd, 192, 108, 16)
    assert ref[..., :3].sum() > 0
    assert_matches_oracle(got, ref, "many-lights")


def test_partition_scene_matches_oracle(gpu_ctx):
    print("[case] partition scene vs oracle, 160x90x16", flush=True)
    sd = scenes.many_lights(partition=True)
    got, ref = render_both(gpu_ctx, sd, 160, 90, 16)
    assert ref[..., :3].sum() > 0
    assert_matches_oracle(got, ref, "partition")


def test_batch_of_64_uneven_scenes(gpu_ctx):
    print("[case] 64-scene batch: batched (launch sequence and one persistent launch) == separate, smallest view "
          "vs oracle", flush=True)
    work = scenes.batch_c3(64)
    scs, cams, npix = [], [], []
    for sd, w, h, spp in work:
        s = dr.Scene(gpu_ctx, sd.vertices, sd.indices, sd.tri_material, sd.dr_materials())
        s.build_bvh(leaf_max=4)
        scs.append(s)
        c = sd.camera
        cams.append(dr.camera_lookat(c["eye"], c["target"], c["up"], c["fov"], w, h))
        npix.append(w * h)
    total = int(sum(npix))
    offs = (np.concatenate([[0], np.cumsum(npix)[:-1]]) * 32).astype(np.int64)
    sep = gpu_ctx.alloc(32 * total)
    bat = gpu_ctx.alloc(32 * total)
    sep.copy_from(np.zeros(4 * total))
    bat.copy_from(np.zeros(4 * total))
    for k, (s, cam) in enumerate(zip(scs, cams)):
        s.render_dev(cam, dr.DrRenderParams(work[k][3], 0, 12, 3, k << 32, 0, 0), sep.ptr + int(offs[k]))
    b = dr.Batch(gpu_ctx, scs)
    views = [(k, cams[k], work[k][3], 0, bat.ptr + int(offs[k])) for k in range(len(scs))]
    b.render_dev(views, max_depth=12, rr_depth=3, seed=0, chunk_paths=1 << 22, trace_sched=1, use_graph=True)
    gpu_ctx.synchronize()
    a_sep = sep.to_numpy(np.float64, 4 * total)
    a_bat = bat.to_numpy(np.float64, 4 * total)
    print(f"  {len(work)} scenes, {sum(sd.indices.shape[0] for sd, *_ in work)} tris, {total} pixels", flush=True)
    assert np.array_equal(a_sep, a_bat), "batched views differ from their single-scene renders"
    # the same batch in 2^24-path chunks dealt to four render lanes (streams)
    print("  render lanes", flush=True)
    bat.copy_from(np.zeros(4 * total))
    b.render_dev(views, max_depth=12, rr_depth=3, seed=0, chunk_paths=1 << 24, trace_sched=1, use_graph=False)
    gpu_ctx.synchronize()
    assert np.array_equal(a_sep, bat.to_numpy(np.float64, 4 * total)), "render lanes differ"
    # the same batch in one persistent launch
    print("  one persistent launch", flush=True)
    bat.copy_from(np.zeros(4 * total))
    b.render_dev(views, max_depth=12, rr_depth=3, seed=0, chunk_paths=1 << 22, trace_sched=1,
                 launch=dr.DR_BATCH_LAUNCH_PERSISTENT)
    gpu_ctx.synchronize()
    assert np.array_equal(a_sep, bat.to_numpy(np.float64, 4 * total)), "persistent launch differs"
    # the view with the fewest samples, against the oracle at its full size and spp
    k = int(np.argmin([n * w[3] for n, w in zip(npix, work)]))
    sd, w, h, spp = work[k]
    got = a_bat[4 * int(offs[k] // 32):4 * (int(offs[k] // 32) + npix[k])].reshape(h, w, 4)
    ref, _ = oracle.render(oracle.BVH(sd.vertices, sd.indices), sd.tri_material, oracle.materials_array(sd.materials),
                           oracle.Camera.from_buffer_copy(bytes(cams[k])), spp=spp, max_depth=12, rr_depth=3,
                           seed=k << 32)
    assert_matches_oracle(got, ref, f"batch view {k} ({w}x{h}, {spp} spp, {sd.indices.shape[0]} tris)")
    b.close()
    for s in scs:
        s.close()
    sep.free()
    bat.free()
