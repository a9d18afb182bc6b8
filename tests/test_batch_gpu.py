"""Batched rendering (dr_batch) against single-scene dr_render: a view of a
batch equals the single render of its scene with seed ^ (view << 32), bit for
bit, with the bounce loop run eagerly or recorded once and replayed, and with
chunks of paths that straddle views."""
import numpy as np
import pytest

import devilray_amd as dr
from devilray_amd import scenes

pytestmark = pytest.mark.gpu


def _scene(ctx, sd):
    s = dr.Scene(ctx, sd.vertices, sd.indices, sd.tri_material, sd.dr_materials())
    s.build_bvh()
    return s


def _cam(sd, w, h):
    c = sd.camera
    return dr.camera_lookat(c["eye"], c["target"], c["up"], c["fov"], w, h)


@pytest.fixture(scope="module")
def three_scenes(gpu_ctx):
    sds = [scenes.cornell_c0(), scenes.many_lights(n_spheres=12, subdiv=2, n_emissive=300, seed=5, extent=6.0),
           scenes.hidden_light_reflector()]
    scs = [_scene(gpu_ctx, sd) for sd in sds]
    yield sds, scs
    for s in scs:
        s.close()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_of_one_equals_render(gpu_ctx, three_scenes, graph):
    sds, scs = three_scenes
    cam = _cam(sds[0], 48, 40)
    ref = scs[0].render(cam, spp=8, max_depth=8, seed=11, sample_offset=3)
    b = dr.Batch(gpu_ctx, [scs[0]])
    acc = gpu_ctx.alloc(48 * 40 * 32)
    acc.copy_from(np.zeros(48 * 40 * 4))
    b.render_dev([(0, cam, 8, 3, acc.ptr)], max_depth=8, seed=11, chunk_paths=4096, use_graph=graph)
    gpu_ctx.synchronize()
    got = acc.to_numpy(np.float64, 48 * 40 * 4).reshape(40, 48, 4)
    b.close()
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("graph,sched", [(False, 1), (True, 1), (True, 3)])
def test_batch_views_equal_single_renders(gpu_ctx, three_scenes, graph, sched):
    sds, scs = three_scenes
    views = [(0, 40, 32, 4, 0), (1, 64, 36, 3, 5), (2, 32, 32, 6, 0), (1, 20, 12, 16, 2)]
    b = dr.Batch(gpu_ctx, scs)
    accs, cams = [], []
    for sc, w, h, spp, off in views:
        a = gpu_ctx.alloc(w * h * 32)
        a.copy_from(np.zeros(w * h * 4))
        accs.append(a)
        cams.append(_cam(sds[sc], w, h))
    seed = 0x1234
    b.render_dev([(v[0], cams[i], v[3], v[4], accs[i].ptr) for i, v in enumerate(views)], max_depth=8, seed=seed,
                 chunk_paths=5000, trace_sched=sched, use_graph=graph)
    gpu_ctx.synchronize()
    for i, (sc, w, h, spp, off) in enumerate(views):
        got = accs[i].to_numpy(np.float64, w * h * 4).reshape(h, w, 4)
        ref = scs[sc].render(cams[i], spp=spp, max_depth=8, seed=seed ^ (i << 32), sample_offset=off)
        assert np.array_equal(got, ref), i
    b.close()


@pytest.mark.parametrize("sched,chunk,bpc", [(1, 5000, 0), (0, 5000, 1), (1, 1 << 20, 2), (3, 4096, 0)])
def test_persistent_launch_equals_single_renders(gpu_ctx, three_scenes, sched, chunk, bpc):
    """DR_BATCH_LAUNCH_PERSISTENT: the whole render in one launch (chunks,
    bounces and stages separated by grid barriers) gives every view's
    accumulation bit for bit, as the single renders do; with several chunks,
    one chunk, and different numbers of resident blocks per CU."""
    sds, scs = three_scenes
    views = [(0, 40, 32, 4, 0), (1, 64, 36, 3, 5), (2, 32, 32, 6, 0), (1, 20, 12, 16, 2)]
    print(f"[case] persistent launch sched {sched} chunk {chunk} blocks/CU {bpc}", flush=True)
    b = dr.Batch(gpu_ctx, scs)
    accs, cams = [], []
    for sc, w, h, spp, off in views:
        a = gpu_ctx.alloc(w * h * 32)
        a.copy_from(np.zeros(w * h * 4))
        accs.append(a)
        cams.append(_cam(sds[sc], w, h))
    seed = 0x5151
    b.render_dev([(v[0], cams[i], v[3], v[4], accs[i].ptr) for i, v in enumerate(views)], max_depth=8, seed=seed,
                 chunk_paths=chunk, trace_sched=sched, launch=dr.DR_BATCH_LAUNCH_PERSISTENT, blocks_per_cu=bpc)
    gpu_ctx.synchronize()
    for i, (sc, w, h, spp, off) in enumerate(views):
        got = accs[i].to_numpy(np.float64, w * h * 4).reshape(h, w, 4)
        ref = scs[sc].render(cams[i], spp=spp, max_depth=8, seed=seed ^ (i << 32), sample_offset=off)
        assert np.array_equal(got, ref), i
    b.close()


def test_persistent_launch_rejects_bad_options(gpu_ctx, three_scenes):
    sds, scs = three_scenes
    b = dr.Batch(gpu_ctx, scs[:1])
    a = gpu_ctx.alloc(16 * 16 * 32)
    for kw in ({"launch": 2}, {"launch": 1, "blocks_per_cu": 17}):
        with pytest.raises(dr.DevilRayError):
            b.render_dev([(0, _cam(sds[0], 16, 16), 1, 0, a.ptr)], **kw)
    b.close()


def test_graph_renders_back_to_back_then_destroy(gpu_ctx, three_scenes):
    """Two graph renders queued without a synchronisation in between (the
    second records a new graph while the first one's replays may still be
    queued), then the batch destroyed at once: the executable graphs must
    outlive their launches, and both images equal the eager render."""
    sds, scs = three_scenes
    cam = _cam(sds[1], 64, 36)
    ref = scs[1].render(cam, spp=6, max_depth=8, seed=4)
    b = dr.Batch(gpu_ctx, [scs[1]])
    accs = []
    for chunk in (4096, 5000):
        a = gpu_ctx.alloc(64 * 36 * 32)
        a.copy_from(np.zeros(64 * 36 * 4))
        accs.append(a)
        b.render_dev([(0, cam, 6, 0, a.ptr)], max_depth=8, seed=4, chunk_paths=chunk, use_graph=True)
    b.close()
    gpu_ctx.synchronize()
    for a in accs:
        assert np.array_equal(a.to_numpy(np.float64, 64 * 36 * 4).reshape(36, 64, 4), ref)
        a.free()


def test_batch_rejects_bad_views(gpu_ctx, three_scenes):
    sds, scs = three_scenes
    b = dr.Batch(gpu_ctx, scs[:2])
    a = gpu_ctx.alloc(16 * 16 * 32)
    with pytest.raises(dr.DevilRayError):
        b.render_dev([(2, _cam(sds[0], 16, 16), 1, 0, a.ptr)])
    with pytest.raises(dr.DevilRayError):
        b.render_dev([(0, _cam(sds[0], 16, 16), 1, 0, a.ptr)], chunk_paths=100)
    b.close()
