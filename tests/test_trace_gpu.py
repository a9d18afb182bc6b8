"""GPU parity of closest-hit / any-hit traversal against the CPU oracle.

Bar: triangle ids bit-exact; t, u, v bit-identical (the device and the oracle
evaluate the same watertight expression with contraction disabled, which is
stricter than the 2-ulp tolerance); occlusion flags equal.
"""
import numpy as np
import pytest

import devilray_amd as dr
import oracle

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def assert_hits_equal(got, ref, label=""):
    bad = np.nonzero(got["id"] != ref["id"])[0]
    assert bad.size == 0, f"{label}: {bad.size} id mismatches, first {bad[:5]} got {got[bad[:5]]} ref {ref[bad[:5]]}"
    hit = ref["id"] >= 0
    for k in ("t", "u", "v"):
        diff = np.nonzero(bits(got[k][hit]) != bits(ref[k][hit]))[0]
        assert diff.size == 0, f"{label}: {k} differs on {diff.size} hits"
    assert np.all(np.isinf(got["t"][~hit])), f"{label}: misses must report t = inf"


@pytest.mark.parametrize("leaf_max", [1, 2, 4, 8])
def test_small_sphere_vs_brute_force(gpu_ctx, leaf_max):
    print(f"[case] small sphere leaf_max={leaf_max}", flush=True)
    v, f = dr.gen_icosphere(3, 0.05, 7)
    rays = dr.gen_rays_surface(v, f, 8192, seed=3, inward=True)
    out = dr.gen_rays_surface(v, f, 4096, seed=4, inward=False)
    allr = np.concatenate([rays, out])
    scene = dr.Scene(gpu_ctx, v, f)
    scene.build_bvh(leaf_max=leaf_max)
    assert_hits_equal(scene.trace_closest(allr), oracle.brute_closest(v, f, allr), f"leaf{leaf_max}")
    assert np.array_equal(scene.trace_any(allr), oracle.brute_any(v, f, allr))
    assert scene.info().overflow_flag == 0
    scene.close()


# split: cap on extra references per triangle; ratio 0.5 splits every
# triangle of this scene (box/ideal area ratios are 2 and more), so the cap of
# 1.0 exercises the rescaling and 8.0 the full counts.
@pytest.mark.parametrize("builder,radius,split", [("lbvh", 0, 0.0), ("ploc", 1, 0.0), ("ploc", 3, 0.0),
                                                   ("ploc", 64, 0.0), ("ploc", 16, 1.0), ("ploc", 16, 8.0),
                                                   ("lbvh", 0, 2.0)])
def test_builders_vs_brute_force(gpu_ctx, builder, radius, split):
    print(f"[case] builder {builder} radius {radius} split {split}", flush=True)
    v, f = dr.gen_icosphere(3, 0.05, 11)
    rays = np.concatenate([dr.gen_rays_surface(v, f, 4096, seed=5, inward=True),
                           dr.gen_rays_surface(v, f, 2048, seed=6, inward=False)])
    scene = dr.Scene(gpu_ctx, v, f)
    for leaf_max in (1, 4):
        scene.build_bvh(leaf_max=leaf_max, builder=builder, radius=radius, split=split,
                        split_ratio=0.5 if split else None)
        refs = scene.info().n_references
        assert (refs == f.shape[0]) if split == 0 else (f.shape[0] < refs <= (1 + split) * f.shape[0] + 1)
        label = f"{builder} r{radius} split {split} leaf {leaf_max}"
        assert_hits_equal(scene.trace_closest(rays), oracle.brute_closest(v, f, rays), label)
        assert np.array_equal(scene.trace_any(rays), oracle.brute_any(v, f, rays))
    scene.close()


def test_split_follows_triangle_shape(gpu_ctx):
    """Default build: a smooth sphere's well-shaped triangles stay whole, the
    slivers of a strongly displaced fine sphere are cut into pieces."""
    print("[case] split follows shape", flush=True)
    v, f = dr.gen_icosphere(4, 0.0, 0)
    scene = dr.Scene(gpu_ctx, v, f)
    scene.build_bvh()
    assert scene.info().n_references == f.shape[0]
    scene.close()
    v, f = dr.gen_icosphere(6, 0.05, 3)
    scene = dr.Scene(gpu_ctx, v, f)
    scene.build_bvh()
    refs = scene.info().n_references
    assert 1.5 * f.shape[0] < refs <= 9 * f.shape[0]
    rays = dr.gen_rays_surface(v, f, 8192, seed=9, inward=True)
    assert_hits_equal(scene.trace_closest(rays), oracle.brute_closest(v, f, rays), "displaced subdiv 6")
    scene.close()


@pytest.mark.parametrize("builder", ["ploc", "lbvh"])
def test_nonfinite_and_huge_coordinates(gpu_ctx, builder):
    """Triangles with a NaN or infinite vertex can never be hit; triangles
    with coordinates near 1e25 overflow the box areas to infinity. The build
    must still give leaves of at most leaf_max references and the same hits as
    brute force."""
    print(f"[case] non-finite and huge coordinates, {builder}", flush=True)
    v, f = dr.gen_icosphere(3, 0.05, 13)
    rng = np.random.default_rng(31)
    extra_v, extra_f = [], []
    base = v.shape[0]
    for k in range(40):
        tri = rng.normal(size=(3, 3)).astype(np.float32)
        if k % 4 == 0:
            tri[rng.integers(3), rng.integers(3)] = np.nan
        elif k % 4 == 1:
            tri[rng.integers(3), rng.integers(3)] = np.inf if k % 8 == 1 else -np.inf
        else:
            tri = (tri * np.float32(1e24) + np.float32(3e25)).astype(np.float32)
        extra_v.append(tri)
        extra_f.append(base + 3 * k + np.arange(3))
    v2 = np.concatenate([v] + extra_v).astype(np.float32)
    f2 = np.concatenate([f, np.array(extra_f, np.uint32)]).astype(np.uint32)
    rays = np.concatenate([dr.gen_rays_surface(v, f, 4096, seed=8, inward=True),
                           dr.gen_rays_surface(v, f, 2048, seed=9, inward=False)])
    ref = oracle.brute_closest(v2, f2, rays)
    ref_any = oracle.brute_any(v2, f2, rays)
    scene = dr.Scene(gpu_ctx, v2, f2)
    for leaf_max, split in ((1, None), (4, None), (8, 0.0), (4, 2.0)):
        scene.build_bvh(leaf_max=leaf_max, builder=builder, split=split)
        assert_hits_equal(scene.trace_closest(rays), ref, f"{builder} leaf {leaf_max} split {split}")
        assert np.array_equal(scene.trace_any(rays), ref_any)
        assert scene.info().overflow_flag == 0
    scene.close()


@pytest.mark.parametrize("n_tris", [2, 3, 5, 4000])
def test_ploc_identical_boxes(gpu_ctx, n_tris):
    """Every triangle the same (exact ties in every clustering round) plus a
    far-away one: the build must finish and the lowest id must win each tie."""
    print(f"[case] identical boxes n={n_tris}", flush=True)
    tri = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], np.float32)
    v = np.concatenate([np.tile(tri, (n_tris - 1, 1)), tri + np.float32([5.0, 5.0, 5.0])]).astype(np.float32)
    f = np.arange(3 * n_tris, dtype=np.uint32).reshape(-1, 3)
    rng = np.random.default_rng(n_tris)
    rays = np.zeros(2048, dr.RAY_DTYPE)
    o = np.stack([rng.uniform(-0.2, 1.2, 2048), rng.uniform(-0.2, 1.2, 2048), np.full(2048, 1.0)], 1)
    o[1024:] += 5.0
    rays["ox"], rays["oy"], rays["oz"] = o[:, 0], o[:, 1], o[:, 2]
    rays["dz"] = -1.0
    rays["tmax"] = np.inf
    scene = dr.Scene(gpu_ctx, v, f)
    scene.build_bvh(leaf_max=2)
    got = scene.trace_closest(rays)
    assert_hits_equal(got, oracle.brute_closest(v, f, rays), f"identical n={n_tris}")
    assert np.all(got["id"][:1024][got["id"][:1024] >= 0] == 0)
    scene.close()


@pytest.fixture(scope="module")
def c1_oracle():
    v, f = dr.gen_icosphere(8, 0.05, 0)
    return v, f, oracle.BVH(v, f)


@pytest.mark.parametrize("split", [None, 0.0, 2.0])
def test_c1_full_size_incoherent_and_coherent(gpu_ctx, c1_oracle, split):
    """Full traversal benchmark size: 1,310,720 triangles, 2^20 rays per batch."""
    print(f"[case] C1 full size split {split}", flush=True)
    v, f, bvh = c1_oracle
    rays = dr.gen_rays_surface(v, f, 1 << 20, seed=0, inward=True, offset=1e-5)
    cam = dr.gen_rays_camera((0.0, 0.0, 3.5), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, 1024, 1024)
    scene = dr.Scene(gpu_ctx, v, f)
    scene.build_bvh(leaf_max=4, split=split)
    ref = bvh.closest(rays)
    got = scene.trace_closest(rays)
    assert_hits_equal(got, ref, "incoherent")
    assert_hits_equal(scene.trace_closest(cam), bvh.closest(cam), "coherent")
    # any-hit with tmax = inf and with tmax = 0.5 * t_closest
    assert np.array_equal(scene.trace_any(rays), bvh.any(rays))
    half = rays.copy()
    hit = ref["id"] >= 0
    half["tmax"][hit] = 0.5 * ref["t"][hit]
    assert np.array_equal(scene.trace_any(half), bvh.any(half))
    if split is None:
        # the counted traversal (the benchmark's HitTests path) returns the same hits
        counted, box, tri = scene.trace_closest_counted(rays)
        assert_hits_equal(counted, ref, "counted incoherent")
        assert np.all(tri[hit] >= 1) and np.all(box >= 1)
        info = scene.info()
        assert int(tri.sum()) == info.simd_stats[3]
        assert 0 < info.stack_high_water <= 3 * info.depth
    assert scene.info().overflow_flag == 0
    scene.close()
