"""Instancing on the device (dr_instances_*) against the oracle's model-space
brute force and against the flattened scene.

  * general transforms (non-uniform scales, mirrors, arbitrary translations):
    closest hits equal the oracle bit for bit (ids, t, u, v) and occlusion
    equals it exactly, on every traversal schedule;
  * exact transforms (one power-of-two magnitude per instance, lattice
    translations and origins): the instanced trace equals the device trace of
    the flattened world copies bit for bit, also for 3000 instances
    (0.96M instanced triangles);
  * edge cases: one instance, misses, NaN rays, empty segments, coincident
    instances (ties go to the lower id), and the argument checks.
"""
import ctypes as C

import numpy as np
import pytest

import devilray_amd as dr
import oracle
from instance_cases import (exact_instances, flatten, general_instances, rays_at_targets, rays_in_box,
                            scenes_small)

pytestmark = pytest.mark.gpu

SCHEDS = [dr.SCHED_PER_RAY, dr.SCHED_PER_RAY_SPECULATIVE, dr.SCHED_PERSISTENT_SPECULATIVE,
          dr.SCHED_PERSISTENT_SPECULATIVE | dr.SCHED_REORDER | dr.SCHED_XCD_SLICES]


@pytest.fixture(scope="module")
def small(gpu_ctx):
    meshes = scenes_small()
    scs = []
    for v, f in meshes:
        s = dr.Scene(gpu_ctx, v, f)
        s.build_bvh()
        scs.append(s)
    yield meshes, scs
    for s in scs:
        s.close()


def _box(v, pad=1.0):
    return v.min(0) - pad, v.max(0) + pad


def _same(a, b):
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("sched", SCHEDS)
def test_general_transforms_equal_oracle(gpu_ctx, small, sched):
    meshes, scs = small
    inst = general_instances(40, len(meshes), seed=11, spread=10.0)
    wv, wf = flatten(meshes, inst)
    lo, hi = _box(wv)
    rays = np.concatenate([rays_at_targets(wv, wf, lo, hi, 6000, seed=12), rays_in_box(lo, hi, 4000, seed=13)])
    print(f"instances: 40 general, {rays.shape[0]} rays, sched {sched}", flush=True)
    I = dr.Instances(gpu_ctx, scs, inst)
    I.set_trace_schedule(sched, sched)
    got = I.trace_closest(rays)
    ref = oracle.instances_closest(meshes, inst, rays)
    assert (ref["id"] >= 0).mean() > 0.5
    bad = np.flatnonzero((got.view(np.uint8).reshape(-1, 16) != ref.view(np.uint8).reshape(-1, 16)).any(axis=1))
    assert bad.size == 0, (bad[:5], got[bad[:5]], ref[bad[:5]])
    short = rays.copy()
    short["tmax"] = np.where(np.arange(rays.shape[0]) % 3 == 0, 2.5, np.inf).astype(np.float32)
    assert np.array_equal(I.trace_any(short), oracle.instances_any(meshes, inst, short))
    info = I.info()
    assert info.overflow_flag == 0
    assert info.n_triangles == sum(meshes[int(k)][1].shape[0] for k in inst["scene"])
    I.close()


@pytest.mark.parametrize("seed", [0, 1])
def test_exact_transforms_equal_flattened_scene(gpu_ctx, small, seed):
    meshes, scs = small
    inst = exact_instances(64, len(meshes), seed=seed, spread=12.0)
    wv, wf = flatten(meshes, inst)
    lo, hi = _box(wv)
    rays = np.concatenate([rays_at_targets(wv, wf, lo, hi, 20000, seed=seed), rays_in_box(lo, hi, 8000, seed=5)])
    print(f"instances: 64 exact, seed {seed}", flush=True)
    I = dr.Instances(gpu_ctx, scs, inst)
    flat = dr.Scene(gpu_ctx, wv, wf)
    flat.build_bvh()
    a, b = I.trace_closest(rays), flat.trace_closest(rays)
    assert (a["id"] >= 0).mean() > 0.6
    assert _same(a, b)
    assert _same(a, oracle.instances_closest(meshes, inst, rays))
    assert np.array_equal(I.trace_any(rays), flat.trace_any(rays))
    flat.close()
    I.close()


def test_many_instances_equal_flattened_scene(gpu_ctx):
    # 3000 placements of a 320-triangle sphere on a lattice: 0.96M instanced
    # triangles, a top-level tree of several levels
    v, f = dr.gen_icosphere(2, 0.05, 4)
    from instance_cases import quantize
    mesh = (quantize(v), f)
    sc = dr.Scene(gpu_ctx, *mesh)
    sc.build_bvh()
    rng = np.random.default_rng(21)
    n = 3000
    inst = np.zeros(n, dr.INSTANCE_DTYPE)
    inst["scale"] = (rng.choice([0.5, 1.0, 2.0], n)[:, None] * rng.choice([-1.0, 1.0], (n, 3), p=[0.2, 0.8])
                     ).astype(np.float32)
    inst["translate"] = (np.round(rng.uniform(-60, 60, (n, 3)) * 4) / 4).astype(np.float32)
    wv, wf = flatten([mesh], inst)
    lo, hi = _box(wv)
    rays = np.concatenate([rays_at_targets(wv, wf, lo, hi, 1 << 17, seed=22),
                           rays_in_box(lo, hi, 1 << 16, seed=23)])
    print(f"instances: {n} of {f.shape[0]} triangles, {rays.shape[0]} rays", flush=True)
    I = dr.Instances(gpu_ctx, [sc], inst)
    flat = dr.Scene(gpu_ctx, wv, wf)
    flat.build_bvh()
    a, b = I.trace_closest(rays), flat.trace_closest(rays)
    assert (a["id"] >= 0).mean() > 0.5
    assert _same(a, b)
    assert np.array_equal(I.trace_any(rays), flat.trace_any(rays))
    info = I.info()
    assert info.overflow_flag == 0 and info.n_triangles == n * f.shape[0]
    assert info.depth >= 2
    flat.close()
    I.close()
    sc.close()


def test_edge_cases(gpu_ctx, small):
    meshes, scs = small
    # one instance at identity == the scene itself
    one = np.zeros(1, dr.INSTANCE_DTYPE)
    one["scale"] = 1.0
    I = dr.Instances(gpu_ctx, scs[:1], one)
    v, f = meshes[0]
    lo, hi = _box(v)
    rays = rays_at_targets(v, f, lo, hi, 5000, seed=31)
    # misses, NaN rays and empty segments
    rays[:50]["dx"], rays[:50]["dy"], rays[:50]["dz"] = 0.0, 0.0, 1.0
    rays[:50]["ox"], rays[:50]["oy"], rays[:50]["oz"] = 100.0, 100.0, 0.0
    rays[50:60]["ox"] = np.nan
    rays[60:70]["tmax"] = 0.0
    a = I.trace_closest(rays)
    assert _same(a, scs[0].trace_closest(rays))
    assert (a["id"][:70] == -1).all()
    assert np.array_equal(I.trace_any(rays), scs[0].trace_any(rays))
    I.close()
    # coincident instances: every hit ties and goes to the first one
    two = np.zeros(2, dr.INSTANCE_DTYPE)
    two["scale"] = 2.0
    two["translate"] = (1.0, -2.0, 0.5)
    I = dr.Instances(gpu_ctx, scs[:1], two)
    wv, wf = flatten(meshes[:1], two)
    lo, hi = _box(wv)
    rays = rays_at_targets(wv, wf, lo, hi, 4000, seed=32)
    a = I.trace_closest(rays)
    hit = a["id"] >= 0
    assert hit.mean() > 0.9 and (a["id"][hit] < f.shape[0]).all()
    assert _same(a, oracle.instances_closest(meshes[:1], two, rays))
    I.close()


def test_argument_checks(gpu_ctx, small):
    meshes, scs = small
    good = np.zeros(1, dr.INSTANCE_DTYPE)
    good["scale"] = 1.0
    for field, value in (("scale", 0.0), ("scale", np.inf), ("translate", np.nan), ("scene", 7)):
        bad = good.copy()
        bad[field] = value
        with pytest.raises(dr.DevilRayError):
            dr.Instances(gpu_ctx, scs, bad)
    unbuilt = dr.Scene(gpu_ctx, *meshes[1])
    with pytest.raises(dr.DevilRayError):
        dr.Instances(gpu_ctx, [unbuilt], good)
    unbuilt.close()
    with pytest.raises(dr.DevilRayError):
        dr.Instances(gpu_ctx, scs, np.zeros(0, dr.INSTANCE_DTYPE))
    h = C.c_void_p()
    assert dr.lib().dr_instances_create(None, None, 0, None, 0, C.byref(h)) != 0
    with pytest.raises(dr.DevilRayError):
        I = dr.Instances(gpu_ctx, scs, good)
        try:
            I.set_trace_schedule(99, 0)
        finally:
            I.close()
