"""Pin the oracle's instanced brute force (or_trace_instances_brute).

The reference tests a ray against each object in model space, with the ray
moved by Transform::applyInverse (include/tracing/IntersectionTestsImpl.hpp
:143-151, include/Transform.hpp:14-21), and keeps the nearest hit over all
objects (:234-251). Checks:
  * exact transforms (power-of-two scales with one magnitude per instance,
    sign flips, lattice translations and origins): the model-space test
    equals brute force over the flattened world copies bit for bit (ids,
    t, u, v; occlusion), which pins the transform's order of operations,
    the id numbering in instance order and the tie rule;
  * one identity instance equals the plain brute force;
  * general transforms (non-uniform scales): t within rounding of the world
    copies' brute force, ids equal except where two surfaces are within
    rounding of each other;
  * a planted mistake (the translation taken in model units, t / s) fails
    the comparison.
"""
import numpy as np
import pytest

import oracle
from instance_cases import (exact_instances, flatten, general_instances, rays_at_targets, rays_in_box,
                            scenes_small)

SCENES = scenes_small()


def _box(v, pad=2.0):
    return v.min(0) - pad, v.max(0) + pad


def test_identity_instance_equals_brute_force():
    v, f = SCENES[0]
    inst = exact_instances(1, 1)
    inst["scale"] = 1.0
    inst["translate"] = 0.0
    lo, hi = _box(v)
    rays = rays_at_targets(v, f, lo, hi, 3000, seed=1)
    a = oracle.instances_closest([SCENES[0]], inst, rays)
    b = oracle.brute_closest(v, f, rays)
    assert (a["id"] >= 0).mean() > 0.9
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("seed", [0, 1])
def test_exact_transforms_equal_flattened(seed):
    inst = exact_instances(24, len(SCENES), seed=seed, spread=8.0)
    wv, wf = flatten(SCENES, inst)
    lo, hi = _box(wv, 1.0)
    rays = np.concatenate([rays_at_targets(wv, wf, lo, hi, 4000, seed=seed),
                           rays_in_box(lo, hi, 2000, seed=seed + 7)])
    a = oracle.instances_closest(SCENES, inst, rays)
    b = oracle.brute_closest(wv, wf, rays)
    assert (a["id"] >= 0).mean() > 0.6
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    # shortened segments and occlusion
    rays["tmax"] = np.where(np.arange(rays.shape[0]) % 2 == 0, 3.0, np.inf).astype(np.float32)
    assert np.array_equal(oracle.instances_any(SCENES, inst, rays), oracle.brute_any(wv, wf, rays))


def test_coincident_instances_tie_to_the_lower_id():
    # the same scene placed twice at the same transform: every hit is a tie
    # and goes to the first instance's triangle
    inst = exact_instances(2, 1, seed=3)
    inst["scene"] = 0
    inst[1] = inst[0]
    n0 = SCENES[0][1].shape[0]
    wv, wf = flatten(SCENES, inst)
    lo, hi = _box(wv)
    rays = rays_at_targets(wv, wf, lo, hi, 2000, seed=4)
    a = oracle.instances_closest(SCENES, inst, rays)
    hit = a["id"] >= 0
    assert hit.mean() > 0.9 and (a["id"][hit] < n0).all()
    assert np.array_equal(a.view(np.uint8), oracle.brute_closest(wv, wf, rays).view(np.uint8))


def test_general_transforms_close_to_flattened():
    inst = general_instances(20, len(SCENES), seed=5, spread=8.0)
    wv, wf = flatten(SCENES, inst)
    lo, hi = _box(wv, 1.0)
    rays = rays_at_targets(wv, wf, lo, hi, 5000, seed=6)
    a = oracle.instances_closest(SCENES, inst, rays)
    b = oracle.brute_closest(wv, wf, rays)
    both = (a["id"] >= 0) & (b["id"] >= 0)
    assert both.mean() > 0.8
    assert (a["id"] == b["id"]).mean() > 0.995
    same = both & (a["id"] == b["id"])
    assert np.allclose(a["t"][same], b["t"][same], rtol=1e-4, atol=1e-5)
    assert np.abs(a["u"][same] - b["u"][same]).max() < 1e-3


def test_planted_transform_order_fails():
    # the translation taken in model units: (o - t / s) / s instead of
    # (o - t) / s
    inst = exact_instances(12, len(SCENES), seed=8, spread=6.0)
    inst["scale"] = np.where(np.abs(inst["scale"]) == 1.0, 2.0 * inst["scale"], inst["scale"])
    wv, wf = flatten(SCENES, inst)
    lo, hi = _box(wv, 1.0)
    rays = rays_at_targets(wv, wf, lo, hi, 2000, seed=9)
    wrong = inst.copy()
    wrong["translate"] = inst["translate"] / inst["scale"]
    a = oracle.instances_closest(SCENES, wrong, rays)
    b = oracle.brute_closest(wv, wf, rays)
    assert (a["id"] != b["id"]).mean() > 0.2
