"""Write tests/golden/roofline_counts.json: the per-ray interior-node visits
and triangle tests of the oracle's canonical binary BVH (binned SAH, 32
bins, C_trav = C_isect = 1, <= 4 triangles per leaf; front-to-back, prune on
t_entry > t_best) on the traversal benchmark's ray batches. These counts fix
the algorithmic bytes per ray (48 + 32*N_node + 40*N_tri closest-hit,
36 + 32*N_node + 40*N_tri any-hit, N_node = interior-node visits)
independently of the GPU's BVH. Box tests (two per visit, plus the root's)
are recorded beside them.

For the 2,048,000-triangle many-lights render (depth 12, roulette from
vertex 3, seed 0), the build integrator's extension and shadow rays over the
same kind of canonical BVH, sampled at 240 x 135 x 4 spp with the render's
field of view: the algorithmic bytes of one path sample are the sum of its
rays' bytes plus 64 B of FP64 accumulation read-modify-write.

Also records the reference-semantics BBH counters (median split, one
triangle per leaf, stackless walk) as context.

Run: python tests/golden/make_roofline_counts.py
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import devilray_amd as dr  # noqa: E402  (generators only)
import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "roofline_counts.json")


def stats(nt, tt, base):
    """nt: box tests per ray (1 for the root + 2 per interior-node visit)."""
    visits = float(np.mean((np.asarray(nt, np.int64) - 1) // 2))
    n_tri = float(np.mean(tt))
    return {"mean_node_visits": round(visits, 4), "mean_box_tests": round(float(np.mean(nt)), 4),
            "mean_tri_tests": round(n_tri, 4), "bytes_per_ray": round(base + 32 * visits + 40 * n_tri, 2)}


def render_counts():
    from devilray_amd import scenes
    sc = scenes.many_lights()
    bvh = oracle.BVH(sc.vertices, sc.indices, max_leaf=4, bins=32)
    c = sc.camera
    w, h, spp = 240, 135, 4
    cam = oracle.Camera.from_buffer_copy(bytes(dr.camera_lookat(c["eye"], c["target"], c["up"], c["fov"], w, h)))
    _, _, cnt = oracle.render(bvh, sc.tri_material, oracle.materials_array(sc.materials), cam, spp=spp,
                              max_depth=12, rr_depth=3, seed=0, counts=True)
    ns = w * h * spp
    ext, ev, et, sh, sv, st = (int(x) for x in cnt)
    per = (48 * ext + 32 * ev + 40 * et + 36 * sh + 32 * sv + 40 * st) / ns + 64
    return {"scene": "many_lights: 2,048,000 tris, 100k emitters; depth 12, roulette from 3, seed 0",
            "sample": f"{w}x{h}x{spp} spp, the 1920x1080 view's field of view",
            "extension_rays_per_sample": round(ext / ns, 4),
            "extension_node_visits_per_ray": round(ev / ext, 4), "extension_tri_tests_per_ray": round(et / ext, 4),
            "shadow_rays_per_sample": round(sh / ns, 4),
            "shadow_node_visits_per_ray": round(sv / max(sh, 1), 4),
            "shadow_tri_tests_per_ray": round(st / max(sh, 1), 4),
            "bytes_per_sample": round(per, 2)}


def main():
    t0 = time.time()
    v, f = dr.gen_icosphere(8, 0.05, 0)
    rays = dr.gen_rays_surface(v, f, 1 << 20, seed=0, inward=True, offset=1e-5)
    cam = dr.gen_rays_camera((0.0, 0.0, 3.5), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, 1024, 1024)
    bvh = oracle.BVH(v, f, max_leaf=4, bins=32)
    print(f"built canonical BVH: {bvh.node_count()} nodes in {time.time() - t0:.1f}s", flush=True)
    out = {"config": "icosphere subdiv 8, sigma 0.05, seed 0; 2^20 rays seed 0 inward offset 1e-5; "
                     "camera (0,0,3.5)->(0,0,0) fov 40 1024x1024",
           "bvh": "binary binned SAH, 32 bins, Ct=Ci=1, max leaf 4", "nodes": bvh.node_count()}
    h, nt, tt = bvh.closest(rays, counts=True)
    out["c1_incoherent_closest"] = stats(nt, tt, 48) | {"hit_fraction": float(np.mean(h["id"] >= 0))}
    o, nt, tt = bvh.any(rays, counts=True)
    out["c1_incoherent_any"] = stats(nt, tt, 36) | {"occluded_fraction": float(np.mean(o))}
    h, nt, tt = bvh.closest(cam, counts=True)
    out["c1_coherent_closest"] = stats(nt, tt, 48) | {"hit_fraction": float(np.mean(h["id"] >= 0))}
    out["c2_render_path_sample"] = render_counts()
    sub = rays[:1 << 16].copy()
    ref = oracle.RefBBH(v, f)
    hr, bt, tr = ref.closest(sub, counts=True)
    out["c1_incoherent_reference_bbh"] = {
        "rays": int(sub.shape[0]), "depth": ref.depth(),
        "mean_box_tests": round(float(np.mean(bt)), 4), "mean_tri_tests": round(float(np.mean(tr)), 4),
        "hit_fraction": float(np.mean(hr["id"] >= 0))}
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))
    print(f"done in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    main()
