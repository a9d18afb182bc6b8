"""Write tests/golden/roofline_counts.json: the per-ray node and triangle test
counts of the oracle's canonical binary BVH (binned SAH, 32 bins, C_trav =
C_isect = 1, <= 4 triangles per leaf; front-to-back, prune on
t_entry > t_best) on the traversal benchmark's ray batches. These counts fix
the algorithmic bytes per ray (48 + 32*N_node + 40*N_tri closest-hit,
36 + 32*N_node + 40*N_tri any-hit) independently of the GPU's BVH.

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
    n_node = float(np.mean(nt))
    n_tri = float(np.mean(tt))
    return {"mean_node_tests": round(n_node, 4), "mean_tri_tests": round(n_tri, 4),
            "bytes_per_ray": round(base + 32 * n_node + 40 * n_tri, 2)}


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
