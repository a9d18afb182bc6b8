"""Benchmark: incoherent closest-hit traversal throughput (Mrays/s) on one MI355X.

Workload (one "step" = one dr_trace_closest launch over the whole batch):
  1,310,720-triangle icosphere (subdivision 8) with per-vertex radius
  1 + 0.05*N(0,1) (Philox, seed 0), BVH built on the device; 2^20 rays with
  origins area-uniform on the surface and directions cosine-distributed about
  the inward face normal (seed 0), origin offset 1e-5 along the direction.
Inputs are resident in HBM before the timed region. The coherent 1024x1024
camera batch and any-hit are reported beside it in "extra".

Prints one JSON line. `roofline.achieved` uses the algorithmic bytes per ray
48 + 32*N_node + 40*N_tri with N from the canonical binary SAH BVH counts
committed in tests/golden/roofline_counts.json, divided by the kernel's mean
HIP-event duration on its stream. `cpu_baseline` times the oracle's
reference-semantics traversal (median-split BBH, stackless walk) on a bounded
sample of the same rays on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import devilray_amd as dr  # noqa: E402
from devilray_amd.hip_events import Event  # noqa: E402

HBM_PEAK_GBS = 8000.0
SUBDIV, SIGMA, SEED, NRAYS = 8, 0.05, 0, 1 << 20
COUNTS_FIXTURE = os.path.join(ROOT, "tests", "golden", "roofline_counts.json")
TRAFFIC_FIXTURE = os.path.join(ROOT, "profiles", "trace_closest_c1_incoherent_traffic.json")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def timed_launches(fn, steps: int, warmup: int, ctx) -> tuple[float, list[float]]:
    for _ in range(warmup):
        fn()
    ctx.synchronize()
    evs = [Event() for _ in range(steps + 1)]
    t0 = time.perf_counter()
    evs[0].record(None)
    for k in range(steps):
        fn()
        evs[k + 1].record(None)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    per = [evs[k].elapsed_ms(evs[k + 1]) for k in range(steps)]
    return wall, per


def cpu_baseline(v, f, rays_host, sample: int):
    import oracle  # CPU baseline leg only
    nthreads = max(1, min(16, os.cpu_count() or 1))
    log(f"[bench] cpu baseline: building reference-semantics BBH over {f.shape[0]} tris")
    bbh = oracle.RefBBH(v, f)
    sub = rays_host[:sample].copy()
    bbh.closest(sub[:1024], nthreads=nthreads)  # warm-up
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        bbh.closest(sub, nthreads=nthreads)
        times.append(time.perf_counter() - t0)
    med = sorted(times)[len(times) // 2]
    return {
        "value": round(sample / med / 1e6, 4),
        "unit": "Mrays/s",
        "cores": nthreads,
        "kind": "port",
        "sample": f"{sample} of the {NRAYS} incoherent rays, closest-hit through the oracle's restatement of "
                  f"the reference median-split BBH + stackless walk + plane triangle test, median of 3 runs, "
                  f"{nthreads} threads",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=1 << 16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--leaf-max", type=int, default=4)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU figures are not produced",
                          "n_gpus": args.gpus}))
        return 2

    log("[bench] generating scene and rays")
    v, f = dr.gen_icosphere(SUBDIV, SIGMA, SEED)
    rays = dr.gen_rays_surface(v, f, NRAYS, seed=SEED, inward=True, offset=1e-5)
    cam = dr.gen_rays_camera((0.0, 0.0, 3.5), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, 1024, 1024)

    ctx = dr.Context(0)
    scene = dr.Scene(ctx, v, f)
    log("[bench] building BVH on device")
    scene.build_bvh(leaf_max=args.leaf_max)
    scene.build_bvh(leaf_max=args.leaf_max)  # second build: warm, timed
    info = scene.info()
    d_rays = ctx.upload(rays)
    d_cam = ctx.upload(cam)
    d_hits = ctx.alloc(NRAYS * 16)
    d_occ = ctx.alloc(NRAYS * 4)

    log("[bench] timing incoherent closest-hit")
    wall, per = timed_launches(lambda: scene.trace_closest_dev(d_rays.ptr, NRAYS, d_hits.ptr),
                               args.steps, args.warmup, ctx)
    kernel_ms = float(np.mean(per))
    value = NRAYS * args.steps / wall / 1e6

    _, per_cam = timed_launches(lambda: scene.trace_closest_dev(d_cam.ptr, NRAYS, d_hits.ptr),
                                args.steps, args.warmup, ctx)
    _, per_any = timed_launches(lambda: scene.trace_any_dev(d_rays.ptr, NRAYS, d_occ.ptr),
                                args.steps, args.warmup, ctx)
    hits = d_hits.to_numpy(dr.HIT_DTYPE, NRAYS)  # last launch: camera batch
    occ = d_occ.to_numpy(np.uint32, NRAYS)
    scene.trace_closest_dev(d_rays.ptr, NRAYS, d_hits.ptr)
    ctx.synchronize()
    hits = d_hits.to_numpy(dr.HIT_DTYPE, NRAYS)
    info2 = scene.info()

    roofline = None
    counts = None
    if os.path.exists(COUNTS_FIXTURE):
        with open(COUNTS_FIXTURE) as fh:
            counts = json.load(fh).get("c1_incoherent_closest")
    if counts:
        bytes_per_ray = 48.0 + 32.0 * counts["mean_node_tests"] + 40.0 * counts["mean_tri_tests"]
        achieved = bytes_per_ray * NRAYS / (kernel_ms * 1e-3) / 1e9
        traffic = None
        if os.path.exists(TRAFFIC_FIXTURE):
            with open(TRAFFIC_FIXTURE) as fh:
                traffic = json.load(fh).get("hbm_bytes_per_launch")
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "bytes_per_ray": round(bytes_per_ray, 2), "kernel_ms": round(kernel_ms, 4)}

    base = None
    if not args.no_cpu_baseline:
        base = cpu_baseline(v, f, rays, args.cpu_sample)

    out = {
        "metric": "Mrays/s incoherent closest-hit",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "traversal kernel in isolation: 1,310,720-tri displaced icosphere (subdiv 8, radius "
                        "1+0.05*N(0,1), seed 0), BVH built on device, 2^20 incoherent rays (area-uniform "
                        "surface origins, cosine about the inward normal, seed 0), closest hit {id,t,u,v}",
            "n_triangles": int(f.shape[0]),
            "n_rays": NRAYS,
            "bvh": f"Morton LBVH + SAH leaf collapse, 4-wide, leaf_max {args.leaf_max}",
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "cpu_baseline": base,
        "extra": {
            "kernel_ms_incoherent": round(kernel_ms, 4),
            "coherent_closest_mrays": round(NRAYS / (float(np.mean(per_cam)) * 1e-3) / 1e6, 2),
            "incoherent_any_mrays": round(NRAYS / (float(np.mean(per_any)) * 1e-3) / 1e6, 2),
            "bvh_build_ms": round(info.build_ms, 3),
            "bvh_nodes": info.n_nodes, "bvh_leaves": info.n_leaves, "bvh_depth": info.depth,
            "bvh_sah_cost": round(info.sah_cost, 3),
            "hit_fraction": round(float(np.mean(hits["id"] >= 0)), 6),
            "occluded_fraction": round(float(np.mean(occ)), 6),
            "stack_overflow": int(info2.overflow_flag),
        },
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
