"""Benchmark: incoherent closest-hit traversal throughput (Mrays/s) on one MI355X.

Workload (one "step" = one dr_trace_closest launch over the whole batch):
  1,310,720-triangle icosphere (subdivision 8) with per-vertex radius
  1 + 0.05*N(0,1) (Philox, seed 0), BVH built on the device; 2^20 rays with
  origins area-uniform on the surface and directions cosine-distributed about
  the inward face normal (seed 0), origin offset 1e-5 along the direction.
Inputs are resident in HBM before the timed region. "extra" reports the
coherent 1024x1024 camera batch, any-hit, the BVH build, and the two render
workloads (Cornell box 256x256 x 16 spp; 2,048,000-triangle many-lights scene
at 1920x1080, timed over a few samples per pixel).

Prints one JSON line. `roofline.achieved` uses the algorithmic bytes per ray
48 + 32*N_node + 40*N_tri with N from the canonical binary SAH BVH counts
committed in tests/golden/roofline_counts.json, divided by the kernel's mean
HIP-event duration on its stream; `traffic` is the measured HBM bytes per
launch from the committed rocprofv3 summary. `cpu_baseline` times the
oracle's reference-semantics traversal (median-split BBH, stackless walk) on a
bounded sample of the same rays on the host cores.

--probe: runs only the traversal launches and a 256 MiB device copy (the
known-size kernel that calibrates FETCH_SIZE / WRITE_SIZE); for rocprofv3.
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
from devilray_amd import scenes  # noqa: E402
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


def render_rate(ctx, sc, spp, max_depth, rr_depth, reps, width=None, height=None):
    """Samples per second of dr_render on a scene (FP64 accumulation)."""
    s = dr.Scene(ctx, sc.vertices, sc.indices, sc.tri_material, sc.dr_materials())
    s.build_bvh(leaf_max=4)
    c = sc.camera
    w, h = width or c["width"], height or c["height"]
    cam = dr.camera_lookat(c["eye"], c["target"], c["up"], c["fov"], w, h)
    acc = ctx.alloc(w * h * 32)
    acc.copy_from(np.zeros(w * h * 4))
    p = dr.DrRenderParams(spp, 0, max_depth, rr_depth, 0, 0, 0)
    s.render_dev(cam, p, acc.ptr)  # builds the light table, warms up
    ctx.synchronize()
    e0, e1 = Event(), Event()
    e0.record()
    for k in range(reps):
        p.sample_offset = spp * (k + 1)
        s.render_dev(cam, p, acc.ptr)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    img = acc.to_numpy(np.float64, w * h * 4).reshape(h, w, 4)
    mean = float((img[..., :3] / np.maximum(img[..., 3:], 1)).mean())
    s.close()
    return {"ms_per_pass": round(ms, 3), "spp_per_pass": spp, "samples_per_s": round(w * h * spp / ms * 1e3, 1),
            "image_mean": round(mean, 6), "resolution": [w, h], "max_depth": max_depth, "rr_depth": rr_depth}


def probe() -> int:
    """Traversal launches + a calibration copy, for rocprofv3 counter runs."""
    v, f = dr.gen_icosphere(SUBDIV, SIGMA, SEED)
    rays = dr.gen_rays_surface(v, f, NRAYS, seed=SEED, inward=True, offset=1e-5)
    ctx = dr.Context(0)
    scene = dr.Scene(ctx, v, f)
    scene.build_bvh(leaf_max=4)
    d_rays = ctx.upload(rays)
    d_hits = ctx.alloc(NRAYS * 16)
    for _ in range(3):
        scene.trace_closest_dev(d_rays.ptr, NRAYS, d_hits.ptr)
    ctx.synchronize()
    a, b = ctx.alloc(256 << 20), ctx.alloc(256 << 20)
    dr.check(dr.lib().dr_memcpy_dtod(ctx.handle, dr.C.c_void_p(b.ptr), dr.C.c_void_p(a.ptr), 256 << 20, None))
    ctx.synchronize()
    print("probe done", flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=1 << 16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-render", action="store_true")
    ap.add_argument("--leaf-max", type=int, default=4)
    ap.add_argument("--probe", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU figures are not produced",
                          "n_gpus": args.gpus}))
        return 2
    if args.probe:
        return probe()

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

    renders = {}
    if not args.no_render:
        log("[bench] render: Cornell box 256x256")
        renders["cornell_256_16spp"] = render_rate(ctx, scenes.cornell_c0(), 16, 8, 3, 5)
        log("[bench] render: many-lights scene 1920x1080")
        renders["many_lights_1080p"] = render_rate(ctx, scenes.many_lights(), 2, 12, 3, 2)

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
            "render": renders,
        },
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
