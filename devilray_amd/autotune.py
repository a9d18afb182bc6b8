"""Launch-shape autotuner for batched rendering: the size-class cache around
the library's own sweep (dr_batch_autotune).

The sweep renders the caller's batch with each view's spp cut to a small
budget once per candidate launch shape -- the launch sequence over queue
sizes (paths in flight per chunk), traversal schedules, refill thresholds
and graph recording, and the one persistent launch over queue sizes,
schedules, refill thresholds and resident blocks per CU -- and keeps the
fastest. Results are cached in a JSON file keyed by a size class (log2 of
the paths, of the triangles and of the view count), so a later batch of the
same class reuses the choice without a sweep.
"""
from __future__ import annotations

import json
import math
import os

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "devilray_amd", "autotune.json")


def size_class(total_paths: int, total_tris: int, n_views: int) -> str:
    return "p{}_t{}_v{}".format(int(round(math.log2(max(total_paths, 1)))), int(round(math.log2(max(total_tris, 1)))),
                                int(round(math.log2(max(n_views, 1)))))


def _load(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _save(path, cache):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune(ctx, batch, views, total_tris: int, max_depth: int = 12, rr_depth: int = 3, budget_spp: int = 4,
         cache_path: str | None = DEFAULT_CACHE, reps: int = 1, log=None) -> dict:
    """views: [(scene index, camera, spp, sample_offset, accum pointer)] as for
    Batch.render_dev (the trial renders go to the library's scratch buffer,
    not to the views' own). Returns the chosen options (a dict of render_dev
    keyword arguments) and, under "_sweep", the timings when a sweep ran."""
    total = sum(c.width * c.height * spp for _, c, spp, _, _ in views)
    key = size_class(total, total_tris, len(views))
    cache = _load(cache_path) if cache_path else {}
    if key in cache:
        return dict(cache[key], _cached=True)
    best, sweep = batch.autotune(views, max_depth=max_depth, rr_depth=rr_depth, budget_spp=budget_spp, reps=reps)
    if log:
        for ms, o in sweep:
            log(f"[autotune] {o} {ms:.2f} ms")
    if cache_path:
        cache[key] = best
        _save(cache_path, cache)
    return dict(best, _sweep=[[round(ms, 3), o] for ms, o in sweep])
