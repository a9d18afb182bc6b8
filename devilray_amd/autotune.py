"""Launch-shape autotuner for batched rendering.

Sweeps the batch options that set how the wavefront runs -- paths in flight
per chunk (the queue size), the traversal schedule of the bounce traces, the
persistent refill threshold, and whether the bounce loop is recorded as a
graph -- on the caller's own batch with each view's spp cut to a small
budget, and keeps the fastest. Results are cached in a JSON file keyed by a
size class (log2 of the paths, of the triangles and of the view count), so a
later batch of the same class reuses the choice without a sweep.
"""
from __future__ import annotations

import itertools
import json
import math
import os

import numpy as np

from .hip_events import Event

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "devilray_amd", "autotune.json")

CHUNKS = (1 << 20, 1 << 21, 1 << 22, 1 << 23)
SCHEDS = (1, 3)           # per-ray speculative, persistent speculative
REFILLS = (16,)
GRAPHS = (False, True)


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
    Batch.render_dev (the trial renders go to scratch accumulation buffers,
    not to the views' own). Returns the chosen options (a dict of render_dev
    keyword arguments) and, under "_sweep", the timings when a sweep ran."""
    total = sum(c.width * c.height * spp for _, c, spp, _, _ in views)
    key = size_class(total, total_tris, len(views))
    cache = _load(cache_path) if cache_path else {}
    if key in cache:
        return dict(cache[key], _cached=True)
    npix = [c.width * c.height for _, c, _, _, _ in views]
    scratch = ctx.alloc(32 * sum(npix))
    offs = np.concatenate([[0], np.cumsum(npix)[:-1]]) * 32
    tv = [(s, c, min(spp, budget_spp), off, scratch.ptr + int(o))
          for (s, c, spp, off, _), o in zip(views, offs)]
    sweep = []
    for chunk, sched, refill, graph in itertools.product(CHUNKS, SCHEDS, REFILLS, GRAPHS):
        opts = dict(chunk_paths=chunk, trace_sched=sched, refill_min=refill, use_graph=graph)
        times = []
        for _ in range(reps + 1):  # the first run sizes buffers and is not timed
            ctx.synchronize()
            e0, e1 = Event(), Event()
            e0.record()
            batch.render_dev(tv, max_depth=max_depth, rr_depth=rr_depth, **opts)
            e1.record()
            e1.synchronize()
            times.append(e0.elapsed_ms(e1))
        ms = float(np.median(times[1:]))
        sweep.append((ms, opts))
        if log:
            log(f"[autotune] {opts} {ms:.2f} ms")
    scratch.free()
    best = min(sweep, key=lambda x: x[0])[1]
    if cache_path:
        cache[key] = best
        _save(cache_path, cache)
    return dict(best, _sweep=[[round(ms, 3), o] for ms, o in sweep])
