"""Autotuner bookkeeping (CPU): size classes and the JSON cache."""
import json

from devilray_amd import autotune


def test_size_class_buckets():
    assert autotune.size_class(1 << 20, 1 << 16, 64) == "p20_t16_v6"
    assert autotune.size_class(3 << 20, 1000, 1) == "p22_t10_v0"


def test_cache_hit_skips_the_sweep(tmp_path):
    path = str(tmp_path / "tune.json")
    key = autotune.size_class(100, 100, 1)
    chosen = {"chunk_paths": 1 << 21, "trace_sched": 1, "refill_min": 16, "use_graph": True}
    with open(path, "w") as fh:
        json.dump({key: chosen}, fh)

    class Cam:
        width, height = 10, 10

    # a cached class returns without touching the batch or the context
    got = autotune.tune(None, None, [(0, Cam(), 1, 0, 0)], 100, cache_path=path)
    assert got["_cached"] and {k: got[k] for k in chosen} == chosen
