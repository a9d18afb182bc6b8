"""The headless benchmark command line (devilray_amd/cpp/devil_ray_benchmark),
which keeps the reference's arguments, messages and report
(benchmark/src/main.cpp:31-109): argument handling on the CPU, a run on the
reference's suzanne model on the GPU."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "devilray_amd", "cpp", "devil_ray_benchmark")
SUZANNE = os.path.join(ROOT, "tests", "golden", "models", "suzanne.obj")


def run(*args):
    return subprocess.run([BIN, *args], capture_output=True, text=True, timeout=300)


def test_usage_and_argument_errors():
    assert os.access(BIN, os.X_OK), "build with `python -m devilray_amd.build`"
    r = run("--help")
    assert r.returncode == 0 and "Usage:" in r.stdout and "ray_count" in r.stdout
    r = run("a.obj", "10", "extra")
    assert r.returncode == 1 and "too many arguments" in r.stderr
    for bad in ("abc", "-5", "0", "12x"):
        r = run(SUZANNE, bad)
        assert r.returncode == 1 and "ray_count must be" in r.stderr, bad
    r = run(os.path.join(ROOT, "no", "such.obj"), "10")
    assert r.returncode == 1 and "Benchmark failed" in r.stderr and "does not exist" in r.stderr


@pytest.mark.gpu
def test_benchmark_on_suzanne():
    print("[case] benchmark CLI on suzanne", flush=True)
    r = run(SUZANNE, "300000")
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "Loaded mesh 'Monkey_On_Cube':" in r.stdout and "967 tris" in r.stdout
    rays = int(re.search(r"Number of rays: (\d+)", r.stdout).group(1))
    hits = int(re.search(r"Triangle hits: (\d+)", r.stdout).group(1))
    tri = int(re.search(r"Triangle tests: (\d+)", r.stdout).group(1))
    box = int(re.search(r"BBox tests: (\d+)", r.stdout).group(1))
    assert rays == 300000
    assert 0 < hits < rays and tri >= hits and box > rays
    assert "== Benchmark Finished in" in r.stdout
