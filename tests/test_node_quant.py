"""The traversal's node boxes are quantised to one byte per plane
(devilray_amd/csrc/node_quant.hpp). Hits stay exact only if every quantised
box contains the float box it replaces and the planes decode without
rounding; tests/quant_check.cpp checks both, that every plane is the
tightest grid point that bounds it and that the step is within its stated
slack of span / 255, on random and edge-case boxes (among them planes that
only the exact grid-point comparisons keep inside their boxes), compiled
with the host compiler."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_quantised_boxes_contain_and_decode_exactly(tmp_path):
    exe = str(tmp_path / "quant_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "quant_check.cpp")], check=True)
    out = subprocess.run([exe, "200000"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("ok "), out.stdout + out.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_box_test_is_conservative_and_culls(tmp_path):
    """The traversal's box test on quantised nodes (exactly decoded planes,
    robust slab test): tests/slab_check.cpp shoots random and grazing rays
    (aimed at the edges and corners of the boxes; zero and tiny direction
    components) and checks that every child the exact ray meets is taken and
    every clear miss is culled."""
    exe = str(tmp_path / "slab_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "slab_check.cpp")],
                   check=True)
    out = subprocess.run([exe, "40000"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("ok "), out.stdout + out.stderr
