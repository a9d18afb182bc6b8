"""The C++ host facade (devilray_amd/cpp: Renderer / Scene / cast over the C
ABI) running the reference's own intersection and render test cases."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "devilray_amd", "cpp", "test_facade")


def test_facade_library_loads():
    import ctypes
    lib = os.path.join(ROOT, "devilray_amd", "libdevilray_cpp.so")
    if not os.path.exists(lib):
        pytest.fail("libdevilray_cpp.so missing: run `python -m devilray_amd.build`")
    ctypes.CDLL(lib)
    assert os.access(BIN, os.X_OK)


@pytest.mark.gpu
def test_facade_reference_cases():
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(res.stdout)
    lines = [ln for ln in res.stdout.splitlines() if ln.startswith(("PASS", "FAIL"))]
    assert len(lines) >= 15
    assert res.returncode == 0, res.stdout + res.stderr
    assert all(ln.startswith("PASS") for ln in lines)
