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
def test_facade_reference_cases(tmp_path):
    models = os.path.join(ROOT, "tests", "golden", "models")
    res = subprocess.run([BIN, str(tmp_path), models], capture_output=True, text=True, timeout=240)
    print(res.stdout)
    lines = [ln for ln in res.stdout.splitlines() if ln.startswith(("PASS", "FAIL"))]
    assert len(lines) >= 76
    assert res.returncode == 0, res.stdout + res.stderr
    assert all(ln.startswith("PASS") for ln in lines)
    # Renderer::saveImage: a PNG whose pixels are the display pixels
    import numpy as np
    from devilray_amd import image
    raw = np.fromfile(tmp_path / "winding.rgba", np.uint8).reshape(64, 64, 4)
    png = image.read_png(str(tmp_path / "winding.png"))
    assert png.shape == (64, 64, 4) and np.array_equal(png, raw)
    import struct
    import zlib
    data = (tmp_path / "winding.png").read_bytes()
    pos, types = 8, []
    while pos < len(data):
        ln, typ = struct.unpack(">I4s", data[pos:pos + 8])
        crc = struct.unpack(">I", data[pos + 8 + ln:pos + 12 + ln])[0]
        assert crc == zlib.crc32(data[pos + 4:pos + 8 + ln]) & 0xFFFFFFFF, typ
        types.append(typ)
        pos += 12 + ln
    assert types == [b"IHDR", b"IDAT", b"IEND"]
