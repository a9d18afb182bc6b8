"""Display conversion and PNG I/O for accumulation buffers.

create_pixels restates Renderer::createPixels (src/Renderer.cpp:51-90):
mean = sum / count per pixel (0 where no samples), optional sRGB transfer
(clamped), x255, round half away from zero, pack r | g<<8 | b<<16 | a<<24,
rows flipped so that the last raster row comes first. PNG reading/writing
stands in for stb (src/Image.cpp:12-23): 8-bit RGBA, zlib via the standard
library.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np


def to_srgb(linear: np.ndarray) -> np.ndarray:
    r = np.clip(linear, 0.0, 1.0).astype(np.float32)
    return np.where(r < 0.0031308, r * np.float32(12.92),
                    np.power(r, np.float32(1 / 2.4)) * np.float32(1.055) - np.float32(0.055))


def create_pixels(accum: np.ndarray, gamma_corrected: bool = True) -> np.ndarray:
    """accum (H, W, 4) float64 sums -> (H, W) uint32 RGBA8, y-flipped."""
    cnt = accum[..., 3:4]
    mean = np.where(cnt > 0, accum[..., :3] / np.where(cnt > 0, cnt, 1), 0.0).astype(np.float32)
    c = to_srgb(mean) if gamma_corrected else mean
    q = np.floor(c * np.float32(255) + np.float32(0.5)).astype(np.int64)  # std::round for x >= 0
    q = np.clip(q, 0, 0xFFFFFFFF).astype(np.uint64)
    px = (q[..., 0] | (q[..., 1] << 8) | (q[..., 2] << 16) | (np.uint64(255) << 24)) & 0xFFFFFFFF
    return px.astype(np.uint32)[::-1].copy()


def unpack_rgba(px: np.ndarray) -> np.ndarray:
    return np.stack([(px >> s) & 0xFF for s in (0, 8, 16, 24)], axis=-1).astype(np.uint8)


def write_png(path: str, rgba: np.ndarray) -> None:
    rgba = np.ascontiguousarray(rgba, np.uint8)
    h, w, _ = rgba.shape
    raw = b"".join(b"\x00" + rgba[y].tobytes() for y in range(h))

    def chunk(t, body):
        return struct.pack(">I", len(body)) + t + body + struct.pack(">I", zlib.crc32(t + body) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) +
                chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


def read_png(path: str) -> np.ndarray:
    """8-bit RGBA (colour type 6) or RGB (type 2), non-interlaced -> (H, W, 4) uint8."""
    data = open(path, "rb").read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG file")
    pos, idat, w, h, ct = 8, b"", 0, 0, 6
    while pos < len(data):
        ln, typ = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + ln]
        pos += 12 + ln
        if typ == b"IHDR":
            w, h, bd, ct, _, _, inter = struct.unpack(">IIBBBBB", body)
            if bd != 8 or ct not in (2, 6) or inter != 0:
                raise ValueError("unsupported PNG format")
        elif typ == b"IDAT":
            idat += body
        elif typ == b"IEND":
            break
    bpp = 4 if ct == 6 else 3
    stride = w * bpp
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, stride + 1)
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        ft, line = raw[y, 0], raw[y, 1:].astype(np.int32)
        if ft == 0:
            cur = line
        elif ft == 2:
            cur = (line + prev) & 255
        else:
            cur = np.zeros(stride, np.int32)
            for x in range(stride):
                a = cur[x - bpp] if x >= bpp else 0
                b = prev[x]
                c = prev[x - bpp] if x >= bpp else 0
                if ft == 1:
                    p = a
                elif ft == 3:
                    p = (a + b) // 2
                else:
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                cur[x] = (line[x] + p) & 255
        out[y] = cur
        prev = cur.astype(np.int32)
    img = out.reshape(h, w, bpp)
    if bpp == 3:
        img = np.concatenate([img, np.full((h, w, 1), 255, np.uint8)], axis=2)
    return img
