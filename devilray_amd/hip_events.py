"""Minimal HIP event timing through ctypes (libamdhip64), for measuring
kernels on the stream they are launched on without pulling in torch."""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os

_hip = None


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        cands = [os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"),
                 ctypes.util.find_library("amdhip64") or "libamdhip64.so"]
        last = None
        for c in cands:
            try:
                _hip = C.CDLL(c)
                break
            except OSError as e:  # try the next candidate
                last = e
        if _hip is None:
            raise OSError(f"libamdhip64 not found: {last}")
        _hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipEventSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip.hipEventDestroy.argtypes = [C.c_void_p]
    return _hip


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        rc = hip().hipEventCreate(C.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"hipEventCreate failed: {rc}")

    def record(self, stream=None):
        rc = hip().hipEventRecord(self.h, stream)
        if rc != 0:
            raise RuntimeError(f"hipEventRecord failed: {rc}")

    def synchronize(self):
        hip().hipEventSynchronize(self.h)

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        rc = hip().hipEventElapsedTime(C.byref(ms), self.h, end.h)
        if rc != 0:
            raise RuntimeError(f"hipEventElapsedTime failed: {rc}")
        return float(ms.value)

    def __del__(self):
        try:
            if self.h:
                hip().hipEventDestroy(self.h)
        except Exception:
            pass
