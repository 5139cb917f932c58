"""HIP event timing on an arbitrary stream (ctypes on libamdhip64, no torch needed)."""
from __future__ import annotations

import ctypes as C
import os

_hip = None


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        for cand in ("libamdhip64.so", os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")):
            try:
                _hip = C.CDLL(cand, mode=C.RTLD_GLOBAL)
                break
            except OSError:
                continue
        if _hip is None:
            raise ImportError("libamdhip64.so not found")
        _hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipEventSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip.hipEventDestroy.argtypes = [C.c_void_p]
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        _ok(hip().hipEventCreate(C.byref(self.h)), "hipEventCreate")

    def record(self, stream: int) -> None:
        _ok(hip().hipEventRecord(self.h, C.c_void_p(stream)), "hipEventRecord")

    def synchronize(self) -> None:
        _ok(hip().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _ok(hip().hipEventElapsedTime(C.byref(ms), self.h, end.h), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            hip().hipEventDestroy(self.h)
        except Exception:
            pass


class DeviceBuffer:
    """Raw device allocation for device-resident inputs of timing runs."""

    def __init__(self, nbytes: int):
        self.ptr = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(self.ptr), max(1, nbytes)), "hipMalloc")
        self.nbytes = nbytes

    def upload(self, arr) -> None:
        _ok(hip().hipMemcpy(self.ptr, arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1), "hipMemcpy H2D")

    def download(self, arr) -> None:
        _ok(hip().hipMemcpy(arr.ctypes.data_as(C.c_void_p), self.ptr, arr.nbytes, 2), "hipMemcpy D2H")

    def __del__(self):
        try:
            hip().hipFree(self.ptr)
        except Exception:
            pass
