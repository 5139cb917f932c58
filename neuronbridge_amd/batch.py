"""Host wrapper of the batched multi-instance engine (nbx_batch_* in include/nbx_core.h)."""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib as L


def bf16_to_f32(w: np.ndarray) -> np.ndarray:
    return (w.astype(np.uint32) << 16).view(np.float32)


def dense_normal_bf16(n: int, sigma: float, seed: int, threads: int = 0) -> np.ndarray:
    out = np.empty((n, n), dtype=np.uint16)
    L.call("nbx_build_dense_normal_bf16", n, float(sigma), seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF, threads,
           out.ctypes.data_as(C.POINTER(C.c_uint16)))
    return out


def poisson_kmax(lam: float) -> int:
    """Largest count the inverse-CDF table can return (same rule as the library's table)."""
    p, cdf = math.exp(-lam), 0.0
    for k in range(64):
        cdf += p
        if math.floor(cdf * 4294967296.0) >= 4294967295.0:
            return k
        p = p * lam / (k + 1)
    return 63


def shift_for_bound(bound: float) -> int:
    s = 24
    while s > 0 and bound * math.ldexp(1.0, s) + 1e6 >= 2147483647.0:
        s -= 1
    return s


@dataclass
class Instance:
    n: int
    tau: float
    pois_lambda: float
    seed: int
    fixed_shift: int
    w_bf16: np.ndarray  # uint16 [n][n], pre-major
    v0: np.ndarray | None = None


class Batch:
    def __init__(self, instances: list[Instance], integrator: int, dt: float, t_ref: int, v_rest: float,
                 v_th: float, r: float, v_reset: float, w_bg: float, device: int = 0):
        self.instances = instances
        arr = (L.BatchInstance * len(instances))()
        self._keep = []
        for k, ins in enumerate(instances):
            w = np.ascontiguousarray(ins.w_bf16, dtype=np.uint16)
            v0 = None if ins.v0 is None else np.ascontiguousarray(ins.v0, dtype=np.float32)
            self._keep += [w, v0]
            arr[k] = L.BatchInstance(n=ins.n, tau=ins.tau, pois_lambda=ins.pois_lambda,
                                     seed_lo=ins.seed & 0xFFFFFFFF, seed_hi=(ins.seed >> 32) & 0xFFFFFFFF,
                                     fixed_shift=ins.fixed_shift, w_bf16=w.ctypes.data_as(C.POINTER(C.c_uint16)),
                                     v0=v0.ctypes.data_as(C.POINTER(C.c_float)) if v0 is not None else None)
        d = L.BatchDesc(count=len(instances), integrator=integrator, dt=dt, t_ref=t_ref, v_rest=v_rest, v_th=v_th,
                        r=r, v_reset=v_reset, w_bg=w_bg, inst=arr)
        h = C.c_void_p()
        L.call("nbx_batch_create", C.byref(d), device, C.byref(h))
        self._keep = []  # the library copied everything it needs
        self.h = h
        self._rec = (0, 0)

    def tune(self, steps: int) -> int:
        thr = C.c_int32()
        L.call("nbx_batch_tune", self.h, int(steps), C.byref(thr))
        return thr.value

    def set_schedule(self, threads: int, chunk: int) -> None:
        L.call("nbx_batch_set_schedule", self.h, int(threads), int(chunk))

    def schedule(self) -> tuple[int, int]:
        t, c = C.c_int32(), C.c_int32()
        L.call("nbx_batch_get_schedule", self.h, C.byref(t), C.byref(c))
        return t.value, c.value

    def set_parts(self, parts) -> None:
        """Split schedule: parts[i] column parts for instance i (see nbx_batch_set_parts)."""
        arr = np.ascontiguousarray(parts, dtype=np.int32)
        if arr.shape != (len(self.instances),):
            raise ValueError("one part count per instance")
        L.call("nbx_batch_set_parts", self.h, arr.ctypes.data_as(C.POINTER(C.c_int32)), arr.size)

    def parts(self) -> np.ndarray:
        """Parts per instance in use (all 0 under the one-workgroup-per-instance schedule)."""
        out = np.zeros(len(self.instances), np.int32)
        L.call("nbx_batch_get_parts", self.h, out.ctypes.data_as(C.POINTER(C.c_int32)), out.size)
        return out

    def record(self, t_begin: int, nsteps: int) -> None:
        L.call("nbx_batch_record", self.h, int(t_begin), int(nsteps))
        self._rec = (t_begin, nsteps)

    def run(self, nsteps: int) -> None:
        L.call("nbx_batch_run", self.h, int(nsteps))

    def run_async(self, nsteps: int) -> None:
        L.call("nbx_batch_run_async", self.h, int(nsteps))

    def sync(self) -> None:
        L.call("nbx_batch_sync", self.h)

    def raster(self, k: int) -> np.ndarray:
        words = (self.instances[k].n + 31) // 32
        out = np.zeros((self._rec[1], words), dtype=np.uint32)
        L.call("nbx_batch_get_raster", self.h, int(k), out.ctypes.data_as(C.POINTER(C.c_uint32)), out.size)
        return out

    def state(self, k: int) -> tuple[np.ndarray, np.ndarray]:
        n = self.instances[k].n
        v = np.zeros(n, np.float32)
        last = np.zeros(n, np.int32)
        L.call("nbx_batch_get_state", self.h, int(k), v.ctypes.data_as(C.POINTER(C.c_float)),
               last.ctypes.data_as(C.POINTER(C.c_int32)))
        return v, last

    def events(self) -> int:
        e = C.c_int64()
        L.call("nbx_batch_events", self.h, C.byref(e))
        return e.value

    @property
    def stream(self) -> int:
        return int(L.lib().nbx_batch_stream(self.h) or 0)

    def close(self) -> None:
        if getattr(self, "h", None):
            L.lib().nbx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
