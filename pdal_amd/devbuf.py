"""Float64 device arrays on the HIP runtime the libraries load (by soname, so
it is the same runtime instance as theirs): inputs kept resident in HBM for
the entry points that take device addresses (pdal_cp_al_batch_set_vectors_device,
pdal_pod_expand_device)."""
from __future__ import annotations

import ctypes as C

import numpy as np

_H2D, _D2H = 1, 2
_L = None


def _hip():
    global _L
    if _L is None:
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipMalloc.restype = C.c_int
        L.hipFree.argtypes = [C.c_void_p]
        L.hipFree.restype = C.c_int
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemcpy.restype = C.c_int
        L.hipDeviceSynchronize.argtypes = []
        L.hipDeviceSynchronize.restype = C.c_int
        _L = L
    return _L


class DeviceArray:
    """`size` float64 values in device memory of the current device."""

    def __init__(self, size: int):
        self.size = int(size)
        p = C.c_void_p()
        rc = _hip().hipMalloc(C.byref(p), max(self.size, 1) * 8)
        if rc != 0:
            raise RuntimeError(f"hipMalloc of {self.size} doubles failed ({rc})")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        out = cls(a.size)
        out.upload(a)
        return out

    def upload(self, a, offset: int = 0) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        if offset < 0 or offset + a.size > self.size:
            raise ValueError(f"{a.size} values at {offset} do not fit {self.size}")
        rc = _hip().hipMemcpy(C.c_void_p(self.ptr + 8 * offset), a.ctypes.data, a.size * 8, _H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy to the device failed ({rc})")

    def download(self) -> np.ndarray:
        out = np.empty(self.size, dtype=np.float64)
        L = _hip()
        if L.hipDeviceSynchronize() != 0 or L.hipMemcpy(out.ctypes.data, C.c_void_p(self.ptr), self.size * 8,
                                                         _D2H) != 0:
            raise RuntimeError("hipMemcpy from the device failed")
        return out

    def close(self) -> None:
        if self.ptr:
            _hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
