"""The allocation-free POD expansion (pdal_pod_expand_device_noalloc) against
pdal_pod_expand_device, through the C ABI, on two calls in a row (the second
with fewer members, so stale staging contents would show), with and without
clipping bounds; and its refusal of a call beyond the staging area.

Device output buffers come from the HIP runtime the libraries themselves
load (looked up by soname, so it is the same runtime instance)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def alloc_filled(self, n: int, value: float) -> int:
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), n * 8) == 0
        host = np.full(n, value, dtype=np.float64)
        assert self.L.hipMemcpy(p, host.ctypes.data, n * 8, H2D) == 0
        return p.value

    def read(self, ptr: int, n: int) -> np.ndarray:
        host = np.empty(n, dtype=np.float64)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(host.ctypes.data, C.c_void_p(ptr), n * 8, D2H) == 0
        return host

    def free(self, ptr: int) -> None:
        self.L.hipFree(C.c_void_p(ptr))


def _case(rng, k, dim):
    basis = rng.standard_normal((k, dim)).astype(np.float32)
    mean = rng.standard_normal(dim).astype(np.float32)
    return basis, mean


@pytest.mark.parametrize("bounds", [False, True])
def test_noalloc_expansion_equals_allocating_one(bounds):
    from pdal_amd.warmstart import DevicePod, load_ws
    ws = load_ws()
    rng = np.random.default_rng(11)
    k, dim = 96, 3001
    basis, mean = _case(rng, k, dim)
    lo = np.full(dim, -0.5) if bounds else None
    hi = np.full(dim, 2.0) if bounds else None
    pod = DevicePod(ws, basis, mean, lo, hi)
    hip = _Hip()
    cap = 300 * dim
    a, b = hip.alloc_filled(cap, np.nan), hip.alloc_filled(cap, np.nan)
    try:
        for batch in (300, 17):  # second call smaller: every entry must come from this call
            print(f"[case] noalloc expansion k={k} dim={dim} batch={batch} bounds={bounds}", flush=True)
            coef = rng.standard_normal((batch, k)).astype(np.float32)
            pod.expand(coef, a, cap)
            assert pod.expand_noalloc(coef, b, cap) == 0
            ha, hb = hip.read(a, cap), hip.read(b, cap)
            np.testing.assert_array_equal(ha[: batch * dim], hb[: batch * dim])
            ref = mean.astype(np.float64) + coef.astype(np.float64) @ basis.astype(np.float64)
            if bounds:
                ref = np.clip(ref, lo, hi)
            np.testing.assert_allclose(hb[: batch * dim].reshape(batch, dim), ref, rtol=1e-4, atol=1e-4)
            if batch == 300:
                assert not np.isnan(hb).any()
            else:  # entries past the second expansion still hold the first call's values
                np.testing.assert_array_equal(hb[batch * dim:], prev[batch * dim:])
            prev = hb
    finally:
        hip.free(a)
        hip.free(b)
        pod.close()


def test_noalloc_expansion_refuses_beyond_staging_area():
    from pdal_amd.warmstart import DevicePod, load_ws
    ws = load_ws()
    rng = np.random.default_rng(12)
    k, dim = 256, 64
    basis, mean = _case(rng, k, dim)
    pod = DevicePod(ws, basis, mean)
    hip = _Hip()
    batch = 513  # 513 x 256 > 131072 staged floats
    out = hip.alloc_filled(batch * dim, 7.0)
    try:
        coef = np.zeros((batch, k), dtype=np.float32)
        print("[case] noalloc expansion refusals", flush=True)
        assert pod.expand_noalloc(coef, out, batch * dim) == -3
        assert pod.expand_noalloc(coef[:512], out, 10) == -1  # output capacity too small
        assert (hip.read(out, batch * dim) == 7.0).all()
        assert pod.expand_noalloc(coef[:512], out, batch * dim) == 0  # 512 x 256 fits exactly
        got = hip.read(out, batch * dim)
        np.testing.assert_array_equal(got[: 512 * dim].reshape(512, dim), np.tile(mean.astype(np.float64), (512, 1)))
        assert (got[512 * dim:] == 7.0).all()
    finally:
        hip.free(out)
        pod.close()
