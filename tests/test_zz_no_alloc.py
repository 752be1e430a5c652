"""nbg_accel_direct_noalloc (no counterpart in the reference: the direct sum for callers
that cannot allocate device memory) against the context entry point nbg_accel with
NBG_SOLVER_DIRECT, through the C ABI: the same kernels, so the same bits.  Two calls in a
row on one device, the second with fewer particles than the first (the shared staging
array still holds the first call's records beyond the second's n), plus the argument
checks, which return before any device work."""
import ctypes
import os

import numpy as np
import pytest

from conftest import plummer_f32

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIP_MEMCPY_H2D, HIP_MEMCPY_D2H = 1, 2


def _hip():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


class _DevArrays:
    """Caller-owned device arrays (the caller allocates; the entry point does not)."""

    def __init__(self, hip, n, k):
        self.hip, self.n = hip, n
        self.ptrs = []
        for _ in range(k):
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), 4 * n) == 0
            self.ptrs.append(p)

    def put(self, i, a):
        a = np.ascontiguousarray(a, np.float32)
        assert self.hip.hipMemcpy(self.ptrs[i], a.ctypes.data, 4 * self.n, HIP_MEMCPY_H2D) == 0

    def get(self, i):
        a = np.empty(self.n, np.float32)
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptrs[i], 4 * self.n, HIP_MEMCPY_D2H) == 0
        return a

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)


def test_noalloc_direct_equals_context_direct_twice_second_smaller(gpu_ctx_factory):
    import pynbg
    L = pynbg.load()
    L.nbg_accel_direct_noalloc.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(pynbg.Params),
                                           ctypes.c_int64] + [ctypes.c_void_p] * 7
    hip = _hip()
    for n, seed, kw in ((20000, 1, dict(eps=0.01)), (7001, 2, dict(eps=0.02, softening="spline"))):
        s = plummer_f32(n, seed)
        p = pynbg.make_params("direct", G=1.3, **kw)
        ctx = gpu_ctx_factory(n)
        ctx.upload(s["x"], s["y"], s["z"], s["m"])
        ctx.accel(p)
        want = ctx.accelerations()
        d = _DevArrays(hip, n, 7)
        try:
            for i, k in enumerate(("x", "y", "z", "m")):
                d.put(i, s[k])
            r = L.nbg_accel_direct_noalloc(0, None, ctypes.byref(p), n, *d.ptrs)
            assert r == 0, r
            assert hip.hipDeviceSynchronize() == 0
            got = np.stack([d.get(4), d.get(5), d.get(6)], axis=1)
        finally:
            d.free()
        print(f"n={n}: max |diff| {np.abs(got - want).max():.3e}")
        np.testing.assert_array_equal(got, want)


def test_noalloc_rejects_bad_arguments():
    import pynbg
    L = pynbg.load()
    L.nbg_accel_direct_noalloc.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(pynbg.Params),
                                           ctypes.c_int64] + [ctypes.c_void_p] * 7
    hip = _hip()
    d = _DevArrays(hip, 16, 7)
    try:
        p = pynbg.make_params("direct", G=1.0, eps=0.01)
        for n in (0, -1, (1 << 20) + 1):
            assert L.nbg_accel_direct_noalloc(0, None, ctypes.byref(p), n, *d.ptrs) == -1, n
        bh = pynbg.make_params("barnes_hut", G=1.0, eps=0.01)
        assert L.nbg_accel_direct_noalloc(0, None, ctypes.byref(bh), 16, *d.ptrs) == -1
        assert L.nbg_accel_direct_noalloc(0, None, None, 16, *d.ptrs) == -1
        assert L.nbg_accel_direct_noalloc(0, None, ctypes.byref(p), 16, None, *d.ptrs[1:]) == -1
        assert L.nbg_accel_direct_noalloc(999, None, ctypes.byref(p), 16, *d.ptrs) == -1
    finally:
        d.free()
