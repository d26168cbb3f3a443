"""The allocation-free entry point psh_dot_noalloc, through the C-ABI, against psh_dot.

psh_dot_noalloc takes the caller's device arrays and keeps its partial sums in a static device
array of the library; its result must be psh_dot's (the same reduction of the same rounded
products) in every reduction mode, on two calls in a row, the second on a shorter input (the
scratch keeps the first call's partials). Device buffers for the test come from the HIP runtime
the library itself links (libamdhip64.so.7); the entry point allocates none.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    from ps26119_amd import psh
    psh.lib()  # loads libps26119_hip.so and with it the HIP runtime this test allocates through
    H = ctypes.CDLL("libamdhip64.so.7")
    P = ctypes.c_void_p
    H.hipMalloc.argtypes = [ctypes.POINTER(P), ctypes.c_size_t]
    H.hipFree.argtypes = [P]
    H.hipMemcpy.argtypes = [P, P, ctypes.c_size_t, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    for f in (H.hipMalloc, H.hipFree, H.hipMemcpy, H.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return H


class DevArray:
    H2D, D2H = 1, 2

    def __init__(self, H, host):
        self.H = H
        self.n = len(host)
        self.p = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(self.p), max(8, 8 * self.n)) == 0
        if self.n:
            h = np.ascontiguousarray(host, dtype=np.float64)
            assert H.hipMemcpy(self.p, h.ctypes.data_as(ctypes.c_void_p), 8 * self.n, self.H2D) == 0

    def get(self):
        out = np.empty(self.n, dtype=np.float64)
        assert self.H.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, 8 * self.n, self.D2H) == 0
        return out

    def free(self):
        self.H.hipFree(self.p)


def test_dot_noalloc_equals_psh_dot(gpu_available, oracle):
    from ps26119_amd import lpgen, psh
    L = psh.lib()
    H = _hip()
    lp = lpgen.random_lp(150_000, seed=2)  # n = 300,000 primal, m = 150,000 dual entries
    ctx = psh.Context(psh.scaled_problem(oracle.scale(lp), lp))
    rng = np.random.default_rng(5)
    x1, x2 = rng.standard_normal(lp.n) * 1e3, rng.standard_normal(lp.n)
    y1, y2 = rng.standard_normal(lp.m), rng.standard_normal(lp.m) * 1e-3
    ctx.upload(psh.X, x1)
    ctx.upload(psh.XH, x2)
    ctx.upload(psh.Y, y1)
    ctx.upload(psh.Y0, y2)
    big = (DevArray(H, x1), DevArray(H, x2), psh.X, psh.XH, lp.n)
    small = (DevArray(H, y1), DevArray(H, y2), psh.Y, psh.Y0, lp.m)
    out = DevArray(H, np.full(1, np.nan))
    try:
        for mode in (psh.RED_TREE, psh.RED_CPU, psh.RED_DECISION):
            for a, b, sa, sb, n in (big, small):  # two calls in a row, the second on the shorter input
                rc = L.psh_dot_noalloc(a.p, b.p, n, mode, out.p, None)
                assert rc == 0, L.psh_noalloc_error().decode()
                assert H.hipDeviceSynchronize() == 0
                got = out.get()[0]
                exp = ctx.dot(sa, sb, mode)
                print(f"mode {mode} n={n}: noalloc {got!r} psh_dot {exp!r}", flush=True)
                assert float(got).hex() == float(exp).hex(), (mode, n)
        # sizes beyond the stated maximum and bad arguments are refused with a message
        assert L.psh_dot_noalloc(big[0].p, big[1].p, -1, psh.RED_TREE, out.p, None) == -3
        assert b"PSH_NOALLOC_MAX_N" in L.psh_noalloc_error()
        assert L.psh_dot_noalloc(big[0].p, big[1].p, 1 << 31, psh.RED_TREE, out.p, None) == -3
        assert L.psh_dot_noalloc(big[0].p, big[1].p, 10, 7, out.p, None) == -3
        assert L.psh_dot_noalloc(None, big[1].p, 10, psh.RED_TREE, out.p, None) == -3
        # an empty input sums to +0.0
        rc = L.psh_dot_noalloc(None, None, 0, psh.RED_CPU, out.p, None)
        assert rc == 0 and H.hipDeviceSynchronize() == 0
        assert float(out.get()[0]).hex() == (0.0).hex()
    finally:
        for d in (big[0], big[1], small[0], small[1], out):
            d.free()
        ctx.close()
