"""The allocation-free broadening entry point (nrgk_broaden_noalloc) through the C ABI.  The caller's
device buffers are allocated here with the HIP runtime (test plumbing); the entry point enqueues on
the caller's stream and allocates nothing.  Checked bit for bit against nrgk_broaden (the allocating
entry point: the same kernels and summation order) on two calls in a row, the second smaller, so
that partial sums a larger call left in the fixed scratch cannot leak into a smaller one; a size
beyond the fixed scratch is rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


class _Dev:
    """a device buffer of doubles owned by the caller"""

    def __init__(self, hip, a):
        self.hip, self.n = hip, len(a)
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(8, 8 * self.n)) == 0
        if self.n:
            assert hip.hipMemcpy(self.p, np.ascontiguousarray(a, np.float64).ctypes.data, 8 * self.n, 1) == 0

    def get(self):
        out = np.zeros(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, 8 * self.n, 2) == 0
        return out

    def free(self):
        assert self.hip.hipFree(self.p) == 0


def test_noalloc_broadening_equals_the_allocating_entry(nk):
    L = nk.lib()
    L.nrgk_broaden_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_double,
                                       C.c_void_p, C.c_void_p]
    L.nrgk_broaden_noalloc.restype = C.c_int
    ctx = nk.default_context()
    hip = _hip()
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    rng = np.random.default_rng(17)
    for n, no in ((20000, 1500), (700, 300)):
        print("noalloc broadening n=%d n_omega=%d" % (n, no), flush=True)
        dele = (n - 1.0) / np.log(10 / 1e-10)
        x = 1e-10 * np.exp(np.arange(n) / dele)
        w = rng.uniform(0, 1, n) * (rng.uniform(0, 1, n) < 0.3)
        om = np.logspace(-9, 0.5, no)
        dx, dw, do, dout = _Dev(hip, x), _Dev(hip, w), _Dev(hip, om), _Dev(hip, np.full(no, np.nan))
        rc = L.nrgk_broaden_noalloc(ctx._h, dx.p, dw.p, n, do.p, no, 0.6, dout.p, stream)
        assert rc == 0, L.nrgk_last_error().decode()
        assert hip.hipStreamSynchronize(stream) == 0
        got = dout.get()
        ref = nk.broaden(x, w, om, 0.6)
        assert np.all(np.isfinite(got)) and np.abs(ref).max() > 0
        np.testing.assert_array_equal(got, ref)
        for d in (dx, dw, do, dout):
            d.free()
    # ceil(600000 / 512) * 2000 partial sums exceed the fixed scratch: rejected, nothing enqueued
    dx = _Dev(hip, np.ones(4))
    rc = L.nrgk_broaden_noalloc(ctx._h, dx.p, dx.p, 600000, dx.p, 2000, 0.6, dx.p, stream)
    assert rc == -1 and "fixed scratch" in L.nrgk_last_error().decode()
    dx.free()
    assert hip.hipStreamDestroy(stream) == 0
