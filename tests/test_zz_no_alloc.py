"""The allocation-free validation entry point (pivx_median_filter_smoothing_noalloc)
against the handle's pivx_median_filter_smoothing_p2 on the same raw
candidates: bit-identical filtered fields on two calls in a row, the second
on a smaller field, plus its size limits.  Device buffers come from the HIP
runtime the library itself uses (libamdhip64 through ctypes)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.ptrs = []

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.ptrs.append(p)
        return p.value

    def h2d(self, dst, a):
        a = np.ascontiguousarray(a)
        assert self.L.hipMemcpy(C.c_void_p(dst), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def d2h(self, a, src):
        assert self.L.hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(src), C.c_size_t(a.nbytes), 2) == 0

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free_all(self):
        for p in self.ptrs:
            self.L.hipFree(p)
        self.ptrs = []


def _candidates(rng, nt, K):
    dx = (2.0 + 0.3 * rng.standard_normal((nt, K))).astype(np.float32)
    dy = (-1.0 + 0.3 * rng.standard_normal((nt, K))).astype(np.float32)
    out = rng.random(nt) < 0.1
    dx[out, 0] += rng.uniform(-8, 8, out.sum()).astype(np.float32)
    dy[out, 0] += rng.uniform(-8, 8, out.sum()).astype(np.float32)
    v = (rng.random((nt, K)) > 0.2).astype(np.uint8)
    return dx, dy, v


@pytest.mark.parametrize("npm,nps,fill", [(2, 1, 1), (1, 2, 1), (3, 0, 0), (0, 1, 1)])
def test_noalloc_filter_equals_handle_filter(pivx, npm, nps, fill):
    """Two calls in a row (a 6600x4400 grid-2 field of 113,300 vectors, then a
    640x480 one of 1,200 vectors), each checked bit for bit against the
    handle's grid-2 filter on the same candidates."""
    hip = Hip()
    rng = np.random.default_rng(90 + npm)
    try:
        for nx, ny in ((6600, 4400), (640, 480)):
            p = pivx.init_params(nx=nx, ny=ny, npass1=1, npass2=1, wx2_median=5, wy2_median=5, wx2_smooth=3,
                                 wy2_smooth=3)
            with pivx.Xcorr2dMulti(p) as x:
                g = x.geometry(2)
                nt, K = g.nt_vec, 4
                dx, dy, v = _candidates(rng, nt, K)
                x.set_raw(2, dx, dy, v)
                x.MedianFilter_Smoothing_p2(npm, 5, 2.0, 2.0, nps, fill)
                want = x.filtered(2)
            cx, cy, cv = hip.malloc(dx.nbytes), hip.malloc(dy.nbytes), hip.malloc(v.nbytes)
            ox, oy, ov = hip.malloc(nt * 4), hip.malloc(nt * 4), hip.malloc(nt)
            hip.h2d(cx, dx)
            hip.h2d(cy, dy)
            hip.h2d(cv, v)
            pivx.median_filter_smoothing_noalloc(cx, cy, cv, g.nx_vec, g.ny_vec, K, 5, 5, npm, 5, 2.0, 2.0, 3, 3,
                                                 nps, fill, ox, oy, ov, 0)
            hip.sync()
            got = (np.empty(nt, np.float32), np.empty(nt, np.float32), np.empty(nt, np.uint8))
            for a, d in zip(got, (ox, oy, ov)):
                hip.d2h(a, d)
            print(f"[noalloc] {nx}x{ny} ({nt} vectors) npm={npm} nps={nps} fill={fill}: "
                  f"invalid={int((got[2] == 0).sum())}", flush=True)
            for a, b in zip(got, want):
                np.testing.assert_array_equal(a, b)
            hip.free_all()
    finally:
        hip.free_all()


def test_noalloc_filter_rejects_what_it_cannot_hold(pivx):
    hip = Hip()
    try:
        d = hip.malloc(1 << 20)
        with pytest.raises(pivx.PivError, match="allocation-free limit"):
            pivx.median_filter_smoothing_noalloc(d, d, d, 400, 400, 4, 5, 5, 2, 5, 2.0, 2.0, 3, 3, 1, 1, d, d, d, 0)
        with pytest.raises(pivx.PivError, match="64 vectors"):
            pivx.median_filter_smoothing_noalloc(d, d, d, 40, 40, 4, 9, 9, 2, 5, 2.0, 2.0, 3, 3, 1, 1, d, d, d, 0)
        with pytest.raises(pivx.PivError, match="null"):
            pivx.median_filter_smoothing_noalloc(0, d, d, 40, 40, 4, 5, 5, 2, 5, 2.0, 2.0, 3, 3, 1, 1, d, d, d, 0)
    finally:
        hip.free_all()
