"""fl3h_density_estimate_noalloc: density estimation on the caller's device
buffers with the library's fixed scratch array instead of allocations.

Its output must equal fl3h_density_estimate's (same kernels, exact mode) bit
for bit, on two calls in a row (the second with a smaller histogram, so the
scratch still holds the first call's data), and histograms beyond the
scratch's stated capacity are rejected.  Device buffers for the test come
from the HIP runtime the library itself is linked against.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def _buckets(rng, W, H, bits):
    b = np.zeros((H, W, 5), dtype=np.float64)
    hits = rng.poisson(rng.uniform(0.0, 3.0, (H, W)) ** 3)
    hits[: H // 6] = 0
    b[:, :, 4] = hits * 255
    b[:, :, 3] = hits * 255
    for k in range(3):
        b[:, :, k] = np.floor(hits * rng.uniform(0, 255, (H, W)))
    return b if bits == 64 else b.astype(np.uint32)


def _de_tables(ref, rad, curve, ss):
    mfc, mfi, ks = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    wp, cp_ = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_double)()
    assert ref.ref().ref_de_filters(rad, 0.0, curve, ss, ctypes.byref(mfc), ctypes.byref(mfi), ctypes.byref(ks),
                                    ctypes.byref(wp), ctypes.byref(cp_)) == 0
    return mfc.value, mfi.value, ks.value, wp, cp_


def _noalloc(prod, hip, b, acc0, W, H, bits, ss, k1, k2, tables):
    mfc, mfi, ks, wp, cp_ = tables
    db, da = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(db), b.nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(da), acc0.nbytes) == 0
    try:
        assert hip.hipMemcpy(db, b.ctypes.data, b.nbytes, H2D) == 0
        assert hip.hipMemcpy(da, acc0.ctypes.data, acc0.nbytes, H2D) == 0
        L = prod.lib()
        rc = L.fl3h_density_estimate_noalloc(db, W, H, bits, ss, k1, k2, 0.4, mfc, mfi, ks, mfi + 1, wp, cp_, da)
        assert rc == 0, prod.last_error()
        out = np.zeros_like(acc0)
        assert hip.hipMemcpy(out.ctypes.data, da, out.nbytes, D2H) == 0
        return out
    finally:
        hip.hipFree(db)
        hip.hipFree(da)


@pytest.mark.parametrize("bits", [33, 64])
def test_noalloc_density_estimate_equals_allocating_call(prod, ref, bits):
    hip = _hip()
    L = prod.lib()
    L.fl3h_density_estimate_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.fl3h_density_estimate_noalloc_max_cells.restype = ctypes.c_longlong
    adt = np.float64 if bits == 64 else np.float32
    k1, k2 = 4.0 * 255 * 268 / 256, 1.0 / 2000.0
    # two calls in a row, the second smaller, with different estimator radii
    for (W, H, ss, rad, seed) in [(600, 400, 2, 9.0, 1), (230, 130, 1, 5.0, 2)]:
        rng = np.random.default_rng(seed)
        b = _buckets(rng, W, H, bits)
        acc0 = rng.uniform(0.0, 3.0, (H, W, 4)).astype(adt)
        tables = _de_tables(ref, rad, 0.4, ss)
        mfc, mfi, ks, wp, cp_ = tables
        want = acc0.copy()
        assert L.fl3h_density_estimate(b.ctypes.data_as(ctypes.c_void_p), W, H, bits, ss, k1, k2, 0.4, mfc, mfi,
                                       ks, mfi + 1, wp, cp_, want.ctypes.data_as(ctypes.c_void_p)) == 0, \
            prod.last_error()
        got = _noalloc(prod, hip, b, acc0, W, H, bits, ss, k1, k2, tables)
        assert np.array_equal(got, want), (W, H, np.abs(got - want).max())
        assert np.abs(want - acc0).max() > 0
    # beyond the scratch: rejected before any device work
    cap = L.fl3h_density_estimate_noalloc_max_cells(bits)
    assert cap > 1_000_000
    W = 4096
    H = int(cap // W) + 1
    mfc, mfi, ks, wp, cp_ = _de_tables(ref, 9.0, 0.4, 1)
    rc = L.fl3h_density_estimate_noalloc(ctypes.c_void_p(16), W, H, bits, 1, k1, k2, 0.4, mfc, mfi, ks, mfi + 1,
                                         wp, cp_, ctypes.c_void_p(16))
    assert rc != 0
    assert "exceed" in prod.last_error()
