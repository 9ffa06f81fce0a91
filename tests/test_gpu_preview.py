"""The preview loop's clients of the device path against the reference:

* fl3h_estimate_bounding_box vs flam3_estimate_bounding_box (flam3.c:3574)
  with the same ISAAC context: one FP64 trajectory on the GPU (replay
  iteration), the same quantiles and the same random-number consumption.
  For a contracting genome the trajectories agree to rounding, so the boxes
  agree within 1e-9; for an expanding one they part after a few dozen steps
  (ulp-level libm differences grow), so the boxes agree statistically.
* fl3h_improve_colors vs flam3_improve_colors (palettes.c:476): the same
  colouring and the same sequence of C-library random() calls, with every
  trial image rendered in one batched call.
"""
import ctypes

import numpy as np
import pytest

from flam3_amd import abi
from flam3_amd import genome as G

pytestmark = pytest.mark.gpu

libc = ctypes.CDLL(None)
libc.srandom.argtypes = [ctypes.c_uint]
libc.random.restype = ctypes.c_long


def _bbox(L, gptr, eps, n, rc):
    bmin = (ctypes.c_double * 2)()
    bmax = (ctypes.c_double * 2)()
    bv = L(gptr, eps, n, bmin, bmax, ctypes.byref(rc))
    return bv, np.array(list(bmin) + list(bmax))


@pytest.mark.parametrize("mk,eps,n,exact", [
    (lambda: G.synthetic_small(seed=31, width=64, height=48), 0.01, 100000, True),
    (lambda: G.synthetic_small(seed=32, width=64, height=48), 0.0, 20000, True),
    (lambda: G.synthetic_1080p(seed=33, width=64, height=36), 0.01, 100000, False),
], ids=["contracting", "contracting_extremes", "expanding"])
def test_bounding_box_matches_reference(prod, ref, oracle, mk, eps, n, exact):
    g = mk()
    R = ref.ref()
    R.flam3_estimate_bounding_box.argtypes = [ctypes.POINTER(abi.flam3_genome), ctypes.c_double, ctypes.c_int,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(abi.randctx)]
    R.flam3_estimate_bounding_box.restype = ctypes.c_int
    L = prod.lib()
    L.fl3h_estimate_bounding_box.argtypes = R.flam3_estimate_bounding_box.argtypes
    L.fl3h_estimate_bounding_box.restype = ctypes.c_int
    with ref.Parsed(G.flame_document([g])) as want_g, ref.Parsed(G.flame_document([g])) as got_g:
        rc_w, rc_g = oracle.seeded_rc("bbox"), oracle.seeded_rc("bbox")
        bw, want = _bbox(R.flam3_estimate_bounding_box, want_g.ptr, eps, n, rc_w)
        bg, got = _bbox(L.fl3h_estimate_bounding_box, got_g.ptr, eps, n, rc_g)
    print("bbox: reference %s (%d bad), gpu %s (%d bad)" % (want, bw, got, bg), flush=True)
    size = np.abs(want[2:] - want[:2]).max()
    if exact:
        assert bg == bw
        assert np.abs(got - want).max() <= 1e-9 * max(1.0, np.abs(want).max())
        assert list(rc_g.randrsl) == list(rc_w.randrsl) and rc_g.randcnt == rc_w.randcnt
    else:
        assert np.abs(got - want).max() <= 0.05 * size, (got, want)


def test_improve_colors_matches_reference(prod, ref, monkeypatch):
    monkeypatch.setenv("isaac_seed", "colors")
    g = G.synthetic_preview(41, size=256)
    R = ref.ref()
    R.flam3_improve_colors.argtypes = [ctypes.POINTER(abi.flam3_genome), ctypes.c_int, ctypes.c_int, ctypes.c_int]
    R.flam3_improve_colors.restype = None
    L = prod.lib()
    L.fl3h_improve_colors.argtypes = R.flam3_improve_colors.argtypes
    L.fl3h_improve_colors.restype = ctypes.c_int
    with ref.Parsed(G.flame_document([g])) as want_g, ref.Parsed(G.flame_document([g])) as got_g:
        libc.srandom(1234)
        R.flam3_improve_colors(want_g.ptr, 12, 0, 10)
        after_w = [libc.random() for _ in range(4)]
        libc.srandom(1234)
        assert L.fl3h_improve_colors(got_g.ptr, 12, 0, 10) == 0
        after_g = [libc.random() for _ in range(4)]
        assert after_g == after_w
        xw = ctypes.cast(want_g[0].xform, abi.XFORM_P)
        xg = ctypes.cast(got_g[0].xform, abi.XFORM_P)
        for i in range(want_g[0].num_xforms):
            assert xg[i].color == xw[i].color, i
        # palette changes need the palette library
        assert L.fl3h_improve_colors(got_g.ptr, 2, 1, 10) == 1
