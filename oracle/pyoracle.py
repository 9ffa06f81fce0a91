"""Test infrastructure: ctypes access to the C restatement in
oracle/liboracle.so (see flam3_oracle.h).  Genomes reach the oracle through
the reference's own XML parser (pyref.Parsed), never through the product
library.  Used only by tests, smoke() and bench.py's CPU-baseline leg.
"""
import ctypes
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, _HERE)
sys.path.insert(0, os.path.dirname(_HERE))
from flam3_amd import abi  # noqa: E402  (the reference's struct layout)
from flam3_amd._lib import fl3h_camera, fl3h_tonemap_params  # noqa: E402  (plain C structs of fl3h.h)
import pyref  # noqa: E402

LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise RuntimeError("oracle not built: %s (make -C oracle)" % LIB)
    L = ctypes.CDLL(LIB)
    P = ctypes.POINTER
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.or_irandinit.argtypes = [P(abi.randctx), ci]
    L.or_irand.argtypes = [P(abi.randctx)]
    L.or_irand.restype = ctypes.c_ulong
    L.or_rand01.argtypes = [P(abi.randctx)]
    L.or_rand01.restype = cd
    L.or_rand11.argtypes = [P(abi.randctx)]
    L.or_rand11.restype = cd
    L.or_apply_xform.argtypes = [P(abi.flam3_genome), ci, vp, vp, P(abi.randctx)]
    L.or_apply_xform.restype = ci
    L.or_iterate.argtypes = [P(abi.flam3_genome), ci, ci, vp, vp, P(abi.randctx)]
    L.or_iterate.restype = ci
    L.or_deposit.argtypes = [P(fl3h_camera), vp, ctypes.c_long, ci, ci, vp]
    L.or_deposit.restype = None
    L.or_density.argtypes = [vp, ci, ci, ci, ci, cd, cd, cd, ci, ci, ci, vp, vp, vp, ci]
    L.or_density.restype = None
    L.or_logscale.argtypes = [vp, ci, ci, ci, cd, cd, vp]
    L.or_logscale.restype = None
    L.or_tonemap.argtypes = [vp, ci, ci, ci, P(fl3h_tonemap_params), vp, vp, ci]
    L.or_tonemap.restype = None
    L.or_render.argtypes = [P(abi.flam3_frame), vp, ci, ci, ci, P(abi.stat_struct), ci]
    L.or_render.restype = ci
    _lib = L
    return L


def seeded_rc(seed):
    """Frame ISAAC context seeded from a string like flam3_init_frame."""
    rc = abi.randctx()
    raw = seed.encode()[:128]
    buf = (ctypes.c_char * 128).from_buffer(rc, abi.randctx.randrsl.offset)
    ctypes.memset(buf, 0, 128)
    ctypes.memmove(buf, raw, len(raw))
    lib().or_irandinit(ctypes.byref(rc), 1)
    return rc


def render(genomes, *, nchan=3, transp=0, bits=33, bpc=1, earlyclip=0, time=0.0, sub_batch_size=10000,
           seed="flam3", field=0, phantom=0, xml=None):
    """Oracle render (one iteration thread) of Genome descriptions or XML."""
    if xml is None:
        from flam3_amd.genome import flame_document
        if not isinstance(genomes, (list, tuple)):
            genomes = [genomes]
        xml = flame_document(genomes)
    os.environ.setdefault("flam3_palettes", pyref.PALETTES)
    with pyref.Parsed(xml) as p:
        g0 = p[0]
        f = pyref.make_frame(p.ptr, p.n, bits=bits, bpc=bpc, earlyclip=earlyclip, time=time, nthreads=1,
                             sub_batch_size=sub_batch_size, seed=seed)
        f.rc = seeded_rc(seed)
        dtype = np.uint16 if bpc == 2 else np.uint8
        img = np.zeros((g0.height, g0.width, nchan), dtype=dtype)
        st = abi.stat_struct()
        rc = lib().or_render(ctypes.byref(f), img.ctypes.data_as(ctypes.c_void_p), field, nchan, transp,
                             ctypes.byref(st), phantom)
        if rc != 0:
            raise RuntimeError("oracle render failed")
        return img, {"badvals": st.badvals, "num_iters": st.num_iters, "rc": f.rc}
