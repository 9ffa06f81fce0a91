"""Test infrastructure: ctypes access to the reference flam3 library built in
place by oracle/Makefile (oracle/_ref/*.so).  Used only by tests, by
``__graft_entry__.smoke()`` and by bench.py's CPU-baseline leg, never by the
product library.

Two builds exist: ``strict`` (no -ffast-math, no FP contraction; pins the C
restatement bit for bit) and ``fast`` (the reference's own -O3 -ffast-math
flags; the CPU baseline).
"""
import ctypes
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(_HERE))
from flam3_amd import abi  # noqa: E402  (the reference's struct layout)

REF_DIR = os.path.join(_HERE, "_ref")
PALETTES = "/root/reference/flam3-palettes.xml"

_libs = {}


def available(kind="strict"):
    return os.path.exists(os.path.join(REF_DIR, "libflam3ref_%s.so" % kind))


def ref(kind="strict"):
    if kind in _libs:
        return _libs[kind]
    path = os.path.join(REF_DIR, "libflam3ref_%s.so" % kind)
    if not os.path.exists(path):
        raise RuntimeError("reference library not built: %s (make -C oracle)" % path)
    L = ctypes.CDLL(path)
    P = ctypes.POINTER
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.flam3_parse_xml2.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ci, P(ci)]
    L.flam3_parse_xml2.restype = P(abi.flam3_genome)
    L.flam3_render.argtypes = [P(abi.flam3_frame), vp, ci, ci, ci, P(abi.stat_struct)]
    L.flam3_render.restype = ci
    L.flam3_iterate.argtypes = [P(abi.flam3_genome), ci, ci, vp, vp, P(abi.randctx)]
    L.flam3_iterate.restype = ci
    L.flam3_xform_preview.argtypes = [P(abi.flam3_genome), ci, cd, ci, ci, vp, P(abi.randctx)]
    L.flam3_xform_preview.restype = ci
    L.flam3_create_xform_distrib.argtypes = [P(abi.flam3_genome)]
    L.flam3_create_xform_distrib.restype = P(ctypes.c_ushort)
    L.prepare_precalc_flags.argtypes = [P(abi.flam3_genome)]
    L.prepare_precalc_flags.restype = ci
    L.xform_precalc.argtypes = [P(abi.flam3_genome), ci]
    L.xform_precalc.restype = None
    L.ref_irand.argtypes = [P(abi.randctx)]
    L.ref_irand.restype = ctypes.c_ulong
    L.ref_seed_rc.argtypes = [P(abi.randctx), ctypes.c_char_p]
    L.ref_seed_rc.restype = None
    L.irandinit.argtypes = [P(abi.randctx), ci]
    L.irandinit.restype = None
    L.ref_de_filters.argtypes = [cd, cd, cd, ci, P(ci), P(ci), P(ci), P(P(cd)), P(P(cd))]
    L.ref_de_filters.restype = ci
    L.flam3_create_spatial_filter.argtypes = [P(abi.flam3_frame), ci, P(P(cd))]
    L.flam3_create_spatial_filter.restype = ci
    L.flam3_create_temporal_filter.argtypes = [ci, ci, cd, cd, P(P(cd)), P(P(cd))]
    L.flam3_create_temporal_filter.restype = cd
    L.flam3_calc_alpha.argtypes = [cd, cd, cd]
    L.flam3_calc_alpha.restype = cd
    L.flam3_calc_newrgb.argtypes = [vp, cd, cd, vp]
    L.flam3_calc_newrgb.restype = None
    L.ref_free.argtypes = [vp]
    L.ref_free_genomes.argtypes = [P(abi.flam3_genome), ci]
    L.ref_interpolate.argtypes = [P(abi.flam3_genome), ci, cd]
    L.ref_interpolate.restype = P(abi.flam3_genome)
    L.ref_interpolate_stagger.argtypes = [P(abi.flam3_genome), ci, cd, cd]
    L.ref_interpolate_stagger.restype = P(abi.flam3_genome)
    L.flam3_render_memory_required.argtypes = [P(abi.flam3_frame)]
    L.flam3_render_memory_required.restype = cd
    _libs[kind] = L
    return L


class Parsed:
    """Genomes parsed by the reference parser (freed on close)."""

    def __init__(self, xml, kind="strict"):
        self.L = ref(kind)
        n = ctypes.c_int(0)
        self.ptr = self.L.flam3_parse_xml2(xml.encode(), b"test.flam3", 1, ctypes.byref(n))
        if not self.ptr or n.value < 1:
            raise RuntimeError("reference parser rejected the genome XML")
        self.n = n.value

    def __getitem__(self, i):
        return self.ptr[i]

    def close(self):
        if self.ptr:
            self.L.ref_free_genomes(self.ptr, self.n)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def seeded_rc(seed, kind="strict"):
    rc = abi.randctx()
    ref(kind).ref_seed_rc(ctypes.byref(rc), seed.encode())
    return rc


def make_frame(gptr, n, *, bits=33, bpc=1, earlyclip=0, time=0.0, nthreads=1, sub_batch_size=10000,
               seed="flam3", kind="strict"):
    f = abi.flam3_frame()
    f.pixel_aspect_ratio = 1.0
    f.genomes = ctypes.cast(gptr, ctypes.c_void_p)
    f.ngenomes = n
    f.verbose = 0
    f.bits = bits
    f.bytes_per_channel = bpc
    f.earlyclip = earlyclip
    f.time = time
    f.progress = None
    f.progress_parameter = None
    f.rc = seeded_rc(seed, kind)
    f.nthreads = nthreads
    f.sub_batch_size = sub_batch_size
    return f


def render(xml, *, nchan=3, transp=0, bits=33, bpc=1, earlyclip=0, time=0.0, nthreads=1,
           sub_batch_size=10000, seed="flam3", kind="strict", field=0):
    """Reference flam3_render of the genomes in `xml`; returns (image, stats)."""
    if os.path.exists(PALETTES):  # only palette-index genomes need it; the test genomes carry colours
        os.environ.setdefault("flam3_palettes", PALETTES)
    with Parsed(xml, kind) as p:
        g0 = p[0]
        f = make_frame(p.ptr, p.n, bits=bits, bpc=bpc, earlyclip=earlyclip, time=time, nthreads=nthreads,
                       sub_batch_size=sub_batch_size, seed=seed, kind=kind)
        dtype = np.uint16 if bpc == 2 else np.uint8
        img = np.zeros((g0.height, g0.width, nchan), dtype=dtype)
        st = abi.stat_struct()
        rc = p.L.flam3_render(ctypes.byref(f), img.ctypes.data_as(ctypes.c_void_p), field, nchan, transp,
                              ctypes.byref(st))
        if rc != 0:
            raise RuntimeError("reference flam3_render failed")
        return img, {"badvals": st.badvals, "num_iters": st.num_iters, "rc": f.rc}
