"""Genome ingestion on a host thread pool (fl3h_ingest_*) with the
reference's own parser (flam3_parse_xml2 of the reference library built by
oracle/Makefile): documents parsed on eight threads come out exactly as the
reference parses them one at a time, and documents the parser rejects or
that break what rendering relies on fail their ticket."""
import ctypes

import numpy as np
import pytest

from flam3_amd import abi
from flam3_amd import genome as G


def _docs():
    docs = []
    for k in range(24):
        docs.append(G.flame_document([G.synthetic_preview(500 + k)]))
    docs.append(G.flame_document([G.synthetic_1080p(seed=7, width=64, height=36)]))
    docs.append(G.flame_document([G.synthetic_8k(seed=8, width=64, height=36)]))
    docs.append(G.flame_document(G.synthetic_animation(seed=9, width=64, height=36, frames=5)))
    # the final xform first in the document: the parser moves it last
    # (parser.c:808-821)
    g = G.synthetic_1080p(seed=10, width=48, height=32)
    x = g.to_xml()
    fin = [l for l in x.split("\n") if l.startswith("<finalxform")][0]
    x = x.replace(fin + "\n", "").replace("<xform ", fin + "\n<xform ", 1)
    docs.append("<flames>\n" + x + "\n</flames>\n")
    return docs


def _fields(g):
    n = g.num_xforms
    xf = ctypes.cast(g.xform, abi.XFORM_P)
    out = [g.width, g.height, g.spatial_oversample, g.sample_density, g.nbatches, g.ntemporal_samples, n,
           g.final_xform_index, g.final_xform_enable, g.time, g.pixels_per_unit, g.estimator]
    for i in range(n):
        out += [xf[i].density, xf[i].color, xf[i].opacity] + [xf[i].c[a][b] for a in range(3) for b in range(2)]
        out += list(xf[i].var)
    nstd = n - (1 if g.final_xform_index >= 0 else 0)
    if g.chaos:
        rows = ctypes.cast(g.chaos, ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))
        out += [rows[i][j] for i in range(nstd) for j in range(nstd)]
    out += [g.palette[i].color[k] for i in range(256) for k in range(4)]
    return np.array(out, dtype=np.float64)


def _lib(prod_lib):
    L = prod_lib
    L.fl3h_ingest_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.fl3h_ingest_create.restype = ctypes.c_void_p
    L.fl3h_ingest_submit.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p]
    L.fl3h_ingest_submit.restype = ctypes.c_longlong
    L.fl3h_ingest_wait.argtypes = [ctypes.c_void_p, ctypes.c_longlong,
                                   ctypes.POINTER(ctypes.POINTER(abi.flam3_genome)), ctypes.POINTER(ctypes.c_int)]
    L.fl3h_ingest_wait.restype = ctypes.c_int
    L.fl3h_ingest_destroy.argtypes = [ctypes.c_void_p]
    return L


def test_parallel_parse_equals_serial_reference_parse(prod_lib, ref):
    L = _lib(prod_lib)
    R = ref.ref()
    parse = ctypes.cast(R.flam3_parse_xml2, ctypes.c_void_p)
    docs = _docs()
    q = L.fl3h_ingest_create(parse, 8, 1)
    assert q
    try:
        tickets = [L.fl3h_ingest_submit(q, d.encode(), b"doc%d" % k) for k, d in enumerate(docs)]
        assert min(tickets) >= 0
        for k in reversed(range(len(docs))):   # waits in another order than submission
            gp = ctypes.POINTER(abi.flam3_genome)()
            n = ctypes.c_int(0)
            assert L.fl3h_ingest_wait(q, tickets[k], ctypes.byref(gp), ctypes.byref(n)) == 0, k
            with ref.Parsed(docs[k]) as want:
                assert n.value == want.n
                for i in range(n.value):
                    assert np.array_equal(_fields(gp[i]), _fields(want[i])), (k, i)
                    if gp[i].final_xform_index >= 0:
                        assert gp[i].final_xform_index == gp[i].num_xforms - 1
            R.ref_free_genomes(gp, n.value)
    finally:
        L.fl3h_ingest_destroy(q)


def test_rejected_documents_fail_their_ticket(prod_lib, ref):
    L = _lib(prod_lib)
    parse = ctypes.cast(ref.ref().flam3_parse_xml2, ctypes.c_void_p)
    good = G.flame_document([G.synthetic_small(seed=3, width=32, height=24)])
    bad_xml = "<flames><flame size='32 24'><xform weight='1' coefs='1 0 0 1 0 0'"   # not XML
    zero_size = good.replace('size="32 24"', 'size="0 24"')
    q = L.fl3h_ingest_create(parse, 4, 1)
    try:
        t = [L.fl3h_ingest_submit(q, d.encode(), b"x") for d in (good, bad_xml, zero_size, good)]
        rcs = []
        for tk in t:
            gp = ctypes.POINTER(abi.flam3_genome)()
            n = ctypes.c_int(0)
            rcs.append(L.fl3h_ingest_wait(q, tk, ctypes.byref(gp), ctypes.byref(n)))
            if gp:
                ref.ref().ref_free_genomes(gp, n.value)
        assert rcs == [0, 1, 1, 0]
        assert L.fl3h_ingest_wait(q, t[0], None, None) == 1   # a ticket is handed over once
    finally:
        L.fl3h_ingest_destroy(q)
