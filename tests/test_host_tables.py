"""Host-side logic of the product library (libflam3_amd.so) against the
reference: per-xform precalculation, xform selection tables and their
breakpoint form, the filter banks, ISAAC seeding, genome construction, and
the C-ABI surface.  All CPU; no GPU calls.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from flam3_amd import abi
from flam3_amd import genome as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class DEF(ctypes.Structure):
    _fields_ = [("max_filtered_counts", ctypes.c_int), ("max_filter_index", ctypes.c_int),
                ("kernel_size", ctypes.c_int), ("filter_widths", ctypes.POINTER(ctypes.c_double)),
                ("filter_coefs", ctypes.POINTER(ctypes.c_double)), ("nfilters", ctypes.c_int)]


@pytest.fixture(scope="module")
def host(prod):
    """Internal host functions of the product library (exported symbols)."""
    L = prod.lib()
    P = ctypes.POINTER
    L.fl3_prepare_precalc_flags.argtypes = [P(abi.flam3_genome)]
    L.fl3_prepare_precalc_flags.restype = ctypes.c_int
    L.fl3_xform_precalc.argtypes = [P(abi.flam3_xform)]
    L.fl3_create_xform_distrib.argtypes = [P(abi.flam3_genome), P(ctypes.c_int)]
    L.fl3_create_xform_distrib.restype = P(ctypes.c_ushort)
    L.fl3_distrib_breakpoints.argtypes = [P(ctypes.c_ushort), ctypes.c_int, ctypes.c_int, P(ctypes.c_int)]
    L.fl3_xform_starts.argtypes = [P(abi.flam3_genome), P(ctypes.c_int), P(ctypes.c_int)]
    L.fl3_xform_starts.restype = ctypes.c_int
    L.fl3_create_de_filters.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, P(DEF)]
    L.fl3_create_de_filters.restype = ctypes.c_int
    L.fl3_create_spatial_filter.argtypes = [P(abi.flam3_genome), ctypes.c_double, ctypes.c_int,
                                            P(P(ctypes.c_double))]
    L.fl3_create_spatial_filter.restype = ctypes.c_int
    L.fl3_create_temporal_filter.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             P(P(ctypes.c_double)), P(P(ctypes.c_double))]
    L.fl3_create_temporal_filter.restype = ctypes.c_double
    L.fl3_adjust_percentage.argtypes = [ctypes.c_double]
    L.fl3_adjust_percentage.restype = ctypes.c_double
    L.fl3_genome_copy.argtypes = [P(abi.flam3_genome), P(abi.flam3_genome)]
    L.fl3_genome_free.argtypes = [P(abi.flam3_genome)]
    return L


def _xforms(g):
    return ctypes.cast(g.xform, abi.XFORM_P)


GENOMES = {
    "small": lambda: G.synthetic_small(seed=41),
    "1080p": lambda: G.synthetic_1080p(seed=42),
    "xaos": lambda: G.synthetic_8k(seed=43),
    "preview": lambda: G.synthetic_preview(44),
}


def _zero_weight_variant():
    g = G.synthetic_1080p(seed=45)
    g.xforms[2].weight = 0.0
    g.xforms[5].weight = 0.0
    return g


@pytest.mark.parametrize("name", list(GENOMES) + ["zero_weight"])
def test_precalc_and_selection_tables_bit_exact(host, ref, name):
    g = GENOMES[name]() if name in GENOMES else _zero_weight_variant()
    mine = g.to_struct()
    try:
        with ref.Parsed(G.flame_document([g])) as p:
            R = ref.ref()
            assert R.prepare_precalc_flags(p.ptr) == 0
            assert host.fl3_prepare_precalc_flags(mine) == 0
            a, b = _xforms(p[0]), _xforms(mine.contents)
            n = mine.contents.num_xforms
            for i in range(n):
                if a[i].density == 0.0 and i != mine.contents.final_xform_index:
                    continue
                R.xform_precalc(p.ptr, i)
                host.fl3_xform_precalc(ctypes.byref(b[i]))
                for f in ["vis_adjusted", "precalc_angles_flag", "precalc_atan_xy_flag", "precalc_atan_yx_flag",
                          "has_preblur", "has_post", "num_active_vars", "julian_rN", "julian_cn", "juliascope_rN",
                          "juliascope_cn", "radialBlur_spinvar", "radialBlur_zoomvar", "waves_dx2", "waves_dy2",
                          "disc2_sinadd", "disc2_cosadd", "disc2_timespi", "super_shape_pm_4",
                          "super_shape_pneg1_n1", "wedgeJulia_cf", "wedgeJulia_rN", "wedgeJulia_cn",
                          "persp_vsin", "persp_vfcos"]:
                    assert getattr(a[i], f) == getattr(b[i], f), (i, f)
                k = a[i].num_active_vars
                assert list(a[i].varFunc)[:k] == list(b[i].varFunc)[:k]
                assert list(a[i].active_var_weights)[:k] == list(b[i].active_var_weights)[:k]
            want = R.flam3_create_xform_distrib(p.ptr)
            nrows = ctypes.c_int()
            got = host.fl3_create_xform_distrib(mine, ctypes.byref(nrows))
            rows = nrows.value
            assert mine.contents.chaos_enable == p[0].chaos_enable
            used_rows = rows if p[0].chaos_enable else 1
            w = np.ctypeslib.as_array(want, shape=(rows * 16384,))
            h = np.ctypeslib.as_array(got, shape=(rows * 16384,))
            assert np.array_equal(w[:used_rows * 16384], h[:used_rows * 16384])
            # breakpoints reproduce every slot
            nstd = n - (mine.contents.final_xform_index >= 0)
            st = (ctypes.c_int * (used_rows * nstd))()
            host.fl3_distrib_breakpoints(got, used_rows, nstd, st)
            starts = np.array(st[:]).reshape(used_rows, nstd)
            slots = np.arange(16384)
            for r in range(used_rows):
                rebuilt = (slots[:, None] >= starts[r, 1:][None, :]).sum(axis=1)
                assert np.array_equal(rebuilt, h[r * 16384:(r + 1) * 16384])
            ref.ref().ref_free(want)
            ref.ref().ref_free(got)
    finally:
        G.free_struct(mine)


def test_xform_starts_equal_table_breakpoints(host):
    """fl3_xform_starts (the breakpoints the jobs hold, without the table,
    with a cache of recent inputs) gives what fl3_distrib_breakpoints gives
    for fl3_create_xform_distrib's table, with and without xaos: on a first
    call, on a repeated one (from the cache), and after a weight changes (a
    new entry)."""
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    for name, mk in GENOMES.items():
        g = mk()
        for changed in (False, True):
            if changed:
                g.xforms[0].weight *= 1.37
            mine = g.to_struct()
            try:
                assert host.fl3_prepare_precalc_flags(mine) == 0
                n = mine.contents.num_xforms
                nstd = n - (mine.contents.final_xform_index >= 0)
                nrows = ctypes.c_int()
                t = host.fl3_create_xform_distrib(mine, ctypes.byref(nrows))
                used = nrows.value if mine.contents.chaos_enable else 1
                want = (ctypes.c_int * (used * nstd))()
                host.fl3_distrib_breakpoints(t, used, nstd, want)
                libc.free(ctypes.cast(t, ctypes.c_void_p))
                for _ in range(2):
                    got = (ctypes.c_int * ((nstd + 1) * nstd))()
                    rows = ctypes.c_int()
                    mine.contents.chaos_enable = -1
                    assert host.fl3_xform_starts(mine, got, ctypes.byref(rows)) == 0
                    assert rows.value == used, (name, rows.value, used)
                    assert mine.contents.chaos_enable == (1 if used > 1 else 0)
                    assert list(got[:used * nstd]) == list(want[:]), (name, changed)
            finally:
                G.free_struct(mine)


def test_selection_table_weight_shares(host):
    """Slot counts follow the weights to within one slot (closed form)."""
    g = G.synthetic_1080p(seed=46)
    mine = g.to_struct()
    try:
        assert host.fl3_prepare_precalc_flags(mine) == 0
        rows = ctypes.c_int()
        t = host.fl3_create_xform_distrib(mine, ctypes.byref(rows))
        h = np.ctypeslib.as_array(t, shape=(16384,))
        w = np.array([x.weight for x in g.xforms])
        counts = np.bincount(h, minlength=len(w))
        assert np.all(np.abs(counts - 16384 * w / w.sum()) <= 1.0)
        assert np.all(np.diff(h.astype(int)) >= 0)
    finally:
        G.free_struct(mine)


@pytest.mark.parametrize("args", [(9.0, 0.0, 0.4, 1), (9.0, 0.0, 0.4, 2), (7.0, 0.0, 0.4, 2), (15.0, 0.0, 0.4, 3),
                                  (9.0, 1.0, 0.6, 1), (20.0, 0.0, 0.2, 2), (3.0, 2.0, 1.0, 1)])
def test_de_filters_bit_exact(host, ref, args):
    mfc, mfi, ks = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    wp, cp_ = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_double)()
    rerr = ref.ref().ref_de_filters(*args, ctypes.byref(mfc), ctypes.byref(mfi), ctypes.byref(ks),
                                    ctypes.byref(wp), ctypes.byref(cp_))
    d = DEF()
    merr = host.fl3_create_de_filters(*args, ctypes.byref(d))
    assert bool(rerr) == bool(merr)   # same rejections (filters.c:284-308)
    if rerr:
        return
    assert (d.max_filtered_counts, d.max_filter_index, d.kernel_size) == (mfc.value, mfi.value, ks.value)
    nf = mfi.value + 1
    assert np.array_equal(np.ctypeslib.as_array(d.filter_widths, (nf,)), np.ctypeslib.as_array(wp, (nf,)))
    assert np.array_equal(np.ctypeslib.as_array(d.filter_coefs, (nf * ks.value,)),
                          np.ctypeslib.as_array(cp_, (nf * ks.value,)))
    # every kernel, expanded from its octant, sums to 1 over the disc
    half = int(np.sqrt(2 * ks.value)) - 1
    while (half + 1) * (half + 2) // 2 < ks.value:
        half += 1
    coefs = np.ctypeslib.as_array(d.filter_coefs, (nf, ks.value))
    for f in range(nf):
        tot, idx = 0.0, 0
        for jj in range(half + 1):
            for ii in range(jj + 1):
                mult = 1 if jj == 0 else (4 if (ii == 0 or ii == jj) else 8)
                tot += mult * coefs[f, idx]
                idx += 1
        assert abs(tot - 1.0) < 1e-12


def test_de_filter_cache_hits_evictions_and_threads(host, ref):
    """fl3_create_de_filters keeps the last eight filter banks (filters.c's
    bank is rebuilt per call in the reference): over ten parameter sets,
    each asked for twice in a row and again after others have evicted it,
    and from eight threads at once, every returned bank equals the
    reference's bit for bit, and every returned bank can be freed."""
    import threading
    host.fl3_free_de_filters.argtypes = [ctypes.POINTER(DEF)]
    sets = [(9.0, 0.0, 0.4, 2), (9.0, 0.0, 0.4, 1), (7.0, 0.0, 0.4, 2), (15.0, 0.0, 0.4, 3), (5.0, 1.0, 0.6, 1),
            (9.0, 0.0, 0.3, 2), (11.0, 0.5, 0.4, 1), (3.0, 0.0, 0.4, 4), (20.0, 0.0, 0.5, 1), (9.0, 2.0, 0.4, 2)]
    want = {}
    for a in sets:
        mfc, mfi, ks = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        wp, cp_ = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_double)()
        assert ref.ref().ref_de_filters(*a, ctypes.byref(mfc), ctypes.byref(mfi), ctypes.byref(ks),
                                        ctypes.byref(wp), ctypes.byref(cp_)) == 0
        nf = mfi.value + 1
        want[a] = (mfc.value, mfi.value, ks.value, np.ctypeslib.as_array(wp, (nf,)).copy(),
                   np.ctypeslib.as_array(cp_, (nf * ks.value,)).copy())

    def check(a):
        d = DEF()
        assert host.fl3_create_de_filters(*a, ctypes.byref(d)) == 0
        mfc, mfi, ks, w, c = want[a]
        nf = mfi + 1
        ok = ((d.max_filtered_counts, d.max_filter_index, d.kernel_size) == (mfc, mfi, ks) and
              np.array_equal(np.ctypeslib.as_array(d.filter_widths, (nf,)), w) and
              np.array_equal(np.ctypeslib.as_array(d.filter_coefs, (nf * ks,)), c))
        host.fl3_free_de_filters(ctypes.byref(d))
        return ok

    for a in sets:              # fill, hit right away
        assert check(a) and check(a), a
    for a in sets + sets[::-1]:  # the first sets were evicted by the later ones
        assert check(a), a
    results = []

    def worker(k):
        rng = np.random.default_rng(k)
        results.append(all(check(sets[int(i)]) for i in rng.integers(0, len(sets), 40)))

    th = [threading.Thread(target=worker, args=(k,)) for k in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert results == [True] * 8


@pytest.mark.parametrize("kernel", range(14))
@pytest.mark.parametrize("ss,radius,aspect,field", [(1, 0.5, 1.0, 0), (2, 0.5, 1.0, 0), (3, 0.7, 1.0, 1),
                                                    (2, 1.3, 0.9, 2)])
def test_spatial_filter_bit_exact(host, ref, kernel, ss, radius, aspect, field):
    g = G.Genome(supersample=ss, filter_radius=radius, filter_shape=kernel, xforms=[G.Xform(weight=1.0)])
    mine = g.to_struct()
    try:
        with ref.Parsed(G.flame_document([g])) as p:
            f = ref.make_frame(p.ptr, 1)
            f.pixel_aspect_ratio = aspect
            fr = ctypes.POINTER(ctypes.c_double)()
            fw_ref = ref.ref().flam3_create_spatial_filter(ctypes.byref(f), field, ctypes.byref(fr))
            fm = ctypes.POINTER(ctypes.c_double)()
            fw = host.fl3_create_spatial_filter(mine, aspect, field, ctypes.byref(fm))
            assert fw == fw_ref
            if fw > 0:
                a = np.ctypeslib.as_array(fr, (fw * fw,))
                b = np.ctypeslib.as_array(fm, (fw * fw,))
                assert np.array_equal(a, b)
                assert abs(b.sum() - 1.0) < 1e-12
    finally:
        G.free_struct(mine)


@pytest.mark.parametrize("n,ftype,fexp,width", [(1, 0, 0.0, 1.0), (2, 0, 0.0, 1.0), (100, 0, 0.0, 1.0),
                                                 (100, 1, 0.0, 0.8), (37, 2, 1.5, 1.0), (37, 2, -2.0, 1.2)])
def test_temporal_filter_bit_exact(host, ref, n, ftype, fexp, width):
    P = ctypes.POINTER
    fa, da = P(ctypes.c_double)(), P(ctypes.c_double)()
    fb, db = P(ctypes.c_double)(), P(ctypes.c_double)()
    sa = ref.ref().flam3_create_temporal_filter(n, ftype, fexp, width, ctypes.byref(fa), ctypes.byref(da))
    sb = host.fl3_create_temporal_filter(n, ftype, fexp, width, ctypes.byref(fb), ctypes.byref(db))
    assert sa == sb
    assert np.array_equal(np.ctypeslib.as_array(fa, (n,)), np.ctypeslib.as_array(fb, (n,)))
    assert np.array_equal(np.ctypeslib.as_array(da, (n,)), np.ctypeslib.as_array(db, (n,)))
    assert max(np.ctypeslib.as_array(fb, (n,))) == 1.0


def test_adjust_percentage_closed_form(host):
    for v in [0.0, 0.25, 0.5, 1.0, 0.9]:
        want = 0.0 if v == 0 else 10.0 ** (np.log2(v))
        assert abs(host.fl3_adjust_percentage(v) - want) < 1e-14


def test_frame_seeding_matches_reference(prod, ref):
    """fl3h_init_frame seeds f->rc exactly as flam3_init_frame would."""
    for seed in ["flam3", "cfg2", "q" * 140]:
        f = abi.flam3_frame()
        prod.lib().fl3h_init_frame(ctypes.byref(f), seed.encode())
        want = ref.seeded_rc(seed)
        assert list(f.rc.randrsl) == list(want.randrsl)
        assert list(f.rc.randmem) == list(want.randmem)
        assert (f.rc.randa, f.rc.randb, f.rc.randc, f.rc.randcnt) == (want.randa, want.randb, want.randc,
                                                                      want.randcnt)


def test_genome_builder_matches_parser(prod, ref):
    """to_struct() produces what the reference parser makes of to_xml()."""
    skip_g = {"xform", "chaos", "edits", "input_image", "parent_fname", "genome_index"}
    uninit = {"waves_dx2", "waves_dy2", "super_shape_pm_4", "super_shape_pneg1_n1", "motion", "varFunc",
              "active_var_weights", "num_active_vars", "precalc_angles_flag", "precalc_atan_xy_flag",
              "precalc_atan_yx_flag", "has_preblur", "has_post", "vis_adjusted"}
    for mk in [lambda: G.synthetic_1080p(seed=47), lambda: G.synthetic_8k(seed=48)]:
        g = mk()
        gp = g.to_struct()
        try:
            with ref.Parsed(G.flame_document([g])) as p:
                a, b = p[0], gp.contents
                for name, _ in abi.flam3_genome._fields_:
                    if name in skip_g:
                        continue
                    va, vb = bytes(memoryview(getattr(type(a), name).__get__(a)) if False else b""), None
                    assert _fbytes(a, name) == _fbytes(b, name), name
                xa, xb = _xforms(a), _xforms(b)
                for i in range(a.num_xforms):
                    for name, _ in abi.flam3_xform._fields_:
                        if name in uninit:
                            continue
                        assert _fbytes(xa[i], name) == _fbytes(xb[i], name), (i, name)
                ns = a.num_xforms - (a.final_xform_index >= 0)
                ca = ctypes.cast(a.chaos, ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))
                cb = ctypes.cast(b.chaos, ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))
                for i in range(ns):
                    assert [ca[i][j] for j in range(ns)] == [cb[i][j] for j in range(ns)]
        finally:
            G.free_struct(gp)


def _fbytes(obj, name):
    f = getattr(type(obj), name)
    return ctypes.string_at(ctypes.addressof(obj) + f.offset, f.size)


def test_abi_offsets_match_reference_header(tmp_path):
    """Every field of the mirrored structs sits at the reference's offset
    (compiled against /root/reference/flam3.h)."""
    if not os.path.exists("/root/reference/flam3.h"):
        pytest.skip("reference sources not present")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "flam3.h"', "int main(void){"]
    for cname, cls in [("flam3_xform", abi.flam3_xform), ("flam3_genome", abi.flam3_genome),
                       ("flam3_frame", abi.flam3_frame), ("randctx", abi.randctx), ("stat_struct", abi.stat_struct)]:
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for name, _ in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, name, cname, name))
    lines.append("return 0;}")
    src = tmp_path / "off.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "off"
    subprocess.run(["gcc", "-I/root/reference", "-I/usr/include/libxml2", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(l.split() for l in out if l.strip())
    for cname, cls in [("flam3_xform", abi.flam3_xform), ("flam3_genome", abi.flam3_genome),
                       ("flam3_frame", abi.flam3_frame), ("randctx", abi.randctx), ("stat_struct", abi.stat_struct)]:
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for name, _ in cls._fields_:
            assert int(got["%s.%s" % (cname, name)]) == getattr(cls, name).offset, (cname, name)


def test_library_exports_every_declared_entry_point(prod):
    """libflam3_amd.so loads and exports every function include/fl3h.h declares."""
    text = open(os.path.join(ROOT, "include", "fl3h.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b((?:flam3|fl3h)_[a-z0-9_]+)\s*\(", text))
    assert {"flam3_render", "flam3_render_memory_required", "fl3h_render"} <= names
    out = subprocess.run(["nm", "-D", "--defined-only", prod.LIB_PATH], check=True, capture_output=True,
                         text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    missing = names - exported
    assert not missing, missing


def test_render_memory_required_matches_reference(prod, ref):
    g = G.synthetic_1080p(seed=49)
    for bits in (32, 33, 64):
        with ref.Parsed(G.flame_document([g])) as p:
            f = ref.make_frame(p.ptr, 1, bits=bits)
            want = ref.ref().flam3_render_memory_required(ctypes.byref(f))
            got = prod.lib().flam3_render_memory_required(ctypes.byref(f))
            assert got == want


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_render_without_gpu_fails_loudly(prod):
    """No CPU fallback: without a device flam3_render returns 1."""
    with pytest.raises(RuntimeError):
        prod.render(G.synthetic_small(seed=1, width=16, height=16, quality=1))
