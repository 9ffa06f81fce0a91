"""Genome interpolation between control points (flam3_interpolate,
flam3.c:717-802; flam3_align and flam3_interpolate_n, interpolation.c)
against the reference library: every field of the interpolated genome,
its xforms and its xaos matrix must be bit-identical."""
import ctypes
import math

import numpy as np
import pytest

from flam3_amd import abi
from flam3_amd import genome as G

XFP = ctypes.POINTER(abi.flam3_xform)
# Xform fields derived from the others by prepare_precalc_flags and
# xform_precalc (variations.c:1942-2126) before every use; the reference
# leaves them uninitialised in an interpolated genome.
DERIVED = {"vis_adjusted", "precalc_angles_flag", "precalc_atan_xy_flag", "precalc_atan_yx_flag", "has_preblur",
           "has_post", "persp_vsin", "persp_vfcos", "julian_rN", "julian_cn", "juliascope_rN", "juliascope_cn",
           "wedgeJulia_rN", "wedgeJulia_cn", "wedgeJulia_cf", "radialBlur_spinvar", "radialBlur_zoomvar",
           "waves_dx2", "waves_dy2", "disc2_sinadd", "disc2_cosadd", "disc2_timespi", "super_shape_pm_4",
           "super_shape_pneg1_n1", "num_active_vars", "active_var_weights", "varFunc"}
ROWS = ctypes.POINTER(ctypes.POINTER(ctypes.c_double))


def _flat(v):
    if isinstance(v, ctypes.Array):
        return [_flat(x) for x in v]
    if isinstance(v, ctypes.Structure):
        return [_flat(getattr(v, f)) for f, _ in v._fields_]
    return v


def _same(a, b):
    if isinstance(a, list):
        return len(a) == len(b) and all(_same(x, y) for x, y in zip(a, b))
    if isinstance(a, float) and isinstance(b, float) and math.isnan(a) and math.isnan(b):
        return True
    return a == b


def _diff_genomes(a, b):
    out = []
    for f, _ in abi.flam3_genome._fields_:
        if f in ("xform", "chaos", "edits"):
            continue
        if not _same(_flat(getattr(a, f)), _flat(getattr(b, f))):
            out.append(f)
    if a.num_xforms != b.num_xforms or a.final_xform_index != b.final_xform_index:
        return out + ["xform layout"]
    xa, xb = ctypes.cast(a.xform, XFP), ctypes.cast(b.xform, XFP)
    for i in range(a.num_xforms):
        for f, _ in abi.flam3_xform._fields_:
            if f == "motion" or f in DERIVED:
                continue
            if not _same(_flat(getattr(xa[i], f)), _flat(getattr(xb[i], f))):
                out.append("xform[%d].%s" % (i, f))
    ns = a.num_xforms - (a.final_xform_index >= 0)
    ca, cb = ctypes.cast(a.chaos, ROWS), ctypes.cast(b.chaos, ROWS)
    for i in range(ns):
        for j in range(ns):
            if ca[i][j] != cb[i][j]:
                out.append("chaos[%d][%d]" % (i, j))
    return out


def _xf(rng, names, **kw):
    picks = rng.choice(len(names), size=min(len(names), int(rng.integers(1, 4))), replace=False)
    x = G.Xform(weight=float(rng.uniform(0.2, 1.0)), color=float(rng.uniform(0, 1)),
                coefs=tuple(float(c) for c in rng.uniform(-1.2, 1.2, 6)),
                variations={names[p]: float(rng.uniform(0.2, 1.0)) for p in picks}, **kw)
    return x


PARAMETRIC = ["blob", "pdj", "fan2", "rings2", "perspective", "julian", "juliascope", "ngon", "curl", "rectangles",
              "super_shape", "flower", "conic", "wedge_julia", "cpow", "mobius"]
PLAIN = ["linear", "spherical", "swirl", "polar", "fan", "rings", "heart", "disc", "spiral", "julia"]


def _params(rng, x):
    for name in list(x.variations):
        for f, _ in abi.flam3_xform._fields_:
            if f.startswith(name + "_") and f not in DERIVED:
                x.params[f] = float(rng.uniform(0.3, 2.0))


def _control_points(seed, n, *, counts, finals, itype="log", interp="linear", pal="hsv_circular", blend=0.0,
                    post=False, xaos=False):
    rng = np.random.default_rng(seed)
    cps = []
    for k in range(n):
        xs = []
        for _ in range(counts[k]):
            x = _xf(rng, PLAIN + PARAMETRIC)
            _params(rng, x)
            if post and rng.uniform() < 0.5:
                x.post = tuple(float(c) for c in rng.uniform(-1, 1, 6))
            if rng.uniform() < 0.3:
                x.animate = 0.0
            xs.append(x)
        if xaos:
            for x in xs:
                x.chaos = [float(c) for c in rng.uniform(0, 2, counts[k])]
        fin = _xf(rng, PLAIN) if finals[k] else None
        g = G.Genome(width=64 + 8 * k, height=48, time=float(10 * k), scale=20.0 + k, rotate=15.0 * k,
                     center=(0.1 * k, -0.2 * k), quality=10.0 + k, supersample=1 + (k % 2), passes=1 + k % 2,
                     temporal_samples=3 + k, estimator_radius=5.0 + k, gamma=2.0 + 0.5 * k,
                     brightness=3.0 + k, vibrancy=0.5 + 0.1 * k, highlight_power=0.5 * k,
                     background=(0.1 * k, 0.2, 0.0), palette=G.hsv_palette(rng), xforms=xs, final=fin,
                     interpolation_type=itype, interpolation=interp, palette_interpolation=pal,
                     hsv_rgb_palette_blend=blend)
        cps.append(g)
    return cps


CASES = [
    ("same_layout_log", dict(n=2, counts=[3, 3], finals=[1, 1])),
    ("pad_xforms_log", dict(n=2, counts=[2, 4], finals=[0, 0])),
    ("pad_final_log", dict(n=2, counts=[3, 3], finals=[1, 0], post=True)),
    ("linear_type", dict(n=2, counts=[4, 2], finals=[0, 1], itype="linear", post=True)),
    ("compat_type", dict(n=2, counts=[2, 3], finals=[1, 1], itype="old")),
    ("older_type", dict(n=2, counts=[3, 2], finals=[0, 1], itype="older")),
    ("rgb_palette", dict(n=2, counts=[3, 3], finals=[0, 0], pal="rgb")),
    ("hsv_palette", dict(n=2, counts=[3, 3], finals=[0, 0], pal="hsv")),
    ("sweep_palette", dict(n=2, counts=[3, 3], finals=[0, 0], pal="sweep")),
    ("circular_blend", dict(n=2, counts=[2, 2], finals=[1, 1], blend=0.35)),
    ("xaos_three_points", dict(n=3, counts=[3, 4, 3], finals=[1, 0, 1], xaos=True)),
    ("five_points_log", dict(n=5, counts=[2, 3, 3, 4, 2], finals=[0, 1, 1, 0, 1], post=True)),
]


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_interpolation_matches_reference(prod_lib, ref, name, kw):
    n = kw["n"]
    cps = _control_points(hash(name) % 1000, **kw)
    times = [-3.0, 0.0, 2.5, 5.0, 10.0 * (n - 1), 10.0 * (n - 1) + 4.0]
    times += [10.0 * k + f for k in range(n - 1) for f in (1.0, 3.3, 7.7)]
    with ref.Parsed(G.flame_document(cps)) as p:
        assert p.n == n
        for t in times:
            want = ref.ref().ref_interpolate(p.ptr, n, t)
            got = abi.flam3_genome()
            assert prod_lib.fl3h_interpolate(p.ptr, n, t, 0.0, ctypes.byref(got)) == 0
            d = _diff_genomes(want.contents, got)
            prod_lib.fl3h_genome_clear(ctypes.byref(got))
            ref.ref().ref_free_genomes(want, 1)
            assert not d, (t, d[:12])


def test_smooth_interpolation_interior_segments(prod_lib, ref):
    """Catmull-Rom blending over four control points.  The reference reads
    outside the control-point array on the first and last segments
    (flam3.c:748-766), so only interior segments are compared."""
    cps = _control_points(7, 5, counts=[2, 3, 3, 2, 3], finals=[1, 0, 1, 1, 0], interp="smooth", post=True)
    with ref.Parsed(G.flame_document(cps)) as p:
        for t in (10.0, 12.5, 17.0, 20.0, 23.9, 29.5):
            want = ref.ref().ref_interpolate(p.ptr, 5, t)
            got = abi.flam3_genome()
            assert prod_lib.fl3h_interpolate(p.ptr, 5, t, 0.0, ctypes.byref(got)) == 0
            d = _diff_genomes(want.contents, got)
            prod_lib.fl3h_genome_clear(ctypes.byref(got))
            ref.ref().ref_free_genomes(want, 1)
            assert not d, (t, d[:12])


def test_single_genome_is_a_copy(prod_lib, ref):
    cps = _control_points(3, 1, counts=[3], finals=[1])
    with ref.Parsed(G.flame_document(cps)) as p:
        want = ref.ref().ref_interpolate(p.ptr, 1, 4.0)
        got = abi.flam3_genome()
        assert prod_lib.fl3h_interpolate(p.ptr, 1, 4.0, 0.0, ctypes.byref(got)) == 0
        d = _diff_genomes(want.contents, got)
        prod_lib.fl3h_genome_clear(ctypes.byref(got))
        ref.ref().ref_free_genomes(want, 1)
        assert not d, d


@pytest.mark.parametrize("stagger", [0.25, 0.8, 1.0])
def test_staggered_interpolation(prod_lib, ref, stagger):
    """Xforms that start their transitions one after another
    (interpolation.c:339-369, flam3_interpolate's stagger argument)."""
    cps = _control_points(11, 2, counts=[4, 5], finals=[1, 0], post=True)
    with ref.Parsed(G.flame_document(cps)) as p:
        for t in (0.5, 2.0, 5.0, 8.5):
            want = ref.ref().ref_interpolate_stagger(p.ptr, 2, t, stagger)
            got = abi.flam3_genome()
            assert prod_lib.fl3h_interpolate(p.ptr, 2, t, stagger, ctypes.byref(got)) == 0
            d = _diff_genomes(want.contents, got)
            prod_lib.fl3h_genome_clear(ctypes.byref(got))
            ref.ref().ref_free_genomes(want, 1)
            assert not d, (t, d[:12])
