"""Pins of the C restatement (oracle/liboracle.so) against the reference.

The reference ships no tests or golden vectors, so the pins are: the
reference library itself (strict build, bit-exact comparisons), golden
fixtures written from it (tests/golden/make_golden.py), and closed-form
invariants that a plausible slip (dropped term, wrong index, swapped
operand) would break.
"""
import ctypes
import json
import math
import os

import numpy as np
import pytest

from flam3_amd import abi
from flam3_amd import genome as G

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- ISAAC

def test_isaac_matches_golden_streams(oracle):
    with open(os.path.join(GOLDEN, "isaac_streams.json")) as f:
        streams = json.load(f)
    L = oracle.lib()
    for seed, want in streams.items():
        if seed.startswith("int:"):
            rc = abi.randctx()
            for k in range(16):
                rc.randrsl[k] = int(seed[4:])
            L.or_irandinit(ctypes.byref(rc), 1)
        else:
            rc = oracle.seeded_rc(seed)
        got = [int(L.or_irand(ctypes.byref(rc))) for _ in range(len(want))]
        assert got == want, seed


def test_isaac_matches_reference_long_stream(oracle, ref):
    L, R = oracle.lib(), ref.ref()
    for seed in ["s", "another seed", "z" * 127]:
        a, b = oracle.seeded_rc(seed), ref.seeded_rc(seed)
        for _ in range(5000):
            assert L.or_irand(ctypes.byref(a)) == R.ref_irand(ctypes.byref(b))


def _isaac_py(seed_bytes, word_bits):
    """Textbook ISAAC (RANDSIZL = 4) in Python with the seeding scramble run
    on `word_bits`-bit words; the refill is always 32-bit."""
    M32, MW = 0xffffffff, (1 << word_bits) - 1
    raw = seed_bytes.ljust(128, b"\0")[:128]
    rsl = [int.from_bytes(raw[8 * k:8 * k + 8], "little") & MW for k in range(16)]

    def mix(s):
        a, b, c, d, e, f, g, h = s
        a = (a ^ (b << 11)) & MW; d = (d + a) & MW; b = (b + c) & MW
        b = (b ^ (c >> 2)) & MW; e = (e + b) & MW; c = (c + d) & MW
        c = (c ^ (d << 8)) & MW; f = (f + c) & MW; d = (d + e) & MW
        d = (d ^ (e >> 16)) & MW; g = (g + d) & MW; e = (e + f) & MW
        e = (e ^ (f << 10)) & MW; h = (h + e) & MW; f = (f + g) & MW
        f = (f ^ (g >> 4)) & MW; a = (a + f) & MW; g = (g + h) & MW
        g = (g ^ (h << 8)) & MW; b = (b + g) & MW; h = (h + a) & MW
        h = (h ^ (a >> 9)) & MW; c = (c + h) & MW; a = (a + b) & MW
        return [a, b, c, d, e, f, g, h]
    s = [0x9e3779b9] * 8
    for _ in range(4):
        s = mix(s)
    mm = [0] * 16
    for src in ("rsl", "mm"):
        for i in range(0, 16, 8):
            words = rsl if src == "rsl" else mm
            s = [(s[k] + words[i + k]) & MW for k in range(8)]
            s = mix(s)
            mm[i:i + 8] = s
    st = {"a": 0, "b": 0, "c": 0, "mm": mm, "r": [0] * 16}

    def refill():
        a, b = st["a"], (st["b"] + st["c"] + 1) & M32
        st["c"] += 1
        m, r = st["mm"], st["r"]
        for i in range(16):
            x = m[i]
            sh = [(a << 13), (a >> 6), (a << 2), (a >> 16)][i % 4]
            a = ((a ^ sh) + m[(i + 8) % 16]) & M32
            y = (m[(x >> 2) & 15] + a + b) & M32
            m[i] = y
            b = (m[(y >> 6) & 15] + x) & M32
            r[i] = b
        st["a"], st["b"] = a, b
    refill()
    out, cnt = [], 16
    for _ in range(48):
        if cnt == 0:
            refill()
            cnt = 16
        cnt -= 1
        out.append(st["r"][cnt])
    return out


def test_isaac_textbook_restatement_with_lp64_seeding(oracle):
    """An independent Python ISAAC reproduces the oracle's stream only when
    its seeding scramble runs on 64-bit words (isaacs.h:17 ub4 is unsigned
    long); with 32-bit words it does not, which pins the LP64 quirk and
    the ind() index (y >> 6) of the refill."""
    for seed in ["flam3", "cfg1"]:
        rc = oracle.seeded_rc(seed)
        got = [int(oracle.lib().or_irand(ctypes.byref(rc))) for _ in range(48)]
        assert _isaac_py(seed.encode(), 64) == got
        assert _isaac_py(seed.encode(), 32) != got


# ---------------------------------------------------------------- iteration

def _prepared(ref, g):
    p = ref.Parsed(G.flame_document([g]))
    R = ref.ref()
    assert R.prepare_precalc_flags(p.ptr) == 0
    dist = R.flam3_create_xform_distrib(p.ptr)
    return p, dist


GENOMES = [
    lambda: G.synthetic_small(seed=21, width=64, height=64),
    lambda: G.synthetic_1080p(seed=22, width=64, height=64),
    lambda: G.synthetic_8k(seed=23, width=64, height=64),
]


@pytest.mark.parametrize("mk", GENOMES)
def test_iterate_bit_exact(oracle, ref, mk):
    p, dist = _prepared(ref, mk())
    try:
        n = 500
        for s in range(4):
            a = np.zeros(4 * n)
            a[:4] = [0.1 * s, -0.2, 0.3, 1.0]
            b = a.copy()
            ra, rb = oracle.seeded_rc("it%d" % s), ref.seeded_rc("it%d" % s)
            ba = oracle.lib().or_iterate(p.ptr, n, 15, a.ctypes.data_as(ctypes.c_void_p), dist, ctypes.byref(ra))
            bb = ref.ref().flam3_iterate(p.ptr, n, 15, b.ctypes.data_as(ctypes.c_void_p), dist, ctypes.byref(rb))
            assert ba == bb
            assert np.array_equal(a, b)
            assert list(ra.randrsl) == list(rb.randrsl) and ra.randcnt == rb.randcnt
    finally:
        p.close()


def _single_variation_genome(vi, rng):
    name = abi.VARIATION_NAMES[vi]
    x = G.Xform(weight=1.0, color=0.5, coefs=tuple(rng.uniform(-1, 1, 6)),
                variations={name: float(rng.uniform(0.3, 1.2))})
    if name == "pre_blur":
        x.variations["linear"] = 1.0
    # non-default parameters for the parametric variations
    params = [f for f, _ in abi.flam3_xform._fields_ if f.startswith(name.split("_")[0] + "_")]
    for f in params:
        x.params[f] = float(rng.uniform(0.2, 2.0))
    return G.Genome(width=32, height=32, xforms=[x])


@pytest.mark.parametrize("vi", range(99))
def test_each_variation_bit_exact(oracle, ref, vi):
    """or_apply_xform on a grid vs the reference's flam3_xform_preview."""
    rng = np.random.default_rng(vi)
    g = _single_variation_genome(vi, rng)
    with ref.Parsed(G.flame_document([g])) as p:
        R = ref.ref()
        nv, depth = 6, 3
        want = np.zeros(2 * (2 * nv + 1) ** 2)
        rcb = ref.seeded_rc("v%d" % vi)
        assert R.flam3_xform_preview(p.ptr, 0, 2.0, nv, depth, want.ctypes.data_as(ctypes.c_void_p),
                                     ctypes.byref(rcb)) == 0
        rca = oracle.seeded_rc("v%d" % vi)
        got = []
        incr = 2.0 / nv
        for xx in range(-nv, nv + 1):
            for yy in range(-nv, nv + 1):
                pt = np.array([xx * incr, yy * incr, 0.0, 0.0])
                q = np.zeros(4)
                for _ in range(depth):
                    oracle.lib().or_apply_xform(p.ptr, 0, pt.ctypes.data_as(ctypes.c_void_p),
                                                q.ctypes.data_as(ctypes.c_void_p), ctypes.byref(rca))
                    pt[:] = q
                got += [pt[0], pt[1]]
        assert np.array_equal(np.array(got), want, equal_nan=True)


# ---------------------------------------------------------------- full renders

RENDER_CASES = [
    ("small", lambda: G.synthetic_small(seed=31, width=80, height=60, quality=15), {}),
    ("ss2_final", lambda: G.synthetic_1080p(seed=32, width=64, height=36, quality=15, supersample=2), {}),
    ("ss3_xaos", lambda: G.synthetic_8k(seed=33, width=48, height=27, quality=8, supersample=3), {}),
    ("earlyclip_rgba", lambda: G.synthetic_small(seed=34, width=48, height=36, quality=10),
     dict(earlyclip=1, nchan=4, transp=1)),
    ("bits64_16bpc", lambda: G.synthetic_small(seed=35, width=48, height=36, quality=10), dict(bits=64, bpc=2)),
    ("bits32", lambda: G.synthetic_small(seed=36, width=48, height=36, quality=10), dict(bits=32)),
    ("field_odd", lambda: G.synthetic_small(seed=37, width=48, height=36, quality=10), dict(field=2)),
]


def _variant(name, g):
    if name == "small":
        g.rotate = 17.0
        g.palette_mode = "linear"
        g.highlight_power = 1.5
        g.vibrancy = 0.6
        g.background = (10.0, 20.0, 30.0)
    return g


@pytest.mark.parametrize("name,mk,kw", RENDER_CASES, ids=[c[0] for c in RENDER_CASES])
def test_render_byte_identical(oracle, ref, name, mk, kw):
    g = _variant(name, mk())
    xml = G.flame_document([g])
    want, ws = ref.render(xml, seed="pin", **kw)
    got, gs = oracle.render(None, xml=xml, seed="pin", phantom=1, **kw)
    assert np.array_equal(got, want)
    assert gs["badvals"] == ws["badvals"] and gs["num_iters"] == ws["num_iters"]
    assert list(gs["rc"].randrsl) == list(ws["rc"].randrsl)


@pytest.mark.parametrize("name", ["small_q10", "ss2_final_q8", "b64_rgba16"])
def test_render_matches_golden(oracle, name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        meta = json.load(f)
    with open(os.path.join(GOLDEN, name + ".flam3")) as f:
        xml = f.read()
    want = np.load(os.path.join(GOLDEN, name + ".npy"))
    got, st = oracle.render(None, xml=xml, seed=meta["seed"], phantom=1, **meta["settings"])
    assert np.array_equal(got, want)
    assert st["num_iters"] == meta["num_iters"]


# ---------------------------------------------------------------- stage invariants

def _camera(width=40, height=30):
    from flam3_amd._lib import fl3h_camera
    cam = fl3h_camera()
    cam.bounds[0], cam.bounds[1], cam.bounds[2], cam.bounds[3] = -2.0, -1.5, 2.0, 1.5
    cam.rot[0][0] = cam.rot[1][1] = 1.0
    cam.width, cam.height = width, height
    cam.ws0 = width / 4.0
    cam.wb0s0 = cam.ws0 * -2.0
    cam.hs1 = height / 3.0
    cam.hb1s1 = cam.hs1 * -1.5
    cam.cmap_size = 256
    for j in range(256):
        for k in range(4):
            cam.dmap[j][k] = (j / 255.0 if k < 3 else 1.0) * 255.0
    return cam


def test_deposit_hit_count_closed_form(oracle):
    """Sum of the hit channel = 255 x (in-bounds visible points), and every
    in-bounds point lands in the cell its coordinates name."""
    rng = np.random.default_rng(5)
    n = 20000
    pts = np.zeros((n, 4))
    pts[:, 0] = rng.uniform(-2.5, 2.5, n)
    pts[:, 1] = rng.uniform(-2.0, 2.0, n)
    pts[:, 2] = rng.uniform(0, 1, n)
    pts[:, 3] = 1.0
    pts[::7, 3] = 0.0
    cam = _camera()
    b = np.zeros((30, 40, 5), dtype=np.uint32)
    oracle.lib().or_deposit(ctypes.byref(cam), pts.ctypes.data_as(ctypes.c_void_p), n, 33, 1,
                            b.ctypes.data_as(ctypes.c_void_p))
    inb = (pts[:, 0] >= -2) & (pts[:, 0] <= 2) & (pts[:, 1] >= -1.5) & (pts[:, 1] <= 1.5) & (pts[:, 3] != 0)
    ix = np.floor((pts[:, 0] + 2) * 10).astype(int)
    iy = np.floor((pts[:, 1] + 1.5) * 10).astype(int)
    keep = inb & (ix < 40) & (iy < 30)
    want = np.zeros((30, 40), dtype=np.int64)
    np.add.at(want, (iy[keep], ix[keep]), 255)
    assert b[:, :, 4].sum() == 255 * keep.sum()
    assert np.array_equal(b[:, :, 4].astype(np.int64), want)
    # colour index = floor(c * 256) clamped (step palette): channel 0 adds floor(ci/255*255)
    c0 = np.zeros((30, 40), dtype=np.int64)
    ci = np.clip((pts[:, 2] * 256).astype(int), 0, 255)
    np.add.at(c0, (iy[keep], ix[keep]), np.floor(ci[keep] / 255.0 * 255.0).astype(np.int64))
    assert np.array_equal(b[:, :, 0].astype(np.int64), c0)


def test_density_estimation_conserves_log_density(oracle, ref):
    """Each DE kernel sums to 1 over its disc, so away from the edges the
    accumulator's alpha channel totals sum_cells k1*log(1+c3*k2) exactly as
    plain log scaling does."""
    rng = np.random.default_rng(6)
    W, H = 80, 60
    b = np.zeros((H, W, 5), dtype=np.uint32)
    b[20:40, 25:55, 4] = rng.integers(0, 4, (20, 30)) * 255
    b[:, :, 3] = b[:, :, 4]
    b[:, :, 0] = b[:, :, 4] // 2
    mfc, mfi, ks = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    wp, cp_ = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_double)()
    assert ref.ref().ref_de_filters(9.0, 0.0, 0.4, 1, ctypes.byref(mfc), ctypes.byref(mfi), ctypes.byref(ks),
                                    ctypes.byref(wp), ctypes.byref(cp_)) == 0
    acc_de = np.zeros((H, W, 4), dtype=np.float64)
    b64 = b.astype(np.float64)
    oracle.lib().or_density(b64.ctypes.data_as(ctypes.c_void_p), W, H, 64, 1, 1.3, 0.02, 0.4, mfc.value, mfi.value,
                            ks.value, wp, cp_, acc_de.ctypes.data_as(ctypes.c_void_p), 0)
    acc_ls = np.zeros((H, W, 4), dtype=np.float64)
    oracle.lib().or_logscale(b64.ctypes.data_as(ctypes.c_void_p), W, H, 64, 1.3, 0.02,
                             acc_ls.ctypes.data_as(ctypes.c_void_p))
    assert math.isclose(acc_de[:, :, 3].sum(), acc_ls[:, :, 3].sum(), rel_tol=1e-12)
    assert math.isclose(acc_de[:, :, 0].sum(), acc_ls[:, :, 0].sum(), rel_tol=1e-12)
    # and the spread is real: DE output is wider than the hit region
    assert (acc_de[:, :, 3] > 0).sum() > (acc_ls[:, :, 3] > 0).sum()


def test_tonemap_identity_filter_closed_form(oracle):
    """With a 1x1 unit filter, vibrancy 1, gamma 1 and no highlight
    control, a cell with all channels equal to a maps to
    min(255, 256*alpha*a/255 / (a/255)...) = min(255, 256 * k * a / 255)."""
    from flam3_amd._lib import fl3h_tonemap_params
    W, H = 4, 1
    acc = np.zeros((H, W, 4), dtype=np.float32)
    vals = [0.0, 10.0, 100.0, 254.0]
    for i, v in enumerate(vals):
        acc[0, i, :] = v
    tp = fl3h_tonemap_params(image_width=W, image_height=H, oversample=1, de_offset=0, filter_width=1, nchan=3,
                             transp=0, bytes_per_channel=1, earlyclip=0, g=1.0, linrange=0.0, vibrancy=1.0,
                             highpow=-1.0)
    out = np.zeros((H, W, 3), dtype=np.uint8)
    filt = np.ones(1)
    oracle.lib().or_tonemap(acc.ctypes.data_as(ctypes.c_void_p), W, H, 33, ctypes.byref(tp),
                            filt.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), W)
    for i, v in enumerate(vals):
        # alpha = v/255 (gamma 1, linrange 0); ls = 256*alpha/(v/255) = 256; colour = 256*v/255
        want = 0 if v == 0 else min(255, int(256.0 * v / 255.0))
        assert int(out[0, i, 0]) == want
