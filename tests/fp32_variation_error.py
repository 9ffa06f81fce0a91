"""Error of the walkers' FP32 xform step against the FP64 oracle, per
variation (shared by tests/test_gpu_stages.py and
tests/golden/make_fp32_bounds.py, which writes the per-variation bounds).

Each deterministic variation is measured on a 41x41 grid over [-2, 2]^2 at
depth 1 (one application of a seeded one-variation xform, the genome of
``variation_genome``).  The error of a point is
max over x, y of |fp32 - fp64| / max(1, |fp64|).

Points are excluded only where the FP64 map itself is ill-conditioned: a
point is "near a pole or a discontinuity" when moving the input by 1e-6
(in x or in y, either sign) changes the FP64 output by more than 1e-3
relative (a local gain above 1e3: rounding the input to FP32 alone, a 6e-8
relative change, then moves the output by more than 6e-5), or when one of
those neighbours is not finite.  Points whose FP64 output is not finite or
beyond 1e9 are counted as non-finite and excluded too.
"""
import ctypes

import numpy as np

from flam3_amd import abi
from flam3_amd import genome as G

GRID = 41
SPAN = 2.0
STEP = 1e-6
GAIN = 1e3

# Variations that draw random numbers (variations.c: the ones that call
# flam3_random_isaac_*): the FP32 walkers use their own streams, so these are
# compared by the distribution of many draws, not point by point.
RANDOM_VARS = {"julia", "noise", "julian", "juliascope", "blur", "gaussian_blur", "radial_blur", "pie", "arch",
               "square", "rays", "blade", "twintrian", "super_shape", "flower", "conic", "parabola", "boarders",
               "cpow", "pre_blur", "wedge_julia"}


def variation_genome(vi):
    rng = np.random.default_rng(1000 + vi)
    name = abi.VARIATION_NAMES[vi]
    x = G.Xform(weight=1.0, color=0.5, coefs=tuple(rng.uniform(-1, 1, 6)),
                variations={name: float(rng.uniform(0.3, 1.2))})
    if name == "pre_blur":
        x.variations["linear"] = 1.0
    for f, _ in abi.flam3_xform._fields_:
        if f.startswith(name.split("_")[0] + "_"):
            x.params[f] = float(rng.uniform(0.2, 2.0))
    return G.Genome(width=32, height=32, xforms=[x])


def apply_oracle(oracle, g, pts, seed):
    """One FP64 application of xform 0 to each point (oracle restatement of
    apply_xform, variations.c:2129-2399)."""
    import pyref
    out = np.zeros_like(pts)
    with pyref.Parsed(G.flame_document([g])) as p:
        pyref.ref().prepare_precalc_flags(p.ptr)
        pyref.ref().xform_precalc(p.ptr, 0)
        rco = oracle.seeded_rc(seed)
        q = np.zeros(4)
        for i in range(len(pts)):
            pt = np.ascontiguousarray(pts[i])
            oracle.lib().or_apply_xform(p.ptr, 0, pt.ctypes.data_as(ctypes.c_void_p),
                                        q.ctypes.data_as(ctypes.c_void_p), ctypes.byref(rco))
            out[i] = q
    return out


def apply_fp32(prod, g, pts, tier=-1):
    """The production walkers' FP32 xform step (fl3h_apply_xform_fp32)."""
    out = np.zeros_like(pts)
    gp = g.to_struct()
    try:
        assert prod.lib().fl3h_apply_xform_fp32(gp, 0, pts.ctypes.data_as(ctypes.c_void_p), len(pts),
                                                out.ctypes.data_as(ctypes.c_void_p), tier) == 0, prod.last_error()
    finally:
        G.free_struct(gp)
    return out


def grid_points():
    xs = np.linspace(-SPAN, SPAN, GRID)
    return np.array([[x, y, 0.5, 0.0] for x in xs for y in xs])


def well_conditioned(oracle, g, pts, want, seed):
    """Mask of the points where the FP64 map's local gain is at most GAIN."""
    ok = np.ones(len(pts), dtype=bool)
    scale = np.maximum(1.0, np.abs(want[:, :2]))
    for k in (0, 1):
        for sgn in (-1.0, 1.0):
            q = pts.copy()
            q[:, k] += sgn * STEP
            nb = apply_oracle(oracle, g, q, seed)
            fin = np.isfinite(nb[:, :2]).all(axis=1)
            with np.errstate(invalid="ignore"):
                gain = (np.abs(nb[:, :2] - want[:, :2]) / scale).max(axis=1) / STEP
            ok &= fin & (gain <= GAIN)
    return ok


def measure(prod, oracle, vi):
    """Error statistics of deterministic variation vi: dict with p50, p99,
    max (relative), finite (fraction of grid points with a finite FP64
    output below 1e9 that the FP32 step also reports as good), excluded
    (fraction left out as ill-conditioned)."""
    g = variation_genome(vi)
    seed = "fp32-%d" % vi
    pts = grid_points()
    got = apply_fp32(prod, g, pts)
    want = apply_oracle(oracle, g, pts, seed)
    finite = np.isfinite(want[:, :2]).all(axis=1) & (np.abs(want[:, :2]) < 1e9).all(axis=1)
    good = finite & (got[:, 3] == 0)
    cond = well_conditioned(oracle, g, pts, want, seed)
    use = good & cond
    with np.errstate(invalid="ignore"):
        rel = (np.abs(got[:, :2] - want[:, :2]) / np.maximum(1.0, np.abs(want[:, :2]))).max(axis=1)
    r = rel[use]
    colour = float(np.abs(got[use, 2] - want[use, 2]).max()) if use.any() else 0.0
    return {"name": abi.VARIATION_NAMES[vi], "points": len(pts), "used": int(use.sum()),
            "finite": float(good.mean()), "fp64_finite": float(finite.mean()),
            "excluded": float((finite & ~cond).mean()),
            "p50": float(np.percentile(r, 50)) if len(r) else 0.0,
            "p99": float(np.percentile(r, 99)) if len(r) else 0.0,
            "max": float(r.max()) if len(r) else 0.0, "colour_max": colour}
