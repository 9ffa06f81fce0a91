"""Algorithmic FP32 operation count of one plotted sample.

The walker kernel has no compulsory memory traffic (genome in LDS, state
in registers), so its roofline is the FP32 vector ALU.  This module counts
the floating-point operations the reference formulas ask for per plotted
sample of a genome: add/sub/mul/neg = 1, fused multiply-add = 2 (the
formulas are written unfused, so each mul and add counts once), and
div/sqrt/sin/cos/atan2/log/exp/pow = 1 each; sincos = 2.  Comparisons,
integer work (ISAAC/xoshiro, xform selection, the cell index conversion)
and moves are not counted.

Per xform application (variations.c:2129-2399):
  colour blend  s*c + (1-s)*p2                         4
  affine        two of a*x + b*y + e                    8
  sumsq + sqrt  (always computed)                       4
  atan2         once if any variation needs precalc_atan          1
  angles        sin a = tx/r, cos a = ty/r if any needs them      2
  each variation: its formula, including `p += weight * n`
  post affine   when present                            8
Per plotted sample: the chosen xform (by selection probability), the final
xform (probability = its opacity), and the deposit's cell coordinates
(two of ws0*x - wb0s0 = 4) plus the palette index (1 mul) = 5; rotation
adds 10 when the camera is rotated.
"""
from __future__ import annotations

import numpy as np

# ops of each variation's formula (file:line of the formula in variations.c),
# and which precalc it needs: "atan" (precalc_atan), "angles" (sina/cosa)
VARIATION_OPS = {
    "linear": (4, None),          # :142  p += w*tx, p += w*ty
    "sinusoidal": (6, None),      # :153  2 sin + 2 mul + 2 add
    "spherical": (6, None),       # :164  add, div, 2 mul, 2 add
    "swirl": (12, None),          # :178  sincos 2, nx/ny 6, p 4
    "horseshoe": (11, None),      # :202  add+div 2, p0 5, p1 4
    "polar": (6, "atan"),         # :218  mul, sub, p 4
    "handkerchief": (10, "atan"),  # :232  (add, sin, 2 mul, add) x 2
    "heart": (9, "atan"),         # :246  2 mul, sincos 2, p0 2, p1 3
    "disc": (10, "atan"),         # :264  2 mul, sincos 2, p 6
    "spiral": (10, "angles"),     # :282  add, div, sincos 2, p 6
    "hyperbolic": (7, "angles"),  # :298  add, p0 3, p1 3
    "diamond": (8, "angles"),     # :311  sincos 2, p 6
    "ex": (16, "atan"),           # :326  2 (add+trig), 6 mul (cubes x r), p 6
    "julia": (9.5, "atan"),       # :350  mul, half-probability add, sqrt+mul, sincos 2, p 4
}

XFORM_BASE = 4 + 8 + 4
DEPOSIT = 5
ROTATION = 10


def xform_ops(x) -> float:
    """FP ops of one application of Xform x."""
    ops = XFORM_BASE
    needs = set()
    for name, w in x.variations.items():
        if w == 0:
            continue
        if name not in VARIATION_OPS:
            raise KeyError("no operation count for variation %r" % name)
        n, need = VARIATION_OPS[name]
        ops += n
        if need:
            needs.add(need)
    ops += ("atan" in needs) * 1 + ("angles" in needs) * 2
    post = getattr(x, "post", None)
    if post is not None and tuple(post) != (1.0, 0.0, 0.0, 1.0, 0.0, 0.0):
        ops += 8
    return float(ops)


def ops_per_sample(g) -> float:
    """Expected FP ops per plotted sample of Genome g (no xaos: xform
    chosen in proportion to its weight)."""
    ws = np.array([x.weight for x in g.xforms], dtype=np.float64)
    if ws.sum() <= 0:
        raise ValueError("genome has no positive xform weight")
    p = ws / ws.sum()
    ops = float(sum(pi * xform_ops(x) for pi, x in zip(p, g.xforms)))
    if g.final is not None:
        ops += getattr(g.final, "opacity", 1.0) * xform_ops(g.final)
    ops += DEPOSIT
    if getattr(g, "rotate", 0.0):
        ops += ROTATION
    return ops
