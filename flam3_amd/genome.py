"""Genome descriptions, their flam3 XML form and their C struct form.

``Genome.to_xml()`` produces a ``<flame>`` element the reference parser reads
(parser.c:288-1256 attribute names); ``Genome.to_struct()`` builds the same
``flam3_genome`` directly through libflam3_amd (defaults as the parser leaves
them), so renders need no XML parser.  Numbers are written with ``repr`` so
strtod reads back the exact doubles.

The ``synthetic_*`` builders produce the seeded workloads the benchmarks and
parity tests use (random affines, variation mixes and HSV palettes).
"""
import ctypes
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import abi
from ._lib import lib

VAR_INDEX = {n: i for i, n in enumerate(abi.VARIATION_NAMES)}

# Variations of the 1080p / animation / preview workloads.
WORKLOAD_VARS = ["julia", "swirl", "horseshoe", "polar", "handkerchief", "heart", "disc", "spiral",
                 "hyperbolic", "diamond", "ex", "linear"]

# Spatial filter names by flam3.h kernel id (flam3.h:673-686, parser.c:418-449).
_KERNELS = ["gaussian", "hermite", "box", "triangle", "bell", "bspline", "lanczos3", "lanczos2",
            "mitchell", "blackman", "catrom", "hamming", "hanning", "quadratic"]


@dataclass
class Xform:
    weight: float = 0.0
    color: float = 0.0
    color_speed: float = 0.5
    opacity: float = 1.0
    coefs: tuple = (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)   # c00 c01 c10 c11 c20 c21 (XML order)
    post: Optional[tuple] = None
    variations: Dict[str, float] = field(default_factory=dict)
    params: Dict[str, float] = field(default_factory=dict)
    chaos: Optional[List[float]] = None
    animate: float = 1.0


@dataclass
class Genome:
    width: int = 100
    height: int = 100
    center: tuple = (0.0, 0.0)
    scale: float = 50.0
    rotate: float = 0.0
    zoom: float = 0.0
    quality: float = 1.0
    supersample: int = 1
    passes: int = 1
    temporal_samples: int = 1
    estimator_radius: float = 9.0
    estimator_minimum: float = 0.0
    estimator_curve: float = 0.4
    gamma: float = 4.0
    gamma_threshold: float = 0.01
    brightness: float = 4.0
    vibrancy: float = 1.0
    highlight_power: float = -1.0
    background: tuple = (0.0, 0.0, 0.0)
    filter_radius: float = 0.5
    filter_shape: int = 0
    palette_mode: str = "step"
    palette: Optional[np.ndarray] = None      # 256 x 3, 0..255
    time: float = 0.0
    interpolation_type: str = "log"
    interpolation: str = "linear"             # between control points: linear / smooth
    palette_interpolation: str = "hsv_circular"
    hsv_rgb_palette_blend: float = 0.0
    temporal_filter_type: str = "box"
    temporal_filter_width: float = 1.0
    temporal_filter_exp: float = 0.0
    xforms: List[Xform] = field(default_factory=list)
    final: Optional[Xform] = None

    # ------------------------------------------------------------- XML
    def to_xml(self):
        r = repr
        attrs = [
            ("time", r(self.time)), ("size", "%d %d" % (self.width, self.height)),
            ("center", "%s %s" % (r(self.center[0]), r(self.center[1]))), ("scale", r(self.scale)),
            ("rotate", r(self.rotate)), ("zoom", r(self.zoom)), ("supersample", str(self.supersample)),
            ("filter", r(self.filter_radius)), ("filter_shape", _KERNELS[self.filter_shape]),
            ("quality", r(self.quality)), ("passes", str(self.passes)),
            ("temporal_samples", str(self.temporal_samples)),
            ("background", " ".join(r(b) for b in self.background)),
            ("brightness", r(self.brightness)), ("gamma", r(self.gamma)), ("vibrancy", r(self.vibrancy)),
            ("highlight_power", r(self.highlight_power)), ("estimator_radius", r(self.estimator_radius)),
            ("estimator_minimum", r(self.estimator_minimum)), ("estimator_curve", r(self.estimator_curve)),
            ("gamma_threshold", r(self.gamma_threshold)), ("palette_mode", self.palette_mode),
            ("interpolation_type", self.interpolation_type), ("interpolation", self.interpolation),
            ("palette_interpolation", self.palette_interpolation),
            ("hsv_rgb_palette_blend", r(self.hsv_rgb_palette_blend)),
            ("temporal_filter_type", self.temporal_filter_type),
            ("temporal_filter_width", r(self.temporal_filter_width)),
            ("temporal_filter_exp", r(self.temporal_filter_exp)),
        ]
        out = ["<flame " + " ".join('%s="%s"' % a for a in attrs) + ">"]
        for x in self.xforms:
            out.append(self._xform_xml("xform", x))
        if self.final is not None:
            out.append(self._xform_xml("finalxform", self.final))
        pal = self.palette if self.palette is not None else np.full((256, 3), 255.0)
        for i in range(256):
            out.append('<color index="%d" rgb="%s %s %s"/>' % (i, r(float(pal[i][0])), r(float(pal[i][1])),
                                                                r(float(pal[i][2]))))
        out.append("</flame>")
        return "\n".join(out)

    @staticmethod
    def _xform_xml(tag, x):
        r = repr
        a = []
        if tag == "xform":
            a.append('weight="%s"' % r(x.weight))
        a.append('color="%s"' % r(x.color))
        a.append('color_speed="%s"' % r(x.color_speed))
        if tag == "xform":
            a.append('opacity="%s"' % r(x.opacity))
        if x.animate != 1.0:
            a.append('animate="%s"' % r(x.animate))
        for name, w in x.variations.items():
            a.append('%s="%s"' % (name, r(float(w))))
        for name, v in x.params.items():
            a.append('%s="%s"' % (name, r(float(v))))
        a.append('coefs="%s"' % " ".join(r(float(c)) for c in x.coefs))
        if x.post is not None:
            a.append('post="%s"' % " ".join(r(float(c)) for c in x.post))
        if x.chaos is not None and tag == "xform":
            a.append('chaos="%s"' % " ".join(r(float(c)) for c in x.chaos))
        return "<%s %s/>" % (tag, " ".join(a))

    # ------------------------------------------------------------- struct
    def to_struct(self):
        """A newly allocated flam3_genome (free with free_struct)."""
        L = lib()
        gp = L.fl3h_genome_new(len(self.xforms), 1 if self.final is not None else 0)
        if not gp:
            raise MemoryError("fl3h_genome_new failed")
        g = gp.contents
        g.time = self.time
        g.width, g.height = self.width, self.height
        g.center[0], g.center[1] = self.center
        g.rot_center[0], g.rot_center[1] = self.center
        g.pixels_per_unit = self.scale
        g.rotate = self.rotate
        g.zoom = self.zoom
        g.spatial_oversample = self.supersample
        g.spatial_filter_radius = self.filter_radius
        g.spatial_filter_select = self.filter_shape
        g.sample_density = self.quality
        g.nbatches = self.passes
        g.ntemporal_samples = self.temporal_samples
        for i in range(3):
            g.background[i] = self.background[i]
        g.brightness = self.brightness
        g.gamma = self.gamma
        g.vibrancy = self.vibrancy
        g.highlight_power = self.highlight_power
        g.estimator = self.estimator_radius
        g.estimator_minimum = self.estimator_minimum
        g.estimator_curve = self.estimator_curve
        g.gam_lin_thresh = self.gamma_threshold
        g.palette_mode = 1 if self.palette_mode == "linear" else 0
        g.interpolation_type = {"linear": 0, "log": 1, "old": 2, "older": 3}[self.interpolation_type]
        g.interpolation = {"linear": 0, "smooth": 1}[self.interpolation]
        g.palette_interpolation = {"hsv": 0, "sweep": 1, "rgb": 2, "hsv_circular": 3}[self.palette_interpolation]
        g.hsv_rgb_palette_blend = self.hsv_rgb_palette_blend
        g.temporal_filter_type = {"box": 0, "gaussian": 1, "exp": 2}[self.temporal_filter_type]
        g.temporal_filter_width = self.temporal_filter_width
        g.temporal_filter_exp = self.temporal_filter_exp
        pal = self.palette if self.palette is not None else np.full((256, 3), 255.0)
        for i in range(256):
            e = g.palette[i]
            e.color[3] = 255.0 / 255.0
            for k in range(3):
                e.color[k] = e.color[3] * float(pal[i][k]) / 255.0
            e.index = float(i)
        xf = ctypes.cast(g.xform, abi.XFORM_P)
        nstd = len(self.xforms)
        for i, x in enumerate(self.xforms):
            self._fill_xform(xf[i], x, final=False)
            if x.chaos is not None:
                rows = ctypes.cast(g.chaos, ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))
                for j, c in enumerate(x.chaos[:nstd]):
                    rows[i][j] = float(c)
        if self.final is not None:
            self._fill_xform(xf[nstd], self.final, final=True)
        return gp

    @staticmethod
    def _fill_xform(d, x, final):
        if not final:
            d.density = x.weight
            d.opacity = x.opacity
        d.color = x.color
        d.color_speed = x.color_speed
        if x.animate != 1.0:
            d.animate = x.animate
        for name, w in x.variations.items():
            d.var[VAR_INDEX[name]] = float(w)
        for name, v in x.params.items():
            setattr(d, name, float(v))
        for k in range(3):
            d.c[k][0] = float(x.coefs[2 * k])
            d.c[k][1] = float(x.coefs[2 * k + 1])
        if x.post is not None:
            for k in range(3):
                d.post[k][0] = float(x.post[2 * k])
                d.post[k][1] = float(x.post[2 * k + 1])


def free_struct(gp):
    lib().fl3h_genome_free(gp)


def genome_array(genomes):
    """Contiguous flam3_genome array (shallow copies of the owned structs)
    plus the owned pointers to free afterwards."""
    owned = [g.to_struct() for g in genomes]
    arr = (abi.flam3_genome * len(owned))()
    for i, gp in enumerate(owned):
        ctypes.memmove(ctypes.byref(arr[i]), gp, ctypes.sizeof(abi.flam3_genome))
    return arr, owned


# ------------------------------------------------------------- palettes / builders

def hsv_palette(rng):
    """256 random HSV colours (h in [0,6), s, v in [0,1]) as RGB 0..255."""
    h = rng.uniform(0.0, 6.0, 256)
    s = rng.uniform(0.0, 1.0, 256)
    v = rng.uniform(0.0, 1.0, 256)
    out = np.zeros((256, 3))
    for i in range(256):
        j = int(math.floor(h[i])) % 6
        f = h[i] - math.floor(h[i])
        p, q, t = v[i] * (1 - s[i]), v[i] * (1 - s[i] * f), v[i] * (1 - s[i] * (1 - f))
        rgb = [(v[i], t, p), (q, v[i], p), (p, v[i], t), (p, q, v[i]), (t, p, v[i]), (v[i], p, q)][j]
        out[i] = [c * 255.0 for c in rgb]
    return out


def _mixed_xform(rng, names, k, lo, hi, weight=None):
    picks = rng.choice(len(names), size=k, replace=False)
    ws = rng.uniform(0.0, 1.0, k)
    return Xform(weight=float(rng.uniform(0.2, 1.0)) if weight is None else weight,
                 color=float(rng.uniform(0.0, 1.0)),
                 coefs=tuple(float(c) for c in rng.uniform(lo, hi, 6)),
                 variations={names[p]: float(w) for p, w in zip(sorted(picks), ws[np.argsort(picks)])})


def synthetic_small(seed=1, width=640, height=480, quality=50.0):
    """Three xforms, linear/sinusoidal/spherical with normalised U[0,1]
    blend weights, affines U[-1,1], weights U[0.2,1], random HSV palette;
    640x480 q50 ss1, DE 9/0/0.4, gamma 4 (the CPU reference workload)."""
    rng = np.random.default_rng(seed)
    xs = []
    for _ in range(3):
        w = rng.uniform(0.0, 1.0, 3)
        w = w / w.sum()
        xs.append(Xform(weight=float(rng.uniform(0.2, 1.0)), color=float(rng.uniform(0.0, 1.0)),
                        coefs=tuple(float(c) for c in rng.uniform(-1.0, 1.0, 6)),
                        variations={"linear": float(w[0]), "sinusoidal": float(w[1]),
                                    "spherical": float(w[2])}))
    return Genome(width=width, height=height, scale=height / 3.0, quality=quality, supersample=1,
                  estimator_radius=9.0, estimator_minimum=0.0, estimator_curve=0.4, gamma=4.0,
                  palette=hsv_palette(rng), xforms=xs)


def synthetic_1080p(seed=2, width=1920, height=1080, quality=500.0, supersample=2, nxf=6):
    """nxf xforms plus a final xform, each mixing 3 of the 12 workload
    variations with U[0,1] weights, affines U[-1.2,1.2], random palette;
    1920x1080 ss2 q500 DE 9 (the GPU headline workload)."""
    rng = np.random.default_rng(seed)
    xs = [_mixed_xform(rng, WORKLOAD_VARS, 3, -1.2, 1.2) for _ in range(nxf)]
    fin = _mixed_xform(rng, WORKLOAD_VARS, 3, -1.2, 1.2, weight=0.0)
    fin.color_speed = 0.0
    return Genome(width=width, height=height, scale=height / 3.0, quality=quality, supersample=supersample,
                  estimator_radius=9.0, palette=hsv_palette(rng), xforms=xs, final=fin)


def synthetic_preview(seed, size=None, quality=None):
    """Genome-search preview request: 2-12 xforms from the 12 workload
    variations, size from {64,128,256,512}^2, quality U[10,100], ss1."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 13))
    s = int(rng.choice([64, 128, 256, 512])) if size is None else size
    q = float(rng.uniform(10.0, 100.0)) if quality is None else quality
    xs = [_mixed_xform(rng, WORKLOAD_VARS, int(rng.integers(1, 4)), -1.2, 1.2) for _ in range(n)]
    return Genome(width=s, height=s, scale=s / 3.0, quality=q, supersample=1, estimator_radius=9.0,
                  palette=hsv_palette(rng), xforms=xs)


def synthetic_animation(seed=3, width=1280, height=720, quality=200.0, supersample=2, temporal_samples=100,
                        frames=120, turn=90.0):
    """Two seeded 5-xform control points (3 of the 12 workload variations per
    xform) at times 0 and frames-1; the second's affines are the first's
    rotated by `turn` degrees, so log interpolation turns them a little per
    frame.  1280x720 ss2 q200, 100 temporal samples, DE radius 7."""
    rng = np.random.default_rng(seed)
    xs = [_mixed_xform(rng, WORKLOAD_VARS, 3, -1.2, 1.2) for _ in range(5)]
    pal = hsv_palette(rng)
    c, s_ = math.cos(math.radians(turn)), math.sin(math.radians(turn))
    out = []
    for k in range(2):
        xk = []
        for x in xs:
            a, b, cc, d, e, f = x.coefs
            if k:  # rotate the linear part: [a c; b d] -> R [a c; b d]
                a, b, cc, d = c * a - s_ * b, s_ * a + c * b, c * cc - s_ * d, s_ * cc + c * d
            xk.append(Xform(weight=x.weight, color=x.color, coefs=(a, b, cc, d, e, f), variations=dict(x.variations)))
        out.append(Genome(width=width, height=height, scale=height / 3.0, quality=quality, supersample=supersample,
                          estimator_radius=7.0, temporal_samples=temporal_samples, palette=pal, xforms=xk,
                          time=float(k * (frames - 1))))
    return out


def synthetic_8k(seed=5, width=7680, height=4320, quality=2000.0, supersample=3):
    """8 xforms + final xform + xaos matrix U[0,2]; 7680x4320 ss3 q2000,
    DE radius 15 curve 0.4 (render with bits=64)."""
    rng = np.random.default_rng(seed)
    xs = [_mixed_xform(rng, WORKLOAD_VARS, 3, -1.2, 1.2) for _ in range(8)]
    for x in xs:
        x.chaos = [float(c) for c in rng.uniform(0.0, 2.0, 8)]
    fin = _mixed_xform(rng, WORKLOAD_VARS, 3, -1.2, 1.2, weight=0.0)
    fin.color_speed = 0.0
    return Genome(width=width, height=height, scale=height / 3.0, quality=quality, supersample=supersample,
                  estimator_radius=15.0, estimator_curve=0.4, palette=hsv_palette(rng), xforms=xs, final=fin)


def flame_document(genomes):
    return "<flames>\n" + "\n".join(g.to_xml() for g in genomes) + "\n</flames>\n"
