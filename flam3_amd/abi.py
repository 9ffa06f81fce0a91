"""ctypes mirror of the flam3 data types that cross the render boundary.

The field lists are read from ``include/flam3_abi.h`` (which is
layout-compatible with the reference's flam3.h:185-615), so the Python side
cannot drift from the C declarations.  ``check_layout`` compares the sizes
against the compiled library.
"""
import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
_HEADER = os.path.join(os.path.dirname(_HERE), "include", "flam3_abi.h")

_CONSTS = {"FLAM3_NVARIATIONS": 99, "FLAM3_NAME_LEN": 64, "FLAM3_PARENT_FN_LEN": 30,
           "FLAM3_RANDSIZ": 16}


def _struct_body(text, tag):
    """Return the declaration lines between ``{`` and the closing ``} tag;``."""
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    m = re.search(r"\}\s*" + re.escape(tag) + r"\s*;", text)
    if not m:
        raise RuntimeError("struct %s not found in %s" % (tag, _HEADER))
    end = m.start()
    depth, i = 0, end
    while i > 0:
        i -= 1
        if text[i] == "}":
            depth += 1
        elif text[i] == "{":
            if depth == 0:
                break
            depth -= 1
    body = text[i + 1:end]
    return [ln.strip() for ln in body.split(";") if ln.strip()]


def _ctype(base):
    return {"double": ctypes.c_double, "int": ctypes.c_int, "char": ctypes.c_char,
            "ub4": ctypes.c_ulong, "long int": ctypes.c_long, "void": ctypes.c_void_p,
            "struct xform": ctypes.c_void_p}[base]


def _fields(decls, extra):
    out = []
    for d in decls:
        d = " ".join(d.split())
        m = re.match(r"^(double|int|char|ub4|long int|void|struct xform|flam3_xform|flam3_genome|"
                     r"flam3_palette|randctx|double\s*\*\*|char\s*\*)\s*(.*)$", d)
        if d.startswith("int (*progress)"):
            out.append(("progress", ctypes.c_void_p))
            continue
        if not m:
            raise RuntimeError("cannot parse declaration: " + d)
        base, rest = m.group(1), m.group(2)
        for name in [n.strip() for n in rest.split(",")]:
            ptr = 0
            while name.startswith("*"):
                ptr += 1
                name = name[1:].strip()
            if base in ("double **", "double**") or (base.startswith("double") and "**" in base):
                out.append((name, ctypes.c_void_p))
                continue
            if base.startswith("char") and "*" in base:
                out.append((name, ctypes.c_char_p))
                continue
            dims = [int(_CONSTS.get(x, x)) if not x.isdigit() else int(x)
                    for x in re.findall(r"\[([^\]]+)\]", name)]
            dims = [eval(str(x), {}, _CONSTS) for x in dims]
            name = re.sub(r"\[.*", "", name)
            if base in extra:
                t = extra[base]
            else:
                t = _ctype(base)
            if ptr:
                t = ctypes.c_void_p
            for n in reversed(dims):
                t = t * n
            out.append((name, t))
    return out


def _dims_resolve(text):
    return text.replace("FLAM3_NAME_LEN + 1", "65")


with open(_HEADER) as _f:
    _TEXT = _dims_resolve(_f.read())


class randctx(ctypes.Structure):
    _fields_ = _fields(_struct_body(_TEXT, "randctx"), {})


class stat_struct(ctypes.Structure):
    _fields_ = _fields(_struct_body(_TEXT, "stat_struct"), {})


class flam3_palette_entry(ctypes.Structure):
    _fields_ = [("index", ctypes.c_double), ("color", ctypes.c_double * 4)]


class flam3_xform(ctypes.Structure):
    _fields_ = _fields(_struct_body(_TEXT, "flam3_xform"), {})


class flam3_genome(ctypes.Structure):
    _fields_ = _fields(_struct_body(_TEXT, "flam3_genome"),
                       {"flam3_xform": flam3_xform, "flam3_palette": flam3_palette_entry * 256})


class flam3_frame(ctypes.Structure):
    _fields_ = _fields(_struct_body(_TEXT, "flam3_frame"), {"randctx": randctx, "flam3_genome": flam3_genome})


# The pointer-typed fields are plain void* above; typed views for convenience.
XFORM_P = ctypes.POINTER(flam3_xform)
GENOME_P = ctypes.POINTER(flam3_genome)

SIZES = {"genome": 10728, "xform": 3240, "frame": 360, "randctx": 288}


def check_layout():
    got = {"genome": ctypes.sizeof(flam3_genome), "xform": ctypes.sizeof(flam3_xform),
           "frame": ctypes.sizeof(flam3_frame), "randctx": ctypes.sizeof(randctx)}
    if got != SIZES:
        raise RuntimeError("flam3 ABI mismatch: %r != %r" % (got, SIZES))
    return got


VARIATION_NAMES = [
    "linear", "sinusoidal", "spherical", "swirl", "horseshoe", "polar", "handkerchief", "heart",
    "disc", "spiral", "hyperbolic", "diamond", "ex", "julia", "bent", "waves", "fisheye", "popcorn",
    "exponential", "power", "cosine", "rings", "fan", "blob", "pdj", "fan2", "rings2", "eyefish",
    "bubble", "cylinder", "perspective", "noise", "julian", "juliascope", "blur", "gaussian_blur",
    "radial_blur", "pie", "ngon", "curl", "rectangles", "arch", "tangent", "square", "rays", "blade",
    "secant2", "twintrian", "cross", "disc2", "super_shape", "flower", "conic", "parabola", "bent2",
    "bipolar", "boarders", "butterfly", "cell", "cpow", "curve", "edisc", "elliptic", "escher", "foci",
    "lazysusan", "loonie", "pre_blur", "modulus", "oscilloscope", "polar2", "popcorn2", "scry",
    "separation", "split", "splits", "stripes", "wedge", "wedge_julia", "wedge_sph", "whorl", "waves2",
    "exp", "log", "sin", "cos", "tan", "sec", "csc", "cot", "sinh", "cosh", "tanh", "sech", "csch",
    "coth", "auger", "flux", "mobius"]
assert len(VARIATION_NAMES) == 99
