"""Loader for libflam3_amd.so, the C-ABI library of this package.

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
flam3_amd/csrc``).  There is no fallback: if the library is missing every
entry point raises.
"""
import ctypes
import os

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libflam3_amd.so")

_lib = None

EXPORTS = [
    "flam3_render", "flam3_render_memory_required", "fl3h_render", "fl3h_render_many", "fl3h_server_create",
    "fl3h_server_submit", "fl3h_server_wait", "fl3h_server_done", "fl3h_server_counts", "fl3h_server_destroy",
    "fl3h_iterate",
    "fl3h_xform_preview", "fl3h_deposit_points", "fl3h_density_estimate", "fl3h_tonemap",
    "fl3h_init_frame", "fl3h_genome_new", "fl3h_genome_free", "fl3h_last_error",
]


class fl3h_render_opts(ctypes.Structure):
    _fields_ = [("replay", ctypes.c_int), ("nwalkers", ctypes.c_int), ("deposit", ctypes.c_int),
                ("chunk_samples", ctypes.c_longlong), ("temporal_batch", ctypes.c_int)]


class fl3h_request(ctypes.Structure):
    _fields_ = [("frame", ctypes.c_void_p), ("out", ctypes.c_void_p), ("field", ctypes.c_int),
                ("nchan", ctypes.c_int), ("transp", ctypes.c_int), ("stats", ctypes.c_void_p),
                ("rc", ctypes.c_int)]


class fl3h_server_opts(ctypes.Structure):
    _fields_ = [("hist_budget_bytes", ctypes.c_double), ("max_batch", ctypes.c_int), ("opts", fl3h_render_opts)]


class fl3h_camera(ctypes.Structure):
    _fields_ = [("bounds", ctypes.c_double * 4), ("rot", (ctypes.c_double * 2) * 2),
                ("rot_center", ctypes.c_double * 2), ("rotate", ctypes.c_int),
                ("ws0", ctypes.c_double), ("wb0s0", ctypes.c_double), ("hs1", ctypes.c_double),
                ("hb1s1", ctypes.c_double), ("width", ctypes.c_int), ("height", ctypes.c_int),
                ("cmap_size", ctypes.c_int), ("palette_mode", ctypes.c_int),
                ("dmap", (ctypes.c_double * 4) * 256)]


class fl3h_tonemap_params(ctypes.Structure):
    _fields_ = [("image_width", ctypes.c_int), ("image_height", ctypes.c_int), ("oversample", ctypes.c_int),
                ("de_offset", ctypes.c_int), ("filter_width", ctypes.c_int), ("nchan", ctypes.c_int),
                ("transp", ctypes.c_int), ("bytes_per_channel", ctypes.c_int), ("earlyclip", ctypes.c_int),
                ("g", ctypes.c_double), ("linrange", ctypes.c_double), ("vibrancy", ctypes.c_double),
                ("highpow", ctypes.c_double), ("background", ctypes.c_double * 3)]


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libflam3_amd.so not built (%s); run __graft_entry__.build()" % LIB_PATH)
    abi.check_layout()
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.flam3_render.argtypes = [P(abi.flam3_frame), vp, ci, ci, ci, P(abi.stat_struct)]
    L.flam3_render.restype = ci
    L.flam3_render_memory_required.argtypes = [P(abi.flam3_frame)]
    L.flam3_render_memory_required.restype = cd
    L.fl3h_render.argtypes = [P(abi.flam3_frame), vp, ci, ci, ci, P(abi.stat_struct), P(fl3h_render_opts)]
    L.fl3h_render.restype = ci
    L.fl3h_render_many.argtypes = [P(fl3h_request), ci, P(fl3h_render_opts)]
    L.fl3h_render_many.restype = ci
    L.fl3h_server_create.argtypes = [P(fl3h_server_opts)]
    L.fl3h_server_create.restype = vp
    L.fl3h_server_submit.argtypes = [vp, P(abi.flam3_frame), vp, ci, ci, ci, P(abi.stat_struct)]
    L.fl3h_server_submit.restype = ctypes.c_longlong
    L.fl3h_server_wait.argtypes = [vp, ctypes.c_longlong, P(cd)]
    L.fl3h_server_wait.restype = ci
    L.fl3h_server_done.argtypes = [vp, ctypes.c_longlong]
    L.fl3h_server_done.restype = ci
    L.fl3h_server_counts.argtypes = [vp, P(ctypes.c_longlong), P(ctypes.c_longlong), P(cd), P(cd)]
    L.fl3h_server_counts.restype = None
    L.fl3h_server_destroy.argtypes = [vp]
    L.fl3h_server_destroy.restype = None
    L.fl3h_iterate.argtypes = [P(abi.flam3_genome), ci, ci, vp, vp, P(abi.randctx), ci, vp]
    L.fl3h_iterate.restype = ci
    L.fl3h_xform_preview.argtypes = [P(abi.flam3_genome), ci, cd, ci, ci, vp, P(abi.randctx)]
    L.fl3h_xform_preview.restype = ci
    L.fl3h_apply_xform_fp32.argtypes = [P(abi.flam3_genome), ci, vp, ctypes.c_longlong, vp, ci]
    L.fl3h_apply_xform_fp32.restype = ci
    L.fl3h_interpolate.argtypes = [P(abi.flam3_genome), ci, cd, cd, P(abi.flam3_genome)]
    L.fl3h_interpolate.restype = ci
    L.fl3h_genome_clear.argtypes = [P(abi.flam3_genome)]
    L.fl3h_genome_clear.restype = None
    L.fl3h_deposit_points.argtypes = [P(fl3h_camera), vp, ctypes.c_longlong, ci, vp]
    L.fl3h_deposit_points.restype = ci
    L.fl3h_density_estimate.argtypes = [vp, ci, ci, ci, ci, cd, cd, cd, ci, ci, ci, ci, vp, vp, vp]
    L.fl3h_density_estimate.restype = ci
    L.fl3h_density_estimate_production.argtypes = [vp, ci, ci, ci, ci, cd, cd, cd, ci, ci, ci, ci, vp, vp, vp]
    L.fl3h_density_estimate_production.restype = ci
    L.fl3h_tonemap.argtypes = [vp, ci, ci, ci, P(fl3h_tonemap_params), vp, vp]
    L.fl3h_tonemap.restype = ci
    L.fl3h_init_frame.argtypes = [P(abi.flam3_frame), ctypes.c_char_p]
    L.fl3h_init_frame.restype = None
    L.fl3h_genome_new.argtypes = [ci, ci]
    L.fl3h_genome_new.restype = P(abi.flam3_genome)
    L.fl3h_genome_free.argtypes = [P(abi.flam3_genome)]
    L.fl3h_genome_free.restype = None
    L.fl3h_last_error.argtypes = []
    L.fl3h_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def last_error():
    return lib().fl3h_last_error().decode(errors="replace")


def check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed: %s" % (what, last_error()))
