"""flam3_amd - the flam3 render path (flam3_render) on AMD MI355X.

The product is the C-ABI library ``libflam3_amd.so`` (declared in
``include/fl3h.h`` and ``include/flam3_abi.h``); this package loads it with
ctypes and adds genome construction and a render helper for Python callers.
"""
from . import abi
from ._lib import LIB_PATH, lib, last_error
from .genome import Genome, Xform, flame_document
from .render import make_frame, render, render_frame, render_many
from .serve import Server

__all__ = ["abi", "LIB_PATH", "lib", "last_error", "Genome", "Xform", "flame_document", "make_frame",
           "render", "render_frame"]
