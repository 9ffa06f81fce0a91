"""Python front end of flam3_render (libflam3_amd.so).

``render(genomes, ...)`` mirrors what flam3-render does around the library
call (flam3-render.c:108-425: frame set-up from environment-style settings,
one flam3_render call, the image buffer) and returns the image as a numpy
array of shape (height, width, nchan).
"""
import ctypes

import numpy as np

from . import abi
from ._lib import check, fl3h_render_opts, fl3h_request, last_error, lib
from .genome import Genome, genome_array


def make_frame(genome_ptr, ngenomes, *, bits=33, bpc=1, earlyclip=0, time=0.0, nthreads=1,
               sub_batch_size=10000, pixel_aspect=1.0, seed="flam3", progress=None):
    """A flam3_frame as flam3-render fills it (flam3-render.c:150-230)."""
    f = abi.flam3_frame()
    f.pixel_aspect_ratio = pixel_aspect
    f.genomes = ctypes.cast(genome_ptr, ctypes.c_void_p)
    f.ngenomes = ngenomes
    f.verbose = 0
    f.bits = bits
    f.bytes_per_channel = bpc
    f.earlyclip = earlyclip
    f.time = time
    f.progress = ctypes.cast(progress, ctypes.c_void_p) if progress is not None else None
    f.progress_parameter = None
    f.nthreads = nthreads
    f.sub_batch_size = sub_batch_size
    lib().fl3h_init_frame(ctypes.byref(f), seed.encode() if seed is not None else None)
    return f


def render_frame(frame, width, height, *, nchan=3, transp=0, field=0, replay=False, nwalkers=0, deposit=0,
                 chunk_samples=0, temporal_batch=0):
    """Run fl3h_render on a prepared frame; returns (image, stats)."""
    bpc = frame.bytes_per_channel
    dtype = np.uint16 if bpc == 2 else np.uint8
    img = np.zeros((height, width, nchan), dtype=dtype)
    stats = abi.stat_struct()
    opts = fl3h_render_opts(1 if replay else 0, int(nwalkers), int(deposit), int(chunk_samples), int(temporal_batch))
    rc = lib().fl3h_render(ctypes.byref(frame), img.ctypes.data_as(ctypes.c_void_p), field, nchan, transp,
                           ctypes.byref(stats), ctypes.byref(opts))
    check(rc, "flam3_render")
    return img, {"badvals": stats.badvals, "num_iters": stats.num_iters,
                 "render_seconds": stats.render_seconds}


def render(genomes, *, nchan=3, transp=0, bits=33, bpc=1, earlyclip=0, time=0.0, nthreads=1,
           sub_batch_size=10000, seed="flam3", replay=False, nwalkers=0, field=0, deposit=0, chunk_samples=0,
           temporal_batch=0):
    """Render Genome descriptions (one still, or an animation frame at
    ``time`` when several control points are given)."""
    if isinstance(genomes, Genome):
        genomes = [genomes]
    arr, owned = genome_array(genomes)
    try:
        f = make_frame(ctypes.addressof(arr), len(genomes), bits=bits, bpc=bpc, earlyclip=earlyclip,
                       time=time, nthreads=nthreads, sub_batch_size=sub_batch_size, seed=seed)
        return render_frame(f, genomes[0].width, genomes[0].height, nchan=nchan, transp=transp,
                            field=field, replay=replay, nwalkers=nwalkers, deposit=deposit,
                            chunk_samples=chunk_samples, temporal_batch=temporal_batch)
    finally:
        for gp in owned:
            lib().fl3h_genome_free(gp)


def render_many(items, *, nwalkers=0, deposit=0, temporal_batch=0):
    """Render several independent images with fl3h_render_many (small ones
    share the GPU).  items: dicts with the keys of ``render`` (``genomes``
    required).  Returns [(image, stats)] in order; raises if any failed."""
    frames, imgs, stats, owned, arrs = [], [], [], [], []
    try:
        reqs = (fl3h_request * len(items))()
        for i, it in enumerate(items):
            gs = it["genomes"]
            if isinstance(gs, Genome):
                gs = [gs]
            arr, own = genome_array(gs)
            arrs.append(arr)
            owned.extend(own)
            f = make_frame(ctypes.addressof(arr), len(gs), bits=it.get("bits", 33), bpc=it.get("bpc", 1),
                           earlyclip=it.get("earlyclip", 0), time=it.get("time", 0.0),
                           nthreads=it.get("nthreads", 1), sub_batch_size=it.get("sub_batch_size", 10000),
                           seed=it.get("seed", "flam3"))
            frames.append(f)
            nchan = it.get("nchan", 3)
            dtype = np.uint16 if it.get("bpc", 1) == 2 else np.uint8
            img = np.zeros((gs[0].height, gs[0].width, nchan), dtype=dtype)
            imgs.append(img)
            st = abi.stat_struct()
            stats.append(st)
            reqs[i].frame = ctypes.cast(ctypes.byref(f), ctypes.c_void_p)
            reqs[i].out = img.ctypes.data_as(ctypes.c_void_p)
            reqs[i].field = it.get("field", 0)
            reqs[i].nchan = nchan
            reqs[i].transp = it.get("transp", 0)
            reqs[i].stats = ctypes.cast(ctypes.byref(st), ctypes.c_void_p)
        opts = fl3h_render_opts(0, int(nwalkers), int(deposit), 0, int(temporal_batch))
        nfail = lib().fl3h_render_many(reqs, len(items), ctypes.byref(opts))
        if nfail:
            raise RuntimeError("fl3h_render_many: %d of %d renders failed: %s" % (nfail, len(items), last_error()))
        return [(img, {"badvals": st.badvals, "num_iters": st.num_iters, "render_seconds": st.render_seconds})
                for img, st in zip(imgs, stats)]
    finally:
        for gp in owned:
            lib().fl3h_genome_free(gp)
