"""Python front end of the render server (fl3h_server_*): continuous
batching of independent renders on one GPU.

    with Server() as srv:
        t = srv.submit(genome, seed="a")
        img, stats, latency = srv.wait(t)

A submitted request keeps its frame, genome structs and image buffer alive
until it has been waited for.
"""
import ctypes

import numpy as np

from . import abi
from ._lib import fl3h_render_opts, fl3h_server_opts, last_error, lib
from .genome import Genome, genome_array
from .render import make_frame


class Ticket:
    __slots__ = ("id", "frame", "img", "stats", "arr", "owned")


class Server:
    def __init__(self, hist_budget_bytes=0.0, max_batch=0, nwalkers=0, deposit=0, temporal_batch=0):
        o = fl3h_server_opts(float(hist_budget_bytes), int(max_batch),
                             fl3h_render_opts(0, int(nwalkers), int(deposit), 0, int(temporal_batch)))
        self._s = lib().fl3h_server_create(ctypes.byref(o))
        if not self._s:
            raise RuntimeError("fl3h_server_create failed")

    def submit(self, genomes, *, seed="flam3", time=0.0, bits=33, bpc=1, nchan=3, transp=0, earlyclip=0,
               nthreads=1, field=0):
        if isinstance(genomes, Genome):
            genomes = [genomes]
        t = Ticket()
        t.arr, t.owned = genome_array(genomes)
        t.frame = make_frame(ctypes.addressof(t.arr), len(genomes), bits=bits, bpc=bpc, earlyclip=earlyclip,
                             time=time, nthreads=nthreads, seed=seed)
        t.img = np.zeros((genomes[0].height, genomes[0].width, nchan), dtype=np.uint16 if bpc == 2 else np.uint8)
        t.stats = abi.stat_struct()
        t.id = lib().fl3h_server_submit(self._s, ctypes.byref(t.frame), t.img.ctypes.data_as(ctypes.c_void_p),
                                        field, nchan, transp, ctypes.byref(t.stats))
        if t.id < 0:
            self._free(t)
            raise RuntimeError("fl3h_server_submit failed")
        return t

    def wait(self, t):
        """(image, stats, seconds from submit to done) of a ticket."""
        lat = ctypes.c_double(0.0)
        rc = lib().fl3h_server_wait(self._s, t.id, ctypes.byref(lat))
        self._free(t)
        if rc != 0:
            raise RuntimeError("render %d failed: %s" % (t.id, last_error()))
        return t.img, {"badvals": t.stats.badvals, "num_iters": t.stats.num_iters}, lat.value

    def done(self, t):
        return bool(lib().fl3h_server_done(self._s, t.id))

    def counts(self):
        """(batches, renders) so far."""
        b, r = ctypes.c_longlong(0), ctypes.c_longlong(0)
        lib().fl3h_server_counts(self._s, ctypes.byref(b), ctypes.byref(r), None, None)
        return b.value, r.value

    def busy_seconds(self):
        """(planner host seconds, runner device seconds) so far."""
        p, g = ctypes.c_double(0), ctypes.c_double(0)
        lib().fl3h_server_counts(self._s, None, None, ctypes.byref(p), ctypes.byref(g))
        return p.value, g.value

    @staticmethod
    def _free(t):
        for gp in t.owned:
            lib().fl3h_genome_free(gp)
        t.owned = []

    def close(self):
        if self._s:
            lib().fl3h_server_destroy(self._s)
            self._s = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
