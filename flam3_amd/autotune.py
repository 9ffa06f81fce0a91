"""Autotuner for the library's launch parameters (fl3h_tuning keys).

    python -m flam3_amd.autotune [--out flam3_amd/tuning.conf] [--quick]

Times renders on the GPU for each candidate value of one parameter at a
time (coordinate descent from the defaults, two rounds) and writes the best
setting as "key value" lines, which the library loads at start-up from
tuning.conf beside libflam3_amd.so (or the file FL3H_TUNING names).

Workloads and the score each parameter is chosen on:

* wg_per_cu, chunk_samples, de_halo: frame time of the 1080p still
  (1920x1080 supersample 2, 6 xforms + final xform, DE radius 9) at quality
  500 (several chunks), plus that of one animation frame (1280x720
  supersample 2, 100 temporal samples, DE radius 7) at quality 200.
* wg_per_cu_single (launches of one chunk): the animation frame alone.

min_samples_small (samples per walker of small launches) is not tuned here: it
trades preview latency against preview throughput (fewer samples per
walker: shorter walkers, more fuse steps in all), a policy rather than a
speed optimum; the library's default favours latency.

Each timing is the median of seven repeats after a warm-up render.  A
parameter keeps its current value unless a candidate beats it by more than
3% (the spread of a score between rounds measured here).
"""
import argparse
import ctypes
import os
import statistics
import sys
import time

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(_HERE))

import numpy as np  # noqa: E402

from flam3_amd import abi  # noqa: E402
from flam3_amd import genome as G  # noqa: E402
from flam3_amd._lib import lib  # noqa: E402
from flam3_amd.render import make_frame  # noqa: E402

CANDIDATES = {
    "wg_per_cu": [4, 6, 7, 8, 12],
    "wg_per_cu_single": [4, 6, 7, 8, 12],
    "chunk_samples": [1 << 27, 1 << 28, 1 << 29],
    "de_halo": [0, 1],
}
MARGIN = 0.03


def _set(key, value):
    L = lib()
    L.fl3h_tuning.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    if L.fl3h_tuning(key.encode(), ctypes.byref(ctypes.c_double(float(value))), None) != 0:
        raise ValueError("tuning %s=%r rejected" % (key, value))


def _get(key):
    L = lib()
    L.fl3h_tuning.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    out = ctypes.c_double()
    if L.fl3h_tuning(key.encode(), None, ctypes.byref(out)) != 0:
        raise ValueError(key)
    return out.value


class Frames:
    """The timed render workloads, built once."""

    def __init__(self, quick):
        self.still = G.synthetic_1080p(seed=2, quality=100.0 if quick else 500.0)
        self.anim = G.synthetic_animation(seed=3, quality=50.0 if quick else 200.0)
        self.repeats = 3 if quick else 7

    def _render(self, gs, t=0.0):
        arr, owned = G.genome_array(gs)
        try:
            f = make_frame(ctypes.addressof(arr), len(gs), time=t, seed="tune")
            img = np.zeros((gs[0].height, gs[0].width, 3), np.uint8)
            st = abi.stat_struct()
            t0 = time.perf_counter()
            if lib().flam3_render(ctypes.byref(f), img.ctypes.data_as(ctypes.c_void_p), 0, 3, 0,
                                  ctypes.byref(st)) != 0:
                raise RuntimeError("render failed")
            return time.perf_counter() - t0
        finally:
            for gp in owned:
                G.free_struct(gp)

    def frame_score(self):
        # (warm-up renders of both: the first render of a workload grows
        # buffers and builds filter banks)
        self._render([self.still])
        self._render(self.anim, 37.0)
        a = statistics.median(self._render([self.still]) for _ in range(self.repeats))
        b = statistics.median(self._render(self.anim, 37.0) for _ in range(self.repeats))
        return a + b

    def single_score(self):
        self._render(self.anim, 37.0)
        return statistics.median(self._render(self.anim, 37.0) for _ in range(self.repeats))



def tune(quick=False, log=print):
    frames = Frames(quick)
    best = {k: _get(k) for k in CANDIDATES}
    for rnd in range(2):
        for key, cands in CANDIDATES.items():
            score_fn = {"wg_per_cu_single": frames.single_score}.get(key, frames.frame_score)
            _set(key, best[key])
            base = score_fn()
            results = {best[key]: base}
            for v in cands:
                if v == best[key]:
                    continue
                _set(key, v)
                results[v] = score_fn()
            pick = min(results, key=results.get)
            if results[pick] < base * (1.0 - MARGIN):
                best[key] = pick
            _set(key, best[key])
            log("round %d %s: %s -> %s" % (rnd, key, {k: round(1e3 * s, 2) for k, s in results.items()},
                                           best[key]))
    return best


def write_conf(path, best, note):
    with open(path, "w") as fh:
        fh.write("# launch parameters chosen by flam3_amd/autotune.py (%s)\n" % note)
        for k, v in best.items():
            fh.write("%s %d\n" % (k, int(v)))


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--out", default=os.path.join(_HERE, "tuning.conf"))
    p.add_argument("--quick", action="store_true", help="lower qualities and fewer repeats")
    a = p.parse_args()
    os.environ["FL3H_TUNING"] = os.devnull   # start from the defaults, not an earlier result
    best = tune(a.quick, log=lambda *m: print(*m, flush=True))
    write_conf(a.out, best, "1080p q500 still + animation frame q200")
    print("wrote", a.out, best, flush=True)


if __name__ == "__main__":
    main()
