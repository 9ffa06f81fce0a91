#!/usr/bin/env python3
"""bench.py - flam3_render on one MI355X: chaos-game samples/s and frame time
of the 1080p still (1920x1080, supersample 2, quality 500, 6 xforms + final
xform, DE radius 9, bits 33).

A step is one full flam3_render call through the C ABI (genome in, 8-bit
image out): interpolation and tables on the host, iteration + deposit,
density estimation and the tone map on the GPU.  `value` is plotted samples
per second over the timed steps.

The roofline entry is for the dominant kernel, the walker kernel
k_iter_fast, timed with HIP events on its stream inside the timed region.
It has no compulsory memory traffic (genome in LDS, walker state in
registers, one 4-byte record written per sample), so it is bound by FP32
vector arithmetic: achieved = samples per launch x the genome's counted FP32
operations per sample (flam3_amd.opcount) / average launch time, against
MI355X's FP32 peak of 157.3 TFLOP/s.  The other stages are reported with
their algorithmic bytes against HBM.  The CPU baseline is the reference
library itself (oracle/_ref, its own -O3 -ffast-math flags, all host
threads) on a reduced-quality render of the same genome.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "chaos-game samples/s + 1080p q500 frame time; preview-request p50/p99 latency"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# Algorithmic bytes of the binned deposit per plotted sample of this genome
# (4-byte records): read and written by the scatter, read by the tile
# accumulation.
BIN_BYTES_PER_SAMPLE = 12.0
# Density estimation per histogram cell (bits 33): 20-byte bucket read,
# 16-byte accumulator read and written.
DE_BYTES_PER_CELL = 20.0 + 32.0


# HBM bytes per walker launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE,
# calibrated by k_copy_probe (profiles/k_iter_fast_1080p_q500_hbm_bytes.txt).
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_iter_fast_1080p_q500_hbm_bytes.txt")


def measured_traffic():
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--quality", type=float, default=500.0)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-quality", type=float, default=60.0)
    p.add_argument("--copy-probe", action="store_true",
                   help="after the timed steps, run k_copy_probe over 1 GiB (a plain copy kernel of known traffic: "
                        "1 GiB read, 1 GiB written) to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE")
    return p.parse_args()


def stage_times(L):
    from flam3_amd._lib import lib
    st = fl3h_stage_times()
    lib().fl3h_profile_get(ctypes.byref(st))
    return st


class fl3h_stage_times(ctypes.Structure):
    _fields_ = [("iterate_ms", ctypes.c_double), ("density_ms", ctypes.c_double), ("tonemap_ms", ctypes.c_double),
                ("iterate_launches", ctypes.c_longlong), ("density_launches", ctypes.c_longlong),
                ("tonemap_launches", ctypes.c_longlong), ("samples", ctypes.c_longlong), ("cells", ctypes.c_longlong),
                ("bin_ms", ctypes.c_double), ("bin_launches", ctypes.c_longlong)]


def gpu_bench(args):
    import numpy as np
    import flam3_amd as F
    from flam3_amd import abi
    from flam3_amd import genome as G

    L = F.lib()
    L.fl3h_profile_reset.argtypes = [ctypes.c_int]
    L.fl3h_profile_get.argtypes = [ctypes.POINTER(fl3h_stage_times)]
    g = G.synthetic_1080p(seed=2, quality=args.quality)
    arr, owned = G.genome_array([g])
    img = np.zeros((g.height, g.width, 3), dtype=np.uint8)
    stats = abi.stat_struct()

    def one(seed):
        f = F.make_frame(ctypes.addressof(arr), 1, bits=33, nthreads=1, seed=seed)
        rc = L.flam3_render(ctypes.byref(f), img.ctypes.data_as(ctypes.c_void_p), 0, 3, 0, ctypes.byref(stats))
        if rc != 0:
            raise RuntimeError("flam3_render failed: " + F.last_error())
        return stats.num_iters

    try:
        for w in range(args.warmup):
            one("warmup-%d" % w)
            log("warmup %d done" % w)
        L.fl3h_profile_reset(1)
        samples = 0
        t0 = time.perf_counter()
        for k in range(args.steps):
            samples += one("step-%d" % k)
        t1 = time.perf_counter()
        st = fl3h_stage_times()
        L.fl3h_profile_get(ctypes.byref(st))
        L.fl3h_profile_reset(0)
        log("timed %d steps: %.3f s; iterate %.2f ms/launch (%d), bin %.2f ms/launch (%d), DE %.2f ms, tonemap %.3f ms"
            % (args.steps, t1 - t0, st.iterate_ms / max(1, st.iterate_launches), st.iterate_launches,
               st.bin_ms / max(1, st.bin_launches), st.bin_launches,
               st.density_ms / max(1, st.density_launches), st.tonemap_ms / max(1, st.tonemap_launches)))
        from flam3_amd import opcount
        return {"samples": samples, "seconds": t1 - t0, "stage": st, "image_mean": float(img.mean()),
                "cells": st.cells, "ops_per_sample": opcount.ops_per_sample(g)}
    finally:
        for gp in owned:
            G.free_struct(gp)


def cpu_baseline(args):
    """Reference flam3_render (oracle/_ref fast build) with every host thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyref
    from flam3_amd import genome as G

    if not pyref.available("fast"):
        return None
    # the box's CPU share for one GPU is 16 threads, whatever os.cpu_count() reports
    threads = max(1, min(16, os.cpu_count() or 1))
    g = G.synthetic_1080p(seed=2, quality=args.cpu_quality)
    xml = G.flame_document([g])
    t0 = time.perf_counter()
    _, st = pyref.render(xml, seed="cpu", nthreads=threads, kind="fast")
    dt = time.perf_counter() - t0
    return {"value": st["num_iters"] / dt, "unit": "samples/s", "cores": threads, "kind": "reference",
            "sample": "reference flam3_render of the same 1080p genome at quality %g (%d samples, %d threads, "
                      "%.2f s wall incl. DE and tone map)" % (args.cpu_quality, st["num_iters"], threads, dt)}


def main():
    args = parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    r = gpu_bench(args)
    if args.copy_probe:
        import flam3_amd as F
        L = F.lib()
        L.fl3h_copy_probe.argtypes = [ctypes.c_longlong]
        if L.fl3h_copy_probe(1 << 30) != 0:
            raise RuntimeError("copy probe failed: " + F.last_error())
        log("copy probe: k_copy_probe, 1073741824 bytes read and written")
    st = r["stage"]
    samples_per_step = r["samples"] / args.steps
    ms_per_step = 1000.0 * r["seconds"] / args.steps
    value = r["samples"] / r["seconds"]
    it_ms = st.iterate_ms / max(1, st.iterate_launches)
    per_launch = st.samples / max(1, st.iterate_launches)
    ops = r["ops_per_sample"]
    achieved = per_launch * ops / (it_ms * 1e-3) / 1e12
    bin_ms = st.bin_ms / max(1, st.bin_launches)
    de_ms = st.density_ms / max(1, st.density_launches)
    out = {
        "metric": METRIC, "value": value, "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "1080p still: seeded genome, 6 xforms (3 of 12 variations each) + final xform, "
                               "1920x1080 supersample 2 quality %g, DE radius 9, bits 33, 8-bit RGB out"
                               % args.quality, "samples_per_frame": samples_per_step,
                   "frame_time_ms": ms_per_step},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": measured_traffic() if args.quality == 500.0 else None,
                     "kernel": "k_iter_fast", "kernel_ms": it_ms, "samples_per_launch": per_launch,
                     "ops_per_sample": ops,
                     "algorithmic_bytes_per_launch": 4.0 * per_launch,
                     "note": "FP32 vector-ALU bound (no matrix work on this path); peak is MI355X FP32, "
                             "157.3 TFLOP/s for vector and dense matrix alike"},
        "stages": {
            "iterate_ms_per_launch": it_ms,
            "bin_ms_per_chunk": bin_ms,
            "bin_GBps": (per_launch * BIN_BYTES_PER_SAMPLE / (bin_ms * 1e-3) / 1e9) if bin_ms > 0 else None,
            "density_ms": de_ms,
            "density_GBps": (st.cells * DE_BYTES_PER_CELL / (de_ms * 1e-3) / 1e9) if de_ms > 0 else None,
            "tonemap_ms": st.tonemap_ms / max(1, st.tonemap_launches)},
    }
    if not args.no_cpu_baseline:
        cb = cpu_baseline(args)
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
