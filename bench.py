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
import gc
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


VALU_FILE = os.path.join(ROOT, "profiles", "k_iter_fast_1080p_q500_valu.txt")
CLOCK_HZ = 2.4e9            # MI355X peak engine clock
SIMDS = 1024                # 256 CUs x 4 SIMDs


def profile_value(path, key):
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith(key + ":"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def measured_traffic():
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def valu_issue_frac(samples_per_launch, kernel_ms):
    """Fraction of the VALU issue slots the walker kernel used: its VALU
    wavefront instructions per sample (rocprofv3 SQ_INSTS_VALU, committed
    profile) x 4 cycles each / (SIMDs x clock x the live launch time)."""
    ips = profile_value(VALU_FILE, "valu_instructions_per_sample")
    if ips is None or kernel_ms <= 0:
        return None
    return ips * samples_per_launch / 64.0 * 4.0 / (SIMDS * CLOCK_HZ * kernel_ms * 1e-3)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--quality", type=float, default=500.0)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-quality", type=float, default=60.0)
    p.add_argument("--no-animation", action="store_true", help="skip the 120-frame animation measurement")
    p.add_argument("--no-preview", action="store_true", help="skip the preview-serving measurement")
    p.add_argument("--no-still", action="store_true",
                   help="profiling aid: run only the animation and preview legs and print their results")
    p.add_argument("--no-config5", action="store_true", help="skip the 8K bits-64 still")
    p.add_argument("--config5-quality", type=float, default=2000.0)
    p.add_argument("--anim-frames", type=int, default=120)
    p.add_argument("--preview-requests", type=int, default=10000)
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
                ("bin_ms", ctypes.c_double), ("bin_launches", ctypes.c_longlong), ("de_taps", ctypes.c_longlong)]


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
        # the DE footprint (taps) of the same genome, counted in one more
        # render outside the timed steps (the count slows the DE pass)
        L.fl3h_profile_reset(2)
        one("taps")
        tp = fl3h_stage_times()
        L.fl3h_profile_get(ctypes.byref(tp))
        st.de_taps = tp.de_taps // max(1, tp.density_launches) * max(1, st.density_launches)
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


def animation_bench(args):
    """The 1280x720 ss2 q200 animation (100 temporal samples per frame, two
    5-xform control points with turning affines): frames rendered one
    flam3_render call at a time, then the same frames through the render
    server (the host work of frame k+1 overlaps the GPU work of frame k)."""
    import numpy as np
    import flam3_amd as F
    from flam3_amd import abi
    from flam3_amd import genome as G

    from flam3_amd import opcount

    L = F.lib()
    gs = G.synthetic_animation(seed=3, frames=args.anim_frames)
    arr, owned = G.genome_array(gs)
    n = args.anim_frames
    h, w = gs[0].height, gs[0].width
    imgs = [np.zeros((h, w, 3), np.uint8) for _ in range(2)]
    try:
        def frame(t):
            return F.make_frame(ctypes.addressof(arr), 2, time=float(t), seed="anim-%d" % t)
        st = abi.stat_struct()
        f = frame(0)
        L.flam3_render(ctypes.byref(f), imgs[0].ctypes.data_as(ctypes.c_void_p), 0, 3, 0, ctypes.byref(st))
        samples = 0.0
        t0 = time.perf_counter()
        for t in range(n):
            f = frame(t)
            if L.flam3_render(ctypes.byref(f), imgs[t & 1].ctypes.data_as(ctypes.c_void_p), 0, 3, 0,
                              ctypes.byref(st)) != 0:
                raise RuntimeError("flam3_render failed: " + F.last_error())
            samples += st.num_iters
        plain = time.perf_counter() - t0
        log("animation: %d frames one call at a time: %.3f s" % (n, plain))
        # the reference's loop structure: one walker launch per temporal sample
        from flam3_amd._lib import fl3h_render_opts
        o = fl3h_render_opts(0, 0, 0, 0, 1)
        nper = min(n, 6)
        t0 = time.perf_counter()
        for t in range(nper):
            f = frame(t)
            if L.fl3h_render(ctypes.byref(f), imgs[t & 1].ctypes.data_as(ctypes.c_void_p), 0, 3, 0, ctypes.byref(st),
                             ctypes.byref(o)) != 0:
                raise RuntimeError("fl3h_render failed: " + F.last_error())
        per_sample = (time.perf_counter() - t0) / nper
        log("animation: one launch per temporal sample: %.1f ms/frame (%d frames)" % (1e3 * per_sample, nper))
        frames = [frame(t) for t in range(n)]
        outs = [np.zeros((h, w, 3), np.uint8) for _ in range(n)]
        stats = [abi.stat_struct() for _ in range(n)]
        srv = F.Server()
        t0 = time.perf_counter()
        ids = [L.fl3h_server_submit(srv._s, ctypes.byref(frames[t]), outs[t].ctypes.data_as(ctypes.c_void_p), 0, 3,
                                    0, ctypes.byref(stats[t])) for t in range(n)]
        for i in ids:
            if L.fl3h_server_wait(srv._s, i, None) != 0:
                raise RuntimeError("served frame failed: " + F.last_error())
        served = time.perf_counter() - t0
        srv.close()
        log("animation: %d frames through the server: %.3f s" % (n, served))
        return {"workload": "1280x720 ss2 q200, 100 temporal samples per frame, 2 control points of 5 xforms, "
                            "DE radius 7; %d frames" % n,
                "samples_per_frame": samples / n,
                "per_sample_launch_ms_per_frame": 1e3 * per_sample,
                "plain_frames_per_s": n / plain, "plain_ms_per_frame": 1e3 * plain / n,
                "served_frames_per_s": n / served, "served_ms_per_frame": 1e3 * served / n,
                "served_samples_per_s": samples / served,
                "served_vs_plain": plain / served,
                "roofline": served_roofline(samples / served, opcount.ops_per_sample(gs[0])),
                "note": "per_sample_launch: one walker launch per temporal sample (the reference's loop, first "
                        "frames only); plain: one flam3_render call per frame, its 100 temporal samples in one "
                        "walker launch; served: the render server, ~8 frames per shared launch with the host work "
                        "of the next batch overlapping the GPU"}
    finally:
        for gp in owned:
            G.free_struct(gp)


def config5_bench(args):
    """The 8K still: 8 xforms + final xform + xaos, 7680x4320 supersample 3
    (a 23134 x 13054 histogram), DE radius 15, bits 64 (FP64 buckets and
    accumulator), one render at the requested quality with stage times."""
    import numpy as np
    import flam3_amd as F
    from flam3_amd import abi
    from flam3_amd import genome as G

    L = F.lib()
    g = G.synthetic_8k(seed=5, quality=args.config5_quality)
    arr, owned = G.genome_array([g])
    img = np.zeros((g.height, g.width, 3), dtype=np.uint8)
    stats = abi.stat_struct()
    try:
        f = F.make_frame(ctypes.addressof(arr), 1, bits=64, nthreads=1, seed="8k")
        L.fl3h_profile_reset(1)
        t0 = time.perf_counter()
        if L.flam3_render(ctypes.byref(f), img.ctypes.data_as(ctypes.c_void_p), 0, 3, 0, ctypes.byref(stats)) != 0:
            raise RuntimeError("flam3_render failed: " + F.last_error())
        dt = time.perf_counter() - t0
        st = fl3h_stage_times()
        L.fl3h_profile_get(ctypes.byref(st))
        L.fl3h_profile_reset(0)
        log("8K q%g bits 64: %.2f s, iterate %.1f ms (%d launches), bin %.1f ms (%d chunks), DE %.1f ms, "
            "tonemap %.1f ms" % (args.config5_quality, dt, st.iterate_ms, st.iterate_launches, st.bin_ms,
                                 st.bin_launches, st.density_ms, st.tonemap_ms))
        return {"workload": "8K still: 8 xforms + final xform + xaos, 7680x4320 ss3, DE radius 15, bits 64, "
                            "quality %g" % args.config5_quality,
                "samples": stats.num_iters, "frame_s": dt, "samples_per_s": stats.num_iters / dt,
                "iterate_ms": st.iterate_ms, "iterate_launches": st.iterate_launches,
                "bin_ms": st.bin_ms, "bin_chunks": st.bin_launches,
                "density_ms": st.density_ms, "tonemap_ms": st.tonemap_ms, "cells": st.cells,
                "deposit": "tile-binned: 8-byte records grouped by band (2^24 cells), then per band counted, "
                           "sorted by tile and summed in LDS tiles as 64-bit fixed point (2^-32), "
                           "then added to the double buckets",
                "accumulation_hbm_frac": stats.num_iters * 80.0 / (st.bin_ms * 1e-3) / (HBM_PEAK_GBS * 1e9)
                if st.bin_ms > 0 else None}
    finally:
        for gp in owned:
            G.free_struct(gp)


ACCUM_BYTES_PER_SAMPLE = 40.0   # 5 channels x (4 B read + 4 B written) per deposited sample (bits 33)


def served_roofline(samples_per_s, ops_per_sample):
    """A served leg's throughput as fractions of the two bounds of its
    dominant stages: the deposit's algorithmic bytes (40 B per sample at bits
    33) against HBM, and the iteration's counted FP32 ops against the FP32
    vector peak."""
    return {"accumulation_hbm_frac": samples_per_s * ACCUM_BYTES_PER_SAMPLE / (HBM_PEAK_GBS * 1e9),
            "iteration_fp32_frac": samples_per_s * ops_per_sample / (FP32_PEAK_TFLOPS * 1e12),
            "ops_per_sample": ops_per_sample}


def preview_pool(n):
    """The genome-search preview workload: n seeded random genomes (2-12
    xforms of the 12 workload variations, 64^2..512^2, quality U[10,100],
    ss1), one per request."""
    from flam3_amd import genome as G
    return [G.synthetic_preview(1000 + k) for k in range(n)]


def preview_bench(args):
    """Genome-search preview serving: a stream of distinct seeded preview
    genomes with Poisson arrivals at several rates into the render server;
    latency from submit to image ready."""
    import numpy as np
    import flam3_amd as F
    from flam3_amd import abi
    from flam3_amd import genome as G
    from flam3_amd import opcount

    L = F.lib()
    n = args.preview_requests
    t0 = time.perf_counter()
    pool = preview_pool(n)
    arrs = [G.genome_array([g]) for g in pool]
    log("preview: %d genomes built in %.1f s" % (n, time.perf_counter() - t0))
    samples = np.array([g.width * g.height * g.quality for g in pool])
    ops = float(np.sum([opcount.ops_per_sample(g) * s for g, s in zip(pool[:500], samples[:500])])
                / samples[:500].sum())
    rng = np.random.default_rng(4)

    def build(k):
        g = pool[k]
        f = F.make_frame(ctypes.addressof(arrs[k][0]), 1, seed="req-%d" % k)
        return f, np.zeros((g.height, g.width, 3), np.uint8), abi.stat_struct()

    try:
        f, img, st = build(0)
        L.flam3_render(ctypes.byref(f), img.ctypes.data_as(ctypes.c_void_p), 0, 3, 0, ctypes.byref(st))
        nseq = min(n, 300)
        t0 = time.perf_counter()
        for k in range(nseq):
            f, img, st = build(k)
            L.flam3_render(ctypes.byref(f), img.ctypes.data_as(ctypes.c_void_p), 0, 3, 0, ctypes.byref(st))
        alone = (time.perf_counter() - t0) / nseq
        sweeps = []
        for rate in (1000.0, 2000.0, 4000.0, None):
            arrive = np.cumsum(rng.exponential(1.0 / rate, n)) if rate else np.zeros(n)
            inputs = [build(k) for k in range(n)]
            srv = F.Server()
            ids = []
            # (the load generator must not stall: a garbage-collection pause
            # would submit the requests due meanwhile in one burst)
            gc.collect()
            gc.disable()
            t0 = time.perf_counter()
            for k in range(n):
                d = t0 + arrive[k] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                f, img, st = inputs[k]
                ids.append(L.fl3h_server_submit(srv._s, ctypes.byref(f), img.ctypes.data_as(ctypes.c_void_p), 0, 3,
                                                0, ctypes.byref(st)))
            lat = []
            for i in ids:
                lv = ctypes.c_double(0.0)
                if L.fl3h_server_wait(srv._s, i, ctypes.byref(lv)) != 0:
                    raise RuntimeError("preview request failed: " + F.last_error())
                lat.append(lv.value)
            wall = time.perf_counter() - t0
            gc.enable()
            batches, _ = srv.counts()
            srv.close()
            lat = np.array(lat)
            sps = float(samples.sum()) / wall
            sweeps.append({"offered_per_s": rate if rate else "all at once", "requests": n,
                           "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
                           "requests_per_s": n / wall, "samples_per_s": sps, "batches": batches,
                           "roofline": served_roofline(sps, ops)})
            log("preview rate %s: p50 %.2f ms p99 %.2f ms, %.0f req/s" % (rate, sweeps[-1]["p50_ms"],
                                                                         sweeps[-1]["p99_ms"], n / wall))
        return {"workload": "preview requests: %d distinct seeded genomes (2-12 xforms, 64^2..512^2, q U[10,100], "
                            "ss1), Poisson arrivals" % n, "mean_samples_per_request": float(samples.mean()),
                "alone_ms_per_request": 1e3 * alone, "alone_requests_per_s": 1.0 / alone,
                "alone_sample": "first %d requests, one flam3_render call each" % nseq, "sweep": sweeps}
    finally:
        for arr, owned in arrs:
            for gp in owned:
                G.free_struct(gp)


def host_threads():
    """Threads for the CPU reference: the host's CPUs available to this
    process, at most 16 (the CPU share of one GPU on the GPU hosts, which
    report more CPUs than the share)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def cpu_preview_baseline(args):
    """Reference flam3_render (oracle/_ref fast build, host_threads()) on
    the first preview requests, one after another."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyref
    from flam3_amd import genome as G

    if not pyref.available("fast"):
        return None
    threads = host_threads()
    pool = preview_pool(60)
    docs = [G.flame_document([g]) for g in pool]
    t0 = time.perf_counter()
    for k, xml in enumerate(docs):
        pyref.render(xml, seed="req-%d" % k, nthreads=threads, kind="fast")
    dt = time.perf_counter() - t0
    return {"requests_per_s": len(docs) / dt, "ms_per_request": 1e3 * dt / len(docs), "cores": threads,
            "kind": "reference", "sample": "first %d preview requests, serial flam3_render calls (%.1f s)"
                                          % (len(docs), dt)}


def cpu_baseline(args):
    """Reference flam3_render (oracle/_ref fast build) with host_threads()
    threads on the same genome at a reduced quality, once with density
    estimation and once without, so that the DE share of the CPU time is
    known."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyref
    from flam3_amd import genome as G

    if not pyref.available("fast"):
        return None
    threads = host_threads()
    g = G.synthetic_1080p(seed=2, quality=args.cpu_quality)
    xml = G.flame_document([g])
    t0 = time.perf_counter()
    _, st = pyref.render(xml, seed="cpu", nthreads=threads, kind="fast")
    dt = time.perf_counter() - t0
    g.estimator_radius = 0.0
    t0 = time.perf_counter()
    pyref.render(G.flame_document([g]), seed="cpu", nthreads=threads, kind="fast")
    dt0 = time.perf_counter() - t0
    return {"value": st["num_iters"] / dt, "unit": "samples/s", "cores": threads, "kind": "reference",
            "de_share": max(0.0, (dt - dt0) / dt),
            "sample": "reference flam3_render of the same 1080p genome at quality %g (%d samples, %d threads, "
                      "%.2f s wall incl. DE and tone map; %.2f s with DE off)"
                      % (args.cpu_quality, st["num_iters"], threads, dt, dt0)}


def main():
    args = parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.no_still:
        legs = {}
        if not args.no_animation:
            legs["animation"] = animation_bench(args)
        if not args.no_preview:
            legs["preview"] = preview_bench(args)
        if not args.no_config5:
            legs["config5"] = config5_bench(args)
        print(json.dumps(legs), flush=True)
        return
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
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": measured_traffic() if args.quality == 500.0 else None,
                     "kernel": "k_iter_fast", "kernel_ms": it_ms, "samples_per_launch": per_launch,
                     "ops_per_sample": ops,
                     "algorithmic_bytes_per_launch": 4.0 * per_launch,
                     "valu_issue_frac": valu_issue_frac(per_launch, it_ms) if args.quality == 500.0 else None,
                     "valu_instructions_per_sample": profile_value(VALU_FILE, "valu_instructions_per_sample"),
                     "note": "bound by FP32 vector ALU work (no matrix work on this path); peak is MI355X's "
                             "FP32 vector peak, 157.3 TFLOP/s. achieved counts only the formula ops of each "
                             "sample's chosen xform and the final xform; the kernel also issues RNG, selection, "
                             "trig range reduction, the point exchange and the record "
                             "(valu_instructions_per_sample, rocprofv3 SQ_INSTS_VALU); valu_issue_frac is the "
                             "share of VALU issue slots it uses"},
        "stages": {
            "iterate_ms_per_launch": it_ms,
            "bin_ms_per_chunk": bin_ms,
            "bin_GBps": (per_launch * BIN_BYTES_PER_SAMPLE / (bin_ms * 1e-3) / 1e9) if bin_ms > 0 else None,
            "density_ms": de_ms,
            "density_GBps": (st.cells * DE_BYTES_PER_CELL / (de_ms * 1e-3) / 1e9) if de_ms > 0 else None,
            "density_hbm_frac": (st.cells * DE_BYTES_PER_CELL / (de_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if de_ms > 0
            else None,
            # the gather's work: each tap of a source's (2r+1)^2 footprint is
            # 4 FP32 multiply-adds (8 flops) into the output cell
            "density_taps": st.de_taps / max(1, st.density_launches),
            "density_tap_fp32_frac": (st.de_taps / max(1, st.density_launches) * 8.0 / (de_ms * 1e-3)
                                      / (FP32_PEAK_TFLOPS * 1e12)) if de_ms > 0 else None,
            "tonemap_ms": st.tonemap_ms / max(1, st.tonemap_launches)},
    }
    if not args.no_config5:
        out["config5"] = config5_bench(args)
    if not args.no_animation:
        out["animation"] = animation_bench(args)
    if not args.no_preview:
        out["preview"] = preview_bench(args)
    if not args.no_cpu_baseline:
        cb = cpu_baseline(args)
        out["cpu_baseline"] = cb
        if not args.no_preview and cb is not None:
            out["preview"]["cpu_baseline"] = cpu_preview_baseline(args)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
