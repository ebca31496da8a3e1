#!/usr/bin/env python3
"""Benchmark of the NL-SAR hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-served]
                    [--no-extra]

A step is one NL-SAR InSAR filtering of the seeded 2048x2048 interferogram (search window
21, patch sizes 3/5/7, scale sizes 1/3, h 15, c 49) as the reference computes it: every
(patch, scale) key's statistics with their bias-corrected ENL and alpha maps, then the
selected key's weighted mean and decomposition, inputs and outputs resident in HBM.  Prints
one JSON line with Mpixel/s, the roofline of the dominant kernel (the last scale pass), the
same output from the selected key's statistics alone, the NL-InSAR 4096^2 x 10 summary, the
64-image NL-SAR and NL-InSAR batch summaries, the served stream's p99 latency (the
2000-request mixed stream) and the CPU baseline (the C restatement in oracle/, timed on this host's cores on a
1536x1536 crop).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

H = W = 2048
SW, PATCHES, SCALES = 21, [3, 5, 7], [1, 3]
HPAR, CPAR = 15.0, 49.0
PEAK_FP32_TFLOPS = 157.3
# HBM-side bytes per launch of the every-key last pass (the roofline's kernel), per pass
# variant (1: two rows per thread, 2: one row): rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on
# this benchmark with --tile 1 / --tile 2 (KiB per dispatch: the mean FETCH and the WRITE of
# the dispatches that write the ENL / alpha maps), scaled by what the same counters read for
# a 256 MiB device-to-device copy (FETCH x 2.000, WRITE x 1.000);
# profiles/nlsar_2048_scale_pass_pmc.md
TRAFFIC_LAST_PASS_BYTES = {1: round((5141421 * 2.000 + 156482 * 1.000) * 1024),
                           2: round((864173 * 2.000 + 148032 * 1.000) * 1024)}
# the same for a later NL-InSAR weights launch at 4096^2 (18 x 54 blocks): the mean over
# iterations 2-10 of one call (profiles/nlinsar_per_call.md)
TRAFFIC_NLINSAR_WEIGHTS_BYTES = round((6.12e10 + 6.12e10 + 6.14e10 + 3.92e10 + 3.54e10 + 3.44e10 + 3.39e10 +
                                       3.35e10 + 3.33e10) / 9)
# algorithmic FLOP per output pixel (each transcendental / divide = 1, FMA = 2):
# pixel dissimilarity 220*27 per scale, patch sums 220*2*(p-1) summed over p, LUT weights
# 220*8 per (p, s), nonlocal statistics 441*11 per (p, s), alpha/ENL 22 per (p, s),
# weighted mean 441 + 8*441*2 + 32, decomposition 12.
# Every key (what the reference computes): 69,953 for the whole filter, of which the last
# scale pass (three patch sizes) does FLOP_PER_PX_LAST_PASS_ALL.
FLOP_PER_PX_TOTAL = 69953
FLOP_PER_PX_LAST_PASS_ALL = 220 * 27 + 220 * 2 * (2 + 4 + 6) + 3 * 220 * 8 + 3 * 441 * 11 + 3 * 22 \
    + (441 + 8 * 441 * 2 + 32) + 12
# The output needs only the key the selection returns (patch 7 at scale 3): the default
# entry point runs one pass with that key; its work per pixel, and the whole filter's
# (covariance 14 + one scale average 80 + that pass)
FLOP_PER_PX_LAST_KEY_PASS = 220 * 27 + 220 * 2 * 6 + 220 * 8 + 441 * 11 + 22 + (441 + 8 * 441 * 2 + 32) + 12
FLOP_PER_PX_LAST_KEY_TOTAL = 14 + 80 + FLOP_PER_PX_LAST_KEY_PASS


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(a1, a2, ph, stats, size=1536):
    """Oracle (C restatement, OpenMP on all host cores) on a size x size crop (about ten
    seconds of CPU work on 16 cores)."""
    from oracle import oracle as O
    O.build()
    c = [np.ascontiguousarray(x[:size, :size]) for x in (a1, a2, ph)]
    t0 = time.perf_counter()
    O.nlsar_insar(*c, SW, PATCHES, SCALES, stats, HPAR, CPAR)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(size * size / dt / 1e6, 4), "unit": "Mpixel/s", "cores": cores,
            "kind": "port",
            "sample": f"NL-SAR on the {size}x{size} top-left crop of the same interferogram, "
                      f"{dt:.1f} s wall on {cores} OpenMP threads"}


def batched(args):
    print(json.dumps(batched_run(args.tiles, args.steps, args.warmup)), flush=True)


def batched_run(tiles, steps, warmup):
    """Batched form: `tiles` seeded interferograms with sides log-uniform in [256, 2048]
    (aspect <= 4:1), the same NL-SAR parameters; plain per-call loop vs one batch (eager
    launches and HIP-graph replay), on one stream with inputs resident in HBM."""
    import despeckcl_amd as D
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S

    stats = D.load_nlsar_stats_collection(os.path.join(ROOT, "tests", "golden", "train32_stats.txt"))
    shapes = S.tile_shapes(tiles, S.SEED_BASE + 3)
    stream = HIP.Stream()
    imgs = []
    for i, (h, w) in enumerate(shapes):
        host = S.interferogram(h, w, S.SEED_BASE + 3000 + i, n_regions=16, coh_range=(0.2, 0.95))
        inp = [HIP.DeviceArray.from_host(x) for x in host]
        out = [HIP.DeviceArray((h, w)) for _ in range(3)]
        imgs.append((inp, out, h, w))
    log(f"batched: {len(imgs)} images generated, {sum(h * w for _, _, h, w in imgs) / 1e6:.1f} Mpixel")
    px = sum(h * w for _, _, h, w in imgs)

    def timed(fn, steps):
        for _ in range(warmup):
            fn()
        stream.synchronize()
        e0, e1 = HIP.Event(), HIP.Event()
        e0.record(stream.handle)
        for _ in range(steps):
            fn()
        e1.record(stream.handle)
        e1.synchronize()
        return e0.elapsed_ms(e1) / steps

    b = DV.NlsarBatch()
    b.set([(*[d.ptr for d in inp], *[d.ptr for d in out], h, w) for inp, out, h, w in imgs],
          SW, PATCHES, SCALES, stats, HPAR, CPAR)
    tile, tune_ms = b.autotune(stream.handle, reps=1)
    log(f"batched: autotune picked tile {tile} ({tune_ms})")
    DV.set_tile(0, tile)

    def plain():
        for inp, out, h, w in imgs:
            DV.nlsar_insar(*[d.ptr for d in inp], *[d.ptr for d in out], h, w, SW, PATCHES, SCALES,
                           stats, HPAR, CPAR, stream=stream.handle)

    plain_ms = timed(plain, steps)
    eager_ms = timed(lambda: b.enqueue(stream.handle), steps)
    graph_ms = timed(lambda: b.graph_launch(stream.handle), steps)
    best = min(eager_ms, graph_ms)
    out = {
        "metric": "Mpixel/s batched NL-SAR (64 interferograms, sides 256-2048, 1 GPU)",
        "value": round(px / 1e6 / (best / 1e3), 3), "unit": "Mpixel/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": round(best, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded speckle model of insarsim_simu.cpp, 16 regions, coherence U[0.2,0.95])",
        "config": {"workload": f"NL-SAR InSAR, {len(imgs)} interferograms per batch", "pixels": px,
                   "shapes_seed": S.SEED_BASE + 3, "search_window": SW, "patch_sizes": PATCHES,
                   "scale_sizes": SCALES, "tile": tile},
        "plain_per_call": {"ms": round(plain_ms, 3), "mpx_s": round(px / 1e6 / (plain_ms / 1e3), 3)},
        "batched_eager": {"ms": round(eager_ms, 3), "mpx_s": round(px / 1e6 / (eager_ms / 1e3), 3)},
        "batched_graph": {"ms": round(graph_ms, 3), "mpx_s": round(px / 1e6 / (graph_ms / 1e3), 3)},
        "speedup_vs_plain": round(plain_ms / best, 3),
        "autotune_ms_per_tile": [round(x, 3) for x in tune_ms],
    }
    b.close()
    for inp, o, _h, _w in imgs:
        for d in inp + o:
            d.free()
    DV.set_tile(0, -1)
    return out


def nlinsar_form(args):
    print(json.dumps(nlinsar_run(args.size or 4096, args.niter, args.steps, args.warmup, args.tile,
                                 not args.no_cpu_baseline)), flush=True)


def nlinsar_batched_run(tiles, steps, warmup, lo=128, hi=1024, niter=5):
    """Batched NL-InSAR: `tiles` seeded interferograms with sides log-uniform in [lo, hi]
    (aspect <= 4:1), search window 21, patch 7, `niter` iterations, lmin 10, estimates from
    zero; plain per-call loop vs one batch (eager launches and HIP-graph replay), on one
    stream with inputs resident in HBM.  Every timed form first clears the estimate arrays
    (hipMemsetAsync on the stream), so each run starts from zero estimates."""
    import despeckcl_amd as D
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S

    shapes = S.tile_shapes(tiles, S.SEED_BASE + 6, lo=lo, hi=hi)
    stream = HIP.Stream()
    imgs = []
    for i, (h, w) in enumerate(shapes):
        host = S.interferogram(h, w, S.SEED_BASE + 6000 + i, n_regions=8, coh_range=(0.3, 0.9))
        imgs.append(([HIP.DeviceArray.from_host(x) for x in host], [HIP.DeviceArray((h, w)) for _ in range(3)],
                     h, w))
    px = sum(h * w for _i, _o, h, w in imgs)
    log(f"nlinsar batched: {len(imgs)} images, {px / 1e6:.1f} Mpixel")

    def zero_all():
        for _i, out, h, w in imgs:
            for d in out:
                HIP.hip().hipMemsetAsync(C_void(d.ptr), 0, C_size(h * w * 4), C_void(stream.handle))

    def timed(fn):
        for _ in range(warmup):
            fn()
        stream.synchronize()
        e0, e1 = HIP.Event(), HIP.Event()
        e0.record(stream.handle)
        for _ in range(steps):
            fn()
        e1.record(stream.handle)
        e1.synchronize()
        return e0.elapsed_ms(e1) / steps

    def plain():
        zero_all()
        for inp, out, h, w in imgs:
            DV.nlinsar(*[d.ptr for d in inp], *[d.ptr for d in out], h, w, 21, 7, niter, 10, stream=stream.handle)

    plain_ms = timed(plain)
    D.lib().dkx_release_workspace()
    b = DV.NlinsarBatch()
    try:
        b.set([(*[d.ptr for d in inp], *[d.ptr for d in out], h, w) for inp, out, h, w in imgs], 21, 7, niter, 10)

        def eager():
            zero_all()
            b.enqueue(stream.handle)

        def graph():
            zero_all()
            b.graph_launch(stream.handle)

        eager_ms = timed(eager)
        graph_ms = timed(graph)
    finally:
        b.close()
        for inp, out, _h, _w in imgs:
            for d in inp + out:
                d.free()
    best = min(eager_ms, graph_ms)
    return {
        "metric": f"Mpixel/s batched NL-InSAR ({len(imgs)} interferograms, sides {lo}-{hi}, {niter} iterations, 1 GPU)",
        "value": round(px / 1e6 / (best / 1e3), 3), "unit": "Mpixel/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": round(best, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded speckle model of insarsim_simu.cpp, 8 regions, coherence U[0.3,0.9])",
        "config": {"workload": f"NL-InSAR, {len(imgs)} interferograms per batch", "pixels": px,
                   "shapes_seed": S.SEED_BASE + 6, "search_window": 21, "patch": 7, "niter": niter, "lmin": 10},
        "plain_per_call": {"ms": round(plain_ms, 3), "mpx_s": round(px / 1e6 / (plain_ms / 1e3), 3)},
        "batched_eager": {"ms": round(eager_ms, 3), "mpx_s": round(px / 1e6 / (eager_ms / 1e3), 3)},
        "batched_graph": {"ms": round(graph_ms, 3), "mpx_s": round(px / 1e6 / (graph_ms / 1e3), 3)},
        "speedup_vs_plain": round(plain_ms / best, 3),
    }


def nlinsar_run(n, niter, steps, warmup, tile=None, with_cpu=False):
    """NL-InSAR on one seeded interferogram per call (default: 4096^2, 32 regions plus 5 %
    point scatterers at x50, search window 21, patch 7, 10 iterations, lmin 10), inputs and
    the estimates resident in HBM; the estimates start from zero, as the Python API's."""
    import ctypes as C
    import despeckcl_amd as D
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S

    big = n >= 1024
    if tile is not None:
        DV.set_tile(1, tile)
    a1, a2, ph = S.interferogram(n, n, S.SEED_BASE + (2 if big else 0), n_regions=32 if big else 8,
                                 coh_range=(0.3, 0.9), scatterer_fraction=0.05 if big else 0.0)
    stream = HIP.Stream()
    d_in = [HIP.DeviceArray.from_host(x) for x in (a1, a2, ph)]
    d_out = [HIP.DeviceArray((n, n)) for _ in range(3)]

    def run(want=False):
        for d in d_out:
            d.zero()
        return DV.nlinsar(*[d.ptr for d in d_in], *[d.ptr for d in d_out], n, n, 21, 7, niter, 10,
                          stream=stream.handle, want_count=want)

    tuned = None
    if tile is None:
        # the autotuner on this image (the estimates are restored before every run)
        for d in d_out:
            d.zero()
        t, tms = DV.nlinsar_autotune(*[d.ptr for d in d_in], *[d.ptr for d in d_out], n, n, 21, 7, niter, 10,
                                     reps=1, stream=stream.handle)
        tuned = {"shape": t, "ms_per_shape": [round(x, 3) for x in tms],
                 "note": "tile shapes 0 / 1 / 2 = 27 x 54, 18 x 54 (512 threads), 54 x 54 (1024 threads); one run each"}
        log(f"nlinsar autotune: {tuned}")
    nsm = run(True)
    stream.synchronize()
    lib = D.lib()
    for _ in range(warmup):
        run()
    stream.synchronize()
    HIP.synchronize()
    e0, e1 = HIP.Event(), HIP.Event()
    e0.record(stream.handle)
    for _ in range(steps):
        run()
    e1.record(stream.handle)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    lib.dkx_set_timing(1)
    lib.dkx_stage_times(None, None, 0)
    run()
    stream.synchronize()
    kinds = (C.c_int * 256)()
    times = (C.c_float * 256)()
    k = lib.dkx_stage_times(kinds, times, 256)
    lib.dkx_set_timing(0)
    wts = [times[i] for i in range(k) if kinds[i] == 11]
    sms = [times[i] for i in range(k) if kinds[i] == 12]
    fls = [times[i] for i in range(k) if kinds[i] == 14]

    px = n * n
    flop = 31791.0 * px * niter
    out = {
        "metric": f"Mpixel/s NL-InSAR {n}x{n}, {niter} iterations (1 GPU)",
        "value": round(px / 1e6 / (ms / 1e3), 3), "unit": "Mpixel/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded speckle model of insarsim_simu.cpp" +
                (", 32 regions, 5% point scatterers x50)" if big else ", 8 regions)"),
        "config": {"workload": f"NL-InSAR, one {n}x{n} interferogram per call", "search_window": 21,
                   "patch": 7, "niter": niter, "lmin": 10, "tile": DV.get_tile(1)},
        "mpx_iter_s": round(px * niter / 1e6 / (ms / 1e3), 3),
        "tflops": round(flop / (ms * 1e-3) / 1e12, 3),
        "weights_ms": [round(x, 3) for x in wts], "smooth_ms": [round(x, 3) for x in sms],
        "fill_ms": [round(x, 3) for x in fls],
        "smoothed_pixel_iterations": int(nsm),
        "autotune": tuned,
    }
    # the dominant kernel: a later iteration's weights launch (similarity cache, KL, weights,
    # ENL and weighted means), 31,791 algorithmic FLOP per pixel; a single call's launch
    # covers every pixel (the interior and the last row / column mappings)
    later = wts[1:] if len(wts) > 1 else wts
    wk_ms = float(np.mean(later))
    ach = 31791.0 * px / (wk_ms * 1e-3) / 1e12
    out["roofline"] = {"bound": "valu", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                       "frac": round(ach / PEAK_FP32_TFLOPS, 4),
                       "traffic": TRAFFIC_NLINSAR_WEIGHTS_BYTES if n == 4096 and tuned and tuned["shape"] == 1 else None,
                       "kernel": "nlinsar_weights (iterations 2 to niter, mean)", "kernel_ms": round(wk_ms, 4),
                       "note": f"31791 algorithmic FLOP/px x {px} px per launch; traffic: the mean HBM-side bytes "
                               "of those launches (all weights stored in iterations 2-4, the predicted "
                               "pixels' after)"}
    for d in d_in + d_out:
        d.free()
    lib.dkx_release_workspace()
    if with_cpu:
        from oracle import oracle as O
        O.build()
        c, it = min(n, 1024), 4
        t0 = time.perf_counter()
        O.nlinsar(a1[:c, :c], a2[:c, :c], ph[:c, :c], 21, 7, it, 10)
        dt = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {"value": round(c * c * it / 1e6 / dt, 4), "unit": "Mpixel-iterations/s",
                               "cores": cores, "kind": "port",
                               "sample": f"NL-InSAR, {it} iterations on the {c}x{c} top-left crop, {dt:.1f} s "
                                         f"on {cores} OpenMP threads"}
    return out


def covmat_form(args):
    """NL-SAR on covariance matrices (default: seeded single-look PolSAR, dimension 3,
    2048^2, search window 21, patch sizes 3/5/7, scale sizes 1/3), statistics trained by the
    library on a 32x32 homogeneous area of the same dimension; the 2*dim*dim interlaced
    planes resident in HBM."""
    import despeckcl_amd as D
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S

    n, dim = args.size or 2048, args.dim
    c = S.polsar(n, n, S.SEED_BASE + 3, dim, n_regions=16)
    inter = np.empty((2 * dim * dim, n, n), np.float32)
    inter[::2] = c.real.reshape(-1, n, n)
    inter[1::2] = c.imag.reshape(-1, n, n)
    del c
    st = D.nlsar_train(S.polsar(32, 32, S.SEED_BASE + 200 + dim, dim, n_regions=0), PATCHES, SCALES)
    stream = HIP.Stream()
    d_in = HIP.DeviceArray.from_host(inter)
    d_out = HIP.DeviceArray(inter.shape)

    def call(want=False):
        return DV.nlsar_covmat(d_in.ptr, d_out.ptr, n, n, dim, SW, PATCHES, SCALES, st, HPAR, CPAR,
                               stream=stream.handle, want_ub=want)

    ub = call(True)
    for _ in range(args.warmup):
        call()
    stream.synchronize()
    HIP.synchronize()
    e0, e1 = HIP.Event(), HIP.Event()
    e0.record(stream.handle)
    for _ in range(args.steps):
        call()
    e1.record(stream.handle)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / args.steps
    out = {
        "metric": f"Mpixel/s NL-SAR covariance dim {dim} {n}x{n} (1 GPU)",
        "value": round(n * n / 1e6 / (ms / 1e3), 3), "unit": "Mpixel/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic single-look covariances (dimension {dim}, 16 regions)",
        "config": {"workload": f"NL-SAR covariance input, one {n}x{n} dim-{dim} image per call",
                   "search_window": SW, "patch_sizes": PATCHES, "scale_sizes": SCALES},
        "ub_pixels": int(ub),
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        O.build()
        cs = min(n, 768)
        crop = np.ascontiguousarray(inter[:, :cs, :cs])
        t0 = time.perf_counter()
        O.nlsar_covmat(crop, dim, SW, PATCHES, SCALES, st, HPAR, CPAR)
        dt = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {"value": round(cs * cs / 1e6 / dt, 4), "unit": "Mpixel/s", "cores": cores,
                               "kind": "port",
                               "sample": f"covariance NL-SAR on the {cs}x{cs} top-left crop, {dt:.1f} s on "
                                         f"{cores} OpenMP threads"}
    print(json.dumps(out), flush=True)


def boxcar_form(args):
    """Boxcar filter (default: seeded 8192^2 interferogram, window 5) with inputs and outputs
    resident in HBM.  Roofline: HBM, 24 algorithmic bytes per pixel (three float planes read,
    three written)."""
    import despeckcl_amd as D
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S

    n, w = args.size or 8192, args.window
    a1, a2, ph = S.interferogram(n, n, S.SEED_BASE + 4, n_regions=16)
    stream = HIP.Stream()
    d_in = [HIP.DeviceArray.from_host(x) for x in (a1, a2, ph)]
    d_out = [HIP.DeviceArray((n, n)) for _ in range(3)]

    def call():
        DV.boxcar(*[d.ptr for d in d_in], *[d.ptr for d in d_out], n, n, w, stream=stream.handle)

    for _ in range(max(1, args.warmup)):
        call()
    stream.synchronize()
    e0, e1 = HIP.Event(), HIP.Event()
    e0.record(stream.handle)
    for _ in range(args.steps):
        call()
    e1.record(stream.handle)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / args.steps
    gbs = 24.0 * n * n / (ms * 1e-3) / 1e9
    out = {
        "metric": f"Mpixel/s boxcar {n}x{n}, window {w} (1 GPU)",
        "value": round(n * n / 1e6 / (ms / 1e3), 3), "unit": "Mpixel/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded speckle model of insarsim_simu.cpp, 16 regions)",
        "config": {"workload": f"boxcar, one {n}x{n} interferogram per call", "window_width": w},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(gbs / 8000.0, 4), "traffic": None,
                     "kernel": "boxcar_tile + boxcar_edge (whole call)",
                     "note": "24 algorithmic bytes per pixel per call"},
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        O.build()
        c = min(n, 2048)
        t0 = time.perf_counter()
        O.boxcar(a1[:c, :c], a2[:c, :c], ph[:c, :c], w)
        dt = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {"value": round(c * c / 1e6 / dt, 3), "unit": "Mpixel/s", "cores": cores,
                               "kind": "port",
                               "sample": f"boxcar on the {c}x{c} top-left crop, {dt:.2f} s on {cores} OpenMP threads"}
    print(json.dumps(out), flush=True)


def goldstein_form(args):
    """Goldstein filter (default: seeded 8192^2 interferogram, patch 32, overlap 4, alpha 0.5,
    the reference example's parameters) with inputs and outputs resident in HBM.  Roofline:
    HBM, 20 algorithmic bytes per pixel (three float planes read, two written)."""
    import despeckcl_amd as D
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S

    n, P, ov, alpha = args.size or 8192, 32, 4, 0.5
    a1, a2, ph = S.interferogram(n, n, S.SEED_BASE + 5, n_regions=16)
    stream = HIP.Stream()
    d_in = [HIP.DeviceArray.from_host(x) for x in (a1, a2, ph)]
    d_out = [HIP.DeviceArray((n, n)) for _ in range(2)]

    def call():
        DV.goldstein(*[d.ptr for d in d_in], *[d.ptr for d in d_out], n, n, P, ov, alpha, stream=stream.handle)

    for _ in range(max(1, args.warmup)):
        call()
    stream.synchronize()
    e0, e1 = HIP.Event(), HIP.Event()
    e0.record(stream.handle)
    for _ in range(args.steps):
        call()
    e1.record(stream.handle)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / args.steps
    gbs = 20.0 * n * n / (ms * 1e-3) / 1e9
    out = {
        "metric": f"Mpixel/s Goldstein {n}x{n}, patch {P}, overlap {ov}, alpha {alpha} (1 GPU)",
        "value": round(n * n / 1e6 / (ms / 1e3), 3), "unit": "Mpixel/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded speckle model of insarsim_simu.cpp, 16 regions)",
        "config": {"workload": f"Goldstein, one {n}x{n} interferogram per call", "patch_size": P,
                   "overlap": ov, "alpha": alpha},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(gbs / 8000.0, 4), "traffic": None,
                     "kernel": "gs_patches + gs_pack (whole call)",
                     "note": "20 algorithmic bytes per pixel per call; the weighted patches "
                             "(8 B x (P/(P-2ov))^2 per pixel) make one extra round trip"},
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        O.build()
        c = min(n, 1024)
        t0 = time.perf_counter()
        O.goldstein(a1[:c, :c], a2[:c, :c], ph[:c, :c], P, ov, alpha)
        dt = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {"value": round(c * c / 1e6 / dt, 3), "unit": "Mpixel/s", "cores": cores,
                               "kind": "port",
                               "sample": f"Goldstein (exact DFT) on the {c}x{c} top-left crop, {dt:.2f} s on "
                                         f"{cores} OpenMP threads"}
    print(json.dumps(out), flush=True)


def served(args):
    out = served_run(args.requests, args.rate, args.budget_gib)
    print(json.dumps(out), flush=True)


def served_run(n_requests, rate, budget_gib):
    """Served stream: requests 70 % NL-SAR (the parameters above) / 30 % NL-InSAR (search
    window 21, patch 7, 5 iterations, lmin 10), sides log-uniform in [128, 4096] (aspect <= 4:1),
    each a seeded crop of one of four 4096^2 interferograms of the same generator, resident on
    the device.  Runs (1) all requests at once through the server (aggregate Mpixel/s),
    (2) the same requests one after another through the plain per-call entry points, (3) an
    open-loop Poisson stream at --rate requests/s through the server (p50 / p99 latency)."""
    import despeckcl_amd as D
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S

    stats = D.load_nlsar_stats_collection(os.path.join(ROOT, "tests", "golden", "train32_stats.txt"))
    n = n_requests
    kinds, shapes, origins, arrivals = S.served_stream(n, rate)
    scenes = []
    for k in range(4):
        host = S.served_scene(k)
        scenes.append([HIP.DeviceArray.from_host(x) for x in host])
        del host
    reqs, bufs, px = [], [], 0
    for i in range(n):
        h, w = shapes[i]
        sc = scenes[i % 4]
        y0, x0 = origins[i]
        inp = [HIP.DeviceArray((h, w)) for _ in range(3)]
        for d, src in zip(inp, sc):
            HIP.copy2d(d.ptr, w * 4, src.ptr + (y0 * 4096 + x0) * 4, 4096 * 4, w * 4, h)
        out = [HIP.DeviceArray((h, w)) for _ in range(3)]
        bufs.append((inp, out))
        reqs.append((int(kinds[i]), *[d.ptr for d in inp], *[d.ptr for d in out], h, w))
        px += h * w
    HIP.synchronize()
    log(f"served: {n} requests, {px / 1e6:.0f} Mpixel, {int(kinds.sum())} NL-InSAR")
    px_sar = sum(h * w for (k, *_r, h, w) in reqs if k == 0)

    srv = DV.Server(budget_gib << 30, SW, PATCHES, SCALES, stats, HPAR, CPAR, 21, 7, 5, 10)
    srv.run_trace(reqs[:8], [0.0] * 8)       # warm-up
    lat_sat, span_sat = srv.run_trace(reqs, [0.0] * n)
    b_sat, nb_sat = srv.counters(), srv.nlinsar_counters()
    log(f"served: saturated {span_sat:.2f} s")
    lat, span = srv.run_trace(reqs, arrivals.tolist())
    b_all, nb_all = srv.counters(), srv.nlinsar_counters()
    srv.close()
    log(f"served: Poisson stream {span:.2f} s")

    stream = HIP.Stream()
    t0 = time.perf_counter()
    by_kind = [0.0, 0.0]
    for (k, a1, a2, ph, r1, r2, r3, h, w) in reqs:
        tk = time.perf_counter()
        if k == 0:
            DV.nlsar_insar(a1, a2, ph, r1, r2, r3, h, w, SW, PATCHES, SCALES, stats, HPAR, CPAR, stream=stream.handle)
        else:
            HIP.hip().hipMemsetAsync(C_void(r1), 0, C_size(h * w * 4), C_void(stream.handle))
            HIP.hip().hipMemsetAsync(C_void(r2), 0, C_size(h * w * 4), C_void(stream.handle))
            HIP.hip().hipMemsetAsync(C_void(r3), 0, C_size(h * w * 4), C_void(stream.handle))
            DV.nlinsar(a1, a2, ph, r1, r2, r3, h, w, 21, 7, 5, 10, stream=stream.handle)
        stream.synchronize()
        by_kind[k] += time.perf_counter() - tk
    serial = time.perf_counter() - t0
    D.lib().dkx_release_workspace()
    lat = np.array(lat)
    out = {
        "metric": "served mixed stream: p50 / p99 latency and aggregate Mpixel/s (1 GPU)",
        "value": round(float(np.percentile(lat, 99)), 3), "unit": "ms (p99 latency)", "n_gpus": 1,
        "higher_is_better": False, "dtype": "f32",
        "data": "synthetic (seeded crops of four 4096^2 interferograms of the insarsim_simu.cpp speckle model)",
        "config": {"workload": "served stream, 70% NL-SAR / 30% NL-InSAR (5 iterations)", "requests": n,
                   "pixels": px, "nlsar_pixels": px_sar, "rate_req_s": rate,
                   "hbm_budget_gib": budget_gib, "sides": [128, 4096]},
        "p50_latency_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_latency_ms": round(float(np.percentile(lat, 99)), 3),
        "mean_latency_ms": round(float(lat.mean()), 3),
        "stream_seconds": round(span, 3),
        "saturated": {"seconds": round(span_sat, 3), "mpx_s": round(px / 1e6 / span_sat, 3),
                      "p50_latency_ms": round(float(np.percentile(lat_sat, 50)), 3)},
        "serial_plain": {"seconds": round(serial, 3), "mpx_s": round(px / 1e6 / serial, 3),
                         "nlsar_seconds": round(by_kind[0], 3), "nlinsar_seconds": round(by_kind[1], 3)},
        "speedup_vs_serial": round(serial / span_sat, 3),
        "nlsar_batches": b_all[0] - b_sat[0], "nlsar_batched_images": b_all[1] - b_sat[1],
        "nlinsar_batches": nb_all[0] - nb_sat[0], "nlinsar_batched_images": nb_all[1] - nb_sat[1],
    }
    return out


def C_void(x):
    import ctypes
    return ctypes.c_void_p(x)


def C_size(x):
    import ctypes
    return ctypes.c_size_t(x)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-served", action="store_true", help="skip the served-stream latency sample")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the NL-InSAR 4096^2 and 64-image batch summaries")
    ap.add_argument("--tile", type=int, default=None)
    ap.add_argument("--form", choices=("plain", "batched", "nlinsar", "nlinsar_batched", "served", "covmat", "boxcar",
                                       "goldstein"), default="plain")
    ap.add_argument("--window", type=int, default=5, help="window width (--form boxcar)")
    ap.add_argument("--dim", type=int, default=3, help="covariance dimension (--form covmat)")
    ap.add_argument("--requests", type=int, default=2000, help="requests (--form served)")
    ap.add_argument("--rate", type=float, default=40.0, help="Poisson arrival rate, requests/s (--form served)")
    ap.add_argument("--budget-gib", type=int, default=150, help="server HBM budget (--form served)")
    ap.add_argument("--size", type=int, default=None,
                    help="image side (--form nlinsar: 4096, covmat: 2048, boxcar / goldstein: 8192)")
    ap.add_argument("--niter", type=int, default=10, help="iterations (--form nlinsar)")
    ap.add_argument("--tiles", type=int, default=64, help="images per batch (--form batched)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build measures the single-GPU NL-SAR path "
              "(multi-GPU sharding is by image bands and is not part of this build)", file=sys.stderr)
        sys.exit(2)
    if args.form == "batched":
        return batched(args)
    if args.form == "nlinsar":
        return nlinsar_form(args)
    if args.form == "nlinsar_batched":
        print(json.dumps(nlinsar_batched_run(args.tiles, args.steps, args.warmup)), flush=True)
        return
    if args.form == "served":
        return served(args)
    if args.form == "covmat":
        return covmat_form(args)
    if args.form == "boxcar":
        return boxcar_form(args)
    if args.form == "goldstein":
        return goldstein_form(args)

    import despeckcl_amd as D
    from despeckcl_amd import device as DV
    from despeckcl_amd import hip as HIP
    from despeckcl_amd import synthetic as S

    stats = D.load_nlsar_stats_collection(os.path.join(ROOT, "tests", "golden", "train32_stats.txt"))
    a1, a2, ph = S.interferogram(H, W, S.SEED_BASE + 1, n_regions=16, coh_range=(0.2, 0.95))
    stream = HIP.Stream()
    d_in = [HIP.DeviceArray.from_host(x) for x in (a1, a2, ph)]
    d_out = [HIP.DeviceArray((H, W)) for _ in range(3)]
    nk = len(PATCHES) * len(SCALES)
    d_enl = HIP.DeviceArray((nk, H, W))
    d_alpha = HIP.DeviceArray((nk, H, W))
    if args.tile is not None:
        DV.set_tile(0, args.tile)
    # a step: the filter as the reference computes it, every (patch, scale) key's statistics
    # with their bias-corrected ENL and alpha maps written out, then the selected key's
    # weighted mean and decomposition
    call_all = lambda: DV.nlsar_insar(*[d.ptr for d in d_in], *[d.ptr for d in d_out], H, W, SW, PATCHES,
                                      SCALES, stats, HPAR, CPAR, enl_maps=d_enl.ptr, alpha_maps=d_alpha.ptr,
                                      stream=stream.handle)
    # the same output from the selected key's statistics alone (best_params.cpp:27-53)
    call_sel = lambda: DV.nlsar_insar(*[d.ptr for d in d_in], *[d.ptr for d in d_out], H, W, SW, PATCHES,
                                      SCALES, stats, HPAR, CPAR, stream=stream.handle)
    # the autotuner: every pass variant timed on this image, for the every-key and the
    # selected-key forms; the automatic choice of the calls below takes its records
    tuned = {}
    for ek in (True, False):
        t, tms = DV.nlsar_insar_autotune(*[d.ptr for d in d_in], *[d.ptr for d in d_out], H, W, SW, PATCHES,
                                         SCALES, stats, HPAR, CPAR, every_key=ek, reps=2, stream=stream.handle)
        tuned["every_key" if ek else "selected_key"] = {"variant": t, "ms_per_variant": [round(x, 4) for x in tms]}
    log(f"autotune: {tuned}")
    ub = DV.nlsar_insar(*[d.ptr for d in d_in], *[d.ptr for d in d_out], H, W, SW, PATCHES, SCALES,
                        stats, HPAR, CPAR, stream=stream.handle, want_ub=True)
    lib = D.lib()
    import ctypes as C
    kinds = (C.c_int * 64)()
    times = (C.c_float * 64)()

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        stream.synchronize()
        HIP.synchronize()
        e0, e1 = HIP.Event(), HIP.Event()
        e0.record(stream.handle)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        e1.record(stream.handle)
        e1.synchronize()
        HIP.synchronize()
        wall = time.perf_counter() - t0
        ms = e0.elapsed_ms(e1) / args.steps
        # the passes' own launches, timed with HIP events on the same stream
        lib.dkx_set_timing(1)
        lib.dkx_stage_times(None, None, 0)
        for _ in range(3):
            fn()
        n = lib.dkx_stage_times(kinds, times, 64)
        lib.dkx_set_timing(0)
        last = [times[i] for i in range(n) if kinds[i] == 2]
        other = [times[i] for i in range(n) if kinds[i] == 1]
        return ms, wall, (float(np.mean(last)) if last else float("nan")), \
            (float(np.mean(other)) if other else float("nan"))

    ms, wall, last_ms, other_ms = timed(call_all)
    ms_sel, _w, sel_ms, _o = timed(call_sel)
    px_main = (H - 1) * (W - 1)
    value = H * W / 1e6 / (ms / 1e3)
    achieved = FLOP_PER_PX_LAST_PASS_ALL * px_main / (last_ms * 1e-3) / 1e12
    whole = FLOP_PER_PX_TOTAL * H * W / (ms * 1e-3) / 1e12
    sel_ach = FLOP_PER_PX_LAST_KEY_PASS * px_main / (sel_ms * 1e-3) / 1e12
    out = {
        "metric": "Mpixel/s NL-SAR 2048² (1 GPU, %FP32 roofline) + p99 latency, mixed stream",
        "value": round(value, 3),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded speckle model of insarsim_simu.cpp, 16 regions, coherence U[0.2,0.95])",
        "config": {"workload": "NL-SAR InSAR, one 2048x2048 interferogram per call, every (patch, scale) "
                               "key's statistics and ENL / alpha maps, as the reference computes them",
                   "height": H, "width": W, "search_window": SW, "patch_sizes": PATCHES,
                   "scale_sizes": SCALES, "h": HPAR, "c": CPAR, "form": "plain per-call",
                   "tile": DV.get_tile(0)},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": TRAFFIC_LAST_PASS_BYTES.get(tuned["every_key"]["variant"]),
                     "kernel": "nlsar_scale_pass, last pass (patches 3/5/7 at scale 3, weighted mean, "
                               "decomposition)",
                     "kernel_ms": round(last_ms, 4),
                     "note": "FP32 vector peak (no MFMA on this path; 78.6 TFLOP/s without packed FP32); "
                             f"{FLOP_PER_PX_LAST_PASS_ALL} algorithmic FLOP/px x {px_main} px per launch"},
        "whole_filter": {"tflops": round(whole, 3), "frac": round(whole / PEAK_FP32_TFLOPS, 4),
                         "flop_per_px": FLOP_PER_PX_TOTAL, "last_pass_ms": round(last_ms, 4),
                         "other_pass_ms": round(other_ms, 4)},
        "selected_key_only": {"note": "the same output from the statistics of the key the selection "
                                      "returns (best_params.cpp:27-53), without the other keys' ENL / alpha",
                              "ms_per_step": round(ms_sel, 4), "value": round(H * W / 1e6 / (ms_sel / 1e3), 3),
                              "pass_ms": round(sel_ms, 4), "pass_tflops": round(sel_ach, 3),
                              "flop_per_px_pass": FLOP_PER_PX_LAST_KEY_PASS},
        "autotune": dict(tuned, note="pass variants 0 / 1 / 2 = 4 / 2 / 1 rows per thread, 32 x 32 tiles; "
                                     "mean of 2 runs each on this image; the pick: the fastest or, within 2 % of it, "
                                     "the fewest rows per thread"),
        "gpu": gpu_info(),
        "p99_latency_ms": None,
        "ub_pixels": int(ub),
        "wall_s_timed": round(wall, 3),
    }
    # the host-pointer entry point (copy in, filter, copy out): PCIe-inclusive rate, reported
    # beside the device-resident value
    t0 = time.perf_counter()
    nh = 3
    for _ in range(nh):
        D.nlsar_insar(a1, a2, ph, SW, PATCHES, SCALES, stats, HPAR, CPAR)
    host_ms = (time.perf_counter() - t0) / nh * 1e3
    out["host_entry"] = {"ms": round(host_ms, 3), "mpx_s": round(H * W / 1e6 / (host_ms / 1e3), 3),
                         "note": "despeckcl_amd.nlsar_insar on host arrays, copies included (selected key)"}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a1, a2, ph, stats)
    for d in d_in + d_out + [d_enl, d_alpha]:
        d.free()
    lib.dkx_release_workspace()
    if not args.no_extra:
        # NL-InSAR 4096^2, 10 iterations with point scatterers, and the 64-image batch
        nl = nlinsar_run(4096, 10, 2, 1)
        out["nlinsar_4096"] = {k: nl[k] for k in ("config", "ms_per_step", "value", "mpx_iter_s", "tflops",
                                                   "roofline", "weights_ms", "smooth_ms", "fill_ms",
                                                   "autotune")}
        bt = batched_run(64, 3, 1)
        out["batched_64"] = {k: bt[k] for k in ("config", "value", "ms_per_step", "plain_per_call",
                                                 "batched_eager", "batched_graph", "speedup_vs_plain")}
        nb = nlinsar_batched_run(64, 2, 1)
        out["nlinsar_batched_64"] = {k: nb[k] for k in ("config", "value", "ms_per_step", "plain_per_call",
                                                         "batched_eager", "batched_graph", "speedup_vs_plain")}
    if not args.no_served:
        # the mixed stream's latency: the 2000-request served workload, as `--form served`
        sv = served_run(2000, 40.0, 150)
        out["p99_latency_ms"] = sv["p99_latency_ms"]
        out["served"] = {k: sv[k] for k in ("config", "p50_latency_ms", "p99_latency_ms", "mean_latency_ms",
                                                   "saturated", "serial_plain", "speedup_vs_serial",
                                                   "nlsar_batches", "nlinsar_batches")}
    print(json.dumps(out), flush=True)


def gpu_info():
    """Compute units and peak engine clock of the GPU agent, from rocminfo."""
    import subprocess
    try:
        txt = subprocess.run(["rocminfo"], capture_output=True, text=True, timeout=60).stdout
    except Exception as e:   # the figure is informational
        return {"error": str(e)}
    info = {}
    for block in txt.split("Agent ")[1:]:
        if "Device Type:             GPU" not in block and "Device Type: GPU" not in block.replace("  ", " "):
            continue
        for line in block.splitlines():
            line = line.strip()
            if line.startswith("Marketing Name:"):
                info["name"] = line.split(":", 1)[1].strip()
            elif line.startswith("Name:") and "gfx" in line:
                info["arch"] = line.split(":", 1)[1].strip()
            elif line.startswith("Compute Unit:"):
                info["compute_units"] = int(line.split(":", 1)[1])
            elif line.startswith("Max Clock Freq. (MHz):"):
                info["max_clock_mhz"] = int(line.split(":", 1)[1])
        break
    if "compute_units" in info and "max_clock_mhz" in info:
        # 64 lanes per CU x 2 FLOP per FMA, x 2 for packed dual-FP32
        info["fp32_peak_tflops_non_packed"] = round(info["compute_units"] * 64 * 2 * info["max_clock_mhz"] / 1e6, 1)
        info["fp32_peak_tflops_packed"] = round(2 * info["fp32_peak_tflops_non_packed"], 1)
    return info


if __name__ == "__main__":
    main()
