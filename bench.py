#!/usr/bin/env python3
"""Benchmark of TDC SAR image formation on one MI355X.

Metric: G pixel*chirp*samples/s (terms/s / 1e9) at rel-L2 <= 1e-4 vs FP64,
plain vs fast (precision) form, and p50/p99 job latency of a serving stream.
Workload (default, "arc"): 1024 chirps x 512 samples along a 1 m arc at 2 m
slant range, 512 x 512 pixels at 2 mm, 2000 scatterers, 20 dB SNR, Hann x
Hann windows pre-applied -- with complex64, FP16 and BF16 raw storage for the
precision form, and the plain form (the drop-in) on the same data. Serving
stream: 1000 seeded Poisson-arrival jobs (64-2048 chirps, 128/256/512
samples, 32^2-512^2 pixels) against a 200 GB HBM budget.

A step is one formation of the full image from inputs already resident in
HBM (correlate-sum kernel plus the fixed-order split reduction). W warmup
steps, then K timed steps between two device synchronisations. `value` is
the precision form (complex64 raw) rate with the library's automatic engine
(the range-profile engine for these 512-sample chirps), the form that meets
the 1e-4 bound on the input as given; the other forms and engines are
reported beside it.

Accuracy: relative L2 error and peak-pixel phase error against the FP64
oracle on a seeded random subset of 4096 pixels plus the 64 brightest
pixels. CPU baseline: the reference method (FP32, per pixel and 128-chirp
chunk, materialised window / reference / product arrays, then a sum) on a
bounded pixel sample using the host cores.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sarlab-tdc_amd"))

METRIC = ("G pixel·chirp·samples/s at rel-L2≤1e-4 vs FP64 (plain vs fast); "
          "p50/p99 job ms")
UNIT = "G pixel·chirp·samples/s"
PEAK_FP32_TFLOPS = 157.3        # MI355X published FP32 vector peak
PEAK_FP16_MFMA_TFLOPS = 2500.0  # MI355X published dense FP16 matrix peak
# The direct sum's FP32 flops per term (the reference method's count): plain =
# 3 sub + 5 norm^2 + sqrt + scale + sin + cos + 8 (complex MAC); precision = 6
# (complex rotation) + 8 (MAC). A kernel that evaluates the same sum with less
# work is credited with these only in the `direct_sum_equiv` fields, never in
# `roofline.achieved`.
DIRECT_FLOP_PER_TERM = {"plain": 20.0, "precision": 14.0}
# FP32 flops each kernel executes per term (roofline.achieved): plain per
# term = range (3 sub, 3 norm^2, sqrt) + the anchored phase (FMA, fract) + sin,
# cos + 8 (MAC), plus the offset from the chirp's first position (num 6, add,
# sqrt, add, rcp, mul, add, FMA) = 27; plain on stop-and-hop input ("range",
# range once per pixel and chirp) = FMA + fract + sin + cos + 8 = 12; vector
# engine = 4 FMA (8 flops, Horner form)
EXEC_FLOP_PER_TERM = {"plain": 27.0, "range": 12.0, "vector": 8.0}
# FP16 MFMA flops the matrix engine executes per term: 8 per product (one
# complex MAC as a real 2x2 block) times hi*hi + hi*lo (+ lo*hi for c64 raw)
MFMA_FLOP_PER_TERM = {"c64": 24.0, "f16": 16.0, "bf16": 16.0}
# Arithmetic the range-profile kernel executes per (pixel, chirp). Every
# pair, FP32: the kernel exponent arguments and exp2 (6), 2 taps - 1 weight
# products, `taps` complex-by-real MACs (4 taps), the anchor sin/cos and the
# rotation into the block sum (10); FP64: the block sums (2 per 8 chirps).
# A chirp whose range is evaluated in FP64 (every chirp, or the first of each
# block of 8 with the incremental range) adds FP64 |pos|^2 - 2 pos.p + |p|^2
# (7), the clamped Newton range step (5), the bin position (2) and the anchor
# phase (1), and with the incremental range FP32 1 / R (1); the block's other
# chirps add FP32 delta / 2 (6), e0 and u (2), the series (4), dR (1), R and
# 1 / R (2), the bin position (3) and the anchor phase (3). Counted as
# FP32-equivalent flops (an FP64 flop issues at half the FP32 rate) against
# the FP32 peak.
PROFILE_BLOCK = 8
PROFILE_FP64_BLOCK_SUMS = 0.25
PROFILE_FP64_EXACT_RANGE = 15.0
PROFILE_FP32_INCR_FIRST = 1.0
PROFILE_FP32_INCR_STEP = 21.0


def profile_fp32_per_pair(taps, incr=False):
    base = 6.0 + (2.0 * taps - 1.0) + 4.0 * taps + 10.0
    if not incr:
        return base
    return base + (PROFILE_FP32_INCR_FIRST + (PROFILE_BLOCK - 1) * PROFILE_FP32_INCR_STEP) / PROFILE_BLOCK


def profile_fp64_per_pair(incr=False):
    exact = PROFILE_FP64_EXACT_RANGE / (PROFILE_BLOCK if incr else 1)
    return exact + PROFILE_FP64_BLOCK_SUMS


def profile_fp32eq_per_pair(taps, incr=False):
    return profile_fp32_per_pair(taps, incr) + 2.0 * profile_fp64_per_pair(incr)


def profile_incremental(sc):
    """Whether the library advances a served job's ranges by FP32 increments
    (the host bound of the range-profile engine, restated: positions and
    pixels from the midpoint of the first and last pixels, the smallest
    distance between their bounding boxes, the largest chirp step, and the
    larger of the anchor wavenumber and |dk| M)."""
    f = sc.f64()
    grid = f["focusGrid"].reshape(-1, 3)
    o = 0.5 * (grid[0] + grid[-1])
    c = f["position"] - o
    g = grid - o
    if len(c) < 2:
        return False
    gap = np.maximum(0.0, np.maximum(g.min(0) - c.max(0), c.min(0) - g.max(0)))
    rmin = float(np.linalg.norm(gap))
    if not rmin > 0.0:
        return False
    dc = float(np.linalg.norm(np.diff(c, axis=0), axis=1).max())
    cn = float(np.linalg.norm(c, axis=1).max())
    pn = float(np.linalg.norm(np.maximum(np.abs(g.min(0)), np.abs(g.max(0)))))
    ns = sc.n_samples
    t = sc.sample_times
    dt = (t[-1] - t[0]) / (ns - 1) if ns > 1 else 0.0
    kh = 2.0 * (sc.f0 + sc.mod_rate * (t[0] + (ns // 2) * dt)) / 299792458.0
    taps, over = served_kernel(ns, sc.n_pixels)
    m = profile_length(ns) * over // 2
    dkm = abs(2.0 * sc.mod_rate * dt / 299792458.0) * m
    eps = 2.0 ** -24
    e0 = dc * (1.0 + dc / rmin)
    err = (eps * (5.0 * cn + 6.0 * pn) * dc / (2.0 * rmin) + 5.0 * eps * e0
           + 0.625 * e0 ** 4 / rmin ** 3)
    return max(abs(kh), dkm) * err <= 2e-6


def profile_length(ns):
    """FFT length of the range-profile engine for ns samples (the library's)."""
    m = 16
    while m < 2 * ns:
        m *= 2
    return m


# Collection time of one aperture: chirps x the 100 us pulse repetition
# interval of every scene. The serving stream's shortest aperture (64 chirps)
# takes 6.4 ms to collect; forming every image within that time keeps a
# sensor that streams back-to-back apertures real-time, so it is the stated
# p99 latency target of the stream.
PRI_S = 100e-6
STREAM_MIN_CHIRPS = 64


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name (Linux /proc/cpuinfo), or the platform's."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or platform.machine()


def make_proc(tdc, s, kind, storage="c64", engine="auto", moving=False):
    p = tdc.TdcProcessor(0)
    p.setLogging(False)
    if kind == "precision":
        f = s.f64()
        p.setRawDataF64(f["rawData"], f["priTimes"], f["sampleTimes"], f["position"],
                        f["modRate"], f["startFreq"])
        p.setFocusGridF64(f["focusGrid"])
        p.setRawStorage(storage)
        p.setEngine(engine)
        keep = f
    else:
        d = s.dropin()
        if moving:
            # the antenna keeps moving along its velocity during each chirp: one
            # position per sample, so the range is evaluated per term
            vel = d["velocity"].astype(np.float64)
            d["position"] = np.ascontiguousarray(
                (d["position"].astype(np.float64) + vel * s.sample_times[None, :, None])
                .astype(np.float32))
        p.setRawData(d["rawData"], d["priTimes"], d["sampleTimes"], d["position"], d["velocity"],
                     d["attitude"], d["modRate"], d["startFreq"], d["bodyBoresight"])
        p.setFocusGrid(d["focusGrid"])
        keep = d
    return p, keep


# The drop-in with the Doppler window on: centroid-centred band of 4 kHz with
# the raised-cosine taper (the platform moves ~10 m/s along the arc, so the
# Doppler spread at 77 GHz is ~5 kHz and the window varies over the aperture).
DOPPLER = dict(center=0.0, bandwidth=4000.0, centroid=True, taper=True)


def setup_variant(tdc, s, kind, storage, engine="auto", doppler=None, moving=False):
    """A processor for one variant with its inputs in HBM: one end-to-end
    start() (host arrays -> H2D -> compute -> D2H; the first call also allocates
    the processor's device buffers), then prepare()."""
    p, keep = make_proc(tdc, s, kind, storage, engine, moving)
    win = ((doppler["center"], doppler["bandwidth"], doppler["centroid"], False,
            doppler["taper"]) if doppler else (0.0, 1e30, False, False, False))
    p.start(*win)
    e2e_first_ms = p.stats()["end_to_end_ms"]
    p.prepare(*win)
    return dict(p=p, keep=keep, win=win, kind=kind, storage=storage, e2e_first_ms=e2e_first_ms)


def time_prepared(v, steps, warmup):
    """`warmup` untimed steps, then exactly `steps` timed ones."""
    p = v["p"]
    if warmup > 0:
        p.run(warmup)
    t0 = time.perf_counter()
    p.run(steps)
    t1 = time.perf_counter()
    v["ms_per_step"] = (t1 - t0) * 1e3 / steps
    v["stats"] = p.stats()


def finish_variant(v):
    """The variant's image and its steady end-to-end time: the median of three
    start() calls on the processor, reusing its device buffers."""
    p, st = v["p"], v["stats"]
    p.fetchImage()
    img = p.getFocusedImage().copy()
    again = []
    for _ in range(3):
        p.start(*v["win"])
        again.append(p.stats()["end_to_end_ms"])
    e2e_ms = sorted(again)[1]
    return dict(kind=v["kind"], storage=v["storage"], ms_per_step=v["ms_per_step"],
                main_kernel_ms=st["main_kernel_ms"],
                taps=st["taps"], oversampling=st["oversampling"], incremental=st["incremental"],
                kernel_ms=st["kernel_ms"], e2e_ms=e2e_ms, e2e_first_ms=v["e2e_first_ms"],
                terms=st["terms"],
                splits=st["chirp_splits"], ppl=st["pixels_per_lane"], engine=st["engine"],
                doppler_active=st["doppler_active"], skipped_chunks=st["skipped_chunks"],
                processed_chunks=st["processed_chunks"], window_per_chirp=st["window_per_chirp"],
                range_per_chirp=st["range_per_chirp"], range_series=st["range_series"],
                workgroups=st["workgroups"], image=img, keep=v["keep"])


def time_variant(tdc, s, kind, storage, steps, warmup, engine="auto", doppler=None,
                 moving=False):
    v = setup_variant(tdc, s, kind, storage, engine, doppler, moving)
    time_prepared(v, steps, warmup)
    return finish_variant(v)


def served_kernel(ns, npx):
    """(taps, oversampling) the server's range-profile tiles run for chirps of
    ns samples and npx pixels: the library's automatic choice."""
    return (6, 4) if profile_length(ns) * 2 <= 16384 and npx >= 32768 else (10, 2)


def serve_stream(tdc, srv, scenes, rate, form, storage="c64", seed=0):
    """Poisson arrivals at `rate` jobs/s into the server `srv` (idle when
    called); returns the stream's record and per-job results."""
    import threading

    rng = np.random.default_rng([5, seed, 1])
    gaps = rng.exponential(1.0 / rate, size=len(scenes))
    arrivals = np.cumsum(gaps) - gaps[0]
    # warm-up: three jobs, not counted
    for s in scenes[:3]:
        f = s.f64()
        srv.wait(srv.submit(f["rawData"], f["sampleTimes"], f["position"], f["focusGrid"],
                            f["modRate"], f["startFreq"], form=form, storage=storage))
    st0 = srv.stats()
    inputs = [s.f64() for s in scenes]
    ids = [None] * len(scenes)
    results = [None] * len(scenes)
    ready = threading.Semaphore(0)

    def waiter():
        for k in range(len(scenes)):
            ready.acquire()
            results[k] = srv.wait(ids[k])

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    for k, f in enumerate(inputs):
        dt = arrivals[k] - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids[k] = srv.submit(f["rawData"], f["sampleTimes"], f["position"], f["focusGrid"],
                            f["modRate"], f["startFreq"], form=form, storage=storage)
        ready.release()
    th.join()
    st = srv.stats()
    lat = np.array([r[1]["latency_ms"] for r in results])
    t_sub = min(r[1]["t_submit"] for r in results)
    t_end = max(r[1]["t_done"] for r in results)
    terms = sum(r[1]["terms"] for r in results)
    makespan = t_end - t_sub
    # the arithmetic the kernels execute for these jobs: per (pixel, chirp) of
    # the range-profile tile (in the steps) and the per-chirp FFT (5 M log2 M)
    # of its table (staged beside the steps, on the upload streams), or per
    # term of the plain tile
    executed = 0.0
    in_steps = 0.0
    aperture_ms = []
    for sc in scenes:
        nc, ns, npx = sc.n_chirps, sc.n_samples, sc.n_pixels
        aperture_ms.append(nc * PRI_S * 1e3)
        if form == "precision":
            taps, over = served_kernel(ns, npx)
            m = profile_length(ns) * over // 2
            pairs = float(npx) * nc * profile_fp32eq_per_pair(taps, profile_incremental(sc))
            executed += pairs + nc * 5.0 * m * np.log2(m)
            in_steps += pairs
        else:
            executed += float(npx) * nc * ns * EXEC_FLOP_PER_TERM["range"]
            in_steps = executed
    direct = terms * DIRECT_FLOP_PER_TERM[form]
    busy_s = (st["device_busy_ms"] - st0["device_busy_ms"]) / 1e3
    realtime = float(np.mean(lat <= np.array(aperture_ms)))
    return dict(form=form, storage=storage, jobs=len(scenes), rate=rate,
                p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), max_ms=float(lat.max()), makespan_s=makespan,
                gterms=terms / makespan / 1e9, terms=terms, steps=st["steps"] - st0["steps"],
                # the server's peak device footprint so far (all streams up to this one)
                device_busy_s=busy_s, bytes_peak=st["bytes_peak"], graphs=st["graphs"],
                # executed FP32(-equivalent) flops over the makespan, and the
                # steps' own flops over the steps' device time, against the FP32
                # vector peak
                executed_tflop=executed / 1e12,
                fp32_peak_frac=executed / makespan / (PEAK_FP32_TFLOPS * 1e12),
                fp32_peak_frac_busy=in_steps / busy_s / (PEAK_FP32_TFLOPS * 1e12),
                # the direct sum's flops for the same images (credited work, not executed)
                direct_sum_equiv_frac=direct / makespan / (PEAK_FP32_TFLOPS * 1e12),
                # share of jobs formed within their own aperture's collection time
                within_aperture_frac=realtime), results


def cpu_refmethod_rate(oracle, s, cores, seconds, seed=99):
    """The reference method (FP32, per pixel and 128-chirp chunk) on a pixel
    sample sized to ~`seconds` of host time; G terms/s and the sample."""
    d = s.dropin()
    per_px = float(s.n_chirps) * s.n_samples
    rng = np.random.default_rng(seed)
    n_cal = min(s.n_pixels, max(cores, 16))
    cal = rng.choice(s.n_pixels, size=n_cal, replace=False)
    t0 = time.perf_counter()
    oracle.refmethod_f32(d, 0.0, 1e30, False, False, False, pixels=cal, nthreads=cores)
    dt = time.perf_counter() - t0
    n_px = int(min(s.n_pixels, max(n_cal, seconds / max(dt, 1e-6) * n_cal)))
    sample = rng.choice(s.n_pixels, size=n_px, replace=False)
    t0 = time.perf_counter()
    oracle.refmethod_f32(d, 0.0, 1e30, False, False, False, pixels=sample, nthreads=cores)
    dt = time.perf_counter() - t0
    return n_px * per_px / dt / 1e9, n_px


def config_leg(tdc, args, cores):
    """The other configurations: plain and precision (complex64) form rates,
    rel-L2 and peak-pixel phase against the FP64 oracle (all pixels up to 4096,
    else a seeded 4096-pixel subset plus the 64 brightest), and the reference
    method's host rate on a bounded pixel sample."""
    from sarlabtdc import scene

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    out = {}
    for name in [x.strip() for x in args.configs.split(",") if x.strip()]:
        s = scene.make(name)
        rec = dict(chirps=s.n_chirps, samples=s.n_samples,
                   pixels=[int(s.grid.shape[0]), int(s.grid.shape[1])], terms=s.terms)
        imgs = {}
        for kind in ("plain", "precision"):
            r = time_variant(tdc, s, kind, "c64", args.config_steps, args.config_warmup)
            imgs[kind] = r["image"].reshape(-1)
            rec[kind] = dict(gterms=round(r["terms"] / (r["ms_per_step"] * 1e-3) / 1e9, 3),
                             gterms_kernel=round(r["terms"] / (r["main_kernel_ms"] * 1e-3) / 1e9, 3),
                             ms_per_step=round(r["ms_per_step"], 4),
                             main_kernel_ms=round(r["main_kernel_ms"], 4),
                             e2e_ms=round(r["e2e_ms"], 3), engine=r["engine"],
                             chirp_splits=r["splits"])
        if not args.no_check:
            npx = s.n_pixels
            if npx <= 4096:
                px = np.arange(npx)
            else:
                rng = np.random.default_rng(4096)
                px = np.unique(np.concatenate([rng.choice(npx, size=4096, replace=False),
                                               np.argsort(np.abs(imgs["precision"]))[-64:]]))
            d = s.dropin()
            refs = {"plain": oracle.image_f64(d["rawData"], d["sampleTimes"], d["position"],
                                              d["focusGrid"], float(d["modRate"]),
                                              float(d["startFreq"]), pixels=px),
                    "precision": oracle.image_f64(s.raw, s.sample_times, s.track, s.grid,
                                                  s.mod_rate, s.f0, pixels=px)}
            for kind in ("plain", "precision"):
                img = imgs[kind][px]
                rec[kind]["rel_l2"] = float(f"{oracle.rel_l2(img, refs[kind]):.4g}")
                rec[kind]["peak_phase_rad"] = float(
                    f"{oracle.peak_phase_error(img, refs[kind]):.4g}")
            rec["checked_pixels"] = int(px.size)
        if not args.no_cpu_baseline:
            rate, n_px = cpu_refmethod_rate(oracle, s, cores, args.config_cpu_seconds)
            rec["cpu_baseline"] = dict(value=round(rate, 5), cores=cores, kind="port",
                                       sample=f"{n_px} random pixels, FP32 reference method")
        rec["precision_vs_plain"] = round(rec["precision"]["gterms"] / rec["plain"]["gterms"], 3)
        log(f"{name}: plain {rec['plain']['gterms']:.1f} G terms/s "
            f"(rel-L2 {rec['plain'].get('rel_l2', float('nan')):.3e}), precision "
            f"{rec['precision']['gterms']:.1f} (rel-L2 "
            f"{rec['precision'].get('rel_l2', float('nan')):.3e})")
        out[name] = rec
    return out


def serve_leg(tdc, args):
    from sarlabtdc import scene

    shapes = scene.stream_shapes(args.serve_jobs)
    t0 = time.perf_counter()
    scenes = []
    for j, sh in enumerate(shapes):
        scenes.append(scene.stream_job(j, sh))
        if (j + 1) % 200 == 0:
            log(f"serving: generated {j + 1}/{len(shapes)} jobs ({time.perf_counter() - t0:.1f} s)")
    out = {}
    # one server for every stream, as deployed, warmed by one unmeasured pass
    # of the stream offered all at once: its pinned-memory and device-arena
    # pools then hold what the streams need (a cold server's first burst pays
    # those allocations, and uploads stalled for tens of ms meanwhile, measured)
    srv = tdc.Server(0, 200e9)
    serve_stream(tdc, srv, scenes, 1e9, "precision", seed=9)
    rec, res = serve_stream(tdc, srv, scenes, args.serve_rate, "precision")
    log(f"serving precision: {rec['jobs']} jobs at {rec['rate']:.0f}/s: p50 {rec['p50_ms']:.2f} ms, "
        f"p99 {rec['p99_ms']:.2f} ms, {rec['gterms']:.1f} G terms/s over {rec['makespan_s']:.2f} s "
        f"(device busy {rec['device_busy_s']:.2f} s, {rec['steps']} steps)")
    out["precision"] = rec
    if args.serve_capacity_rate > 0:
        # the same jobs offered faster, to show throughput and latency near the
        # device's capacity
        rec_c, _ = serve_stream(tdc, srv, scenes, args.serve_capacity_rate, "precision", seed=1)
        log(f"serving precision at {rec_c['rate']:.0f}/s: p50 {rec_c['p50_ms']:.2f} ms, "
            f"p99 {rec_c['p99_ms']:.2f} ms, {rec_c['gterms']:.1f} G terms/s over "
            f"{rec_c['makespan_s']:.2f} s (device busy {rec_c['device_busy_s']:.2f} s)")
        out["precision_capacity"] = rec_c
    if args.serve_saturation_rate > 0:
        # the same jobs offered faster than the device forms them: the makespan
        # is the server's, so the executed fraction is the device-bound one
        rec_s, _ = serve_stream(tdc, srv, scenes, args.serve_saturation_rate, "precision", seed=2)
        log(f"serving precision at {rec_s['rate']:.0f}/s: p50 {rec_s['p50_ms']:.2f} ms, "
            f"p99 {rec_s['p99_ms']:.2f} ms, {rec_s['gterms']:.1f} G terms/s over "
            f"{rec_s['makespan_s']:.3f} s (device busy {rec_s['device_busy_s']:.3f} s, executed "
            f"{rec_s['fp32_peak_frac']:.3f} of the FP32 peak over the makespan)")
        out["precision_saturation"] = rec_s
    if args.serve_resident_jobs > 0:
        # the same jobs formed one at a time on the resident path (a processor
        # per job's inputs, the same engine and kernel): the device time the
        # server's steps are measured against
        n_res = min(len(scenes), args.serve_resident_jobs)
        p = tdc.TdcProcessor(0)
        p.setLogging(False)
        win = (0.0, 1e30, False, False, False)
        main_ms = pass_ms = 0.0
        for sc in scenes[:n_res]:
            f = sc.f64()
            p.setRawDataF64(f["rawData"], f["priTimes"], f["sampleTimes"], f["position"],
                            f["modRate"], f["startFreq"])
            p.setFocusGridF64(f["focusGrid"])
            p.prepare(*win)
            p.run(1)
            p.run(3)
            st = p.stats()
            main_ms += st["main_kernel_ms"]
            pass_ms += st["kernel_ms"]
        res_rec = dict(jobs=n_res, sum_correlate_ms=main_ms, sum_pass_ms=pass_ms)
        if n_res == len(scenes) and "precision_saturation" in out:
            busy_ms = out["precision_saturation"]["device_busy_s"] * 1e3
            res_rec["saturated_steps_ms"] = busy_ms
            res_rec["steps_vs_correlate"] = main_ms / busy_ms
            res_rec["steps_vs_pass"] = pass_ms / busy_ms
        out["resident_same_jobs"] = res_rec
        log(f"serving: the same {n_res} jobs one at a time on the resident path: correlate-sum "
            f"{main_ms:.1f} ms, whole passes {pass_ms:.1f} ms"
            + (f" (saturated steps {res_rec['saturated_steps_ms']:.1f} ms)"
               if "saturated_steps_ms" in res_rec else ""))
    n_plain = min(len(scenes), args.serve_plain_jobs)
    if n_plain > 0:
        rec_p, _ = serve_stream(tdc, srv, scenes[:n_plain], args.serve_rate, "plain")
        log(f"serving plain: {rec_p['jobs']} jobs at {rec_p['rate']:.0f}/s: p50 "
            f"{rec_p['p50_ms']:.2f} ms, p99 {rec_p['p99_ms']:.2f} ms, {rec_p['gterms']:.1f} G terms/s")
        out["plain"] = rec_p
    srv.close()
    # accuracy of a few served jobs against the FP64 oracle (checker only)
    if not args.no_check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        rng = np.random.default_rng(55)
        errs = []
        for k in rng.choice(len(scenes), size=min(4, len(scenes)), replace=False):
            s = scenes[k]
            img = res[k][0].reshape(-1)
            px = np.unique(np.concatenate([rng.choice(img.size, size=min(256, img.size),
                                                      replace=False), [int(np.argmax(np.abs(img)))]]))
            ref = oracle.image_f64(s.raw, s.sample_times, s.track, s.grid, s.mod_rate, s.f0,
                                   pixels=px)
            errs.append(oracle.rel_l2(img[px], ref))
        out["precision"]["rel_l2_max_checked"] = float(max(errs))
        log(f"serving precision: rel-L2 of {len(errs)} checked jobs <= {max(errs):.3e}")
    return out


def main():
    # the library's log lines (the reference's "Completed logging setup" on
    # every processor) go to stdout: send everything to stderr and keep the
    # real stdout for the one JSON line
    json_out = os.fdopen(os.dup(1), "w")
    sys.stdout.flush()
    os.dup2(2, 1)
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    # the GPU's clocks settle under sustained work: from an idle GPU the
    # range-profile kernel on 1024 x 512 x 512^2 took 0.50-0.55 ms for its first
    # steps, 0.43 ms after 50 and 0.418-0.422 ms from 150 to 2000 (measured)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="arc", choices=["line_jitter", "arc", "helix"])
    # variants run in this order: the plain form's legs (~30-60 ms per step)
    # come first and leave the GPU at sustained clocks for the precision
    # forms, whose steps are ~0.5 ms
    ap.add_argument("--variants", default="plain,doppler,plain_moving,vector,matrix,f16,bf16,c64",
                    help="comma list of c64, f16, bf16 (precision form, automatic engine), "
                         "matrix and vector (precision form, c64, that engine), plain, doppler "
                         "(the plain form with the Doppler window on) and plain_moving (the "
                         "plain form with the antenna moving during each chirp: range per "
                         "term)")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 accuracy check")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--configs", default="line_small,line_jitter,helix",
                    help="other workloads measured beside the headline one (plain and "
                         "precision form, accuracy, host rate); empty to skip")
    ap.add_argument("--config-steps", type=int, default=5)
    ap.add_argument("--config-warmup", type=int, default=5)
    ap.add_argument("--config-cpu-seconds", type=float, default=1.5)
    ap.add_argument("--serve-jobs", type=int, default=1000, help="serving-stream jobs (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=2400.0, help="offered jobs/s")
    ap.add_argument("--serve-capacity-rate", type=float, default=9600.0,
                    help="offered jobs/s of a second precision stream of the same jobs (0: skip)")
    ap.add_argument("--serve-saturation-rate", type=float, default=100000.0,
                    help="offered jobs/s of a third precision stream of the same jobs, above the "
                         "device's capacity (0: skip)")
    ap.add_argument("--serve-resident-jobs", type=int, default=1000,
                    help="jobs of the stream also formed one at a time on the resident path, "
                         "for the device time of the same work without the server (0: skip)")
    ap.add_argument("--serve-plain-jobs", type=int, default=200,
                    help="jobs of the stream also served in the plain form")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (the traffic calibration "
                         "kernel for rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        # torch's copy kernel in a process that never loads the TDC library
        import torch

        n = args.copy_probe * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4, "dispatches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the path shards by pixel tiles with no exchange; see the design notes)",
              file=sys.stderr)
        sys.exit(2)

    import sarlabtdc as tdc
    from sarlabtdc import scene

    t0 = time.perf_counter()
    s = scene.make(args.workload)
    log(f"scene {args.workload}: {s.n_chirps} x {s.n_samples}, {s.grid.shape[0]}x"
        f"{s.grid.shape[1]} px, {len(s.amplitudes)} scatterers "
        f"({time.perf_counter() - t0:.1f} s)")

    # one small image first, so that no variant's end-to-end time carries the
    # process's one-time device and module initialisation
    warm = scene.make("line_small")
    for kind in ("precision", "plain"):
        p, _ = make_proc(tdc, warm, kind)
        p.start(0.0, 1e30, False, False, False)

    results = {}
    # every variant is set up first (inputs in HBM, one end-to-end start()),
    # then the variants' warm-up and timed steps run back to back, so that no
    # host-side set-up idles the GPU between them and every variant is timed at
    # the clocks of sustained work; the images and end-to-end times follow
    names = [x.strip() for x in args.variants.split(",") if x.strip()]
    prepared = {}
    for v in names:
        kind = "plain" if v in ("plain", "doppler", "plain_moving") else "precision"
        storage = "c64" if kind == "plain" or v in ("vector", "matrix") else v
        engine = v if v in ("vector", "matrix") else "auto"
        prepared[v] = setup_variant(tdc, s, kind, storage, engine,
                                    DOPPLER if v == "doppler" else None, v == "plain_moving")
    for v in names:
        time_prepared(prepared[v], args.steps, args.warmup)
    for v in names:
        r = finish_variant(prepared.pop(v))
        r["gterms"] = r["terms"] / (r["ms_per_step"] * 1e-3) / 1e9
        r["gterms_kernel"] = r["terms"] / (r["main_kernel_ms"] * 1e-3) / 1e9
        r["gterms_e2e"] = r["terms"] / (r["e2e_ms"] * 1e-3) / 1e9
        results[v] = r
        log(f"{v:5s}: {r['ms_per_step']:.3f} ms/step, kernel {r['main_kernel_ms']:.3f} ms, "
            f"{r['gterms']:.1f} G terms/s (e2e {r['e2e_ms']:.1f} ms), splits {r['splits']}, "
            f"engine {r['engine']}, wgs {r['workgroups']}")

    # ---- accuracy against the FP64 oracle (checker only) ----
    acc = {}
    if not args.no_check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        npx = s.n_pixels
        ref_img = next(iter(results.values()))["image"].reshape(-1)
        rng = np.random.default_rng(1234)
        px = np.unique(np.concatenate([rng.choice(npx, size=min(4096, npx), replace=False),
                                       np.argsort(np.abs(ref_img))[-64:]]))
        t0 = time.perf_counter()
        o64 = None
        o32 = None
        for v, r in results.items():
            if v == "doppler":
                d = r["keep"]
                dop = dict(velocity=d["velocity"], attitude=d["attitude"],
                           boresight=d["bodyBoresight"], center=DOPPLER["center"],
                           bandwidth=DOPPLER["bandwidth"], centroid=DOPPLER["centroid"],
                           taper=DOPPLER["taper"])
                ref = oracle.image_f64(d["rawData"], d["sampleTimes"], d["position"],
                                       d["focusGrid"], float(d["modRate"]),
                                       float(d["startFreq"]), doppler=dop, pixels=px)
            elif v == "plain":
                if o32 is None:
                    d = r["keep"]
                    o32 = oracle.image_f64(d["rawData"], d["sampleTimes"], d["position"],
                                           d["focusGrid"], float(d["modRate"]),
                                           float(d["startFreq"]), pixels=px)
                ref = o32
            elif v == "plain_moving":
                d = r["keep"]
                ref = oracle.image_f64(d["rawData"], d["sampleTimes"], d["position"],
                                       d["focusGrid"], float(d["modRate"]),
                                       float(d["startFreq"]), pixels=px)
            else:
                if o64 is None:
                    o64 = oracle.image_f64(s.raw, s.sample_times, s.track, s.grid, s.mod_rate,
                                           s.f0, pixels=px)
                ref = o64
            img = r["image"].reshape(-1)[px]
            acc[v] = dict(rel_l2=oracle.rel_l2(img, ref), peak_phase_rad=oracle.peak_phase_error(img, ref))
            log(f"{v:5s}: rel-L2 {acc[v]['rel_l2']:.3e}, peak phase {acc[v]['peak_phase_rad']:.3e} rad")
        log(f"oracle on {px.size} px: {time.perf_counter() - t0:.1f} s")

    # ---- CPU baseline: the reference method on the host cores ----
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
        d = s.dropin()
        per_px = float(s.n_chirps) * s.n_samples
        rng = np.random.default_rng(99)
        n_cal = max(cores, 16)
        cal = rng.choice(s.n_pixels, size=n_cal, replace=False)
        t0 = time.perf_counter()
        oracle.refmethod_f32(d, 0.0, 1e30, False, False, False, pixels=cal, nthreads=cores)
        dt = time.perf_counter() - t0
        # three timed runs of a third of the time budget each; the median is reported
        n_px = int(min(s.n_pixels, max(n_cal, args.cpu_seconds / 3 / max(dt, 1e-6) * n_cal)))
        n_px = max(cores, n_px // cores * cores)
        sample = rng.choice(s.n_pixels, size=n_px, replace=False)
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            oracle.refmethod_f32(d, 0.0, 1e30, False, False, False, pixels=sample, nthreads=cores)
            runs.append(time.perf_counter() - t0)
        dt = sorted(runs)[1]
        # the same method on one thread, over the calibration pixels
        t0 = time.perf_counter()
        oracle.refmethod_f32(d, 0.0, 1e30, False, False, False, pixels=cal, nthreads=1)
        dt1 = time.perf_counter() - t0
        cpu = dict(value=n_px * per_px / dt / 1e9, unit=UNIT, cores=cores, kind="port",
                   single_thread=n_cal * per_px / dt1 / 1e9, nproc=os.cpu_count(),
                   cpu_model=cpu_model(),
                   sample=f"{args.workload}: {n_px} random pixels x {s.n_chirps} chirps x "
                          f"{s.n_samples} samples, FP32 reference method, median of 3 runs "
                          f"({', '.join(f'{r:.1f}' for r in runs)} s); single thread over "
                          f"{n_cal} pixels ({dt1:.1f} s)")
        log(f"cpu baseline: {cpu['value']:.4f} G terms/s on {cores} cores (median of "
            f"{', '.join(f'{r:.1f}' for r in runs)} s), {cpu['single_thread']:.4f} on one thread; "
            f"{cpu['cpu_model']}, nproc {cpu['nproc']}")

    cores_cfg = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    configs = config_leg(tdc, args, cores_cfg) if args.configs.strip() else None
    serving = serve_leg(tdc, args) if args.serve_jobs > 0 else None

    head = results.get("c64") or next(iter(results.values()))
    kind = head["kind"]
    fkey = "range" if kind == "plain" and head["range_per_chirp"] else kind
    terms_per_s = head["terms"] / (head["main_kernel_ms"] * 1e-3)
    # what the direct sum would execute for the same image (credited work)
    direct = terms_per_s * DIRECT_FLOP_PER_TERM[kind] / 1e12
    direct_equiv = dict(flop_per_term=DIRECT_FLOP_PER_TERM[kind], tflops=round(direct, 3),
                        frac_of_fp32_peak=round(direct / PEAK_FP32_TFLOPS, 4))
    traffic = None
    simd = None
    sq = {}
    kname = {"matrix": "precision_matrix", "vector": "precision",
             "profile": "profile"}.get(head["engine"], fkey)
    if kind == "plain" and head["doppler_active"]:
        kname = "doppler_chirp" if head["window_per_chirp"] else "doppler"

    tfile = os.path.join(ROOT, "profiles", f"traffic_{kname}_{args.workload}.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            tj = json.load(fh)
        traffic = float(tj["traffic_bytes_per_launch"])
        sq = tj.get("sq", {})
        if sq.get("GRBM_GUI_ACTIVE"):
            # share of the SIMDs' cycles (1024 SIMDs, GRBM summed over 8 XCDs) spent
            # issuing MFMA and other VALU work (2 cycles per wave64 VALU instruction),
            # from the committed rocprofv3 PMC pass of the same kernel and workload
            cyc = sq["GRBM_GUI_ACTIVE"] / 8 * 1024
            mfma = sq.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0) / cyc
            valu = 2.0 * (sq["SQ_INSTS_VALU"] - sq.get("SQ_INSTS_MFMA", 0.0)) / cyc
            simd = dict(mfma_busy_frac=round(mfma, 3), valu_issue_frac=round(valu, 3),
                        source=os.path.relpath(tfile, ROOT))
    if head["engine"] == "profile":
        # the range-profile kernel's own arithmetic per (pixel, chirp), on the
        # vector ALUs (no MFMA), over its measured launch time
        taps = head["taps"]
        incr = bool(head["incremental"])
        pairs_per_s = terms_per_s / s.n_samples
        fp32eq = profile_fp32eq_per_pair(taps, incr)
        achieved = pairs_per_s * fp32eq / 1e12
        if simd is not None and sq.get("SQ_INSTS_VALU"):
            wave_chirps = s.n_pixels / 64.0 * s.n_chirps
            sq = dict(sq, **tj.get("sq64", {}))
            simd = dict(valu_issue_frac=simd["valu_issue_frac"],
                        valu_insts_per_wave_chirp=round(sq["SQ_INSTS_VALU"] / wave_chirps, 2),
                        **{k.lower().replace("sq_insts_", "") + "_per_wave_chirp":
                           round(sq[k] / wave_chirps, 2)
                           for k in ("SQ_INSTS_VALU_TRANS_F32", "SQ_INSTS_VALU_FMA_F64",
                                     "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_ADD_F64",
                                     "SQ_INSTS_VALU_CVT") if k in sq},
                        source=os.path.relpath(tfile, ROOT))
            # a SIMD issues up to two VALU instructions per quad-cycle: the shares
            # of its quad-cycles that issued two, one and none (third PMC pass)
            if "issue" in tj:
                simd["issue"] = tj["issue"]
        roof = dict(bound="valu", achieved=round(achieved, 3), peak=PEAK_FP32_TFLOPS,
                    unit="TFLOP/s", frac=round(achieved / PEAK_FP32_TFLOPS, 4), traffic=traffic,
                    kernel=(f"profile_kernel ({head['storage']}, {taps} taps, "
                            f"{head['oversampling']}x oversampling)"),
                    executed=dict(fp32_flop_per_pixel_chirp=profile_fp32_per_pair(taps, incr),
                                  fp64_flop_per_pixel_chirp=profile_fp64_per_pair(incr),
                                  incremental_range=incr,
                                  fp32_equivalent_per_pixel_chirp=fp32eq,
                                  pixel_chirp_pairs_per_launch=float(s.n_pixels) * s.n_chirps),
                    direct_sum_equiv=direct_equiv, simd_cycles=simd)
        # the same kernel's arithmetic as the hardware counts it
        # (SQ_INSTS_VALU_FLOPS_* of the committed PMC pass, per pair), over this
        # run's kernel time: a check on the hand count above
        hwf = tj.get("hw_flops") if os.path.exists(tfile) else None
        if hwf and hwf.get("fp32_equivalent"):
            hw_ach = pairs_per_s * hwf["fp32_equivalent"] / 1e12
            roof["executed"]["hw_counted"] = dict(
                fp32_per_pixel_chirp=hwf["fp32"], fp32_trans_per_pixel_chirp=hwf["fp32_trans"],
                fp64_per_pixel_chirp=hwf["fp64"],
                fp32_equivalent_per_pixel_chirp=hwf["fp32_equivalent"],
                tflops=round(hw_ach, 3), frac=round(hw_ach / PEAK_FP32_TFLOPS, 4),
                source=os.path.relpath(tfile, ROOT))
    elif head["engine"] == "matrix":
        # FP16 products on the matrix cores; the VALU part (phasors, nested
        # sums, anchors) issues on the same SIMDs in between
        mflop = MFMA_FLOP_PER_TERM[head["storage"]]
        achieved = terms_per_s * mflop / 1e12
        roof = dict(bound="mfma", achieved=round(achieved, 3), peak=PEAK_FP16_MFMA_TFLOPS,
                    unit="TFLOP/s", frac=round(achieved / PEAK_FP16_MFMA_TFLOPS, 4),
                    traffic=traffic, mfma_flop_per_term=mflop,
                    kernel=f"{kname}_kernel ({head['storage']})",
                    direct_sum_equiv=direct_equiv, simd_cycles=simd)
    else:
        ekey = "vector" if head["engine"] == "vector" else fkey
        flop = EXEC_FLOP_PER_TERM[ekey]
        achieved = terms_per_s * flop / 1e12
        roof = dict(bound="valu", achieved=round(achieved, 3), peak=PEAK_FP32_TFLOPS,
                    unit="TFLOP/s", frac=round(achieved / PEAK_FP32_TFLOPS, 4),
                    traffic=traffic, flop_per_term=flop, kernel=f"{kname}_kernel ({head['storage']})",
                    direct_sum_equiv=direct_equiv, simd_cycles=simd)
    # HBM bytes per launch (committed PMC pass) over this run's kernel time
    roof["hbm_gbs"] = (round(traffic / (head["main_kernel_ms"] * 1e-3) / 1e9, 1)
                       if traffic else None)
    out = {
        "metric": METRIC,
        "value": round(head["gterms"], 3),
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.workload}: {s.n_chirps} chirps x {s.n_samples} samples, "
                         f"{s.grid.shape[0]}x{s.grid.shape[1]} pixels, {kind} form, "
                         f"{head['storage']} raw"),
            "chirps": s.n_chirps, "samples": s.n_samples,
            "pixels": [int(s.grid.shape[0]), int(s.grid.shape[1])],
            "terms": head["terms"],
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "forms": {
            v: dict(gterms=round(r["gterms"], 3), gterms_kernel=round(r["gterms_kernel"], 3),
                    ms_per_step=round(r["ms_per_step"], 4),
                    main_kernel_ms=round(r["main_kernel_ms"], 4),
                    e2e_ms=round(r["e2e_ms"], 3), e2e_first_ms=round(r["e2e_first_ms"], 3),
                    gterms_e2e_pcie=round(r["gterms_e2e"], 3),
                    chirp_splits=r["splits"], engine=r["engine"],
                    **({"taps": r["taps"], "oversampling": r["oversampling"],
                        "incremental_range": bool(r["incremental"])}
                       if r["engine"] == "profile" else {}),
                    **({"range_per_chirp": bool(r["range_per_chirp"]),
                        "range_series": bool(r["range_series"])} if r["kind"] == "plain" else {}),
                    **({"skipped_chunks": r["skipped_chunks"],
                        "processed_chunks": r["processed_chunks"],
                        "window_per_chirp": bool(r["window_per_chirp"])}
                       if r["doppler_active"] else {}),
                    **({k: float(f"{x:.4g}") for k, x in acc.get(v, {}).items()}))
            for v, r in results.items()
        },
    }
    if configs:
        out["configs"] = configs
    if serving:
        out["serving"] = {
            k: {kk: (float(f"{vv:.5g}") if isinstance(vv, float) else vv) for kk, vv in v.items()}
            for k, v in serving.items()}
        pr = serving["precision"]
        out["p50_ms"] = round(pr["p50_ms"], 3)
        out["p99_ms"] = round(pr["p99_ms"], 3)
        target = STREAM_MIN_CHIRPS * PRI_S * 1e3
        for k in ("precision", "precision_capacity", "precision_saturation", "plain"):
            if k in out["serving"]:
                out["serving"][k]["p99_target_ms"] = target
                out["serving"][k]["p99_target_met"] = bool(serving[k]["p99_ms"] <= target)
    if "plain" in results and "c64" in results:
        out["precision_vs_plain"] = round(results["c64"]["gterms"] / results["plain"]["gterms"], 3)
    if "vector" in results and "matrix" in results:
        out["matrix_vs_vector"] = round(results["matrix"]["gterms"] / results["vector"]["gterms"], 3)
    if "matrix" in results and "c64" in results and results["c64"]["engine"] == "profile":
        out["profile_vs_matrix"] = round(results["c64"]["gterms"] / results["matrix"]["gterms"], 3)
    json_out.write(json.dumps(out) + "\n")
    json_out.flush()


if __name__ == "__main__":
    main()
