#!/usr/bin/env python3
"""Benchmark of the SOCS aerial-image path and the GPU ILT on one MI355X.

Headline (one step = one batched SOCS pass over the whole batch, inputs
resident in HBM, replayed as one HIP graph launch): 512 seeded Manhattan
clips, 1024^2 px at 1 nm (~40 % clear), EUV 13.5 nm NA 0.33, x-dipole
sectors sigma 0.7-0.9 with a 35 deg opening, 5 focus planes
(-60/-30/0/+30/+60 nm), K = 32 kernels per plane built on the host in FP64
(2560 images per step).

Also reported in the same JSON line (each skipped with --no-single,
--no-rects, --no-uneven, --no-ilt, --no-warm):
  single   one 256^2 @ 4 nm ArF clip (65/130 nm lines truncated at the field
           edge + 8 rectangles), K = 24: latency of the exact-spectrum and
           raster paths, and (with the CPU leg) the oracle on the same clip.
  rects    the headline optics and focus planes on masks given as rectangle
           lists (512 clips of 8 x 8 contacts with sub-pixel size, pitch and
           offset, 64 disjoint cells each): exact band spectrum in FP64 on the
           device, then the same field/coarse/image kernels; Mpix/s including
           the host cell plan and upload, and the rect_band kernel against
           the FP64 vector peak (8 flops per cell and band bin); then the
           same contacts as regular hexagons (simple polygons, edge-sum
           kernel), reported as Mpix/s and edge-bins/s.
  uneven   1000 clips with edges {256, 512, 1024, 2048} px (p .4/.3/.2/.1),
           30-60 % clear, ArF 193i annular 0.7/0.9 at 2 nm, K = 32, one graph
           launch per batch, autotuned launch shapes; aggregate Mpix/s, p50/p99
           per-clip latency (each clip's size imaged alone, graph replay) and
           p50/p99 batch completion time (the clip's size group finishing in an
           eager pass, groups issued smallest first).
  ilt      64 contact-array targets 512^2 at 2 nm, ArF 193i quasar, K = 24,
           resist cost (t 0.25, alpha 50), 50 CG/Armijo iterations; clips/s.
  warm     U-Net warm start (4 levels, 16..128 channels, BF16 on the matrix
           cores): trained on (target, 200-iteration ILT theta) pairs from
           the GPU ILT (seeded contact and line/space targets, 512^2 at 2 nm),
           then on held-out targets plain ILT (200 iterations from the
           target) vs warm start refined to the same per-target cost (no
           convergence stop above it) within the 200-iteration budget, with
           plain ILT from the target as the fallback for targets still above
           it (all timed); clips/s to that fixed cost for both.  The default
           sizes are a reduced version of the stated 32/64/128/256-channel,
           20,000-pair, 20-epoch training (--warm-width, --warm-pairs,
           --warm-epochs, --warm-eval).

`roofline`: the dominant kernel (socs_fields) against the FP32 peak with
algorithmic flops per launch over its HIP-event duration; `traffic` is the
PMC-measured HBM bytes per launch of that kernel (profiles/).
`roofline_path_hbm`: the whole path's algorithmic HBM bytes per step over
the step time against 8 TB/s.  `cpu_baseline`: the FP64 CPU restatement of
the reference algorithm on a bounded sample of the headline workload on this
host's cores.
"""
import argparse
import json
import math
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP32_NONPACKED_TFLOPS = 78.65  # one FP32 op per lane per clock, no packed instructions
METRIC = "SOCS Mpix/s at K=32; ILT clips/s to fixed cost: plain vs NN warm-start"
COUNTERS_SUMMARY = os.path.join(ROOT, "profiles", "socs_batch_1024_euv_k32_counters.md")


def fft_flops(n):
    return 5.0 * n * max(1.0, math.log2(n))


def fields_flops(nimg, K, m):
    """FP32 flops of one socs_fields launch as the kernel executes them.

    Coarse grid m = 32 Q, band box W = 2 ((m - 1) // 4) + 1, 32-point
    transforms (5 N log2 N each).  Per image and kernel:
      x pass, Q W tasks:  W products (6) + W accumulates (2) + (W - 32) fold
                          twiddles (6) + 32 class twists (6) + FFT_32
      y pass, 32 Q^2 tasks: W accumulates (2) + (W - 32) fold twiddles (6)
                          + 32 residue twists (6) + FFT_32 + 32 x lambda|E|^2 (4)
    """
    F = 32
    Q = m // F
    W = 2 * ((m - 1) // 4) + 1
    fold = 6.0 * max(W - F, 0)
    x = Q * W * (8.0 * W + fold + 6.0 * F + fft_flops(F))
    y = F * Q * Q * (2.0 * W + fold + 6.0 * F + fft_flops(F) + 4.0 * F)
    return nimg * K * (x + y)


def fields_algorithmic_flops(Ks, band, m, nclips):
    """Algorithmic FP32 flops of the coarse-field synthesis for a batch:
    per image and kernel, R band rows plus m columns of length-m inverse
    FFTs (5 m log2 m each) and 3 flops per lambda|E|^2 accumulate on the m^2
    grid, R = 2B + 1: K (R + m) 5 m log2 m + 3 K m^2, summed over the planes'
    kernel counts and the clips."""
    R = 2 * band + 1
    per_kernel = (R + m) * fft_flops(m) + 3.0 * m * m
    return nclips * sum(Ks) * per_kernel


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary (None if absent)."""
    try:
        text = open(COUNTERS_SUMMARY).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch\s+" + re.escape(kernel) + r"\s+([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def bench_fixed(args, H, W, np, torch):
    src, opt, grid, K = W.euv_dipole_1024_setup()
    n = grid.size
    t0 = time.time()
    eng = H.AerialImageEngine(src, opt, grid, max_kernels=K)
    # the other focus planes: factor, Gram product and kernels on the device
    eng.build_planes_device(W.EUV_FOCUS_PLANES_NM)
    planes = [eng.device_kernel_set(z) for z in W.EUV_FOCUS_PLANES_NM]
    kernel_build_s = time.time() - t0
    t0 = time.time()
    H.engine.build_kernel_set(opt, eng.source_points, grid, 30.0, src.wavelength_nm, K)
    host_plane_s = time.time() - t0
    sets = [eng.kernel_set(z) for z in W.EUV_FOCUS_PLANES_NM]
    Ks = [s.num_kernels for s in sets]
    S = [len(s.support) for s in sets]
    B, m = planes[0].band, planes[0].coarse
    P = len(planes)
    nclips = args.clips
    masks = W.manhattan_clips(nclips, n, seed=2026, clear_fraction=0.4)
    out = torch.empty((nclips, P, n, n), dtype=torch.float32, device="cuda")
    ctx = eng.context()
    torch.cuda.synchronize()

    if args.copy_calibration_mib > 0:
        nel = args.copy_calibration_mib * (1 << 20) // 4
        src_c = torch.ones(nel, dtype=torch.float32, device="cuda")
        dst_c = torch.empty_like(src_c)
        torch.cuda.synchronize()
        dst_c.copy_(src_c)
        torch.cuda.synchronize()
        del src_c, dst_c

    def step(graph):
        H.engine.socs_images_device(ctx, masks, planes, out, graph=graph)

    for _ in range(max(args.warmup, 1)):
        step(True)  # the first call records the graph
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / args.steps

    # Per-kernel durations: the same launches issued eagerly with HIP events
    # on the context stream (graph replays carry no events).
    ctx.set_profiling(True)
    ctx.stage_times(reset=True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        step(False)
    torch.cuda.synchronize()
    dt_eager = (time.perf_counter() - t) / args.steps
    stages = ctx.stage_times(reset=True)
    ctx.set_profiling(False)

    nimg = nclips * P
    value = nimg * n * n / 1e6 / dt
    algo_bytes = 4.0 * n * n * nclips + 4.0 * n * n * nimg + sum(8.0 * k * s for k, s in zip(Ks, S))
    stage_ms = {k: round(v[0] / max(v[1], 1), 4) for k, v in stages.items() if v[1] > 0}
    # the call is issued as clip ranges: totals per step, launches per step
    step_ms = {k: round(v[0] / args.steps, 4) for k, v in stages.items() if v[1] > 0}
    launches = {k: v[1] // args.steps for k, v in stages.items() if v[1] > 0}
    nl = launches["socs_fields"]
    f_ms = step_ms["socs_fields"]
    exec_flops = fields_flops(nimg, max(Ks), m)
    algo_flops = fields_algorithmic_flops(Ks, B, m, nclips)
    achieved_tf = algo_flops / (f_ms * 1e-3) / 1e12
    exec_tf = exec_flops / (f_ms * 1e-3) / 1e12
    traffic = measured_traffic("socs_fields")
    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic seeded Manhattan clips generated on the GPU (~40% clear); "
                "FP64 host kernel sets",
        "config": {"workload": "batched SOCS, 512 clips x 5 focus planes, 1024x1024 @ 1 nm, "
                               "EUV NA 0.33 x-dipole, K=32, one graph launch per step",
                   "clips": nclips, "grid": f"{n}x{n} @ {grid.pixel_nm} nm",
                   "focus_nm": list(W.EUV_FOCUS_PLANES_NM), "kernels_per_plane": Ks,
                   "support": S, "band": B, "coarse_grid": m,
                   "source_points": int(len(eng.source_points)), "method": sets[0].method,
                   "parallelism": "1 GPU"},
        "images_per_s": round(nimg / dt, 1),
        "eager_ms_per_step": round(dt_eager * 1e3, 4),
        "kernel_build_s": round(kernel_build_s, 3),
        "kernel_build": "engine set-up (source sampling, in-focus plane on the host) + the "
                        "other 4 planes with the factor, Gram product and kernels on the GPU "
                        "(huv_kernels_build_device) + uploads",
        "host_kernel_build_s_per_plane": round(host_plane_s, 3),
        "stage_ms_per_launch": stage_ms,
        "stage_ms_per_step": step_ms,
        "clip_ranges": nl,
        "roofline": {"bound": "valu", "achieved": round(achieved_tf, 2), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved_tf / FP32_PEAK_TFLOPS, 4),
                     "traffic": traffic, "kernel": "socs_fields",
                     "compute": "FP32 on the vector ALUs (no matrix instructions)",
                     "algorithmic_flops_per_launch": algo_flops / nl,
                     "launches_per_step": nl,
                     "algorithmic_formula": "per image: K (R + m) 5 m log2 m + 3 K m^2, "
                                            f"R = 2B+1 = {2 * B + 1}, m = {m}",
                     "executed_flops_per_launch": exec_flops / nl,
                     "frac_executed": round(exec_tf / FP32_PEAK_TFLOPS, 4),
                     "attainable": {
                         "rate_tflops": FP32_NONPACKED_TFLOPS,
                         "why": "packed FP32 instructions are disabled in every device build, "
                                "halving the vector rate",
                         "min_ms_per_step_executed_flops": round(
                             exec_flops / FP32_NONPACKED_TFLOPS / 1e9, 4),
                         "frac_of_attainable": round(exec_tf / FP32_NONPACKED_TFLOPS, 4)},
                     "ms_per_launch": stage_ms["socs_fields"], "ms_per_step": f_ms},
        "roofline_path_hbm": {"bound": "hbm", "achieved": round(algo_bytes / dt / 1e9, 1),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(algo_bytes / dt / 1e9 / HBM_PEAK_GBS, 4),
                              "algorithmic_bytes_per_step": algo_bytes,
                              "band_image_write_GBps": round(
                                  4.0 * n * n * nimg / (step_ms["band_image"] * 1e-3) / 1e9, 1)},
    }
    if not args.no_cpu:
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ncpu = max(1, min(args.cpu_clips, nclips))
        sample = masks[:ncpu].cpu().numpy().astype(np.float64)
        gpu_imgs = out[:ncpu].cpu().numpy()

        def cpu_images(count, nthreads, check):
            worst = 0.0
            t = time.perf_counter()
            for c in range(count):
                spec = oracle.fft2(sample[c].astype(np.complex128))
                for p in range(P):
                    ref = oracle.socs_image(sets[p], spec, eng.flare_fraction(), threads=nthreads)
                    if check:
                        worst = max(worst, float(np.abs(ref - gpu_imgs[c, p]).max()))
            return time.perf_counter() - t, worst

        cdt, worst = cpu_images(ncpu, threads, True)
        cpu_mpix = ncpu * P * n * n / 1e6 / cdt
        cdt1, _ = cpu_images(1, 1, False)
        cpu1_mpix = P * n * n / 1e6 / cdt1
        result["cpu_baseline"] = {
            "value": round(cpu_mpix, 4), "unit": "Mpix/s", "cores": threads, "kind": "port",
            "sample": f"{ncpu} clips x {P} planes of the same workload, FP64, reference "
                      f"structure (row-pruned inverse FFT per kernel, <=16 groups), {cdt:.1f} s "
                      f"on {threads} threads; 1 clip x {P} planes on 1 thread, {cdt1:.1f} s",
            "one_thread_value": round(cpu1_mpix, 4),
            "host": cpu_model(),
            "calibration": calibration_workload(H, np, oracle, threads)}
        result["speedup_vs_cpu"] = round(value / cpu_mpix, 1)
        result["max_abs_err_vs_fp64_sample"] = worst
    return result


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return f"{line.split(':', 1)[1].strip()}, {os.cpu_count()} logical CPUs"
    except OSError:
        pass
    return f"{os.cpu_count()} logical CPUs"


def calibration_workload(H, np, oracle, threads):
    """The reference's own image benchmark (benches/aerial_benchmark.rs:95-108:
    F2 157.63 nm, conventional sigma 0.7, NA 0.75, 2 nm pixels, at most 20
    kernels, a 65/180 nm line/space mask, one compute(mask) per iteration),
    timed on the oracle here beside the published M3 Pro figures (README.md:516:
    2.7 ms at 256^2, 21 ms at 512^2 on 11 cores), so the CPU baseline can be
    read against the reference's own numbers."""
    out = {"workload": "F2 157.63 nm sigma 0.7, NA 0.75, 2 nm, K <= 20, L/S 65/180 nm, "
                       "exact spectrum + SOCS image (oracle, FP64)",
           "published_ms": {"256": 2.7, "512": 21.0}, "published_host": "Apple M3 Pro, 11 cores"}
    for n in (256, 512):
        grid = H.GridConfig(n, 2.0)
        eng = H.AerialImageEngine(H.Source.f2(0.7), H.ProjectionOptics(0.75), grid, max_kernels=20)
        ks = eng.kernel_set(0.0)
        feats = [("lines", 65.0, 180.0, 0, 0.0)]
        for nt, key in ((threads, f"{n}_ms_{threads}_threads"), (1, f"{n}_ms_1_thread")):
            reps = 20 if n == 256 else 5
            t = time.perf_counter()
            for _ in range(reps):
                spec = oracle.mask_spectrum(n, 2.0, feats, dark_field=False)
                oracle.socs_image(ks, spec, eng.flare_fraction(), threads=nt)
            out[key] = round((time.perf_counter() - t) / reps * 1e3, 3)
        out[f"{n}_kernels"] = eng.num_kernels()
        out[f"{n}_ratio_to_published_{threads}_threads"] = round(
            out[f"{n}_ms_{threads}_threads"] / out["published_ms"][str(n)], 3)
    return out


def bench_rects(args, H, W, np, torch):
    src, opt, grid, K = W.euv_dipole_1024_setup()
    n = grid.size
    eng = H.AerialImageEngine(src, opt, grid, max_kernels=K)
    planes = [eng.device_kernel_set(z) for z in W.EUV_FOCUS_PLANES_NM]
    B, P = planes[0].band, len(planes)
    nclips = args.clips
    rng = np.random.default_rng(77)
    L = n * grid.pixel_nm
    clips = []
    for _ in range(nclips):
        pitch = rng.uniform(0.08, 0.11) * L
        size = rng.uniform(0.3, 0.6) * pitch
        x0, y0 = rng.uniform(-0.5, 0.5, 2) * pitch
        clips.append([(x0 + (i - 3.5) * pitch, y0 + (j - 3.5) * pitch, size, size, 1.0)
                      for j in range(8) for i in range(8)])
    cells = 64 * nclips
    arrays = H.engine.rect_arrays(clips, 0.0)
    ctx = eng.context()
    out = torch.empty((nclips, P, n, n), dtype=torch.float32, device="cuda")

    def step():
        H.engine.socs_images_rects_device(ctx, grid, arrays, planes, out)

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / args.steps
    ctx.set_profiling(True)
    ctx.stage_times(reset=True)
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    stages = ctx.stage_times(reset=True)
    ctx.set_profiling(False)
    stage_ms = {k: v[0] / max(v[1], 1) for k, v in stages.items() if v[1] > 0}
    tab_ms, band_ms = stage_ms["rect_tables"], stage_ms["rect_band"]
    flops = 8.0 * cells * (2 * B + 1) ** 2
    tf = flops / (band_ms * 1e-3) / 1e12

    # The same clips with the contacts drawn as regular hexagons (simple
    # polygons, 6 edges each): the spectrum band stage then also runs the
    # polygon edge-sum kernel; timed the same way.
    hexes = []
    for rects in clips:
        hexes.append([([(x + 0.5 * w * np.cos(a), y + 0.5 * w * np.sin(a))
                        for a in np.arange(6) * (np.pi / 3)], 1.0) for (x, y, w, h, _) in rects])
    edges = 6 * 64 * nclips
    no_rects = H.engine.rect_arrays([[] for _ in clips], 0.0)
    hex_arrays = H.engine.PolygonArrays(hexes)

    def poly_step():
        H.engine.socs_images_features_device(ctx, grid, no_rects, hex_arrays, planes, out)

    poly_step()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        poly_step()
    torch.cuda.synchronize()
    dt_poly = (time.perf_counter() - t) / args.steps
    ctx.set_profiling(True)
    ctx.stage_times(reset=True)
    poly_step()
    torch.cuda.synchronize()
    pst = ctx.stage_times(reset=True)
    ctx.set_profiling(False)
    poly_band_ms = sum(pst[k][0] / max(pst[k][1], 1) for k in ("rect_band", "poly_band"))
    poly_only_ms = pst["poly_band"][0] / max(pst["poly_band"][1], 1)
    return {"workload": f"{nclips} clips of 8x8 contacts (sub-pixel), 1024x1024 @ 1 nm, EUV "
                        f"x-dipole, {P} focus planes, K={K}, exact rectangle spectrum",
            "value": round(nclips * P * n * n / 1e6 / dt, 2), "unit": "Mpix/s",
            "ms_per_batch": round(dt * 1e3, 4),
            "rect_tables_ms": round(tab_ms, 4), "rect_band_ms": round(band_ms, 4),
            "rect_band_cells": cells, "band": B,
            "rect_band_roofline": {"bound": "fp64 vector", "achieved": round(tf, 3),
                                   "peak": 78.6, "unit": "TFLOP/s",
                                   "frac": round(tf / 78.6, 4)},
            "stages_ms": {k: round(v, 4) for k, v in stage_ms.items()},
            "polygons": {"workload": f"same clips, contacts as regular hexagons ({edges} edges)",
                         "value": round(nclips * P * n * n / 1e6 / dt_poly, 2),
                         "unit": "Mpix/s (host polygon plan and upload included)",
                         "ms_per_batch": round(dt_poly * 1e3, 4),
                         "band_stage_ms": round(poly_band_ms, 4),
                         "poly_band_ms": round(poly_only_ms, 4),
                         "edge_bins_per_s": round(edges * (2 * B + 1) ** 2 / (poly_only_ms * 1e-3),
                                                  1)}}


def bench_uneven(args, H, W, np, torch):
    edges = W.uneven_edges(args.uneven_clips, seed=3)
    ctx = H.DeviceContext(0)
    ctx.set_autotune(True)
    rng = np.random.default_rng(33)
    groups, pix = [], 0
    t0 = time.time()
    for e in sorted(set(int(x) for x in edges)):
        cnt = int((edges == e).sum())
        src, opt, grid, K = W.arf_annular_2nm_setup(e)
        eng = H.AerialImageEngine(src, opt, grid, max_kernels=K, context=ctx)
        planes = [eng.device_kernel_set(0.0)]
        cf = torch.from_numpy(rng.uniform(0.3, 0.6, cnt))
        masks = W.manhattan_clips(cnt, e, seed=100 + e, clear_fraction=cf,
                                  max_side=max(16, e // 8))
        out = torch.empty((cnt, 1, e, e), dtype=torch.float32, device="cuda")
        groups.append((masks, planes, out, eng))
        pix += cnt * e * e
    host_s = time.time() - t0
    batch = [(g[0], g[1], g[2]) for g in groups]
    # tune each size class on its own inputs, then record the batch graph
    for g in batch:
        H.engine.socs_images_device(ctx, g[0], g[1], g[2])
    for _ in range(2):
        H.engine.socs_images_batch_device(ctx, batch, graph=True)
    torch.cuda.synchronize()
    reps = max(2, args.steps // 2)
    t = time.perf_counter()
    for _ in range(reps):
        H.engine.socs_images_batch_device(ctx, batch, graph=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    # batch-mode completion time of every clip: its size group's completion
    # in an eager pass, groups issued smallest first (clips of a group share it)
    ev0 = torch.cuda.Event(enable_timing=True)
    evs = [torch.cuda.Event(enable_timing=True) for _ in batch]
    ev0.record()
    for g, ev in zip(batch, evs):
        H.engine.socs_images_device(ctx, g[0], g[1], g[2])
        ev.record()
    torch.cuda.synchronize()
    lat = []
    for g, ev in zip(batch, evs):
        lat += [ev0.elapsed_time(ev)] * g[0].shape[0]
    lat = np.array(lat)
    # single-clip latency: one clip of each size imaged on its own (graph
    # replay, median of 20), assigned to every clip of that size
    single = {}
    for masks, planes, out, _ in groups:
        e = masks.shape[-1]
        m1, o1 = masks[:1].contiguous(), out[:1]
        for _ in range(3):
            H.engine.socs_images_device(ctx, m1, planes, o1, graph=True)
        times = []
        for _ in range(20):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            H.engine.socs_images_device(ctx, m1, planes, o1, graph=True)
            b.record()
            b.synchronize()
            times.append(a.elapsed_time(b))
        single[e] = float(np.median(times))
    lat1 = np.concatenate([[single[g[0].shape[-1]]] * g[0].shape[0] for g in batch])
    tuned = {}
    for masks, planes, _, eng in groups:
        e = masks.shape[-1]
        tuned[str(e)] = {"band": planes[0].band, "count": int(masks.shape[0]),
                         "m_n1_xnb_xlanes_chunks": ctx.tuning(e, planes[0].band, 1, planes[0].host.num_kernels)}
    return {"workload": f"{len(edges)} clips, edges 256/512/1024/2048 px at 2 nm, ArF 193i "
                        "annular 0.7/0.9, K=32, one graph launch",
            "mpix_per_batch": round(pix / 1e6, 1), "ms_per_batch": round(dt * 1e3, 3),
            "value": round(pix / 1e6 / dt, 1), "unit": "Mpix/s",
            "p50_clip_latency_ms": round(float(np.percentile(lat1, 50)), 3),
            "p99_clip_latency_ms": round(float(np.percentile(lat1, 99)), 3),
            "clip_latency": "one clip imaged alone (graph replay, median of 20) per size, over "
                            "the batch's clips",
            "single_clip_ms_by_edge": {str(k): round(v, 3) for k, v in sorted(single.items())},
            "p50_batch_completion_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_batch_completion_ms": round(float(np.percentile(lat, 99)), 3),
            "sizes": tuned, "host_kernel_build_s": round(host_s, 2)}


def bench_ilt(args, H, W, np, torch):
    from highuvlith_amd import ilt as L
    src, opt, grid, K = W.arf_quasar_512_setup()
    eng = H.AerialImageEngine(src, opt, grid, max_kernels=K)
    targets = W.contact_array_targets(args.ilt_targets, seed=4, grid=grid)
    cfg = L.ILTConfig(cost="resist", threshold=0.25, steepness=50.0, regularization=0.0,
                      max_iterations=args.ilt_iterations)
    L.optimize_ilt(eng, targets[:2], L.ILTConfig(cost="resist", regularization=0.0,
                                                 max_iterations=2))  # warm-up
    res = L.optimize_ilt(eng, targets, cfg)
    out = {"workload": f"{len(targets)} contact-array targets 512x512 @ 2 nm, ArF 193i quasar, "
                       f"K={K}, resist cost, {args.ilt_iterations} CG/Armijo iterations",
           "value": round(len(targets) / res.wall_s, 3), "unit": "clips/s",
           "wall_s": round(res.wall_s, 3),
           "ms_per_iteration": round(res.wall_s / args.ilt_iterations * 1e3, 3),
           "mean_iterations": float(np.mean(res.iterations)),
           "mean_initial_cost": float(np.mean([h[0] for h in res.cost_history])),
           "mean_final_cost": float(np.mean(res.final_cost)),
           }
    # Compulsory HBM traffic of a CG iteration: theta, gradient and direction
    # read once, 12 B per pixel and target.  The iteration is a chain of
    # dependent launches with host decisions between them, so its bound is
    # launch and synchronisation latency, not HBM; the fraction says how far.
    npix = grid.size * grid.size
    comp = 12.0 * npix * len(targets)
    ms_it = res.wall_s / args.ilt_iterations * 1e3
    out["roofline"] = {"bound": "latency (dependent launches and host round trips)",
                       "compulsory_bytes_per_iteration": comp,
                       "compulsory_ms_at_peak": round(comp / (HBM_PEAK_GBS * 1e9) * 1e3, 4),
                       "achieved_GBps": round(comp / (ms_it * 1e-3) / 1e9, 1),
                       "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(comp / (ms_it * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if not args.no_cpu:
        # the oracle's run_stage (FP64, OpenMP over kernels) on the first target,
        # 3 iterations from the same start
        import oracle
        conds = [(c.dose, c.weight, eng.kernels(c.defocus_nm)) for c in cfg.conditions]
        orc = oracle.IltProblem(grid.size, targets[0], conds, eng.flare_fraction(), cost=cfg.cost,
                                threshold=cfg.threshold, steepness=cfg.steepness,
                                beta=cfg.sigmoid_steepness, tv_weight=cfg.regularization,
                                tv_eps=cfg.tv_epsilon, bin_weight=cfg.binarization_weight)
        t0 = time.perf_counter()
        _, hist, _ = orc.run_stage(targets[0] - 0.5, 3)
        per_it = (time.perf_counter() - t0) / max(len(hist) - 1, 1)
        out["cpu_oracle_s_per_target_iteration"] = round(per_it, 4)
        out["cpu_oracle_clips_per_s"] = round(1.0 / (per_it * args.ilt_iterations), 4)
        out["cpu_oracle_threads"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return out


def bench_single(args, H, W, np, torch):
    """One 256^2 @ 4 nm ArF clip (65 nm lines on a 130 nm pitch truncated at the
    field edge, plus 8 rectangles), K = 24, one focus plane: GPU latency of
    the exact-spectrum (feature list) and raster paths, median of 50 calls;
    with the CPU leg, the oracle (exact spectrum + SOCS image, FP64) on the
    same clip, on 16 threads and on one, and the GPU's error against it."""
    src, opt, grid, K = W.arf_lines_256_setup()
    eng = H.AerialImageEngine(src, opt, grid, max_kernels=K)
    feats = W.arf_lines_256_features(0)
    raster = torch.as_tensor(W.arf_lines_256_raster(0)[None].astype(np.float32), device="cuda")

    def median_ms(fn):
        fn()
        ts = []
        for _ in range(50):
            torch.cuda.synchronize()
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts)) * 1e3

    img = {}
    ms_feat = median_ms(lambda: img.__setitem__("f", eng.compute_features([feats], (0.0,))))
    ms_rast = median_ms(lambda: img.__setitem__("r", eng.compute_batch(raster, (0.0,))))
    out = {"workload": "one clip 256x256 @ 4 nm, ArF 193i annular 0.7/0.9, K=24, 65/130 nm "
                       "lines + 8 rectangles, one focus plane",
           "features_ms": round(ms_feat, 4), "raster_ms": round(ms_rast, 4),
           "features_images_per_s": round(1e3 / ms_feat, 1),
           "raster_images_per_s": round(1e3 / ms_rast, 1),
           "latency": "host call to device completion, median of 50"}
    if not args.no_cpu:
        import oracle
        ks = eng.kernel_set(0.0)
        n = grid.size

        def cpu(threads):
            t = time.perf_counter()
            spec = oracle.mask_spectrum(n, grid.pixel_nm, feats, dark_field=False)
            ref = oracle.socs_image(ks, spec, eng.flare_fraction(), threads=threads)
            return ref, (time.perf_counter() - t) * 1e3
        cpu(16)
        ref, ms16 = cpu(16)
        _, ms1 = cpu(1)
        out["cpu_oracle_ms_16_threads"] = round(ms16, 3)
        out["cpu_oracle_ms_1_thread"] = round(ms1, 3)
        out["max_abs_err_vs_oracle"] = float(np.abs(img["f"][0, 0].cpu().numpy() - ref).max())
    return out


def bench_warm(args, H, W, np, torch):
    from highuvlith_amd import ilt as L
    from highuvlith_amd import unet as UN
    from highuvlith_amd import warmstart as WS
    src, opt, grid, K = W.arf_quasar_512_setup()
    eng = H.AerialImageEngine(src, opt, grid, max_kernels=K)
    cfg = L.ILTConfig(cost="resist", threshold=0.25, steepness=50.0, regularization=0.0)
    n = grid.size
    train = WS.contact_line_targets(args.warm_pairs, n, grid.pixel_nm, seed=500)
    held = WS.contact_line_targets(args.warm_eval, n, grid.pixel_nm, seed=900)
    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    labels, lab_s = WS.ilt_labels(eng, train, cfg, iterations=200, log=log)
    net = UN.UNet(eng.context(), levels=4, base=args.warm_width, seed=7)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = net.train(train, labels, epochs=args.warm_epochs, batch=8, lr=1e-3, log=log)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del train, labels
    torch.cuda.empty_cache()
    res = UN.compare_warm_start(eng, net, held.cpu().numpy(), cfg, budget=200)
    T = args.warm_eval
    steps = args.warm_epochs * (args.warm_pairs // 8)
    stated = (args.warm_pairs >= 20000 and args.warm_epochs >= 20 and args.warm_eval >= 512
              and args.warm_width == 32)
    chans = "/".join(str(args.warm_width << l) for l in range(4))
    return {"workload": f"U-Net 4 levels {chans} ch, {args.warm_pairs} training pairs "
                        f"(512x512 @ 2 nm, 200-iteration ILT labels), {args.warm_epochs} epochs "
                        f"Adam lr 1e-3 batch 8; {T} held-out targets",
            "scale": "stated (32/64/128/256 ch, 20,000 pairs, 20 epochs, 512 held out)" if stated
                     else "reduced from the stated 32/64/128/256 ch, 20,000 pairs x 20 epochs "
                          "(time limit of the default run)",
            "plain_clips_per_s": round(T / res.plain_wall_s, 3),
            "warm_clips_per_s": round(T / res.warm_wall_s, 3),
            "unit": "clips/s to the plain 200-iteration cost",
            "warm_over_plain_wall": round(res.warm_wall_s / res.plain_wall_s, 4),
            "warm_wall_includes": "inference, refinement and plain-ILT fallback",
            "warm_stages_s": res.stages_s,
            "plain_mean_iterations": float(np.mean(res.plain_iterations)),
            "warm_mean_iterations": float(np.mean(res.warm_iterations)),
            "reached_fraction": float(np.mean(res.reached)),
            "reached_by_refinement_fraction": float(np.mean(res.reached_by_refinement)),
            "continued_count": int(np.sum(res.continued)),
            "plain_fallback_count": int(np.sum(res.plain_fallback)),
            "never_reached_count": int(np.sum(~res.reached)),
            "fallback_count": int(np.sum(res.fallback)),
            "mean_plain_cost": float(np.mean(res.plain_cost)),
            "mean_warm_cost": float(np.mean(res.warm_cost)),
            "label_generation_s": round(lab_s, 2), "train_s": round(train_s, 2),
            "train_samples_per_s": round(steps * 8 / train_s, 1),
            "train_loss_first_last": [round(hist[0], 5), round(hist[-1], 5)]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--clips", type=int, default=512)
    ap.add_argument("--cpu-clips", type=int, default=2,
                    help="clips in the CPU-baseline sample (x 5 planes)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-rects", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-ilt", action="store_true")
    ap.add_argument("--no-warm", action="store_true")
    ap.add_argument("--warm-pairs", type=int, default=4096)
    ap.add_argument("--warm-epochs", type=int, default=16)
    ap.add_argument("--warm-eval", type=int, default=512)
    ap.add_argument("--warm-width", type=int, default=16,
                    help="U-Net base channels (the stated network is 32/64/128/256: 32)")
    ap.add_argument("--uneven-clips", type=int, default=1000)
    ap.add_argument("--ilt-targets", type=int, default=64)
    ap.add_argument("--ilt-iterations", type=int, default=50)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB before the steps "
                         "(a known-size reference for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(clips and targets are independent: each GPU would take its own share of the batch)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import highuvlith_amd as H
    from highuvlith_amd import workloads as W

    result = bench_fixed(args, H, W, np, torch)
    print("fixed batch done", file=sys.stderr, flush=True)
    for name, fn, skip in (("single", bench_single, args.no_single),
                           ("rects", bench_rects, args.no_rects),
                           ("uneven", bench_uneven, args.no_uneven),
                           ("ilt", bench_ilt, args.no_ilt),
                           ("warm", bench_warm, args.no_warm)):
        if skip:
            continue
        try:
            result[name] = fn(args, H, W, np, torch)
        except Exception as exc:  # report, keep the headline line
            result[name] = {"error": f"{type(exc).__name__}: {exc}"}
        print(f"{name} done", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
