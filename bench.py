#!/usr/bin/env python3
"""Benchmark of the batched SOCS aerial-image path on one MI355X.

Workload (config 2): 512 seeded Manhattan clips, 1024^2 px at 1 nm (~40 %
clear), EUV 13.5 nm NA 0.33, x-dipole annular sectors sigma 0.7-0.9 with a
35 deg opening, 5 focus planes (-60/-30/0/+30/+60 nm), K = 32 kernels per
plane built on the host in FP64.  One step = one huv_socs_images call over
the whole batch (2560 images), inputs resident in HBM.

Prints one JSON line.  `roofline` is the path's algorithmic HBM bytes per
step over the step time (events on the context stream); `roofline_fp32` is
the FP32 vector rate of the dominant kernel (socs_fields).  `cpu_baseline`
times the FP64 CPU restatement of the reference algorithm on a bounded
sample of the same workload on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_VECTOR_PEAK_TFLOPS = 157.3


def fft_flops(n):
    return 5.0 * n * max(1.0, __import__("math").log2(n))


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--clips", type=int, default=512)
    ap.add_argument("--cpu-clips", type=int, default=2,
                    help="clips in the CPU-baseline sample (x 5 planes)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB before the steps "
                         "(a known-size reference for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(clips are independent, see the multi-GPU note in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import highuvlith_amd as H
    from highuvlith_amd import workloads as W

    src, opt, grid, K = W.c2_setup()
    n = grid.size
    t0 = time.time()
    eng = H.AerialImageEngine(src, opt, grid, max_kernels=K)
    planes = [eng.device_kernel_set(z) for z in W.C2_FOCUS_NM]
    host_build_s = time.time() - t0
    sets = [eng.kernel_set(z) for z in W.C2_FOCUS_NM]
    Ks = [s.num_kernels for s in sets]
    S = [len(s.support) for s in sets]
    B, m = planes[0].band, planes[0].coarse
    P = len(planes)
    nclips = args.clips
    masks = W.manhattan_clips(nclips, n, seed=2026, clear_fraction=0.4)
    out = torch.empty((nclips, P, n, n), dtype=torch.float32, device="cuda")
    ctx = eng.context()
    torch.cuda.synchronize()

    def step():
        H.engine.socs_images_device(ctx, masks, planes, out)

    if args.copy_calibration_mib > 0:
        nel = args.copy_calibration_mib * (1 << 20) // 4
        src_c = torch.ones(nel, dtype=torch.float32, device="cuda")
        dst_c = torch.empty_like(src_c)
        torch.cuda.synchronize()
        dst_c.copy_(src_c)
        torch.cuda.synchronize()
        del src_c, dst_c

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.set_profiling(True)
    ctx.stage_times(reset=True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / args.steps
    stages = ctx.stage_times(reset=True)
    ctx.set_profiling(False)

    nimg = nclips * P
    mpix = nimg * n * n / 1e6
    value = mpix / dt
    algo_bytes = 4.0 * n * n * nclips + 4.0 * n * n * nimg + sum(8.0 * k * s for k, s in zip(Ks, S))
    achieved = algo_bytes / dt / 1e9
    f_ms, f_calls = stages["socs_fields"]
    f_ms_per = f_ms / max(f_calls, 1)
    flops = fields_flops(nimg, max(Ks), m)
    stage_ms = {k: round(v[0] / max(v[1], 1), 4) for k, v in stages.items() if v[1] > 0}
    pipeline_ms = sum(stage_ms.values())

    result = {
        "metric": "SOCS Mpix/s at K=32 (config 2: 512 clips x 5 focus planes, 1024^2, EUV)",
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (complex FP32 fields and FFTs; kernels built in FP64 on the host)",
        "data": "synthetic seeded Manhattan clips generated on the GPU (~40% clear)",
        "config": {"workload": "config 2: batched fixed-size SOCS", "clips": nclips,
                   "grid": f"{n}x{n} @ {grid.pixel_nm} nm", "focus_nm": list(W.C2_FOCUS_NM),
                   "kernels_per_plane": Ks, "support": S, "band": B, "coarse_grid": m,
                   "source_points": int(len(eng.source_points)),
                   "method": sets[0].method, "parallelism": "1 GPU"},
        "images_per_s": round(nimg / dt, 1),
        "host_kernel_build_s": round(host_build_s, 2),
        "stage_ms_per_launch": stage_ms,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "whole SOCS path per step (6 launches)",
                     "algorithmic_bytes": algo_bytes,
                     "event_ms_per_step": round(pipeline_ms, 4)},
        "roofline_fp32": {"kernel": "socs_fields", "achieved": round(flops / (f_ms_per * 1e-3) / 1e12, 2),
                          "peak": FP32_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                          "frac": round(flops / (f_ms_per * 1e-3) / 1e12 / FP32_VECTOR_PEAK_TFLOPS, 4),
                          "flops_per_launch": flops, "ms_per_launch": round(f_ms_per, 4)},
    }

    if not args.no_cpu:
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ncpu = max(1, min(args.cpu_clips, nclips))
        sample = masks[:ncpu].cpu().numpy().astype(np.float64)
        gpu_imgs = out[:ncpu].cpu().numpy()
        t = time.perf_counter()
        worst = 0.0
        for c in range(ncpu):
            spec = oracle.fft2(sample[c].astype(np.complex128))
            for p in range(P):
                ref = oracle.socs_image(sets[p], spec, eng.flare_fraction(), threads=threads)
                worst = max(worst, float(np.abs(ref - gpu_imgs[c, p]).max()))
        cdt = time.perf_counter() - t
        cpu_mpix = ncpu * P * n * n / 1e6 / cdt
        result["cpu_baseline"] = {
            "value": round(cpu_mpix, 4), "unit": "Mpix/s", "cores": threads, "kind": "port",
            "sample": f"{ncpu} clips x {P} planes of the same workload, FP64, reference "
                      f"structure (row-pruned inverse FFT per kernel, <=16 groups), {cdt:.1f} s"}
        result["speedup_vs_cpu"] = round(value / cpu_mpix, 1)
        result["max_abs_err_vs_fp64_sample"] = worst
    print(json.dumps(result))


if __name__ == "__main__":
    main()
