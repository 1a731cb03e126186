"""Learned warm start for the GPU ILT: seeded targets, ILT labels, training
and the comparison against plain ILT from the target.

Targets (generated on the GPU with exact area coverage):
  contact arrays  2x2..6x6 squares, side U(40, 60) nm, pitch U(90, 140) nm,
                  per-contact jitter up to 3 nm (non-overlapping, so the
                  coverage is the sum of separable 1-D overlaps);
  line/space      dark lines of CD U(40, 80) nm on a pitch U(CD + 50, CD + 120)
                  nm, horizontal or vertical, random phase.
Labels are the GPU ILT's theta after a fixed number of CG/Armijo iterations
from the target (ilt.rs:755-951 semantics).
"""
import math
import time

import numpy as np


def _overlap(lo, hi, edges0, pixel):
    """Area fraction of pixels [edges0 + i p, edges0 + (i+1) p) covered by [lo, hi]."""
    import torch
    n = edges0.shape[-1]
    a = edges0
    b = edges0 + pixel
    return (torch.clamp(torch.minimum(hi, b) - torch.maximum(lo, a), min=0.0) / pixel)


def contact_line_targets(count, n, pixel_nm, seed, device="cuda"):
    """(count, n, n) float32 targets on `device`: even indices contact arrays,
    odd indices line/space patterns."""
    import torch
    rng = np.random.default_rng(seed)
    half = n * pixel_nm / 2.0
    edges = torch.as_tensor(-half + np.arange(n) * pixel_nm, dtype=torch.float64, device=device)
    out = torch.zeros((count, n, n), dtype=torch.float64, device=device)
    for t in range(count):
        if t % 2 == 0:
            cx, cy = rng.integers(2, 7, 2)
            side = rng.uniform(40.0, 60.0)
            px, py = rng.uniform(90.0, 140.0, 2)
            xs = ((np.arange(cx) - (cx - 1) / 2) * px)[None, :] + rng.uniform(-3, 3, (cy, cx))
            ys = ((np.arange(cy) - (cy - 1) / 2) * py)[:, None] + rng.uniform(-3, 3, (cy, cx))
            xs = torch.as_tensor(xs.ravel(), device=device)[:, None]
            ys = torch.as_tensor(ys.ravel(), device=device)[:, None]
            ox = _overlap(xs - side / 2, xs + side / 2, edges[None, :], pixel_nm)  # (C, n) along x
            oy = _overlap(ys - side / 2, ys + side / 2, edges[None, :], pixel_nm)  # (C, n) along y
            out[t] = oy.T @ ox  # [y][x] = sum_c oy_c(y) ox_c(x)
        else:
            cd = rng.uniform(40.0, 80.0)
            pitch = cd + rng.uniform(50.0, 120.0)
            phase = rng.uniform(0.0, pitch)
            k0 = math.floor((-half - phase) / pitch) - 1
            k1 = math.ceil((half - phase) / pitch) + 1
            starts = torch.as_tensor(phase + np.arange(k0, k1) * pitch, device=device)[:, None]
            dark = _overlap(starts, starts + cd, edges[None, :], pixel_nm).sum(0).clamp(max=1.0)
            row = 1.0 - dark
            if rng.random() < 0.5:
                out[t] = row[None, :].expand(n, n)
            else:
                out[t] = row[:, None].expand(n, n)
    return out.to(torch.float32).contiguous()


def ilt_labels(engine, targets, config, iterations=200, batch=64, log=None):
    """Theta after `iterations` CG/Armijo steps from theta = target - 1/2, per
    target (torch CUDA), in batches; returns (thetas, wall seconds).  `log`
    (optional) receives a progress line every 16 batches."""
    import torch
    from .ilt import IltProblem
    prob = IltProblem(engine, config)
    out = torch.empty_like(targets)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(0, targets.shape[0], batch):
        tg = targets[i:i + batch].contiguous()
        th = prob.initial_theta(tg)
        prob.run_stage(tg, th, iterations)
        out[i:i + batch] = th
        if log and (i // batch) % 16 == 15:
            log(f"labels: {i + tg.shape[0]} / {targets.shape[0]}")
    torch.cuda.synchronize()
    return out, time.perf_counter() - t0
