"""Seeded synthetic workloads of the imaging and ILT benchmarks.

  arf_lines_256      one 256^2 clip at 4 nm: 65/130 nm lines + 8 Manhattan
                     rectangles (sides U(40, 200) nm), ArF 193i (NA 1.35),
                     annular 0.7/0.9, K <= 24
  euv_dipole_1024    512 Manhattan clips, 1024^2 at 1 nm, ~40 % clear, EUV NA
                     0.33, x-dipole sectors 0.7-0.9 / 35 deg, focus -60..60 nm
                     (5 planes), K = 32
  arf_annular_2nm    1000 clips with edges {256, 512, 1024, 2048} px
                     (p = .4/.3/.2/.1), 30-60 % density, ArF 193i annular
                     0.7/0.9 at 2 nm, K = 32
  arf_quasar_512     64 contact-array targets 512^2 at 2 nm, ArF 193i quasar,
                     K = 24
Masks are generated on the GPU (torch) for the batched workloads, so the
timed region starts with inputs already resident in HBM.
"""
import math

import numpy as np

from . import engine as E

EUV_FOCUS_PLANES_NM = (-60.0, -30.0, 0.0, 30.0, 60.0)


def arf_lines_256_setup():
    grid = E.GridConfig(256, 4.0)
    src = E.Source.arf(E.IlluminationShape.Annular(0.7, 0.9))
    return src, E.ProjectionOptics.immersion_193i(), grid, 24


def arf_lines_256_features(seed=0):
    """Bright-field mask: opaque 65 nm lines on a 130 nm pitch (one centred at
    x = 0, truncated at the 1024 nm field edge) plus 8 opaque rectangles."""
    rng = np.random.default_rng(seed)
    half = 512.0
    feats = [("lines", 65.0, 130.0, 0, 0.0)]
    for _ in range(8):
        w, h = rng.uniform(40.0, 200.0, 2)
        x, y = rng.uniform(-half, half, 2)
        feats.append(("rect", float(x), float(y), float(w), float(h)))
    return feats


def arf_lines_256_raster(seed=0):
    """Area-coverage transmittance (n, n) of the 256^2 lines mask (product raster)."""
    src, opt, grid, _ = arf_lines_256_setup()
    rects = []
    half = grid.field_size_nm() / 2.0
    for f in arf_lines_256_features(seed):
        if f[0] == "lines":
            _, cd, pitch, _, off = f
            k_lo = math.floor((-half - off - cd / 2) / pitch)
            k_hi = math.ceil((half - off + cd / 2) / pitch)
            for k in range(k_lo, k_hi + 1):
                rects.append((off + k * pitch, 0.0, cd, 2 * half, 0.0))
        else:
            _, x, y, w, h = f
            rects.append((x, y, w, h, 0.0))
    return E.rasterize_rects(rects, grid, background=1.0).real


def euv_dipole_1024_setup():
    grid = E.GridConfig(1024, 1.0)
    src = E.Source(E.EUV_WAVELENGTH_NM, E.IlluminationShape.SectorPoles(0.7, 0.9, 35.0, 0.0, 2))
    return src, E.EuvProjectionOptics.nxe_033(), grid, 32


def arf_annular_2nm_setup(edge):
    grid = E.GridConfig(edge, 2.0)
    src = E.Source.arf(E.IlluminationShape.Annular(0.7, 0.9))
    return src, E.ProjectionOptics.immersion_193i(), grid, 32


def arf_quasar_512_setup():
    grid = E.GridConfig(512, 2.0)
    src = E.Source.arf(E.IlluminationShape.SectorPoles(0.7, 0.9, 30.0, 45.0, 4))
    return src, E.ProjectionOptics.immersion_193i(), grid, 24


def manhattan_clips(nclips, n, seed, clear_fraction=0.4, min_side=8, max_side=128,
                    device="cuda"):
    """Binary Manhattan clips (N, n, n) float32 on `device`: clear = 1 inside the
    union of random axis-aligned rectangles, drawn until the expected clear
    fraction is `clear_fraction` (scalar or per-clip tensor)."""
    import torch
    g = torch.Generator(device="cpu")
    g.manual_seed(int(seed))
    mean_side = 0.5 * (min_side + max_side)
    cf = torch.as_tensor(clear_fraction, dtype=torch.float64).expand(nclips)
    # expected union coverage 1 - exp(-R a / n^2) = clear fraction
    nrect = (-torch.log1p(-cf) * n * n / (mean_side * mean_side)).ceil().long()
    R = int(nrect.max())
    w = torch.randint(min_side, max_side + 1, (nclips, R), generator=g)
    h = torch.randint(min_side, max_side + 1, (nclips, R), generator=g)
    x0 = (torch.rand((nclips, R), generator=g) * (n - w + 1)).long()
    y0 = (torch.rand((nclips, R), generator=g) * (n - h + 1)).long()
    live = torch.arange(R)[None, :] < nrect[:, None]
    x1, y1 = x0 + w, y0 + h
    out = torch.empty((nclips, n, n), dtype=torch.float32, device=device)
    chunk = 64
    for c0 in range(0, nclips, chunk):
        c1 = min(nclips, c0 + chunk)
        diff = torch.zeros((c1 - c0, n + 1, n + 1), dtype=torch.int32, device=device)
        ci = torch.arange(c1 - c0)[:, None].expand(-1, R)
        sel = live[c0:c1]
        for ys, xs, val in ((y0, x0, 1), (y0, x1, -1), (y1, x0, -1), (y1, x1, 1)):
            idx = (ci[sel].to(device), ys[c0:c1][sel].to(device), xs[c0:c1][sel].to(device))
            diff.index_put_(idx, torch.full((int(sel.sum()),), val, dtype=torch.int32,
                                            device=device), accumulate=True)
        cov = diff.cumsum(1).cumsum(2)[:, :n, :n]
        out[c0:c1] = (cov > 0).float()
        del diff, cov
    return out


def uneven_edges(nclips=1000, seed=3):
    rng = np.random.default_rng(seed)
    return rng.choice([256, 512, 1024, 2048], size=nclips, p=[0.4, 0.3, 0.2, 0.1])


def contact_array_targets(ntargets, seed=4, grid=None):
    """Contact arrays: 2x2..6x6 squares, side U(40, 60), pitch U(90, 140) nm,
    per-contact jitter up to 3 nm; 4x supersampled area coverage."""
    from .targets import create_target_holes
    grid = grid or arf_quasar_512_setup()[2]
    rng = np.random.default_rng(seed)
    out = np.zeros((ntargets, grid.size, grid.size))
    for t in range(ntargets):
        cx, cy = rng.integers(2, 7, 2)
        side = rng.uniform(40.0, 60.0)
        px, py = rng.uniform(90.0, 140.0, 2)
        centres = []
        for iy in range(cy):
            for ix in range(cx):
                jx, jy = rng.uniform(-3.0, 3.0, 2)
                centres.append(((ix - (cx - 1) / 2) * px + jx, (iy - (cy - 1) / 2) * py + jy))
        out[t] = create_target_holes(grid, centres, side, "square")
    return out
