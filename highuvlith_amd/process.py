"""Through-focus callers of the imaging engine (crates/highuvlith-core/src/process.rs),
routed through the batched GPU entry points.

  focus_images   process.rs:429-432   one image per focus of one mask (the
                                      spectrum once, every plane in one call)
  batch_defocus  process.rs:1300-1311 the same as (focus, image) pairs; the
                 py_sweep.rs:148-164  Python sweep's batch_defocus
  batch_defocus_many                  several masks x several focus values in
                                      one batched call (the process-window and
                                      OPC callers' access pattern)

Masks are feature lists (rectangles, gratings, arrays, simple polygons:
compute(mask) through the exact spectrum, aerial.rs:1775-1777) or
transmittance rasters (compute_from_transmittance, aerial.rs:1789-1799).
The CD metrology of the process window (ProcessWindow, process.rs:444-1290)
runs on the returned images and is not part of this module.
"""
import numpy as np

from ._lib import HuvError


def _is_raster(mask):
    try:
        import torch
        if isinstance(mask, torch.Tensor):
            return True
    except ImportError:
        pass
    return isinstance(mask, np.ndarray)


def batch_defocus_many(engine, masks, focuses, dark_field=False, absorber=0j):
    """Images (N, P, n, n) torch tensor of N masks at P focus values in one
    batched call: masks is a list of feature lists, or an (N, n, n) raster
    batch (real or complex)."""
    focuses = tuple(float(f) for f in focuses)
    if not focuses:
        raise HuvError(1, "focuses: at least one focus value is required")
    if _is_raster(masks):
        cplx = masks.is_complex() if hasattr(masks, "is_complex") else np.iscomplexobj(masks)
        if cplx:
            return engine.compute_batch_complex(masks, focuses)
        return engine.compute_batch(masks, focuses)
    return engine.compute_features(list(masks), focuses, dark_field=dark_field,
                                   absorber=absorber)


def focus_images(engine, mask, focuses, dark_field=False, absorber=0j):
    """One image per focus (process.rs:429-432): list of (n, n) numpy arrays."""
    if _is_raster(mask):
        imgs = batch_defocus_many(engine, mask[None] if mask.ndim == 2 else mask, focuses)
    else:
        imgs = batch_defocus_many(engine, [mask], focuses, dark_field, absorber)
    imgs = imgs[0].cpu().numpy().astype(np.float64)
    return [imgs[p] for p in range(imgs.shape[0])]


def batch_defocus(engine, mask, focuses, dark_field=False, absorber=0j):
    """process.rs:1300-1311: [(focus_nm, image (n, n)), ...] in the given order."""
    focuses = [float(f) for f in focuses]
    return list(zip(focuses, focus_images(engine, mask, focuses, dark_field, absorber)))
