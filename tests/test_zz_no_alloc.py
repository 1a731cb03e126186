"""huv_socs_images_noalloc: the no-allocation variant of huv_socs_images.

Called through the C ABI (ctypes), its images must equal huv_socs_images'
on the same inputs (same kernels and launch shapes, deterministic sums: bit
for bit) and stay within 1e-5 of the clear field of the FP64 oracle, on two
calls in a row, the second with fewer clips.  Calls before the set-up call
and batches larger than the fixed scratch are refused with a status.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import highuvlith_amd as H
from highuvlith_amd import _lib

pytestmark = pytest.mark.gpu


def _noalloc(ctx, masks, planes, out, pixel_transfer=0):
    import torch
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    arr = (C.c_void_p * len(planes))(*[p.handle.value for p in planes])
    return ctx.lib.huv_socs_images_noalloc(ctx.handle, masks.shape[0], masks.shape[-1],
                                           C.c_void_p(masks.data_ptr()), len(planes), arr,
                                           C.c_void_p(out.data_ptr()), int(pixel_transfer))


def test_noalloc_matches_socs_images_on_two_calls():
    import torch
    n = 256
    grid = H.GridConfig(n, 4.0)
    eng = H.AerialImageEngine(H.Source.arf(H.IlluminationShape.Annular(0.7, 0.9)),
                              H.ProjectionOptics.immersion_193i(), grid, max_kernels=24)
    zs = (0.0, 40.0)
    planes = [eng.device_kernel_set(z) for z in zs]
    ctx = eng.context()
    rng = np.random.default_rng(41)
    masks6 = torch.from_numpy((rng.random((6, n, n)) > 0.55).astype(np.float32)).cuda()
    out = torch.empty((6, 2, n, n), dtype=torch.float32, device="cuda")
    # before the set-up call the tables are missing: refused, nothing launched
    st = _noalloc(ctx, masks6, planes, out)
    assert st == _lib.HUV_INVALID_PARAMETER, st
    assert b"huv_socs_images once" in ctx.lib.huv_last_error()
    # set-up: one ordinary call with the same kernel sets
    eng.compute_batch(masks6[:1], zs)
    for count in (6, 3):  # second call smaller than the first
        m = masks6[:count].contiguous()
        ref = eng.compute_batch(m, zs)
        got = torch.full((count, 2, n, n), float("nan"), device="cuda")
        st = _noalloc(ctx, m, planes, got)
        assert st == _lib.HUV_OK, ctx.lib.huv_last_error()
        torch.cuda.synchronize()
        diff = (got - ref).abs().max().item()
        print("noalloc vs socs_images", count, diff)
        assert diff == 0.0, (count, diff)
        spec = oracle.fft2(m[count - 1].cpu().numpy().astype(np.complex128))
        for p, z in enumerate(zs):
            o = oracle.socs_image(eng.kernel_set(z), spec, eng.flare_fraction(), threads=8)
            err = float(np.abs(got[count - 1, p].cpu().numpy() - o).max())
            print("noalloc vs oracle", count, z, err)
            assert err <= 1e-5, (count, z, err)


def test_noalloc_refuses_batches_beyond_its_scratch():
    import torch
    n = 256
    grid = H.GridConfig(n, 4.0)
    eng = H.AerialImageEngine(H.Source.arf(H.IlluminationShape.Annular(0.7, 0.9)),
                              H.ProjectionOptics.immersion_193i(), grid, max_kernels=24)
    planes = [eng.device_kernel_set(0.0)]
    ctx = eng.context()
    eng.compute_batch(torch.zeros((1, n, n), device="cuda"), (0.0,))
    big = 2000  # about 0.13 MB of intermediates per clip at this shape: > 64 MiB
    masks = torch.zeros((big, n, n), device="cuda")
    out = torch.empty((big, 1, n, n), device="cuda")
    st = _noalloc(ctx, masks, planes, out)
    assert st == _lib.HUV_INVALID_PARAMETER, st
    assert b"MiB of scratch" in ctx.lib.huv_last_error()
