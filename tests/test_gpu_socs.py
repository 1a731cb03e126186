"""GPU parity of the batched SOCS image path against the FP64 oracle.

Gate: max |I_gpu - I_oracle| <= 1e-5 of the clear-field
intensity (images are normalised, clear field = 1).  Same kernel sets, same
seeded masks; the oracle computes the reference's structure in FP64
(aerial.rs:1682-1759).
"""
import numpy as np
import pytest

import oracle
import highuvlith_amd as H

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _engine(n, px, lam_kind="arf", max_kernels=24, ill=None, optics=None):
    grid = H.GridConfig(n, px)
    if lam_kind == "arf":
        src = H.Source.arf(ill or H.IlluminationShape.Annular(0.7, 0.9))
        opt = optics or H.ProjectionOptics.immersion_193i()
    else:
        src = H.Source(13.5, ill or H.IlluminationShape.SectorPoles(0.7, 0.9, 35.0, 0.0, 2))
        opt = optics or H.EuvProjectionOptics.nxe_033()
    return H.AerialImageEngine(src, opt, grid, max_kernels=max_kernels)


def _manhattan(n, nrect, rng, lo=8, hi=64):
    m = np.zeros((n, n), np.float32)
    for _ in range(nrect):
        w, h = rng.integers(lo, hi, 2)
        x, y = rng.integers(0, n - w), rng.integers(0, n - h)
        m[y:y + h, x:x + w] = 1.0
    return m


def _oracle_images(eng, masks, zs, pixel_transfer=False):
    out = np.zeros((len(masks), len(zs)) + masks.shape[1:])
    for c, msk in enumerate(masks):
        spec = oracle.fft2(msk.astype(np.complex128))
        if pixel_transfer:
            spec = oracle.pixel_transfer(spec)
        for p, z in enumerate(zs):
            out[c, p] = oracle.socs_image(eng.kernel_set(z), spec, eng.flare_fraction(), threads=8)
    return out


def test_band_spectrum_matches_fft(gpu_ctx):
    import torch
    rng = np.random.default_rng(3)
    for n, band in [(256, 13), (512, 40), (1024, 47), (2048, 54)]:
        masks = rng.random((2, n, n)).astype(np.float32)
        got = H.engine.band_spectrum_device(gpu_ctx, torch.from_numpy(masks).cuda(), band)
        got = got.cpu().numpy()
        for c in range(2):
            full = np.fft.fft2(masks[c].astype(np.float64))
            ks = np.arange(-band, band + 1)
            ref = full[np.ix_(ks % n, ks % n)].T  # [kx][ky]
            err = np.abs(got[c] - ref).max() / np.abs(ref).max()
            print(n, band, err)
            assert err < 2e-6, (n, err)


def test_lines_256_image_matches_oracle():
    """256^2 @ 4 nm, ArF 193i annular 0.7/0.9, K=24 (lines + rectangles clip)."""
    eng = _engine(256, 4.0)
    rng = np.random.default_rng(11)
    masks = np.stack([_manhattan(256, 8, rng) for _ in range(3)])
    got = eng.compute_batch(masks, (0.0,)).cpu().numpy()
    ref = _oracle_images(eng, masks, (0.0,))
    err = np.abs(got - ref).max()
    print("lines-256 max err", err, "K", eng.num_kernels())
    assert err <= TOL


@pytest.mark.parametrize("n,px,kind,K", [
    (256, 2.0, "arf", 16),    # coarse 32
    (512, 2.0, "arf", 24),    # coarse 64
    (1024, 2.0, "arf", 12),   # coarse 128
    (1024, 1.0, "euv", 32),   # coarse 192 (the 1024^2 EUV batch shape)
    (2048, 2.0, "arf", 6),    # coarse 256
])
def test_sizes_match_oracle(n, px, kind, K):
    eng = _engine(n, px, kind, max_kernels=K)
    rng = np.random.default_rng(n + K)
    masks = np.stack([_manhattan(n, 12, rng, n // 32, n // 4) for _ in range(2)])
    zs = (0.0, 30.0)
    got = eng.compute_batch(masks, zs).cpu().numpy()
    ref = _oracle_images(eng, masks, zs)
    ds = eng.device_kernel_set(0.0)
    err = np.abs(got - ref).max()
    print(n, px, kind, "band", ds.band, "coarse", ds.coarse, "K", eng.num_kernels(), "err", err)
    assert err <= TOL


def test_pixel_transfer_and_clear_field():
    eng = _engine(512, 2.0, max_kernels=24)
    rng = np.random.default_rng(5)
    masks = np.stack([_manhattan(512, 10, rng, 16, 128), np.ones((512, 512), np.float32)])
    got = eng.compute_batch(masks, (0.0,), pixel_transfer=True).cpu().numpy()
    ref = _oracle_images(eng, masks, (0.0,), pixel_transfer=True)
    assert np.abs(got - ref).max() <= TOL
    # a clear mask images to the truncated clear field, uniform
    assert np.abs(got[1, 0] - ref[1, 0]).max() <= TOL
    assert got[1, 0].std() < 1e-5


def test_deterministic_and_batch_equals_single():
    eng = _engine(512, 2.0, max_kernels=24)
    rng = np.random.default_rng(9)
    masks = np.stack([_manhattan(512, 10, rng, 16, 128) for _ in range(4)])
    a = eng.compute_batch(masks, (0.0, 50.0)).cpu().numpy()
    b = eng.compute_batch(masks, (0.0, 50.0)).cpu().numpy()
    assert np.array_equal(a, b)
    single = eng.compute_batch(masks[2:3], (50.0,)).cpu().numpy()
    assert np.array_equal(single[0, 0], a[2, 1])


def test_dimension_mismatch_is_an_error():
    eng = _engine(256, 4.0)
    with pytest.raises(H.HuvError) as e:
        eng.compute_batch(np.zeros((1, 128, 128), np.float32), (0.0,))
    assert e.value.kind == "DimensionMismatch"
