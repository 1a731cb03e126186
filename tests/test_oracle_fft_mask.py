"""Oracle FFT and exact mask spectrum against definitions and the reference's
own fixtures (math/fft2d.rs:228-414, mask.rs:1979-2083)."""
import numpy as np

import oracle
import highuvlith_amd as H


def test_fft_matches_naive_dft_and_numpy():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8))
    f = oracle.fft2(x)
    assert np.abs(f - oracle.dft2_naive(x)).max() < 1e-10
    assert np.abs(oracle.fft2(x, True) - oracle.dft2_naive(x, True)).max() < 1e-12
    y = rng.normal(size=(64, 32)) + 1j * rng.normal(size=(64, 32))
    assert np.abs(oracle.fft2(y) - np.fft.fft2(y)).max() < 1e-10


def test_fft_roundtrip_parseval_dc_and_pruned():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32))
    assert np.abs(oracle.fft2(oracle.fft2(x), True) - x).max() < 1e-12
    f = oracle.fft2(x)
    assert abs((np.abs(f) ** 2).sum() / x.size - (np.abs(x) ** 2).sum()) < 1e-9
    assert abs(f[0, 0] - x.sum()) < 1e-10
    # pruned inverse on a spectrum with zero rows equals the full inverse
    s = np.zeros_like(x)
    rows = [0, 1, 5, 31]
    s[rows] = x[rows]
    assert np.abs(oracle.ifft2_pruned(s, rows) - oracle.fft2(s, True)).max() < 1e-14


def test_clear_mask_dc_is_n_squared():
    s = oracle.mask_spectrum(32, 3.0, [], dark_field=False)
    assert s[0, 0] == 1024.0
    assert np.all(s.ravel()[1:] == 0)


def test_rect_spectrum_matches_reference_quadrature_fixture():
    # mask.rs:1989-2012 (dark-field rect with sub-pixel edges, N = 64, 4 nm)
    s = oracle.mask_spectrum(64, 4.0, [("rect", 3.3, -5.2, 41.7, 60.1)], dark_field=True)
    fixtures = [
        ((0, 0), (1.566356250000e2, 0.0)),
        ((0, 1), (-1.498118343911e2, 4.781639045323e0)),
        ((1, 0), (-1.405921075627e2, -2.510656487625e1)),
        ((2, 3), (-6.653606958088e1, -1.753688923051e1)),
        ((63, 5), (2.900396579588e1, -1.013751332586e1)),
        ((60, 61), (-5.385969788449e0, 3.773485996544e0)),
        ((5, 62), (1.074513265635e1, 1.494311273936e1)),
    ]
    for (ky, kx), (re, im) in fixtures:
        assert abs(s[ky, kx] - complex(re, im)) < 1e-9 * 156.6


def test_line_space_spectrum_only_pitch_harmonics():
    # mask.rs:2015-2039: 65/180 L/S on a 2-period field
    s = oracle.mask_spectrum(64, 360.0 / 64, [("lines", 65.0, 180.0, 0, 0.0)], dark_field=False)
    expect = [(0, (2.616888888889e3, 0.0)), (2, (-1.175951712624e3, -1.158211344566e2)),
              (4, (-4.897878251666e2, -9.742485581761e1)),
              (62, (-1.175951712624e3, 1.158211344566e2)),
              (60, (-4.897878251666e2, 9.742485581761e1))]
    for kx, (re, im) in expect:
        assert abs(s[0, kx] - complex(re, im)) < 1e-8 * 2616.9
    mask = np.ones_like(s, dtype=bool)
    mask[0, ::2] = False
    assert np.all(s[mask] == 0)


def test_pixel_aligned_raster_times_pixel_transfer_is_exact():
    # mask.rs:2071-2083; the raster is the product's (host C++), the spectrum the oracle's
    grid = H.GridConfig(64, 4.0)
    t = H.rasterize_rects([(4.0, -8.0, 40.0, 24.0, 1.0)], grid, background=0.0)
    spec = oracle.pixel_transfer(oracle.fft2(t))
    exact = oracle.mask_spectrum(64, 4.0, [("rect", 4.0, -8.0, 40.0, 24.0)], dark_field=True)
    assert np.abs(spec - exact).max() / np.abs(exact).max() < 1e-12


def test_raster_converges_to_exact_spectrum_with_finer_pixels():
    # mask.rs:2041-2069
    def low_order_error(n):
        grid = H.GridConfig(n, 256.0 / n)
        exact = oracle.mask_spectrum(n, 256.0 / n, [("rect", 3.3, -5.2, 41.7, 60.1)], True)
        r = oracle.fft2(H.rasterize_rects([(3.3, -5.2, 41.7, 60.1, 1.0)], grid, 0.0))
        idx = [0, 1, 2, 3, n - 1, n - 2]
        return np.abs(r[np.ix_(idx, idx)] - exact[np.ix_(idx, idx)]).max() / abs(exact[0, 0])
    coarse, fine = low_order_error(64), low_order_error(256)
    assert fine < 5e-4 and fine < coarse / 3.0


def test_overlapping_rects_raster_area_and_union():
    grid = H.GridConfig(32, 2.0)
    t = H.rasterize_rects([(0, 0, 20, 20, 1.0), (5, 5, 20, 20, 1.0)], grid, 0.0).real
    assert abs(t.sum() * 4.0 - (400 + 400 - 15 * 15)) < 1e-9
    s = oracle.mask_spectrum(32, 2.0, [("rect", 0, 0, 20, 20), ("rect", 5, 5, 20, 20)], True)
    assert abs(s[0, 0] - t.sum()) < 1e-9
