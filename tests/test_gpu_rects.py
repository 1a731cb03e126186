"""Exact spectrum of rectangle masks on the GPU (mask.rs:616-657, 1365-1502)
against the FP64 oracle (orc_mask_spectrum), and images from it against the
oracle's SOCS image of the same spectrum.

Gates: band coefficients within 2e-6 of n^2 (the clear-field DC value; the
device stores FP32), images within 1e-5 of the clear-field intensity.  At
full size the pixel-aligned identity S = FFT(raster) sinc(pi mx/n) sinc(pi my/n)
(mask.rs:38) ties the analytic path to the FFT band path.
"""
import numpy as np
import pytest

import oracle
import highuvlith_amd as H

pytestmark = pytest.mark.gpu
SPEC_TOL = 2e-6
IMG_TOL = 1e-5


def _random_rects(rng, L, count, value):
    """Sub-pixel rectangles (centre, size) partly outside the field, overlapping."""
    out = []
    for _ in range(count):
        w, h = rng.uniform(0.02 * L, 0.3 * L, 2)
        x, y = rng.uniform(-0.6 * L, 0.6 * L, 2)
        out.append((float(x), float(y), float(w), float(h), value))
    return out


def _band(spec, band):
    n = spec.shape[0]
    ks = np.arange(-band, band + 1)
    return spec[np.ix_(ks % n, ks % n)].T  # [kx][ky]


@pytest.mark.parametrize("dark_field,absorber", [(True, 0j), (False, 0j),
                                                 (True, 0.245 * np.exp(1j * np.pi))])
def test_band_matches_oracle_spectrum(gpu_ctx, dark_field, absorber):
    rng = np.random.default_rng(11)
    n, px, band = 256, 1.37, 40
    L = n * px
    grid = H.GridConfig(n, px)
    fg, bg = (1.0, absorber) if dark_field else (absorber, 1.0)
    clips = [_random_rects(rng, L, k, fg) for k in (1, 5, 17)] + [[]]
    got = H.engine.band_spectrum_rects_device(gpu_ctx, clips, grid, band, backgrounds=bg)
    got = got.cpu().numpy()
    for c, rects in enumerate(clips):
        feats = [("rect", x, y, w, h) for (x, y, w, h, _) in rects]
        ref = _band(oracle.mask_spectrum(n, px, feats, dark_field=dark_field, absorber=absorber),
                    band)
        err = np.abs(got[c] - ref).max() / n ** 2
        print(c, len(rects), err)
        assert err <= SPEC_TOL, (c, err)


def test_painter_rule_and_complex_values(gpu_ctx):
    """Later rectangles win; values may be complex; the oracle's own overlap
    resolution is fed the same picture as disjoint pieces."""
    n, px, band = 128, 2.0, 30
    grid = H.GridConfig(n, px)
    v1, v2 = 0.5 + 0.25j, -0.3 + 0.8j
    clips = [[(0.0, 0.0, 60.0, 40.0, v1), (10.0, 0.0, 20.0, 100.0, v2)]]
    got = H.engine.band_spectrum_rects_device(gpu_ctx, clips, grid, band, backgrounds=0.0)
    got = got.cpu().numpy()[0]
    # disjoint pieces: v1 on [-30, 0) and [20, 30) over y in [-20, 20); v2 on [0, 20) x [-50, 50)
    s1 = oracle.mask_spectrum(n, px, [("rect", -15.0, 0.0, 30.0, 40.0),
                                      ("rect", 25.0, 0.0, 10.0, 40.0)], dark_field=True)
    s2 = oracle.mask_spectrum(n, px, [("rect", 10.0, 0.0, 20.0, 100.0)], dark_field=True)
    ref = _band(v1 * s1 + v2 * s2, band)
    err = np.abs(got - ref).max() / n ** 2
    print("painter", err)
    assert err <= SPEC_TOL


def test_pixel_aligned_equals_fft_band_with_pixel_transfer(gpu_ctx):
    """Full size (n = 1024): rectangles on pixel edges, analytic band ==
    FFT band of the raster times the pixel transfer function."""
    import torch
    rng = np.random.default_rng(5)
    n, px, band = 1024, 1.0, 47
    grid = H.GridConfig(n, px)
    clips, masks = [], np.zeros((3, n, n), np.float32)
    for c in range(3):
        rects = []
        for _ in range(40):
            w, h = rng.integers(4, 120, 2)
            x0, y0 = rng.integers(0, n - w), rng.integers(0, n - h)
            masks[c, y0:y0 + h, x0:x0 + w] = 1.0
            rects.append((x0 - n / 2 + w / 2, y0 - n / 2 + h / 2, float(w), float(h), 1.0))
        clips.append(rects)
    got = H.engine.band_spectrum_rects_device(gpu_ctx, clips, grid, band).cpu().numpy()
    ref = H.engine.band_spectrum_device(gpu_ctx, torch.from_numpy(masks).cuda(), band,
                                        pixel_transfer=True).cpu().numpy()
    err = np.abs(got - ref).max() / n ** 2
    print("aligned", err)
    assert err <= 4e-6


def test_images_match_oracle():
    rng = np.random.default_rng(8)
    n, px = 256, 1.6
    grid = H.GridConfig(n, px)
    eng = H.AerialImageEngine(H.Source.arf(H.IlluminationShape.Annular(0.7, 0.9)),
                              H.ProjectionOptics.immersion_193i(), grid, max_kernels=24)
    L = n * px
    clips = [_random_rects(rng, L, k, 1.0) for k in (3, 9)]
    zs = (0.0, 40.0)
    got = eng.compute_rects(clips, zs, backgrounds=0.0).cpu().numpy()
    for c, rects in enumerate(clips):
        spec = oracle.mask_spectrum(n, px, [("rect", x, y, w, h) for (x, y, w, h, _) in rects],
                                    dark_field=True)
        for p, z in enumerate(zs):
            ref = oracle.socs_image(eng.kernel_set(z), spec, eng.flare_fraction(), threads=8)
            err = np.abs(got[c, p] - ref).max()
            print(c, z, err)
            assert err <= IMG_TOL, (c, z, err)


def test_clear_and_empty_clips():
    """A clip with no rectangles over a clear background images to the
    (truncated) clear field, uniform; over a dark background to zero."""
    n, px = 512, 2.0
    eng = H.AerialImageEngine(H.Source.arf(H.IlluminationShape.Annular(0.7, 0.9)),
                              H.ProjectionOptics.immersion_193i(), H.GridConfig(n, px),
                              max_kernels=16)
    got = eng.compute_rects([[], []], (0.0,), backgrounds=[1.0, 0.0]).cpu().numpy()
    spec = np.zeros((n, n), np.complex128)
    spec[0, 0] = n * n
    ref = oracle.socs_image(eng.kernel_set(0.0), spec, eng.flare_fraction(), threads=8)
    assert np.abs(got[0, 0] - ref).max() <= IMG_TOL
    assert got[0, 0].std() < 1e-5
    assert np.abs(got[1, 0]).max() <= IMG_TOL


def test_invalid_rectangle_batches_are_rejected(gpu_ctx):
    import ctypes as C
    import torch
    n, px = 256, 1.0
    grid = H.GridConfig(n, px)
    with pytest.raises(H.HuvError):
        H.engine.band_spectrum_rects_device(gpu_ctx, [[(0.0, 0.0, np.inf, 1.0, 1.0)]], grid, 10)
    with pytest.raises(H.HuvError):
        H.engine.band_spectrum_rects_device(gpu_ctx, [[]], grid, 200)  # 2B + 1 > n
    # decreasing offsets
    off = np.array([0, 2, 1], np.int64)
    rects = np.zeros((2, 6))
    bg = np.zeros((2, 2))
    out = torch.zeros((2, 9, 9, 2), device="cuda")
    st = gpu_ctx.lib.huv_band_spectrum_rects(
        gpu_ctx.handle, 2, n, px, 4, off.ctypes.data_as(C.POINTER(C.c_int64)),
        rects.ctypes.data_as(C.POINTER(C.c_double)), bg.ctypes.data_as(C.POINTER(C.c_double)),
        C.c_void_p(out.data_ptr()))
    assert st != 0
    # an edge grid beyond 2^20 cells
    many = [(float(i) * 0.2 - 100.0, float(i) * 0.2 - 100.0, 0.1, 0.1, 1.0) for i in range(1100)]
    with pytest.raises(H.HuvError):
        H.engine.band_spectrum_rects_device(gpu_ctx, [many], grid, 10)


def test_gratings_and_arrays_match_oracle():
    """Line/space gratings and rectangle arrays (mask.rs periodic primitives),
    commensurate or not, mixed with rectangles: band and images from the
    expanded rectangles against the oracle's exact spectrum of the features."""
    n, px, band = 256, 1.25, 40
    grid = H.GridConfig(n, px)
    masks = [[("lines", 37.5, 80.0, 0, 3.3)],
             [("lines", 22.0, 61.7, 1, -8.0), ("rect", 10.0, 20.0, 55.5, 33.3)],
             [("array", 14.2, 18.9, 40.0, 64.0, 2.5, -6.0)]]
    clips, bgs = [], []
    for f in masks:
        r, b = H.engine.mask_feature_rects(f, grid, dark_field=True, absorber=0.2j)
        clips.append(r)
        bgs.append(b)
    ctx = H.DeviceContext(0)
    got = H.engine.band_spectrum_rects_device(ctx, clips, grid, band, backgrounds=bgs)
    got = got.cpu().numpy()
    specs = [oracle.mask_spectrum(n, px, f, dark_field=True, absorber=0.2j) for f in masks]
    for c in range(len(masks)):
        err = np.abs(got[c] - _band(specs[c], band)).max() / n ** 2
        print("grating band", c, err)
        assert err <= SPEC_TOL, (c, err)
    eng = H.AerialImageEngine(H.Source.arf(H.IlluminationShape.Annular(0.7, 0.9)),
                              H.ProjectionOptics.immersion_193i(), grid, max_kernels=16)
    img = eng.compute_features(masks, (0.0,), dark_field=True, absorber=0.2j).cpu().numpy()
    for c in range(len(masks)):
        ref = oracle.socs_image(eng.kernel_set(0.0), specs[c], eng.flare_fraction(), threads=8)
        err = np.abs(img[c, 0] - ref).max()
        print("grating image", c, err)
        assert err <= IMG_TOL, (c, err)


def test_tuned_launch_shapes_give_the_same_images():
    """The rectangle path follows a launch shape set for its kernel sets:
    a larger coarse grid (wider band box) and one or two x lanes agree with
    the default shape."""
    rng = np.random.default_rng(21)
    n, px = 512, 2.0
    grid = H.GridConfig(n, px)
    eng = H.AerialImageEngine(H.Source.arf(H.IlluminationShape.Annular(0.7, 0.9)),
                              H.ProjectionOptics.immersion_193i(), grid, max_kernels=16)
    clips = [_random_rects(rng, n * px, k, 1.0) for k in (4, 12)]
    base = eng.compute_rects(clips, (0.0, 30.0)).cpu().numpy()
    ctx = eng.context()
    planes = [eng.device_kernel_set(z) for z in (0.0, 30.0)]
    B, K = planes[0].band, max(p.host.num_kernels for p in planes)
    for m, xl in [(128, 1), (128, 2), (256, 2)]:
        if m < 4 * B + 1:
            continue
        ctx.set_tuning(n, B, 2, K, m=m, xlanes=xl)
        got = eng.compute_rects(clips, (0.0, 30.0)).cpu().numpy()
        err = np.abs(got - base).max()
        print("shape", m, xl, err)
        assert err <= 2e-6, (m, xl, err)


def test_polygons_match_oracle():
    """Simple polygons (mask.rs:1286-1352: either orientation, concave,
    clipped at the field edge, complex amplitude) next to rectangles: band and
    images against the oracle's edge-sum spectrum (orc_polygon_spectrum)."""
    n, px, band = 256, 1.5, 40
    grid = H.GridConfig(n, px)
    tri = [(-60.0, -50.0), (20.0, -45.0), (-30.0, 40.0)]
    l_shape = [(40.0, -30.0), (230.0, -30.0), (230.0, 10.0), (70.0, 10.0), (70.0, 50.0),
               (40.0, 50.0)]                                   # clipped at x = 192
    cw = [(-150.0, 100.0), (-100.0, 100.0), (-100.0, 60.0), (-150.0, 60.0)][::-1]
    v2 = 0.4 - 0.7j
    rect_clips = [[(-120.0, -120.0, 30.0, 20.0, 1.0)], [], [(0.0, 150.0, 40.0, 20.0, 1.0)]]
    poly_clips = [[(tri, 1.0)], [(l_shape, v2), (cw, 1.0)], [(tri, 1.0), (l_shape, 1.0)]]
    ctx = H.DeviceContext(0)
    got = H.engine.band_spectrum_features_device(ctx, rect_clips, poly_clips, grid, band)
    got = got.cpu().numpy()
    specs = []
    for rects, polys in zip(rect_clips, poly_clips):
        s = oracle.mask_spectrum(n, px, [("rect",) + tuple(r[:4]) for r in rects], dark_field=True)
        for verts, v in polys:
            s = oracle.polygon_spectrum(n, px, verts, v, spectrum=s)
        specs.append(s)
    for c in range(3):
        err = np.abs(got[c] - _band(specs[c], band)).max() / n ** 2
        print("polygon band", c, err)
        assert err <= SPEC_TOL, (c, err)
    eng = H.AerialImageEngine(H.Source.arf(H.IlluminationShape.Annular(0.7, 0.9)),
                              H.ProjectionOptics.immersion_193i(), grid, max_kernels=16)
    masks = [[("polygon", tri), ("rect", -120.0, -120.0, 30.0, 20.0)],
             [("polygon", cw), ("polygon", l_shape)]]
    img = eng.compute_features(masks, (0.0,), dark_field=True).cpu().numpy()
    refs = [specs[0],
            oracle.polygon_spectrum(n, px, l_shape, 1.0, spectrum=oracle.polygon_spectrum(n, px, cw))]
    for c in range(2):
        ref = oracle.socs_image(eng.kernel_set(0.0), refs[c], eng.flare_fraction(), threads=8)
        err = np.abs(img[c, 0] - ref).max()
        print("polygon image", c, err)
        assert err <= IMG_TOL, (c, err)


def test_crossing_or_overlapping_polygons_are_rejected():
    n, px = 256, 1.0
    grid = H.GridConfig(n, px)
    ctx = H.DeviceContext(0)
    bow = [(-40.0, -20.0), (40.0, 20.0), (40.0, -20.0), (-40.0, 20.0)]
    with pytest.raises(H.HuvError):
        H.engine.band_spectrum_features_device(ctx, [[]], [[(bow, 1.0)]], grid, 10)
    tri = [(-30.0, -30.0), (30.0, -30.0), (-30.0, 30.0)]
    with pytest.raises(H.HuvError):
        H.engine.band_spectrum_features_device(ctx, [[(0.0, 0.0, 20.0, 20.0, 1.0)]],
                                               [[(tri, 1.0)]], grid, 10)
    circle = [(60.0 * np.cos(a), 60.0 * np.sin(a)) for a in np.linspace(0, 2 * np.pi, 9000,
                                                                           endpoint=False)]
    with pytest.raises(H.HuvError):  # more than 8192 vertices
        H.engine.band_spectrum_features_device(ctx, [[]], [[(circle, 1.0)]], grid, 10)


def test_polygons_sharing_an_edge_up_to_rounding_are_accepted():
    """Two polygons whose shared edge differs by a few ulps overlap by far less
    than 1e-9 of the field length, which the reference counts as no overlap
    (mask.rs:838-841,1099-1108): accepted, and the band equals the oracle's."""
    n, px, band = 256, 1.0, 30
    grid = H.GridConfig(n, px)
    xe = 10.0
    left = [(-40.0, -20.0), (xe, -20.0), (xe, 20.0), (-40.0, 25.0)]
    xr = float(np.nextafter(np.nextafter(xe, -1.0), -1.0))   # 2 ulps to the left
    right = [(xr, -20.0), (50.0, -25.0), (50.0, 20.0), (xr, 20.0)]
    ctx = H.DeviceContext(0)
    got = H.engine.band_spectrum_features_device(ctx, [[]], [[(left, 1.0), (right, 0.5)]],
                                                 grid, band).cpu().numpy()
    ref = oracle.polygon_spectrum(n, px, right, 0.5, spectrum=oracle.polygon_spectrum(n, px, left))
    err = np.abs(got[0] - _band(ref, band)).max() / n ** 2
    print("shared edge", err)
    assert err <= SPEC_TOL, err
    # a real overlap (1 nm) is still refused
    right2 = [(xe - 1.0, -20.0), (50.0, -25.0), (50.0, 20.0), (xe - 1.0, 20.0)]
    with pytest.raises(H.HuvError):
        H.engine.band_spectrum_features_device(ctx, [[]], [[(left, 1.0), (right2, 0.5)]],
                                               grid, band)


@pytest.mark.parametrize("band", [64, 100])
def test_polygon_band_wider_than_127_columns(band):
    """Band boxes of 129 and 201 columns (more than 1024 tiles of 4 x 4): the
    polygon pass deals a clip over more work-groups; the band must equal the
    oracle's edge-sum spectrum."""
    n, px = 512, 2.0
    grid = H.GridConfig(n, px)
    tri = [(-60.0, -50.0), (20.0, -45.0), (-30.0, 40.0)]
    hexa = [(150.0 + 40.0 * np.cos(a), -100.0 + 40.0 * np.sin(a)) for a in np.arange(6) * np.pi / 3]
    ctx = H.DeviceContext(0)
    got = H.engine.band_spectrum_features_device(
        ctx, [[(-200.0, 200.0, 50.0, 30.0, 1.0)], []], [[(tri, 1.0)], [(hexa, 0.3 + 0.2j)]],
        grid, band).cpu().numpy()
    s0 = oracle.polygon_spectrum(n, px, tri, 1.0, spectrum=oracle.mask_spectrum(
        n, px, [("rect", -200.0, 200.0, 50.0, 30.0)], dark_field=True))
    s1 = oracle.polygon_spectrum(n, px, hexa, 0.3 + 0.2j)
    for c, s in enumerate((s0, s1)):
        err = np.abs(got[c] - _band(s, band)).max() / n ** 2
        print("wide band", band, c, err)
        assert err <= SPEC_TOL, (band, c, err)
    with pytest.raises(H.HuvError):   # beyond the polygon pass's LDS tables
        H.engine.band_spectrum_features_device(ctx, [[]], [[(tri, 1.0)]], H.GridConfig(2048, 1.0),
                                               600)
