"""Host FP64 kernel-set construction (aerial.rs:992-1317) pinned by an
independent Abbe sum, closed forms and the reference's own unit tests
(tests/imaging_validation.rs, tests/analytical_validation.rs, aerial.rs tests)."""
import math

import numpy as np
import pytest

import oracle
import highuvlith_amd as H
from highuvlith_amd import IlluminationShape as I


def conv07(x, y):
    return 1.0 if math.hypot(x, y) <= 0.7 else 0.0


PATTERN_2D = [("rect", -64.0, 16.0, 96.0, 224.0), ("rect", 24.0, -80.0, 176.0, 64.0),
              ("rect", 120.0, 120.0, 48.0, 48.0)]


def socs_abs(ks, spec):
    return oracle.socs_intensity(ks.n, ks.support, ks.nonzero_rows, ks.eigenvalues, ks.values,
                                 spec)


def test_select_kernel_count_fixtures():
    # aerial.rs:2542-2555
    vals = [5.0, 3.0, 1.0, 0.5, 1e-14]
    assert H.select_kernel_count(vals, 9.5, 10, 1.0) == 4
    assert H.select_kernel_count(vals, 9.5, 2, 1.0) == 2
    assert H.select_kernel_count(vals, 9.5, 10, 0.8) == 2
    assert H.select_kernel_count(vals, 9.5, 10, 0.95) == 4
    assert H.select_kernel_count([0.0], 0.0, 10, 1.0) == 0
    pair = [4.0, 2.0, 2.0 * (1 + 1e-9), 1.0]
    assert H.select_kernel_count(pair, 9.0, 2, 1.0) == 1
    assert H.select_kernel_count(pair, 9.0, 10, 0.6) == 3
    assert H.select_kernel_count([3.0, 3.0, 1.0], 7.0, 1, 1.0) == 2


@pytest.mark.parametrize("case", ["conventional", "dipole_gram", "annular_psm"])
def test_socs_with_all_kernels_equals_abbe(case):
    # imaging_validation.rs:200-290: |SOCS - Abbe| / peak <= 1e-8
    if case == "conventional":
        grid, na, ill, ppa, zs = H.GridConfig(64, 8.0), 0.5, I.Conventional(0.7), 13, (0.0, 250.0, -400.0)
        feats, dark, absorber = PATTERN_2D, True, 0j
    elif case == "dipole_gram":
        grid, na, ill, ppa, zs = H.GridConfig(128, 9.375), 0.5, I.Dipole(2.0 / 3.0, 0.05, 0.0), 40, (0.0, 300.0)
        feats = [("rect", -600.0 + (k + 0.5) * 150.0, 0.0, 75.0, 1200.0) for k in range(8)]
        dark, absorber = False, 0j
    else:
        grid, na, ill, ppa, zs = H.GridConfig(64, 8.0), 0.6, I.Annular(0.5, 0.9), 15, (0.0, 150.0)
        feats, dark = PATTERN_2D, True
        absorber = math.sqrt(0.06) * complex(math.cos(math.pi), math.sin(math.pi))
    opt = H.ProjectionOptics(na, flare_fraction=0.0)
    pts = H.sample_source(ill, ppa, grid.freq_step() / (na / 100.0))
    spec = oracle.mask_spectrum(grid.size, grid.pixel_nm, feats, dark, absorber)
    for z in zs:
        ks = H.build_kernel_set(opt, pts, grid, z, 100.0, 100000)
        if case == "dipole_gram":
            assert ks.method == "dense_gram" and len(ks.support) > ks.num_columns
        abbe = oracle.abbe_image(grid.size, grid.pixel_nm, spec, pts, 0, na, 1.0, 0.0, 1.0, z, 100.0)
        err = np.abs(socs_abs(ks, spec) - abbe).max() / abbe.max()
        assert err <= 1e-8, (case, z, err)


def test_euv_optics_socs_equals_abbe_with_obscuration():
    grid = H.GridConfig(64, 2.0)
    opt = H.EuvProjectionOptics(0.55, central_obscuration=0.2)
    pts = H.sample_source(I.Annular(0.3, 0.8), 9, 0.5)
    feats = [("rect", -10.0, 6.0, 22.0, 40.0), ("rect", 20.0, -14.0, 30.0, 12.0)]
    spec = oracle.mask_spectrum(64, 2.0, feats, True)
    for z in (0.0, 40.0):
        ks = H.build_kernel_set(opt, pts, grid, z, 13.5, 100000)
        abbe = oracle.abbe_image(64, 2.0, spec, pts, 1, 0.55, 1.0, 0.2, 1.0, z, 13.5)
        assert np.abs(socs_abs(ks, spec) - abbe).max() / abbe.max() <= 1e-8


def test_clear_field_is_unity_and_truncation_bound():
    # imaging_validation.rs:293-327
    grid = H.GridConfig(32, 12.5)
    opt = H.ProjectionOptics(0.5, flare_fraction=0.0)
    spec = oracle.mask_spectrum(32, 12.5, [], False)
    for ill in (I.Conventional(0.7), I.Annular(0.5, 0.9), I.Dipole(2.0 / 3.0, 0.05, 0.0)):
        pts = H.sample_source(ill, None, grid.freq_step() / (0.5 / 100.0))
        for z in (0.0, 200.0, -1000.0):
            exact = H.build_kernel_set(opt, pts, grid, z, 100.0, 100000)
            img = oracle.socs_image(exact, spec, 0.0)
            assert np.abs(img - 1.0).max() < 1e-12
            trunc = H.build_kernel_set(opt, pts, grid, z, 100.0, 6)
            img = oracle.socs_image(trunc, spec, 0.0)
            kept = (trunc.eigenvalues / trunc.normalization_divisor).sum()
            dropped = kept / trunc.captured_energy_fraction - kept
            assert np.all(img <= 1 + 1e-12) and np.all(1 - img <= dropped + 1e-12)


def test_all_decomposition_paths_agree():
    # aerial.rs:2623-2646
    grid = H.GridConfig(32, 12.5)
    opt = H.ProjectionOptics(0.5, flare_fraction=0.0)
    pts = H.sample_source(I.Annular(0.3, 0.8), 9, grid.freq_step() / (0.5 / 100.0))
    feats = [("rect", -40.0, 10.0, 75.0, 150.0), ("rect", 60.0, -50.0, 50.0, 50.0)]
    spec = oracle.mask_spectrum(32, 12.5, feats, True)
    for z in (0.0, 180.0):
        sets = {m: H.build_kernel_set(opt, pts, grid, z, 100.0, 1000, method=m)
                for m in ("dense_tcc", "dense_gram", "randomized")}
        tcc = sets["dense_tcc"]
        assert tcc.num_columns > len(tcc.support)
        rank = tcc.num_kernels
        l0 = tcc.eigenvalues[0]
        for name, tol in (("dense_gram", 1e-12), ("randomized", 1e-10)):
            o = sets[name]
            assert o.num_kernels == rank
            assert np.abs(o.eigenvalues - tcc.eigenvalues).max() < tol * l0
        assert abs(tcc.captured_energy_fraction - 1.0) < 1e-10
        ref = socs_abs(tcc, spec)
        for name in ("dense_gram", "randomized"):
            assert np.abs(socs_abs(sets[name], spec) - ref).max() < 1e-10


def test_truncated_randomized_matches_dense_leading_pairs():
    # aerial.rs:2649-2675
    grid = H.GridConfig(32, 12.5)
    opt = H.ProjectionOptics(0.5, flare_fraction=0.0)
    pts = H.sample_source(I.Annular(0.3, 0.8), 9, grid.freq_step() / (0.5 / 100.0))
    dense = H.build_kernel_set(opt, pts, grid, 90.0, 100.0, 6, method="dense_tcc")
    rand = H.build_kernel_set(opt, pts, grid, 90.0, 100.0, 6, method="randomized")
    full = H.build_kernel_set(opt, pts, grid, 90.0, 100.0, 1000, method="dense_tcc")
    assert rand.num_kernels == 6
    l = 6 + 16
    for k in range(6):
        bound = max((full.eigenvalues[l] / full.eigenvalues[k]) ** 7, 1e-13)
        assert abs(rand.eigenvalues[k] / dense.eigenvalues[k] - 1) <= 10 * bound
    spec = oracle.mask_spectrum(32, 12.5, [("rect", -40.0, 10.0, 75.0, 150.0)], True)
    assert np.abs(socs_abs(dense, spec) - socs_abs(rand, spec)).max() < 1e-5


def test_coherent_three_beam_closed_form():
    # analytical_validation.rs:228-293 (on a power-of-two field: pitch 320, 4 periods)
    wl, na, pitch, pixel, periods = 100.0, 0.5, 320.0, 10.0, 4
    n = int(pitch * periods / pixel)
    grid = H.GridConfig(n, pixel)
    field = pitch * periods
    feats = [("rect", -field / 2 + (k + 0.5) * pitch, 0.0, pitch / 2, field) for k in range(periods)]
    spec = oracle.mask_spectrum(n, pixel, feats, False)
    c0, c1, cm1 = spec[0, 0] / n ** 2, spec[0, periods] / n ** 2, spec[0, n - periods] / n ** 2
    opt = H.ProjectionOptics(na, flare_fraction=0.0)
    pts = H.sample_source(I.Conventional(0.01), 1, grid.freq_step() / (na / wl))
    for z in (0.0, 400.0, -900.0):
        s = wl / pitch
        ct = math.sqrt(max(1 - s * s, 0.0))
        phi = 2 * math.pi * z * s * s / ((1 + ct) * wl)
        ks = H.build_kernel_set(opt, pts, grid, z, wl, 48)
        img = oracle.socs_image(ks, spec, 0.0)
        x = 2 * math.pi * np.arange(n) / (n / periods)
        e = c0 + (c1 * np.exp(1j * x) + cm1 * np.exp(-1j * x)) * complex(math.cos(phi), math.sin(phi))
        assert np.abs(img[0] - np.abs(e) ** 2).max() < 1e-12
        assert np.abs(img[n // 3] - np.abs(e) ** 2).max() < 1e-12


def test_support_extends_to_one_plus_sigma():
    # aerial.rs:2345-2375: first order at 1.25 NA/lambda is inside (1 + sigma) NA/lambda
    grid = H.GridConfig(64, 10.0)
    eng = H.AerialImageEngine(H.Source(100.0, I.Conventional(0.7)), H.ProjectionOptics(0.5),
                              grid, max_kernels=40)
    dense = eng.kernels(0.0).kernels
    assert np.abs(dense[:, 0, 4]).max() > 1e-6 or np.abs(dense[:, 0, 60]).max() > 1e-6


def test_pupil_sampling_guard_and_settings_validation():
    grid = H.GridConfig(1024, 1.0)
    with pytest.raises(H.HuvError) as e:
        H.AerialImageEngine(H.Source(1.0, I.Conventional(0.5)), H.ProjectionOptics(0.3), grid, 20)
    assert e.value.kind == "PupilSamplingTooDense"
    small = H.GridConfig(32, 4.0)
    src, opt = H.Source.f2(0.5), H.ProjectionOptics(0.75)
    for bad in (H.ImagingSettings(max_kernels=0), H.ImagingSettings(kernel_energy_fraction=0.0),
                H.ImagingSettings(kernel_energy_fraction=float("nan")),
                H.ImagingSettings(kernel_energy_fraction=1.5),
                H.ImagingSettings(source_points_per_axis=0)):
        with pytest.raises(H.HuvError):
            H.AerialImageEngine(src, opt, small, settings=bad)


def test_config_kernel_sets_have_expected_shapes():
    # 256 x 4 nm ArF 193i annular 0.7/0.9, K <= 24: S = 577, randomized path
    eng = H.AerialImageEngine(H.Source.arf(I.Annular(0.7, 0.9)), H.ProjectionOptics.immersion_193i(),
                              H.GridConfig(256, 4.0), max_kernels=24)
    d = eng.kernel_diagnostics()
    assert d["num_frequencies"] == 577 and d["method"] == "randomized"
    assert d["num_kernels"] <= 24 and abs(d["clear_field_intensity"] - 1.0) < 1e-9
