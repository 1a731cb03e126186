"""Kernel sets built with the GPU's help (huv_kernels_build_device: factor,
Gram product and kernels on the device in FP64, eigendecomposition on host
threads) against the host build (KernelBuilder::build, aerial.rs:1041-1317):
same support, kernel count, clear field and pupil peak, eigenvalues within
1e-10 relative, and FP64 oracle images of the two sets within 1e-10 of the
clear field (eigenvector phases are free, so kernels are compared through
images)."""
import time

import numpy as np
import pytest

import oracle
import highuvlith_amd as H
from highuvlith_amd import engine as E
from highuvlith_amd import workloads as W

pytestmark = pytest.mark.gpu


def _compare(eng, host_sets, dev_sets, mask):
    spec = oracle.fft2(mask.astype(np.complex128))
    for h, d in zip(host_sets, dev_sets):
        assert d.method == h.method and d.num_kernels == h.num_kernels
        assert np.array_equal(d.support, h.support)
        assert np.array_equal(d.nonzero_rows, h.nonzero_rows)
        assert d.num_columns == h.num_columns
        assert abs(d.clear_field - h.clear_field) <= 1e-12 * h.clear_field
        assert abs(d.pupil_peak - h.pupil_peak) <= 1e-12 * h.pupil_peak
        rel = np.abs(d.eigenvalues - h.eigenvalues).max() / h.eigenvalues[0]
        ih = oracle.socs_image(h, spec, eng.flare_fraction(), threads=16)
        idev = oracle.socs_image(d, spec, eng.flare_fraction(), threads=16)
        err = float(np.abs(ih - idev).max())
        print("z", h.defocus_nm, h.method, "K", h.num_kernels, "eig rel", rel, "image diff", err)
        assert rel <= 1e-10
        assert err <= 1e-10


def test_device_build_matches_host_dense_gram():
    """The batched EUV workload's five focus planes (1024^2 @ 1 nm, x-dipole,
    K = 32, dense Gram path)."""
    src, opt, grid, K = W.euv_dipole_1024_setup()
    eng = H.AerialImageEngine(src, opt, grid, max_kernels=K)
    zs = list(W.EUV_FOCUS_PLANES_NM)
    t = time.perf_counter()
    host = [E.build_kernel_set(opt, eng.source_points, grid, z, src.wavelength_nm, K)
            for z in zs]
    t_host = time.perf_counter() - t
    ctx = eng.context()
    E.build_kernel_sets_device(ctx, opt, eng.source_points, grid, zs[:1], src.wavelength_nm, K)
    t = time.perf_counter()
    dev = E.build_kernel_sets_device(ctx, opt, eng.source_points, grid, zs, src.wavelength_nm, K)
    t_dev = time.perf_counter() - t
    print(f"host build {t_host:.3f} s, device build {t_dev:.3f} s for {len(zs)} planes")
    mask = W.manhattan_clips(1, grid.size, seed=12, clear_fraction=0.4)[0].cpu().numpy()
    _compare(eng, host[:2] + host[-1:], dev[:2] + dev[-1:], mask)
    # the engine's cache takes device-built planes
    sets = eng.build_planes_device(zs)
    assert [s.defocus_nm for s in sets] == zs


def test_device_build_matches_host_dense_tcc():
    """A set whose factor has fewer candidate frequencies than source points
    (dense TCC path, ArF annular 0.7/0.9 on 256^2 @ 4 nm)."""
    src, opt, grid, _ = W.arf_lines_256_setup()
    eng = H.AerialImageEngine(src, opt, grid, max_kernels=24)
    zs = [0.0, 45.0]
    host = [E.build_kernel_set(opt, eng.source_points, grid, z, src.wavelength_nm, 24,
                               method="dense_tcc") for z in zs]
    dev = E.build_kernel_sets_device(eng.context(), opt, eng.source_points, grid, zs,
                                     src.wavelength_nm, 24, method="dense_tcc")
    rng = np.random.default_rng(3)
    mask = (rng.random((256, 256)) > 0.6).astype(np.float64)
    _compare(eng, host, dev, mask)


def test_device_build_refuses_the_randomized_path():
    src, opt, grid, _ = W.arf_lines_256_setup()
    eng = H.AerialImageEngine(src, opt, grid, max_kernels=24)
    with pytest.raises(H.HuvError) as e:
        E.build_kernel_sets_device(eng.context(), opt, eng.source_points, grid, [0.0],
                                   src.wavelength_nm, 24)
    assert e.value.kind == "Unsupported"
