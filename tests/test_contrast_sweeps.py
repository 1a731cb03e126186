"""End-to-end pins of source sampling + kernel build + SOCS against the
reference's fine-quadrature contrast fixtures
(crates/highuvlith-core/tests/imaging_source_sampling.rs:69-131)."""
import numpy as np

import oracle
import highuvlith_amd as H
from highuvlith_amd import IlluminationShape as I


def ls_contrast(source, optics, pitch):
    grid = H.GridConfig(128, pitch / 64.0)
    eng = H.AerialImageEngine(source, optics, grid, settings=H.ImagingSettings())
    spec = oracle.mask_spectrum(128, grid.pixel_nm, [("lines", pitch / 2.0, pitch, 0, 0.0)], False)
    img = oracle.socs_image(eng.kernel_set(0.0), spec, optics.flare_fraction)
    mx, mn = img.max(), img.min()
    return (mx - mn) / (mx + mn) if mx + mn > 0 else 0.0


def check_sweep(got, ref, tol):
    got, ref = np.asarray(got), np.asarray(ref)
    assert np.abs(got - ref).max() <= tol, np.abs(got - ref).max()
    assert np.all(np.diff(got) >= -1e-9)


def test_krf_sigma_sweep_tracks_reference():
    REF = [0.0] * 13 + [0.00080, 0.00350, 0.00714, 0.01140, 0.01610, 0.02114, 0.02644, 0.03192,
                        0.03754, 0.04326, 0.04901, 0.05485, 0.06070]
    optics = H.ProjectionOptics(0.80)
    sig = [0.60 + 0.01 * k for k in range(26)]
    got = [ls_contrast(H.Source.krf(s), optics, 180.0) for s in sig]
    check_sweep(got, REF, 0.004)
    assert all(c < 1e-12 for c in got[:12])


def test_near_coherent_pitch_sweeps_track_reference():
    GAUSS = [0.00755, 0.01203, 0.01861, 0.02798, 0.04096, 0.05841, 0.08109, 0.10986, 0.14545,
             0.18806, 0.23767, 0.29389, 0.35594, 0.42298, 0.49358, 0.56631, 0.63990, 0.71259,
             0.78333, 0.85071, 0.90397, 0.91755, 0.92669, 0.93422, 0.94052]
    DISK = [0.0] * 12 + [0.00452, 0.08625, 0.20506, 0.33667, 0.46995, 0.59883, 0.72004, 0.83206,
                         0.91057, 0.92620, 0.93832, 0.94785, 0.95540]
    optics = H.ProjectionOptics(0.33, flare_fraction=0.0)
    pitches = [36.0 + 0.25 * k for k in range(25)]
    g = H.Source(13.5, I.CoherentGaussian(0.05))
    check_sweep([ls_contrast(g, optics, p) for p in pitches], GAUSS, 0.025)
    d = H.Source(13.5, I.Conventional(0.05))
    check_sweep([ls_contrast(d, optics, p) for p in pitches], DISK, 0.025)
