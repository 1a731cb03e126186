"""Adaptive source sampling (aerial.rs:941) against the reference's own
invariant tests (aerial.rs:2220-2447)."""
import math

import numpy as np
import pytest

import highuvlith_amd as H
from highuvlith_amd import IlluminationShape as I


def pts_of(shape, ppa=None):
    return H.sample_source(shape, ppa, 0.5)


def d4_closed(p):
    s = set(map(tuple, p[:, :2].tolist()))
    return all((-a, b) in s and (a, -b) in s and (b, a) in s for a, b in s)


def test_conventional_is_symmetric_and_normalized():
    p = pts_of(I.Conventional(0.6))
    assert abs(p[:, 2].sum() - 1.0) < 1e-12
    assert len(p) > 500
    smax = np.hypot(p[:, 0], p[:, 1]).max()
    assert 0.57 < smax <= 0.6
    assert d4_closed(p)


def test_small_gaussian_is_resolved():
    p = pts_of(I.CoherentGaussian(0.02))
    assert len(p) >= 150
    m2 = (p[:, 2] * (p[:, 0] ** 2 + p[:, 1] ** 2)).sum()
    assert abs(m2 / (2 * 0.02 ** 2) - 1) < 0.01
    assert np.hypot(p[:, 0], p[:, 1]).max() < 0.2


def test_coherent_limit_and_override():
    # Some(1) collapses to the centroid; Some(11) puts 11 points across the box
    one = pts_of(I.Conventional(0.7), 1)
    assert len(one) == 1 and abs(one[0, 0]) < 1e-12 and abs(one[0, 1]) < 1e-12
    p = pts_of(I.Conventional(0.7), 11)
    assert len(set(p[:, 0].tolist())) == 11


def test_dipole_points_inside_poles_and_dark_source_rejected():
    p = pts_of(I.Dipole(0.7, 0.08, 0.0))
    assert len(p) >= 150
    assert np.all(np.hypot(np.abs(p[:, 0]) - 0.7, p[:, 1]) <= 0.08 + 1e-12)
    with pytest.raises(H.HuvError) as e:
        pts_of(I.Conventional(0.0) if False else I.Annular(3.0, 3.5))
    assert e.value.kind == "InvalidParameter"


@pytest.mark.parametrize("shape", [I.Annular(0.5, 0.9), I.Quadrupole(0.7, 0.2, 30.0),
                                   I.CoherentGaussian(0.05), I.Conventional(0.05),
                                   I.SectorPoles(0.7, 0.9, 30.0, 45.0, 4)])
def test_symmetric_fills_are_exactly_symmetric(shape):
    p = pts_of(shape)
    assert d4_closed(p)
    assert abs(p[:, 2].sum() - 1.0) < 1e-12


def test_dipole_box_is_d2_only():
    for shape in (I.Dipole(0.7, 0.15, 0.0), I.SectorPoles(0.7, 0.9, 35.0, 0.0, 2)):
        p = pts_of(shape)
        s = set(map(tuple, p[:, :2].tolist()))
        assert all((-a, b) in s and (a, -b) in s for a, b in s)


def test_documented_point_counts():
    # aerial.rs:931-933 typical sizes
    assert abs(len(pts_of(I.Annular(0.5, 0.9))) - 1310) < 30
    assert abs(len(pts_of(I.Conventional(0.7))) - 1180) < 30
    assert abs(len(pts_of(I.Dipole(0.7, 0.15, 0.0))) - 260) < 15


def test_tracks_a_moving_straight_edge():
    s = 0.75
    p = pts_of(I.Conventional(s))
    worst = 0.0
    for k in range(61):
        a = 0.60 + 0.0025 * k
        exact = (s * s * math.acos(a / s) - a * math.sqrt(s * s - a * a)) / (math.pi * s * s)
        worst = max(worst, abs(p[p[:, 0] > a, 2].sum() - exact))
    assert worst < 2.5e-3
    sg = 0.05
    p = pts_of(I.CoherentGaussian(sg))

    def tail(a):
        b, n = a + 12 * sg, 2000
        h = (b - a) / n
        x = a + np.arange(n + 1) * h
        c = np.where((np.arange(n + 1) % 2) == 1, 4.0, 2.0)
        c[0] = c[-1] = 1.0
        return (c * np.exp(-x * x / (2 * sg * sg))).sum() * h / 3 / (math.sqrt(2 * math.pi) * sg)
    worst = max(abs(p[p[:, 0] > a, 2].sum() - tail(a)) for a in -0.15 + 0.005 * np.arange(61))
    assert worst < 1e-2


def test_bounds_point_weight_of_peaked_fills():
    for sg in (0.02, 0.05, 0.3):
        p = pts_of(I.CoherentGaussian(sg))
        assert p[:, 2].max() <= 1.0 / 200.0
        u = 1.0 / (2 * sg * sg)
        exact = 2 * sg * sg * (1 - u * math.exp(-u) / (1 - math.exp(-u)))
        m2 = (p[:, 2] * (p[:, 0] ** 2 + p[:, 1] ** 2)).sum()
        assert abs(m2 / exact - 1) < 0.01
    p = pts_of(I.Conventional(0.05))
    assert 180 <= len(p) <= 320
    m2 = (p[:, 2] * (p[:, 0] ** 2 + p[:, 1] ** 2)).sum()
    assert abs(m2 / (0.05 ** 2 / 2) - 1) < 0.01
