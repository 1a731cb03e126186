"""Oracle polygon spectrum (orc_polygon_spectrum, mask.rs:1222-1352) pinned
against independent results: the closed-form rectangle spectrum, a concave
clipped polygon written as disjoint rectangles (mask.rs:2332-2386), and
Gauss-Legendre quadrature of the Fourier integral over triangles."""
import numpy as np
import pytest

import oracle


def _quad_triangle(n, pixel, tri, mx, my, order=48):
    """S[m] = N^2 c(m) e^{-i pi (mx+my)(N-1)/N} with c(m) = (1/L^2) integral over the
    triangle, by Gauss-Legendre on the Duffy map p = A + u (B - A) + u t (C - B)."""
    L = n * pixel
    g, w = np.polynomial.legendre.leggauss(order)
    g, w = 0.5 * (g + 1.0), 0.5 * w
    A, B, C = (np.asarray(p, float) for p in tri)
    det = abs((B - A)[0] * (C - B)[1] - (B - A)[1] * (C - B)[0])
    u, t = np.meshgrid(g, g, indexing="ij")
    wu, wt = np.meshgrid(w, w, indexing="ij")
    px = A[0] + u * (B - A)[0] + u * t * (C - B)[0]
    py = A[1] + u * (B - A)[1] + u * t * (C - B)[1]
    f = np.exp(-2j * np.pi * (mx * px + my * py) / L)
    integral = np.sum(f * u * wu * wt) * det
    return n * n * integral / L ** 2 * np.exp(-1j * np.pi * (mx + my) * (n - 1) / n)


def test_rectangle_as_polygon_matches_rect_both_orientations():
    n, px = 64, 4.0
    ref = oracle.mask_spectrum(n, px, [("rect", 1.4, 10.4, 39.4, 40.6)], dark_field=True)
    ccw = [(-18.3, -9.9), (21.1, -9.9), (21.1, 30.7), (-18.3, 30.7)]
    for verts in (ccw, ccw[::-1]):
        got = oracle.polygon_spectrum(n, px, verts)
        assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-11


def test_concave_clipped_polygon_matches_two_rects():
    n, px = 64, 4.0
    l_shape = [(-40.0, -30.0), (150.0, -30.0), (150.0, 10.0), (-10.0, 10.0), (-10.0, 50.0),
               (-40.0, 50.0)]
    got = oracle.polygon_spectrum(n, px, l_shape)
    assert abs(got[0, 0].real - 4096.0 * (168.0 * 40.0 + 30.0 * 40.0) / 65536.0) < 1e-9
    ref = oracle.mask_spectrum(n, px, [("rect", 44.0, -10.0, 168.0, 40.0),
                                       ("rect", -25.0, 30.0, 30.0, 40.0)], dark_field=True)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-10


@pytest.mark.parametrize("tri", [[(-30.0, -30.0), (30.0, -30.0), (-30.0, 30.0)],
                                 [(5.5, -41.0), (37.2, 12.9), (-22.4, 33.3)]])
def test_triangle_matches_quadrature(tri):
    n, px = 64, 4.0
    got = oracle.polygon_spectrum(n, px, tri, value=0.7 - 0.2j)
    worst = 0.0
    for mx, my in [(0, 0), (1, 0), (0, 3), (2, -5), (-7, 4), (11, 9), (-16, -13), (31, 1)]:
        q = (0.7 - 0.2j) * _quad_triangle(n, px, tri, mx, my)
        worst = max(worst, abs(got[my % n, mx % n] - q) / (n * n))
    assert worst < 1e-12, worst


def test_bow_tie_is_rejected_and_spectra_add():
    n, px = 32, 4.0
    with pytest.raises(RuntimeError):
        oracle.polygon_spectrum(n, px, [(-40.0, -20.0), (40.0, 20.0), (40.0, -20.0), (-40.0, 20.0)])
    a = oracle.polygon_spectrum(n, px, [(-10.0, -10.0), (10.0, -10.0), (0.0, 12.0)])
    both = oracle.polygon_spectrum(n, px, [(-10.0, -10.0), (10.0, -10.0), (0.0, 12.0)], 2.0,
                                   spectrum=a)
    assert np.abs(both - 3.0 * a).max() < 1e-12 * n * n
