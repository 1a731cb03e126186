"""Reference mask features (rectangles, line/space gratings, rectangle arrays)
as painted rectangles for the GPU rectangle path (engine.mask_feature_rects).

The oracle's exact spectrum (orc_mask_spectrum, mask.rs:616-657) of the
original features must equal the oracle's exact spectrum of the rectangles
they expand to: commensurate gratings (periodic closed form), incommensurate
ones (truncated at the field edge), lines straddling the field edge (split
into two clipped pieces), both orientations, and arrays."""
import numpy as np
import pytest

import oracle
import highuvlith_amd as H
from highuvlith_amd import engine as E

CASES = [
    [("lines", 30.0, 64.0, 0, 5.0)],                      # commensurate, split at the edge
    [("lines", 45.0, 70.0, 1, -12.5)],                    # incommensurate, horizontal lines
    [("array", 10.0, 12.0, 40.0, 50.0, 0.0, 3.0)],        # rectangle array
    [("array", 17.3, 9.1, 32.0, 64.0, 4.4, -7.0), ("rect", 20.0, -30.0, 60.0, 25.0)],
    [("lines", 64.0, 64.0, 0, 0.0)],                      # width >= pitch: clear field
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("dark_field", [True, False])
def test_expansion_has_the_features_spectrum(case, dark_field):
    n, px = 64, 4.0
    feats = CASES[case]
    rects, bg = E.mask_feature_rects(feats, H.GridConfig(n, px), dark_field=dark_field)
    ref = oracle.mask_spectrum(n, px, feats, dark_field=dark_field)
    fg = 1.0 if dark_field else 0.0
    assert all(abs(r[4] - fg) < 1e-15 for r in rects)
    assert abs(bg - (0.0 if dark_field else 1.0)) < 1e-15
    got = oracle.mask_spectrum(n, px, [("rect",) + tuple(r[:4]) for r in rects],
                               dark_field=dark_field)
    err = np.abs(got - ref).max() / n ** 2
    assert err < 1e-12, err


def test_many_periods_become_the_mean_and_unknown_kinds_are_errors():
    g = H.GridConfig(64, 4.0)
    rects, _ = E.mask_feature_rects([("lines", 1e-4, 2e-4, 0, 0.0)], g, dark_field=True)
    assert len(rects) == 1 and abs(rects[0][4] - 0.5) < 1e-12 and rects[0][2] == 256.0
    with pytest.raises(H.HuvError):
        E.mask_feature_rects([("polygon", [(0, 0), (1, 0), (0, 1)])], g)


def test_feature_lists_split_polygons_and_pack_offsets():
    """Host packing for the C ABI: polygons are split from the rectangle-type
    features in order, and the polygon / vertex offsets index the flat arrays."""
    g = H.GridConfig(64, 4.0)
    feats = [("rect", 0.0, 0.0, 10.0, 10.0), ("polygon", [(0, 0), (5, 0), (0, 5)]),
             ("lines", 30.0, 64.0, 0, 5.0), ("polygon", [(1, 1), (9, 1), (9, 9), (1, 9)])]
    rects, polys, bg = E.mask_feature_lists(feats, g, dark_field=True, absorber=0.1j)
    assert bg == 0.1j and len(polys) == 2
    assert [len(v) for v, _ in polys] == [3, 4] and all(val == 1.0 for _, val in polys)
    assert len(rects) == 1 + 5  # the rectangle and the five clipped line pieces
    pa = E.PolygonArrays([polys, [], [polys[1]]])
    assert list(pa.poly_offsets) == [0, 2, 2, 3]
    assert list(pa.vertex_offsets) == [0, 3, 7, 11]
    assert pa.vertices.shape == (11, 2) and pa.values.shape == (3, 2)
    assert np.array_equal(pa.vertices[7:11], np.asarray(polys[1][0], float))
    empty = E.PolygonArrays([[], []])
    assert list(empty.poly_offsets) == [0, 0, 0] and list(empty.vertex_offsets) == [0]


def test_gray_transmittance_range():
    """Gray rectangles: transmittance T in [0, 1] gives amplitude sqrt(T); a
    value outside [0, 1] is an invalid parameter (mask.rs:360) and a non-finite
    one is dropped as the reference's spectrum drops it (mask.rs:1009-1011)."""
    g = H.GridConfig(64, 4.0)
    rects, _ = E.mask_feature_rects([("gray", 0.0, 0.0, 20.0, 20.0, 0.36)], g, dark_field=True)
    assert len(rects) == 1 and abs(rects[0][4] - 0.6) < 1e-15
    rects, _ = E.mask_feature_rects([("gray", 0.0, 0.0, 20.0, 20.0, float("nan")),
                                     ("rect", 5.0, 5.0, 4.0, 4.0)], g, dark_field=True)
    assert len(rects) == 1 and rects[0][4] == 1.0
    for bad in (-0.1, 1.5):
        with pytest.raises(H.HuvError):
            E.mask_feature_rects([("gray", 0.0, 0.0, 20.0, 20.0, bad)], g)


def test_polygon_lists_must_match_the_clips():
    g = H.GridConfig(64, 4.0)
    with pytest.raises(H.HuvError):
        E.band_spectrum_features_device(None, [[], []], [[]], g, 8)
