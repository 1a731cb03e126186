"""Seeded warm-start targets: exact area coverage of contacts and lines (CPU)."""
import numpy as np

from highuvlith_amd import warmstart as WS


def test_contact_and_line_targets_have_exact_areas():
    import torch
    n, px = 512, 2.0   # 1024 nm field: a 6 x 6 array at 140 nm pitch fits
    t = WS.contact_line_targets(6, n, px, seed=9, device="cpu").numpy()
    assert t.shape == (6, n, n) and t.dtype == np.float32
    assert t.min() >= 0.0 and t.max() <= 1.0 + 1e-6
    rng = np.random.default_rng(9)
    for k in range(6):
        if k % 2 == 0:
            cx, cy = rng.integers(2, 7, 2)
            side = rng.uniform(40.0, 60.0)
            rng.uniform(90.0, 140.0, 2)
            rng.uniform(-3, 3, (cy, cx))
            rng.uniform(-3, 3, (cy, cx))
            area = t[k].sum() * px * px
            assert abs(area - cx * cy * side * side) < 1e-6 * area
        else:
            cd = rng.uniform(40.0, 80.0)
            pitch = cd + rng.uniform(50.0, 120.0)
            rng.uniform(0.0, pitch)
            rng.random()
            duty = t[k].mean()
            # a 512 nm field holds a non-integer number of periods
            assert abs(duty - (pitch - cd) / pitch) < 0.5 * pitch / (n * px)
