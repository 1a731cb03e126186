"""The FP64 U-Net oracle (oracle/unet_ref.py) against central finite
differences and BF16 rounding fixtures, so it can serve as the checker of the
device U-Net."""
import numpy as np

from oracle import unet_ref as U


def test_bf16_rounding_fixtures():
    # BF16 keeps 8 significant bits: the spacing above 1.0 is 2^-7.
    x = np.array([1.0, 1.00390625, 1.001953125, 1.005859375, 1.01171875, -3.14159265, 65504.0])
    got = U.bf16(x)
    assert got[0] == 1.0
    assert got[1] == 1.0            # 1 + 2^-8: tie, to even
    assert got[2] == 1.0            # below half a spacing
    assert got[3] == 1.0078125      # 0.75 of a spacing: up
    assert got[4] == 1.015625       # 1 + 3 * 2^-8: tie, to even (1 + 2^-6)
    assert got[5] == -3.140625 and got[6] == 65536.0


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    lay = U.Layout(levels=2, base=16)
    p = rng.normal(0, 0.2, lay.n)
    x = rng.random((1, 32, 32))            # continuous input: no pooling ties
    t = rng.normal(0, 1, (1, 32, 32))
    loss, g = U.loss_grad(p, lay, x, t)
    idx = np.concatenate([rng.choice(lay.n - 17, 24, replace=False), [lay.head_w + 3, lay.head_b]])
    h = 1e-6
    worst = 0.0
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = U.loss_grad(p + e, lay, x, t)
        lm, _ = U.loss_grad(p - e, lay, x, t)
        fd = (lp - lm) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / max(abs(fd), 1e-3 * np.abs(g).max()))
    assert worst < 1e-5, worst


def test_pool_gradient_goes_to_the_first_maximum():
    x = np.zeros((1, 2, 2, 1))
    dy = np.ones((1, 1, 1, 1))
    dx = U._pool_bwd(x, dy)
    assert dx[0, 0, 0, 0] == 1.0 and dx.sum() == 1.0
