"""Device U-Net (BF16 MFMA, FP32 accumulation) against the FP64 oracle with
BF16 emulation (oracle/unet_ref.py), training, and the ILT warm start."""
import numpy as np
import pytest

from oracle import unet_ref as U
import highuvlith_amd as H
from highuvlith_amd import unet as UN

pytestmark = pytest.mark.gpu


def _binary(n, count, seed):
    rng = np.random.default_rng(seed)
    m = np.zeros((count, n, n), np.float32)
    for c in range(count):
        for _ in range(12):
            w, h = rng.integers(4, n // 4, 2)
            x, y = rng.integers(0, n - w), rng.integers(0, n - h)
            m[c, y:y + h, x:x + w] = 1.0
    return m


def _blocks(lay):
    out = []
    for name in ("enc1", "enc2", "dec1", "dec2"):
        for l, s in enumerate(getattr(lay, name)):
            if s is None:
                continue
            (wo, ws), (bo, bs) = s[0], s[1]
            out.append((f"{name}[{l}].w", wo, int(np.prod(ws))))
            out.append((f"{name}[{l}].b", bo, int(np.prod(bs))))
    out.append(("head", lay.head_w, lay.base + 1))
    return out


@pytest.fixture(scope="module")
def ctx():
    return H.DeviceContext(0)


@pytest.mark.parametrize("base", [32, 16])
def test_forward_matches_oracle(ctx, base):
    import torch
    net = UN.UNet(ctx, levels=4, base=base, seed=1)
    lay = U.Layout(4, base)
    assert net.num_params == lay.n
    p = net.get_params().astype(np.float64)
    x = _binary(128, 2, 3)
    out = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    ref, _ = U.forward(p, lay, x.astype(np.float64), bf=True)
    err = np.abs(out - ref).max() / np.abs(ref).max()
    print("forward max rel err", err, "ref scale", np.abs(ref).max())
    assert err <= 1e-2


@pytest.mark.parametrize("base", [32, 16])
def test_gradient_matches_oracle(ctx, base):
    import torch
    net = UN.UNet(ctx, levels=4, base=base, seed=2)
    lay = U.Layout(4, base)
    p = net.get_params().astype(np.float64)
    x = _binary(128, 2, 5)
    t = (x * 1.5 - 0.7).astype(np.float32)
    loss, g = net.loss_and_grad(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda(),
                                want_grad=True)
    lref, gref = U.loss_grad(p, lay, x.astype(np.float64), t.astype(np.float64), bf=True)
    print("loss", loss, lref)
    assert abs(loss - lref) / lref <= 2e-3
    worst = 0.0
    for name, off, cnt in _blocks(lay):
        a, b = g[off:off + cnt].astype(np.float64), gref[off:off + cnt]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"{name:12s} rel {rel:.2e} |g| {np.linalg.norm(b):.3e}")
        worst = max(worst, rel)
    assert worst <= 3e-2


def test_training_reduces_the_loss(ctx):
    import torch
    net = UN.UNet(ctx, levels=4, base=32, seed=3)
    x = torch.from_numpy(_binary(128, 8, 7)).cuda()
    t = (x * 1.2 - 0.6).contiguous()
    hist = net.train(x, t, epochs=15, batch=4, lr=1e-3)
    print("loss per epoch", [round(h, 5) for h in hist])
    assert hist[-1] < 0.2 * hist[0]


def test_invalid_shapes_are_rejected(ctx):
    import torch
    net = UN.UNet(ctx, levels=4, base=32, seed=4)
    with pytest.raises(H.HuvError):
        net.forward(torch.zeros((1, 96, 96), device="cuda"))
    with pytest.raises(H.HuvError):
        UN.UNet(ctx, levels=7, base=32)
