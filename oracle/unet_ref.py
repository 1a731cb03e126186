"""ORACLE — test infrastructure only.  FP64 numpy restatement of the U-Net of
include/huv.h (huv_unet_*): same parameter layout, forward pass and exact
backward pass, used by tests/ as the checker of the device kernels.  The
reference has no U-Net; this follows the network as specified in the C ABI.

With bf16=True every tensor the device stores in BF16 is rounded to BF16
(round to nearest even) at the same point: weights, activations after each
conv / pool / upsample, the head's output gradient, the masked output
gradients, data gradients and skip-gradient sums.  Accumulations stay FP64
(the device accumulates in FP32).
"""
import numpy as np


def bf16(x):
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u << 16).astype(np.uint32).view(np.float32).astype(np.float64)


class Layout:
    """Flat parameter layout (mirrors huv_unet_create)."""

    def __init__(self, levels, base):
        self.levels, self.base = levels, base
        ch = lambda l: base << l
        off = 0
        self.enc1, self.enc2 = [], []

        def add(cin, cout):
            nonlocal off
            w = (off, (cout, 3, 3, cin))
            off += cout * 9 * cin
            b = (off, (cout,))
            off += cout
            return (w, b, cin, cout)

        for l in range(levels):
            self.enc1.append(add(8 if l == 0 else ch(l - 1), ch(l)))
            self.enc2.append(add(ch(l), ch(l)))
        self.dec1, self.dec2 = [None] * (levels - 1), [None] * (levels - 1)
        for l in range(levels - 2, -1, -1):
            self.dec1[l] = add(ch(l + 1) + ch(l), ch(l))
            self.dec2[l] = add(ch(l), ch(l))
        self.head_w = off
        off += base
        self.head_b = off
        off += 1
        self.n = off
        self.ch = ch


def _get(p, spec):
    off, shape = spec
    return p[off:off + int(np.prod(shape))].reshape(shape)


def _conv(x, w, b):
    """x (N,H,W,Cin), w (Cout,3,3,Cin) -> (N,H,W,Cout) and the im2col matrix."""
    N, H, W, C = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    cols = np.stack([xp[:, dy:dy + H, dx:dx + W, :] for dy in range(3) for dx in range(3)], axis=3)
    cols = cols.reshape(N * H * W, 9 * C)
    y = cols @ w.reshape(w.shape[0], -1).T
    if b is not None:
        y = y + b
    return y.reshape(N, H, W, -1), cols


def _conv_bwd(dy, cols, w, xshape):
    N, H, W, C = xshape
    d2 = dy.reshape(N * H * W, -1)
    dw = (d2.T @ cols).reshape(w.shape)
    db = d2.sum(axis=0)
    dcols = (d2 @ w.reshape(w.shape[0], -1)).reshape(N, H, W, 3, 3, C)
    dxp = np.zeros((N, H + 2, W + 2, C))
    for dyy in range(3):
        for dxx in range(3):
            dxp[:, dyy:dyy + H, dxx:dxx + W, :] += dcols[:, :, :, dyy, dxx, :]
    return dw, db, dxp[:, 1:H + 1, 1:W + 1, :]


def _pool(x):
    N, H, W, C = x.shape
    v = x.reshape(N, H // 2, 2, W // 2, 2, C)
    return v.max(axis=(2, 4))


def _pool_bwd(x, dy):
    """Gradient to the first maximum in order (0,0), (0,1), (1,0), (1,1)."""
    N, H, W, C = x.shape
    v = x.reshape(N, H // 2, 2, W // 2, 2, C).transpose(0, 1, 3, 2, 4, 5).reshape(N, H // 2, W // 2, 4, C)
    arg = v.argmax(axis=3)
    dx = np.zeros_like(v)
    np.put_along_axis(dx, arg[:, :, :, None, :], dy[:, :, :, None, :], axis=3)
    return dx.reshape(N, H // 2, W // 2, 2, 2, C).transpose(0, 1, 3, 2, 4, 5).reshape(N, H, W, C)


def _up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def forward(p, layout, x, bf=False):
    """x: (N,H,W) target images.  Returns (out (N,H,W), cache)."""
    r = bf16 if bf else (lambda a: a)
    L = layout.levels
    P = {}
    for name in ("enc1", "enc2", "dec1", "dec2"):
        P[name] = [None if s is None else (r(_get(p, s[0])), _get(p, s[1])) for s in getattr(layout, name)]
    N, H, W = x.shape
    x8 = np.zeros((N, H, W, 8))
    x8[..., 0] = r(x)
    c = {"x8": x8, "e1": [None] * L, "e2": [None] * L, "pool": [None] * L, "cat": [None] * L,
         "d1": [None] * L, "d2": [None] * L, "cols": {}, "P": P}
    for l in range(L):
        src = x8 if l == 0 else c["pool"][l - 1]
        y, cols = _conv(src, *P["enc1"][l])
        c["e1"][l] = r(np.maximum(y, 0.0))
        c["cols"][("enc1", l)] = cols
        y, cols = _conv(c["e1"][l], *P["enc2"][l])
        c["e2"][l] = r(np.maximum(y, 0.0))
        c["cols"][("enc2", l)] = cols
        if l < L - 1:
            c["pool"][l] = _pool(c["e2"][l])
    for l in range(L - 2, -1, -1):
        u = c["e2"][L - 1] if l == L - 2 else c["d2"][l + 1]
        c["cat"][l] = np.concatenate([_up(u), c["e2"][l]], axis=3)
        y, cols = _conv(c["cat"][l], *P["dec1"][l])
        c["d1"][l] = r(np.maximum(y, 0.0))
        c["cols"][("dec1", l)] = cols
        y, cols = _conv(c["d1"][l], *P["dec2"][l])
        c["d2"][l] = r(np.maximum(y, 0.0))
        c["cols"][("dec2", l)] = cols
    hw = p[layout.head_w:layout.head_w + layout.base]
    hb = p[layout.head_b]
    out = c["d2"][0] @ hw + hb
    return out, c


def loss_grad(p, layout, x, target, bf=False):
    """Mean-squared error and its gradient w.r.t. the flat parameters."""
    r = bf16 if bf else (lambda a: a)
    L = layout.levels
    out, c = forward(p, layout, x, bf)
    Pn = out.size
    d = out - target
    loss = float((d * d).sum() / Pn)
    dout = 2.0 * d / Pn
    if bf:
        dout = dout.astype(np.float32).astype(np.float64)
    g = np.zeros_like(p)
    hw = p[layout.head_w:layout.head_w + layout.base]
    g[layout.head_w:layout.head_w + layout.base] = np.einsum("nhw,nhwc->c", dout, c["d2"][0])
    g[layout.head_b] = dout.sum()
    gd2 = [None] * L
    gd2[0] = r(dout[..., None] * hw)
    P = c["P"]

    def conv_back(name, l, spec, xin, yout, dy, need_dx):
        dy = np.where(yout > 0, dy, 0.0)
        w = P[name][l][0]
        dw, db, dx = _conv_bwd(dy, c["cols"][(name, l)], w, xin.shape)
        (wo, ws), (bo, bs) = spec[0], spec[1]
        g[wo:wo + dw.size] = dw.ravel()
        g[bo:bo + db.size] = db
        return r(dx) if need_dx else None

    ge2 = [None] * L
    for l in range(L - 1):
        gd1 = conv_back("dec2", l, layout.dec2[l], c["d1"][l], c["d2"][l], gd2[l], True)
        gcat = conv_back("dec1", l, layout.dec1[l], c["cat"][l], c["d1"][l], gd1, True)
        cu = layout.ch(l + 1)
        gu = gcat[..., :cu]
        N, H, W, _ = gu.shape
        gu = r(gu.reshape(N, H // 2, 2, W // 2, 2, cu).sum(axis=(2, 4)))
        if l == L - 2:
            ge2[L - 1] = gu
        else:
            gd2[l + 1] = gu
        ge2[l] = gcat[..., cu:]
    gpool = [None] * L
    for l in range(L - 1, -1, -1):
        if l < L - 1:
            ge2[l] = r(ge2[l] + _pool_bwd(c["e2"][l], gpool[l]))
        ge1 = conv_back("enc2", l, layout.enc2[l], c["e1"][l], c["e2"][l], ge2[l], True)
        src = c["x8"] if l == 0 else c["pool"][l - 1]
        gsrc = conv_back("enc1", l, layout.enc1[l], src, c["e1"][l], ge1, l > 0)
        if l > 0:
            gpool[l - 1] = gsrc
    return loss, g
