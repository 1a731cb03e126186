"""The trained U-Net warm start against plain ILT from the target (small
scale): targets and labels from the GPU ILT, training on the GPU, then each
held-out target refined from the network's theta until it reaches the plain
ILT cost (SOCS-judged), with the target initialisation as fallback."""
import numpy as np
import pytest

import highuvlith_amd as H
from highuvlith_amd import ilt as L
from highuvlith_amd import unet as UN
from highuvlith_amd import warmstart as WS
from highuvlith_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_warm_start_reaches_the_plain_cost_in_fewer_iterations():
    import torch
    src, opt, _, K = W.arf_quasar_512_setup()
    eng = H.AerialImageEngine(src, opt, H.GridConfig(256, 2.0), max_kernels=K)
    cfg = L.ILTConfig(cost="resist", regularization=0.0)
    targets = WS.contact_line_targets(72, 256, 2.0, seed=11)
    train, held = targets[:64].contiguous(), targets[64:].contiguous()
    labels, lab_s = WS.ilt_labels(eng, train, cfg, iterations=60)
    net = UN.UNet(eng.context(), levels=4, base=32, seed=5)
    hist = net.train(train, labels, epochs=25, batch=8, lr=1e-3)
    res = UN.compare_warm_start(eng, net, held.cpu().numpy(), cfg, budget=60)
    print("labels", round(lab_s, 3), "s; train loss", [round(h, 4) for h in hist[::5]], hist[-1])
    print("plain cost", res.plain_cost, "iters", res.plain_iterations, "wall", res.plain_wall_s)
    print("warm  cost", res.warm_cost, "iters", res.warm_iterations, "wall", res.warm_wall_s)
    print("reached", res.reached, "fallback", res.fallback, "by refinement",
          res.reached_by_refinement, "continued", res.continued, "plain fallback",
          res.plain_fallback)
    assert np.all(np.isfinite(res.warm_cost))
    assert hist[-1] < 0.5 * hist[0]
    # every target ends at its plain cost (fallbacks included, their iterations counted)
    assert res.reached.mean() >= 0.95
    assert res.reached_by_refinement.mean() >= 0.5
    assert res.warm_iterations.mean() < res.plain_iterations.mean()
