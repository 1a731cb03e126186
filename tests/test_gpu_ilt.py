"""GPU ILT cost and exact adjoint gradient against the FP64 oracle
(ilt.rs:648-738) and central finite differences; the GPU CG / Armijo loop
against the oracle's run_stage (ilt.rs:755-866)."""
import math

import numpy as np
import pytest

import oracle
import highuvlith_amd as H
from highuvlith_amd import ilt as L
from highuvlith_amd import workloads as W

pytestmark = pytest.mark.gpu


def quasar_engine(n=512):
    src, opt, grid, K = W.arf_quasar_512_setup()
    grid = H.GridConfig(n, grid.pixel_nm)
    return H.AerialImageEngine(src, opt, grid, max_kernels=K)


def oracle_problem(eng, cfg, target):
    conds = [(c.dose, c.weight, eng.kernels(c.defocus_nm)) for c in cfg.conditions]
    return oracle.IltProblem(eng.grid.size, target, conds, eng.flare_fraction(), cost=cfg.cost,
                             threshold=cfg.threshold, steepness=cfg.steepness,
                             beta=cfg.sigmoid_steepness, tv_weight=cfg.regularization,
                             tv_eps=cfg.tv_epsilon, bin_weight=cfg.binarization_weight,
                             absorber=cfg.absorber, pixel_nm=eng.grid.pixel_nm,
                             min_feature_nm=cfg.min_feature_nm,
                             proxy=cfg.gradient == "local_proxy")


def perturbed_theta(target, seed):
    n = target.shape[-1]
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    return (target - 0.5) + 0.3 * np.sin(i * 0.037 + seed) * np.cos(j * 0.023 - seed)


@pytest.mark.parametrize("case", ["quasar_resist", "multi_condition"])
def test_cost_and_gradient_match_oracle(case):
    if case == "quasar_resist":
        eng = quasar_engine(512)
        cfg = L.ILTConfig(cost="resist", regularization=0.0)
        targets = W.contact_array_targets(2, seed=11, grid=eng.grid)
    else:
        eng = quasar_engine(256)
        cfg = L.ILTConfig(cost="image", regularization=0.05, binarization_weight=0.02,
                          absorber=-math.sqrt(0.06),
                          conditions=[L.ProcessCondition(0.0, 1.0, 1.0),
                                      L.ProcessCondition(60.0, 1.05, 0.5)])
        targets = W.contact_array_targets(2, seed=12, grid=eng.grid)
    prob = L.IltProblem(eng, cfg)
    thetas = np.stack([perturbed_theta(t, s) for s, t in enumerate(targets)])
    cost, grad = prob.cost_and_gradient(targets, thetas)
    grad = grad.cpu().numpy()
    for t in range(len(targets)):
        orc = oracle_problem(eng, cfg, targets[t])
        c_ref, g_ref = orc.cost_and_gradient(thetas[t].astype(np.float32).astype(np.float64))
        rel_c = abs(cost[t] - c_ref) / abs(c_ref)
        rel_g = np.abs(grad[t] - g_ref).max() / np.abs(g_ref).max()
        print(case, t, "cost", cost[t], c_ref, rel_c, "grad rel", rel_g)
        assert rel_c <= 1e-5
        assert rel_g <= 1e-3


def test_gradient_against_finite_differences():
    """Directional central differences of the GPU cost, and the GPU gradient at
    32 random pixels against central differences of the FP64 oracle cost."""
    eng = quasar_engine(512)
    cfg = L.ILTConfig(cost="resist", regularization=0.0)
    target = W.contact_array_targets(1, seed=21, grid=eng.grid)
    theta = perturbed_theta(target[0], 3)[None].astype(np.float32)
    prob = L.IltProblem(eng, cfg)
    _, grad = prob.cost_and_gradient(target, theta)
    g = grad.cpu().numpy()[0]
    rng = np.random.default_rng(5)
    n = eng.grid.size
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    h = 1e-2
    for d in range(3):
        v = (((i * 31 + j * 17 + d * 7) % 13) / 13.0 - 0.5).astype(np.float32)
        cp = prob.cost(target, theta + h * v[None])[0]
        cm = prob.cost(target, theta - h * v[None])[0]
        fd = (cp - cm) / (2 * h)
        an = float((g.astype(np.float64) * v).sum())
        print("directional", fd, an, abs(fd - an) / abs(an))
        assert abs(fd - an) / abs(an) <= 1e-3
    orc = oracle_problem(eng, cfg, target[0])
    th64 = theta[0].astype(np.float64)
    mag = np.abs(g)
    cand = np.flatnonzero(mag.ravel() > 1e-2 * mag.max())
    pix = rng.choice(cand, 32, replace=False)
    hh = 1e-5
    worst = 0.0
    for p in pix:
        e = np.zeros(n * n)
        e[p] = hh
        e = e.reshape(n, n)
        fd = (orc.cost(th64 + e) - orc.cost(th64 - e)) / (2 * hh)
        worst = max(worst, abs(g.ravel()[p] - fd) / abs(fd))
    print("32-pixel FD worst rel", worst)
    assert worst <= 1e-3


def test_gpu_optimizer_tracks_oracle_run_stage():
    eng = quasar_engine(256)
    cfg = L.ILTConfig(cost="resist", regularization=0.0, max_iterations=12)
    targets = W.contact_array_targets(3, seed=31, grid=eng.grid)
    res = L.optimize_ilt(eng, targets, cfg)
    for t in range(len(targets)):
        orc = oracle_problem(eng, cfg, targets[t])
        _, hist, _ = orc.run_stage(targets[t] - 0.5, 12)
        g = res.cost_history[t]
        print(t, g[:4], hist[:4], res.termination[t])
        assert abs(g[0] - hist[0]) / hist[0] < 1e-5
        k = min(len(g), len(hist), 6)
        assert np.all(np.abs(g[:k] - hist[:k]) / hist[:k] < 1e-3)
        assert np.all(np.diff(g) <= 0)
        assert abs(g[-1] - hist[-1]) / hist[-1] < 2e-2


def test_trajectory_does_not_depend_on_the_batch():
    """A target's CG / Armijo run is the same alone, in a batch of five, and in
    a batch whose other targets stop early: bit for bit, theta and history."""
    import torch
    eng = quasar_engine(256)
    cfg = L.ILTConfig(cost="resist", regularization=0.0)
    prob = L.IltProblem(eng, cfg)
    targets = torch.as_tensor(W.contact_array_targets(5, seed=8, grid=eng.grid), device="cuda")
    th5 = prob.initial_theta(targets)
    _, h5, it5, _ = prob.run_stage(targets, th5, 25)
    # the others stop after 3 iterations, target 2 runs on alone
    out = prob.run_stage_traced(targets, prob.initial_theta(targets), 25,
                                iteration_offset=[22, 22, 0, 22, 22])
    for t in (0, 2, 4):
        th1 = prob.initial_theta(targets[t:t + 1])
        _, h1, it1, _ = prob.run_stage(targets[t:t + 1], th1, 25)
        print("target", t, "iterations", it1[0], it5[t], "final", h1[0][-1], h5[t][-1])
        assert it1[0] == it5[t]
        assert np.array_equal(h1[0], h5[t])
        assert torch.equal(th1[0], th5[t])
    assert out["iterations"][2] == it5[2]
    assert np.array_equal(out["cost_history"][2], h5[2])
    assert torch.equal(out["theta"][2], th5[2])


def test_invalid_ilt_configs_rejected():
    eng = quasar_engine(256)
    with pytest.raises(H.HuvError):
        L.IltProblem(eng, L.ILTConfig(conditions=[]))
    with pytest.raises(H.HuvError):   # filter band wider than 127 bins
        L.IltProblem(eng, L.ILTConfig(min_feature_nm=1.0))
    with pytest.raises(H.HuvError):
        L.IltProblem(eng, L.ILTConfig(absorber=complex(float("nan"), 0.0)))
    with pytest.raises(H.HuvError):
        L.IltProblem(eng, L.ILTConfig(gradient="finite_differences"))
    with pytest.raises(H.HuvError):
        L.IltProblem(eng, L.ILTConfig(learning_rate=0.0))
    prob = L.IltProblem(eng, L.ILTConfig())
    with pytest.raises(H.HuvError):
        prob.cost(np.zeros((1, 128, 128)), np.zeros((1, 128, 128)))


def test_benchmark_ilt_run_matches_oracle_run_stage():
    """The ILT benchmark as run: 64 contact-array targets 512^2 @ 2 nm (seed 4),
    ArF 193i quasar, K = 24, sigmoid-resist cost (threshold 0.25, steepness 50),
    no regularisers, 50 CG / Armijo iterations, all targets in one batched
    optimisation.  Two of its targets are re-run by the FP64 oracle's run_stage
    (ilt.rs:755-866) from the same start: the start costs agree to 1e-5 and
    the first iterates to 1e-3.  Over 50 iterations FP32 rounding lets the
    Armijo branches of the two runs drift apart, so the end is checked by what
    the path does not change: the oracle's FP64 cost at the GPU's final theta
    equals the GPU's final cost (1e-4), and both runs end at the same cost
    level (within 10 %)."""
    src, opt, grid, K = W.arf_quasar_512_setup()
    eng = H.AerialImageEngine(src, opt, grid, max_kernels=K)
    targets = W.contact_array_targets(64, seed=4, grid=grid)
    cfg = L.ILTConfig(cost="resist", threshold=0.25, steepness=50.0, regularization=0.0,
                      max_iterations=50)
    res = L.optimize_ilt(eng, targets, cfg)
    assert all(np.all(np.diff(h) <= 0) for h in res.cost_history)
    theta = res.theta.cpu().numpy().astype(np.float64)
    for t in (0, 37):
        orc = oracle_problem(eng, cfg, targets[t])
        _, hist, _ = orc.run_stage(targets[t] - 0.5, 50)
        g = res.cost_history[t]
        k = min(len(g), len(hist), 6)
        at_gpu = orc.cost(theta[t])
        print("benchmark target", t, "gpu", g[0], g[-1], len(g) - 1, "oracle", hist[0], hist[-1],
              len(hist) - 1, "oracle cost at the gpu theta", at_gpu)
        assert abs(g[0] - hist[0]) / hist[0] < 1e-5
        assert np.all(np.abs(g[:k] - hist[:k]) / hist[:k] < 1e-3)
        assert abs(at_gpu - g[-1]) / g[-1] < 1e-4
        assert abs(g[-1] - hist[-1]) / hist[-1] < 0.1


def _annular_engine(n, px, K, kind="arf"):
    if kind == "arf":
        src = H.Source.arf(H.IlluminationShape.Annular(0.7, 0.9))
        opt = H.ProjectionOptics.immersion_193i()
    else:
        src = H.Source(13.5, H.IlluminationShape.SectorPoles(0.7, 0.9, 35.0, 0.0, 2))
        opt = H.EuvProjectionOptics.nxe_033()
    return H.AerialImageEngine(src, opt, H.GridConfig(n, px), max_kernels=K)


ILT_CASES = {
    # coarse grids above 64: the adjoint's team-FFT passes through HBM
    "coarse128": (lambda: _annular_engine(1024, 2.0, 8),
                  dict(cost="resist", regularization=0.0)),
    "coarse192": (lambda: _annular_engine(1024, 1.0, 6, "euv"),
                  dict(cost="image", regularization=0.02)),
    "coarse256": (lambda: _annular_engine(2048, 2.0, 4),
                  dict(cost="resist", regularization=0.0)),
    # attenuated phase-shift absorber with a complex amplitude (ilt.rs:544-552, 672)
    "complex_absorber": (lambda: quasar_engine(256),
                         dict(cost="image", regularization=0.05, binarization_weight=0.02,
                              absorber=0.245 * np.exp(1j * 2.2))),
    # Gaussian density filter (ilt.rs:443-454, 524-535, 733-737)
    "density_filter": (lambda: quasar_engine(512),
                       dict(cost="resist", regularization=0.0, min_feature_nm=40.0)),
    # local-proxy gradient (ilt.rs:719-721)
    "local_proxy": (lambda: quasar_engine(256),
                    dict(cost="resist", regularization=0.01, gradient="local_proxy")),
}


@pytest.mark.parametrize("case", list(ILT_CASES))
def test_ilt_extensions_match_oracle(case):
    """Cost and dC/dtheta of the device ILT against the FP64 oracle (cost 1e-5,
    gradient 1e-3 of its largest entry) on coarse grids 128 / 192 / 256, with a
    complex absorber, with the density filter and with the local proxy."""
    make, kw = ILT_CASES[case]
    eng = make()
    cfg = L.ILTConfig(**kw)
    ds = eng.device_kernel_set(0.0)
    targets = W.contact_array_targets(2, seed=41, grid=eng.grid)
    prob = L.IltProblem(eng, cfg)
    thetas = np.stack([perturbed_theta(t, s) for s, t in enumerate(targets)])
    cost, grad = prob.cost_and_gradient(targets, thetas)
    grad = grad.cpu().numpy()
    for t in range(len(targets)):
        orc = oracle_problem(eng, cfg, targets[t])
        c_ref, g_ref = orc.cost_and_gradient(thetas[t].astype(np.float32).astype(np.float64))
        rel_c = abs(cost[t] - c_ref) / abs(c_ref)
        rel_g = np.abs(grad[t] - g_ref).max() / np.abs(g_ref).max()
        print(case, "band", ds.band, "coarse", ds.coarse, t, "cost", cost[t], c_ref, rel_c,
              "grad rel", rel_g)
        assert rel_c <= 1e-5
        assert rel_g <= 1e-3
    if case == "density_filter":
        m = prob.mask(thetas).cpu().numpy()
        for t in range(len(targets)):
            ref = oracle_problem(eng, cfg, targets[t]).mask(
                thetas[t].astype(np.float32).astype(np.float64))
            assert np.abs(m[t] - ref).max() <= 1e-5


def test_local_proxy_stage_tracks_oracle():
    """Steepest descent on the local proxy with plain-decrease acceptance
    (ilt.rs:767-801) against the oracle's run_stage on the same objective."""
    eng = quasar_engine(256)
    cfg = L.ILTConfig(cost="resist", regularization=0.01, gradient="local_proxy",
                      max_iterations=8)
    targets = W.contact_array_targets(2, seed=51, grid=eng.grid)
    prob = L.IltProblem(eng, cfg)
    _, hist, its, term = prob.run_stage(targets, prob.initial_theta(targets), 8)
    for t in range(2):
        orc = oracle_problem(eng, cfg, targets[t])
        _, ref, _ = orc.run_stage(targets[t] - 0.5, 8, use_cg=False)
        k = min(len(hist[t]), len(ref), 5)
        print("proxy", t, hist[t][:k], ref[:k], term[t])
        assert np.all(np.abs(hist[t][:k] - ref[:k]) / ref[:k] < 1e-3)
        assert np.all(np.diff(hist[t]) < 0)


def test_optimize_ilt_stages_and_result_fields():
    """optimize_ilt with binarisation continuation (ilt.rs:881-951): a first
    stage of 4 iterations without the penalty, then the rest of the 10-iteration
    budget with it; stage starts, per-stage monotone costs, gradient norms,
    snapshots every 3 iterations, the images of the continuous and the binary
    mask and their pattern errors, against the oracle run the same way."""
    eng = quasar_engine(256)
    cfg = L.ILTConfig(cost="resist", regularization=0.0, binarization_weight=0.05,
                      binarization_after=4, max_iterations=10, snapshot_every=3)
    targets = W.contact_array_targets(2, seed=61, grid=eng.grid)
    res = L.optimize_ilt(eng, targets, cfg)
    for t in range(2):
        h, starts = res.cost_history[t], res.stage_starts[t]
        assert starts[0] == 0 and len(starts) == 2
        k1 = starts[1]
        assert np.all(np.diff(h[:k1]) <= 0) and np.all(np.diff(h[k1:]) <= 0)
        assert res.iterations[t] == len(h) - len(starts) <= 10
        assert len(res.gradient_norm_history[t]) == len(h)
        # the oracle: same two stages
        orc = oracle_problem(eng, L.ILTConfig(cost="resist", regularization=0.0,
                                              binarization_weight=0.0), targets[t])
        th1, r1, _ = orc.run_stage(targets[t] - 0.5, 4)
        orc.set_bin_weight(0.05)
        th2, r2, _ = orc.run_stage(th1, 10 - (len(r1) - 1))
        print("stages", t, h[:k1], r1, h[k1:], r2)
        assert np.all(np.abs(h[:k1] - r1[:k1]) / r1[:k1] < 1e-3)
        assert abs(h[k1] - r2[0]) / r2[0] < 1e-3
        assert abs(h[-1] - r2[-1]) / r2[-1] < 2e-2
        # snapshots at 0, 3, 6, 9 and the final iteration
        its = [s[0] for s in res.snapshots[t]]
        done = int(res.iterations[t])
        assert its == sorted(set([0] + list(range(3, done + 1, 3)) + [done]))
        assert np.array_equal(res.snapshots[t][-1][1].cpu().numpy(),
                              res.mask_transmittance[t].cpu().numpy())
        # result images and pattern errors (ilt.rs:918-950)
        m = res.mask_transmittance[t].cpu().numpy().astype(np.float64)
        img_ref = orc.image(m, 0)
        assert np.abs(res.aerial_image[t].cpu().numpy() - img_ref).max() <= 1e-5
        binary = (m > 0.5).astype(np.float64)
        assert np.array_equal(res.binary_mask[t].cpu().numpy(), binary.astype(np.float32))
        bimg_ref = orc.image(binary, 0)
        assert np.abs(res.binary_aerial_image[t].cpu().numpy() - bimg_ref).max() <= 1e-5
        pe = int(np.count_nonzero((res.aerial_image[t].cpu().numpy() > 0.25) != (targets[t] > 0.5)))
        assert res.pattern_error[t] == pe
