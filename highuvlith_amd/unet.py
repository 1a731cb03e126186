"""U-Net warm start for the GPU ILT (include/huv.h, huv_unet_*).

The network maps a target pattern (N, H, W) to an initial theta for
`optimize_ilt`; it is trained on the GPU on (target, ILT-converged theta)
pairs produced by the GPU ILT itself, with BF16 weights/activations and FP32
accumulation on the matrix cores.  The reference has no learned part: its ILT
starts from the target (ilt.rs:495-522), which stays the fallback here.
"""
import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import HuvError, check


class UNet:
    def __init__(self, ctx, levels=4, base=32, seed=0):
        self.ctx, self.levels, self.base = ctx, levels, base
        h = C.c_void_p()
        check(ctx.lib.huv_unet_create(ctx.handle, levels, base, C.c_uint64(seed), C.byref(h)))
        self.handle = h
        n = C.c_int64()
        check(ctx.lib.huv_unet_num_params(h, C.byref(n)))
        self.num_params = n.value

    def __del__(self):
        try:
            if self.handle and self.ctx.handle:
                self.ctx.lib.huv_unet_destroy(self.handle)
        except Exception:
            pass
        self.handle = None

    def get_params(self):
        out = np.empty(self.num_params, np.float32)
        check(self.ctx.lib.huv_unet_get_params(self.handle, out.ctypes.data_as(C.c_void_p)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.num_params:
            raise HuvError(2, f"expected {self.num_params} parameters, got {p.size}")
        check(self.ctx.lib.huv_unet_set_params(self.handle, p.ctypes.data_as(C.c_void_p)))

    @staticmethod
    def _dev(x):
        import torch
        if not isinstance(x, torch.Tensor):
            x = torch.as_tensor(np.ascontiguousarray(x, np.float32), device="cuda")
        x = x.to(torch.float32).contiguous()
        return x[None] if x.dim() == 2 else x

    def forward(self, x, chunk=32):
        """theta_0 = net(target) for a batch (torch CUDA or numpy (N, H, W)),
        `chunk` samples per call (activations are sized per call)."""
        import torch
        x = self._dev(x)
        out = torch.empty_like(x)
        self.ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        N, H, W = x.shape
        for i in range(0, N, chunk):
            xi, oi = x[i:i + chunk], out[i:i + chunk]
            check(self.ctx.lib.huv_unet_forward(self.handle, xi.shape[0], H, W,
                                                C.c_void_p(xi.data_ptr()),
                                                C.c_void_p(oi.data_ptr())))
        return out

    def loss_and_grad(self, x, target, want_grad=False):
        import torch
        x, t = self._dev(x), self._dev(target)
        self.ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        N, H, W = x.shape
        loss = C.c_double()
        g = np.empty(self.num_params, np.float32) if want_grad else None
        check(self.ctx.lib.huv_unet_loss_grad(
            self.handle, N, H, W, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
            C.byref(loss), g.ctypes.data_as(C.c_void_p) if want_grad else None))
        return (loss.value, g) if want_grad else loss.value

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        check(self.ctx.lib.huv_unet_adam(self.handle, lr, beta1, beta2, eps))

    def train(self, inputs, targets, epochs=20, batch=8, lr=1e-3, seed=0, log=None):
        """Adam on shuffled mini-batches; inputs/targets torch CUDA (S, H, W).
        Returns the mean loss per epoch."""
        import torch
        rng = np.random.default_rng(seed)
        S = inputs.shape[0]
        hist = []
        for ep in range(epochs):
            order = rng.permutation(S)
            tot, cnt = 0.0, 0
            for i in range(0, S - batch + 1, batch):
                idx = torch.as_tensor(order[i:i + batch], device=inputs.device)
                loss = self.loss_and_grad(inputs.index_select(0, idx), targets.index_select(0, idx))
                self.adam_step(lr)
                tot += loss
                cnt += 1
            hist.append(tot / max(cnt, 1))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.5f}")
        return hist


@dataclass
class WarmStartResult:
    plain_cost: np.ndarray       # plain ILT cost after the fixed budget (the threshold)
    plain_wall_s: float
    plain_iterations: np.ndarray
    warm_cost: np.ndarray        # warm-started cost at termination (after any fallback)
    warm_wall_s: float           # inference, refinement, continuation and fallback runs
    warm_iterations: np.ndarray  # all iterations spent on the warm side, per target
    reached: np.ndarray          # final cost at or below the plain cost
    fallback: np.ndarray         # warm start worse than the target at iteration 0
    continued: np.ndarray        # the convergence test would have stopped it above the plain cost
    plain_fallback: np.ndarray   # still above after the budget: plain ILT from the target
    reached_by_refinement: np.ndarray  # reached before the convergence test would have fired
    stages_s: dict = None        # wall seconds of each warm-side stage


# "At or below the plain cost" with no slack: a target's ILT run does not
# depend on the other targets in its batch (the field kernels split the
# kernel sum the same way at every batch size), so a plain fallback repeats
# the plain run bit for bit.
REACH_RTOL = 0.0


def convergence_iteration(history, tol):
    """First iteration at which the driver's convergence test (three
    consecutive relative decreases below tol, ilt.rs:755-866) fires on this
    cost history, or -1.  A history that ends at its stop cost is scanned up
    to its last step only, where the stop cost takes precedence."""
    small = 0
    for i in range(1, len(history) - 1):
        prev = history[i - 1]
        rel = (prev - history[i]) / prev if prev > 0.0 else 0.0
        small = small + 1 if rel < tol else 0
        if small >= 3:
            return i
    return -1


def compare_warm_start(engine, net, targets, config, budget=200):
    """Plain ILT from the target for `budget` iterations vs the network warm
    start, each target refined until it reaches its plain cost, with plain ILT
    as the fallback; every stage of the warm side is timed:

    1. inference; a target whose warm cost exceeds its target-initialised cost
       starts from the target instead (`fallback`);
    2. refinement with the same optimiser until the plain cost or the budget,
       all targets in one batched run: the convergence test, which would stop
       a target that stalls above the plain cost, is off (`continued` counts
       the targets it would have stopped there; `reached_by_refinement` those
       that reach the plain cost before it would have fired);
    3. a target still above the plain cost after the budget is re-run as plain
       ILT from the target until it reaches it (`plain_fallback`).
    """
    import torch
    from .ilt import IltProblem
    prob = IltProblem(engine, config)
    tg = torch.as_tensor(np.ascontiguousarray(targets, np.float32), device="cuda")
    T = tg.shape[0]
    # Untimed set-up for both sides: optimiser workspaces and kernel tables
    # for this batch, and the network's activation buffers.
    prob.run_stage(tg, prob.initial_theta(tg), 1)
    net.forward(tg[:32])
    theta_plain = prob.initial_theta(tg)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, hist_p, its_p, _ = prob.run_stage(tg, theta_plain, budget)
    torch.cuda.synchronize()
    plain_wall = time.perf_counter() - t0
    plain_cost = np.array([h[-1] for h in hist_p])
    goal = plain_cost * (1.0 + REACH_RTOL)

    stages = {}

    def lap(name, t_prev):
        torch.cuda.synchronize()
        t = time.perf_counter()
        stages[name] = t - t_prev
        return t

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    theta_w = net.forward(tg)
    c_warm = prob.cost(tg, theta_w)
    c_tgt = prob.cost(tg, prob.initial_theta(tg))
    fallback = c_warm > c_tgt
    if fallback.any():
        sel = torch.as_tensor(fallback, device="cuda")
        theta_w[sel] = prob.initial_theta(tg)[sel]
    t1 = lap("inference", t0)
    out = prob.run_stage_traced(tg, theta_w, budget, stop_cost=goal, convergence_tol=0.0)
    t1 = lap("refinement", t1)
    its = np.asarray(out["iterations"]).copy()
    cost = np.array([h[-1] for h in out["cost_history"]])
    stall = np.array([convergence_iteration(h, config.convergence_tol)
                      for h in out["cost_history"]])
    continued = stall >= 0   # the convergence test would have stopped it above the goal
    by_refinement = (cost <= goal) & ~continued
    # plain ILT from the target for whatever is still above
    plain_fb = cost > goal
    if plain_fb.any():
        idx = np.flatnonzero(plain_fb)
        sub = torch.as_tensor(idx, device="cuda")
        th = prob.initial_theta(tg[sub])
        _, hist_f, its_f, _ = prob.run_stage(tg[sub], th, budget, stop_cost=goal[idx])
        its[idx] += np.asarray(its_f)
        cost[idx] = np.minimum(cost[idx], [h[-1] for h in hist_f])
    lap("plain_fallback", t1)
    warm_wall = sum(stages.values())
    return WarmStartResult(plain_cost, plain_wall, np.asarray(its_p), cost, warm_wall, its,
                           cost <= goal, fallback, continued, plain_fb, by_refinement,
                           {k: round(v, 4) for k, v in stages.items()})
