"""GPU ILT mirror of highuvlith-core's ILT interface (crates/highuvlith-core/src/ilt.rs).

``ILTConfig`` (ilt.rs:159-224), ``IltProblem`` (ilt.rs:369-739) and
``optimize_ilt`` (ilt.rs:881-951) with the same argument meaning; the cost,
the exact adjoint gradient and the CG / Armijo loop run on the GPU through
include/huv.h for a batch of targets at once.
"""
import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import HuvError, check, dptr
from .targets import pattern_error


@dataclass
class ProcessCondition:
    """opc.rs:326-360."""
    defocus_nm: float = 0.0
    dose: float = 1.0
    weight: float = 1.0


@dataclass
class ILTConfig:
    learning_rate: float = 0.5
    regularization: float = 0.01          # TV weight
    max_iterations: int = 100
    convergence_tol: float = 1e-4
    min_feature_nm: float = 0.0           # Gaussian density filter (not on the GPU path)
    optimizer: str = "cg"                 # "cg" | "sd"
    cost: str = "image"                   # "image" | "resist"
    threshold: float = 0.25
    steepness: float = 50.0
    conditions: List[ProcessCondition] = field(default_factory=lambda: [ProcessCondition()])
    sigmoid_steepness: float = 4.0
    binarization_weight: float = 0.0
    tv_epsilon: float = 0.01
    absorber: complex = 0j

    def print_threshold(self):
        return self.threshold if self.cost == "resist" else 0.5


class IltProblem:
    """Batched GPU ILT objective bound to an AerialImageEngine."""

    def __init__(self, engine, config: ILTConfig):
        if config.min_feature_nm > 0.0:
            raise HuvError(8, "min_feature_nm (density filter) is not supported on the GPU path")
        if not config.conditions:
            raise HuvError(1, "conditions: at least one process condition is required")
        self.engine, self.config = engine, config
        self.n = engine.grid.size
        ctx = engine.context()
        self.ctx = ctx
        self._sets = [engine.device_kernel_set(c.defocus_nm) for c in config.conditions]
        arr = (C.c_void_p * len(self._sets))(*[s.handle.value for s in self._sets])
        dose = np.array([c.dose for c in config.conditions], dtype=np.float64)
        weight = np.array([c.weight for c in config.conditions], dtype=np.float64)
        a = complex(config.absorber)
        cfg = _lib.huv_ilt_config(1 if config.cost == "resist" else 0, config.threshold,
                                  config.steepness, config.sigmoid_steepness,
                                  config.regularization, config.tv_epsilon,
                                  config.binarization_weight, a.real, a.imag)
        h = C.c_void_p()
        check(ctx.lib.huv_ilt_create(ctx.handle, self.n, len(self._sets), arr, dptr(dose),
                                     dptr(weight), C.byref(cfg), C.byref(h)))
        self.handle = h

    def __del__(self):
        try:
            if self.handle and self.ctx.handle:
                self.ctx.lib.huv_ilt_destroy(self.handle)
        except Exception:
            pass
        self.handle = None

    def _dev(self, x):
        import torch
        if not isinstance(x, torch.Tensor):
            x = torch.as_tensor(np.ascontiguousarray(x, dtype=np.float32), device="cuda")
        x = x.to(torch.float32).contiguous()
        if x.dim() == 2:
            x = x[None]
        if x.shape[-2:] != (self.n, self.n):
            raise HuvError(2, f"{self.n}x{self.n} target (engine grid) expected")
        return x

    def initial_theta(self, targets, init="target"):
        """ilt.rs:495-522: theta = Z - 1/2 (target) or logit(clamp(m))/beta (mask)."""
        import torch
        t = self._dev(targets)
        if init == "target":
            return (t - 0.5).contiguous()
        m = t.clamp(0.01, 0.99)
        return (torch.log(m / (1 - m)) / self.config.sigmoid_steepness).contiguous()

    def cost_and_gradient(self, targets, theta, gradient=True):
        import torch
        tg, th = self._dev(targets), self._dev(theta)
        T = tg.shape[0]
        cost = np.zeros(T)
        grad = torch.empty_like(th) if gradient else None
        self.ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        check(self.ctx.lib.huv_ilt_cost_grad(
            self.handle, T, C.c_void_p(tg.data_ptr()), C.c_void_p(th.data_ptr()), dptr(cost),
            C.c_void_p(grad.data_ptr()) if gradient else None))
        return cost, grad

    def cost(self, targets, theta):
        return self.cost_and_gradient(targets, theta, gradient=False)[0]

    def mask(self, theta):
        import torch
        return torch.sigmoid(self.config.sigmoid_steepness * self._dev(theta))

    def run_stage(self, targets, theta, max_iterations, stop_cost=None):
        """CG / SD with Armijo on every target; theta (torch, device) updated in place.

        stop_cost: optional per-target costs; a target stops as soon as its
        cost is at or below its value (termination "ReachedCost")."""
        import torch
        tg, th = self._dev(targets), self._dev(theta)
        T = tg.shape[0]
        H = max_iterations + 1
        hist = np.zeros((T, H))
        iters = (C.c_int * T)()
        term = (C.c_int * T)()
        cfg = self.config
        self.ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        if stop_cost is None:
            check(self.ctx.lib.huv_ilt_optimize(
                self.handle, T, C.c_void_p(tg.data_ptr()), C.c_void_p(th.data_ptr()),
                max_iterations, cfg.learning_rate, cfg.convergence_tol,
                int(cfg.optimizer == "cg"), dptr(hist), iters, term))
        else:
            stop = np.ascontiguousarray(np.broadcast_to(np.asarray(stop_cost, np.float64), (T,)))
            check(self.ctx.lib.huv_ilt_optimize_to(
                self.handle, T, C.c_void_p(tg.data_ptr()), C.c_void_p(th.data_ptr()),
                max_iterations, cfg.learning_rate, cfg.convergence_tol,
                int(cfg.optimizer == "cg"), dptr(stop), dptr(hist), iters, term))
        its = np.array(list(iters))
        return th, [hist[t, :its[t] + 1] for t in range(T)], its, np.array(list(term))


TERMINATION = {0: "Converged", 1: "MaxIterations", 2: "LineSearchFailed", 3: "ReachedCost"}


@dataclass
class ILTResult:
    theta: object
    mask_transmittance: object
    cost_history: list
    final_cost: np.ndarray
    iterations: np.ndarray
    termination: list
    wall_s: float


def optimize_ilt(engine, targets, config: ILTConfig, init=None):
    """optimize_ilt (ilt.rs:881-951) for a batch of targets on the GPU.

    ``init``: None (theta = Z - 1/2) or a batch of initial transmission masks
    (IltInit::Mask).  Binarisation continuation is not split into stages here.
    """
    import time
    import torch
    prob = IltProblem(engine, config)
    theta = prob.initial_theta(targets) if init is None else prob.initial_theta(init, "mask")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    theta, hist, its, term = prob.run_stage(targets, theta, config.max_iterations)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    final = np.array([h[-1] for h in hist])
    return ILTResult(theta, prob.mask(theta), hist, final, its,
                     [TERMINATION[int(t)] for t in term], wall)
