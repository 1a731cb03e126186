"""GPU ILT mirror of highuvlith-core's ILT interface (crates/highuvlith-core/src/ilt.rs).

``ILTConfig`` (ilt.rs:159-224), ``IltProblem`` (ilt.rs:369-739) and
``optimize_ilt`` (ilt.rs:881-951) with the same argument meaning; the cost,
the exact adjoint gradient (or the local proxy), the density filter and the
CG / Armijo loop run on the GPU through include/huv.h for a batch of targets
at once.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List

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


GRADIENTS = {"adjoint": 0, "local_proxy": 1}


@dataclass
class ILTConfig:
    """ilt.rs:159-224.  init: "target" (theta = Z - 1/2), ("uniform", m0), or a
    batch of initial transmission masks (IltInit::Mask)."""
    learning_rate: float = 0.5
    regularization: float = 0.01          # TV weight
    max_iterations: int = 100
    convergence_tol: float = 1e-4
    min_feature_nm: float = 0.0           # Gaussian density filter (0 = none)
    gradient: str = "adjoint"             # "adjoint" | "local_proxy"
    optimizer: str = "cg"                 # "cg" | "sd"
    cost: str = "image"                   # "image" | "resist"
    threshold: float = 0.25
    steepness: float = 50.0
    conditions: List[ProcessCondition] = field(default_factory=lambda: [ProcessCondition()])
    sigmoid_steepness: float = 4.0
    binarization_weight: float = 0.0
    binarization_after: int = 0
    tv_epsilon: float = 0.01
    absorber: complex = 0j
    init: object = "target"
    snapshot_every: int = 0

    def print_threshold(self):
        return self.threshold if self.cost == "resist" else 0.5


class IltProblem:
    """Batched GPU ILT objective bound to an AerialImageEngine."""

    def __init__(self, engine, config: ILTConfig):
        if not config.conditions:
            raise HuvError(1, "conditions: at least one process condition is required")
        if config.gradient not in GRADIENTS:
            raise HuvError(1, f"gradient: {config.gradient!r} is not adjoint or local_proxy")
        if not (config.learning_rate > 0.0):
            raise HuvError(1, "learning_rate: must be positive and finite")
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
                                  config.binarization_weight, a.real, a.imag,
                                  config.min_feature_nm, engine.grid.pixel_nm,
                                  GRADIENTS[config.gradient])
        self.handle = None
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
        """ilt.rs:495-522: theta = Z - 1/2 ("target"), logit(m0)/beta everywhere
        (("uniform", m0), m0 in (0, 1)), or logit(clamp(m, 0.01, 0.99))/beta of a
        given mask batch ("mask" with the masks passed as `targets`)."""
        import torch
        t = self._dev(targets)
        beta = self.config.sigmoid_steepness
        if isinstance(init, str) and init == "target":
            return (t - 0.5).contiguous()
        if isinstance(init, tuple) and init[0] == "uniform":
            m0 = float(init[1])
            if not (np.isfinite(m0) and 0.0 < m0 < 1.0):
                raise HuvError(1, "init: uniform initial transmission must be in (0, 1)")
            m0 = min(max(m0, 0.01), 0.99)
            return torch.full_like(t, float(np.log(m0 / (1 - m0)) / beta))
        if isinstance(init, str) and init == "mask":
            m = t.clamp(0.01, 0.99)
            return (torch.log(m / (1 - m)) / beta).contiguous()
        raise HuvError(1, f"init: unknown initialisation {init!r}")

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
        """m = sigma(beta G_s * theta) (ilt.rs:537-542), on the device."""
        import torch
        th = self._dev(theta)
        m = torch.empty_like(th)
        self.ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        check(self.ctx.lib.huv_ilt_mask(self.handle, th.shape[0], C.c_void_p(th.data_ptr()),
                                        C.c_void_p(m.data_ptr())))
        return m

    def set_binarization(self, weight):
        """Binarisation weight of later evaluations (problem.bin_weight, ilt.rs:903-907)."""
        check(self.ctx.lib.huv_ilt_set_binarization(self.handle, float(weight)))

    def image(self, m, condition=0):
        """Aerial image (flare included, not dose-scaled) of masks m (N, n, n) at
        the focus of condition `condition` (ilt.rs:554-562): the transmittance
        a + (1 - a) m through the engine's SOCS path."""
        import torch
        m = self._dev(m)
        a = complex(self.config.absorber)
        z = (self.config.conditions[condition].defocus_nm,)
        if a.imag != 0.0:
            t = (a + (1 - a) * m.to(torch.complex64)).contiguous()
            return self.engine.compute_batch_complex(t, z)[:, 0]
        return self.engine.compute_batch((a.real + (1 - a.real) * m).contiguous(), z)[:, 0]

    def run_stage(self, targets, theta, max_iterations, stop_cost=None):
        """CG / SD with Armijo on every target; theta (torch, device) updated in place.

        stop_cost: optional per-target costs; a target stops as soon as its
        cost is at or below its value (termination "ReachedCost")."""
        out = self.run_stage_traced(targets, theta, max_iterations, stop_cost=stop_cost)
        return out["theta"], out["cost_history"], out["iterations"], out["termination"]

    def run_stage_traced(self, targets, theta, max_iterations, stop_cost=None,
                         snapshot_every=0, iteration_offset=None, nslots=0, snapshots=None,
                         convergence_tol=None):
        """run_stage with its trace: per-target cost and gradient-norm histories
        and, with snapshot_every > 0, the masks at every global iteration that
        is a multiple of it written to snapshots (torch (T, nslots, n, n)).
        iteration_offset: iterations each target ran before (max_iterations is
        then the whole budget); convergence_tol overrides the config's."""
        import torch
        tg, th = self._dev(targets), self._dev(theta)
        T = tg.shape[0]
        H = max_iterations + 1
        hist = np.zeros((T, H))
        gn = np.zeros((T, H))
        iters = (C.c_int * T)()
        term = (C.c_int * T)()
        cfg = self.config
        stop = None
        if stop_cost is not None:
            stop = np.ascontiguousarray(np.broadcast_to(np.asarray(stop_cost, np.float64), (T,)))
        off = None
        if iteration_offset is not None:
            off = (C.c_int * T)(*[int(v) for v in np.broadcast_to(iteration_offset, (T,))])
        if snapshot_every > 0:
            assert snapshots is not None and snapshots.shape == (T, nslots, self.n, self.n)
        self.ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        tol = cfg.convergence_tol if convergence_tol is None else convergence_tol
        check(self.ctx.lib.huv_ilt_run_stage(
            self.handle, T, C.c_void_p(tg.data_ptr()), C.c_void_p(th.data_ptr()),
            max_iterations, cfg.learning_rate, tol, int(cfg.optimizer == "cg"),
            dptr(stop) if stop is not None else None, dptr(hist), dptr(gn), int(snapshot_every),
            off, int(nslots), C.c_void_p(snapshots.data_ptr()) if snapshot_every > 0 else None,
            iters, term))
        its = np.array(list(iters))
        return {"theta": th, "cost_history": [hist[t, :its[t] + 1] for t in range(T)],
                "gradient_norm_history": [gn[t, :its[t] + 1] for t in range(T)],
                "iterations": its, "termination": np.array(list(term))}


TERMINATION = {0: "Converged", 1: "MaxIterations", 2: "LineSearchFailed", 3: "ReachedCost"}


@dataclass
class ILTResult:
    """ilt.rs:228-265, one entry per target (lists / arrays over the batch)."""
    theta: object                   # (T, n, n) torch
    mask_transmittance: object      # m in (0, 1), (T, n, n) torch
    aerial_image: object            # image of m at the first condition's focus (T, n, n)
    cost_history: list
    stage_starts: list              # indices into each cost history where a stage starts
    final_cost: np.ndarray
    iterations: np.ndarray
    converged: np.ndarray
    termination: list
    binary_mask: object             # m > 0.5 as 0 / 1
    binary_aerial_image: object
    pattern_error: np.ndarray       # printed-state mismatches, continuous mask
    binary_pattern_error: np.ndarray
    gradient_norm_history: list
    snapshots: list                 # per target: [(iteration, m (n, n) torch), ...]
    wall_s: float


def optimize_ilt(engine, targets, config: ILTConfig, init=None):
    """optimize_ilt (ilt.rs:881-951) for a batch of targets on the GPU.

    ``init`` overrides ``config.init``: None, "target", ("uniform", m0) or a
    batch of initial transmission masks (IltInit::Mask).  With
    binarization_weight > 0 and binarization_after = k > 0 the first stage runs
    k iterations without the binarisation penalty and a second stage spends
    the rest of the budget with it (continuation, ilt.rs:894-915); every
    target's iteration count carries over between the stages.
    """
    import time
    import torch
    prob = IltProblem(engine, config)
    targets = prob._dev(targets)
    T, n = targets.shape[0], prob.n
    ini = config.init if init is None else init
    if isinstance(ini, (str, tuple)):
        theta = prob.initial_theta(targets, ini)
    else:
        theta = prob.initial_theta(ini, "mask")
    every = int(config.snapshot_every)
    nslots = config.max_iterations // every if every > 0 else 0
    snaps_dev = (torch.empty((T, max(nslots, 1), n, n), device="cuda") if every > 0 else None)
    snapshots = [[] for _ in range(T)]
    if every > 0:
        m0 = prob.mask(theta)
        for t in range(T):
            snapshots[t].append((0, m0[t].clone()))
    warmup = min(config.binarization_after, config.max_iterations) \
        if config.binarization_weight > 0.0 else 0
    hist = [[] for _ in range(T)]
    gnh = [[] for _ in range(T)]
    starts = [[] for _ in range(T)]
    done = np.zeros(T, dtype=np.int64)
    term = np.ones(T, dtype=np.int64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()

    def stage(budget, offset, take):
        # one launch for the batch; target t runs at most budget - offset[t] iterations
        out = prob.run_stage_traced(targets, theta, budget, snapshot_every=every,
                                    iteration_offset=offset, nslots=nslots, snapshots=snaps_dev)
        for t in range(T):
            if take[t]:
                starts[t].append(len(hist[t]))
                hist[t].extend(out["cost_history"][t].tolist())
                gnh[t].extend(out["gradient_norm_history"][t].tolist())
        return out

    if warmup > 0:
        prob.set_binarization(0.0)
        out = stage(warmup, done, np.ones(T, bool))
        done += out["iterations"]
        prob.set_binarization(config.binarization_weight)
    remaining = config.max_iterations - done
    # a second stage runs for targets with budget left (or as the only stage)
    take = (remaining > 0) | (warmup == 0)
    if np.any(take):
        out = stage(config.max_iterations, done.copy(), take)
        done += np.where(take, out["iterations"], 0)
        term = np.where(take, out["termination"], 1)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    m = prob.mask(theta)
    binary = (m > 0.5).to(torch.float32)
    img = prob.image(m, 0)
    bimg = prob.image(binary, 0)
    d0 = config.conditions[0].dose
    thr = config.print_threshold()
    tcpu = targets.cpu().numpy()
    pe = np.array([pattern_error(d0 * img[t].cpu().numpy(), tcpu[t], thr) for t in range(T)])
    bpe = np.array([pattern_error(d0 * bimg[t].cpu().numpy(), tcpu[t], thr) for t in range(T)])
    if every > 0:
        for t in range(T):
            for k in range(1, int(done[t]) // every + 1):
                if k - 1 < nslots:
                    snapshots[t].append((k * every, snaps_dev[t, k - 1].clone()))
            if snapshots[t][-1][0] != int(done[t]):
                snapshots[t].append((int(done[t]), m[t].clone()))
    final = np.array([h[-1] for h in hist])
    names = [TERMINATION[int(x)] for x in term]
    return ILTResult(theta, m, img, [np.array(h) for h in hist], starts, final, done,
                     np.array([x == "Converged" for x in names]), names, binary, bimg, pe, bpe,
                     [np.array(g) for g in gnh], snapshots, wall)
