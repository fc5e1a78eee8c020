"""Differentiation workload (grad_128): adjoint timing, finite-difference check, fit.

128 bubbles (exponential nucleation, beta L = 15, seeded), 32 x 64 surface
grid, 500 time points to percolation, 128 omegas geomspace(0.5, 300), 32
Fibonacci k-hat.  theta = (xi, A) -> coefficients [[A]], powers [[xi + 3]].
Loss: mean over (k-hat, omega) of (log S(theta) - log S(theta*))^2 with
theta* = (2.2, 0.9).
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import configs
from . import many_bubbles as mb
from .autograd import log_spectrum_loss, theta_spectrum

THETA_STAR = (2.2, 0.9)


class GradProblem:
    def __init__(self, wl=None, precision: str = "fp64"):
        self.wl = wl if wl is not None else configs.grad_128()
        wl = self.wl
        lb = mb.LatticeBubbles(mb.EmptyLattice(), wl.interior, wl.exterior)
        self.bf = mb.GeneralizedBulkFlow(lb)
        self.bf.set_resolution(wl.n_cos_thetax, wl.n_phix)
        self.spec = dict(bf=self.bf, w_arr=wl.w_arr, k_hat=wl.k_hat, t_begin=wl.t_begin,
                         t_end=wl.t_end, n_t=wl.n_t, precision=precision)
        with torch.no_grad():
            self.target = theta_spectrum(torch.tensor(THETA_STAR, dtype=torch.float64), **self.spec)

    def loss(self, theta) -> float:
        with torch.no_grad():
            t = torch.as_tensor(theta, dtype=torch.float64)
            return float(log_spectrum_loss(theta_spectrum(t, **self.spec), self.target))

    def loss_and_grad(self, theta):
        t = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
        L = log_spectrum_loss(theta_spectrum(t, **self.spec), self.target)
        L.backward()
        return L.item(), t.grad.numpy().copy()

    def central_fd(self, theta, h: float = 1e-6) -> np.ndarray:
        g = np.zeros(2)
        for j in range(2):
            tp = np.array(theta, dtype=np.float64)
            tm = tp.copy()
            tp[j] += h
            tm[j] -= h
            g[j] = (self.loss(tp) - self.loss(tm)) / (2.0 * h)
        return g

    def fit(self, theta0=(1.5, 1.3), steps: int = 200, lr: float = 0.05):
        """Adam with a cosine-annealed step size; returns (theta, history)."""
        t = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
        opt = torch.optim.Adam([t], lr=lr)
        sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=steps, eta_min=lr * 1e-3)
        hist = []
        for _ in range(steps):
            opt.zero_grad()
            L = log_spectrum_loss(theta_spectrum(t, **self.spec), self.target)
            L.backward()
            opt.step()
            sched.step()
            hist.append((L.item(), t[0].item(), t[1].item()))
        return t.detach().numpy().copy(), hist


def measure(steps: int = 20, fit_steps: int = 200) -> dict:
    """Gradients per second (forward + adjoint), the FD check and the fit.

    The timed gradients and the fit use the blocked FP64 forward (within
    ~1e-11 of the plain FP64 form); the finite-difference check uses the plain
    FP64 forward, whose rounding is smooth enough for h = 1e-6."""
    theta = (1.8, 1.1)
    P64 = GradProblem(precision="fp64")
    _, g = P64.loss_and_grad(theta)
    fd = P64.central_fd(theta, 1e-6)
    rel = float(np.max(np.abs(g - fd) / np.maximum(np.abs(fd), 1e-300)))
    P = GradProblem(P64.wl, precision="fp64_blocked")
    for _ in range(3):  # warm-up
        P.loss_and_grad(theta)
    dev_ms, walls = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        _, gb = P.loss_and_grad(theta)
        walls.append(time.perf_counter() - t0)
        dev_ms.append(P.bf.last_timings()["total_ms"])
    wall = float(np.median(walls))
    t1 = time.perf_counter()
    th, hist = P.fit(steps=fit_steps)
    fit_s = time.perf_counter() - t1
    return {"grad_per_s": 1.0 / wall, "wall_ms_per_grad": wall * 1e3,
            "wall_ms_per_grad_mean": float(np.mean(walls)) * 1e3,
            "adjoint_device_ms": float(np.median(dev_ms)),
            "grad": g.tolist(), "fd_h1e-6": fd.tolist(), "fd_max_rel_err": rel,
            "grad_blocked_forward": gb.tolist(),
            "fit": {"theta0": [1.5, 1.3], "steps": fit_steps, "theta": th.tolist(),
                    "theta_star": list(THETA_STAR), "final_loss": hist[-1][0],
                    "abs_err": [abs(th[0] - THETA_STAR[0]), abs(th[1] - THETA_STAR[1])],
                    "seconds": fit_s},
            "workload": "grad_128: 128 bubbles (beta L=15), 32x64, 500 t, 128 omegas "
                        "geomspace(0.5,300), 32 k-hat"}
