"""torch.autograd binding of the bulk-flow spectrum and its hand-written adjoint.

The forward pass is ``GeneralizedBulkFlow.compute_spectrum`` (FP64 on the
GPU); the backward pass turns dL/dS into the cotangent of C,

    S = 8 pi w^3 (|C+|^2 + |Cx|^2)  =>  dL/dRe C + i dL/dIm C = dL/dS * 16 pi w^3 * C,

and hands it to the adjoint kernel (``spectrum_vjp``), which recomputes the
collision masks and returns dL/dcoefficients and dL/dpowers.  The
parameterisation of the bulk-flow scaling f = sum_k a_k ratio^p_k
(generalized_bulk_flow.rs:555-559) by theta = (xi, A) follows the reference's
post-processing, power = xi + 3 (src/gw_fopt/bubble_dynamics/utils/
data_gw_bulkflow.py:86): coefficients [[A]], powers [[xi + 3]].
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .many_bubbles import GeneralizedBulkFlow


class BulkFlowSpectrum(torch.autograd.Function):
    """S(k, s, w) as a differentiable function of (coefficients, powers).

    ``spec`` holds the fixed inputs: ``bf`` (a GeneralizedBulkFlow with its
    resolution set), ``w_arr``, ``k_hat``, ``t_begin``, ``t_end``, ``n_t``,
    ``time_cutoff`` and the forward ``precision``.
    """

    @staticmethod
    def forward(ctx, coefficients: torch.Tensor, powers: torch.Tensor, spec: dict) -> torch.Tensor:
        bf: GeneralizedBulkFlow = spec["bf"]
        coef = coefficients.detach().to("cpu", torch.float64).numpy()
        pows = powers.detach().to("cpu", torch.float64).numpy()
        bf.set_gradient_scaling_params(coef.tolist(), pows.tolist(), spec.get("damping"))
        S, _, C = bf.compute_spectrum(spec["w_arr"], spec["k_hat"], t_begin=spec.get("t_begin"),
                                      t_end=spec["t_end"], n_t=spec["n_t"],
                                      time_cutoff=spec.get("time_cutoff"),
                                      precision=spec.get("precision", "fp64"), return_c=True)
        ctx.spec = spec
        ctx.c = C
        ctx.params = (coef, pows)
        ctx.device = coefficients.device
        return torch.from_numpy(S).to(coefficients.device)

    @staticmethod
    def backward(ctx, grad_s: torch.Tensor):
        spec = ctx.spec
        bf: GeneralizedBulkFlow = spec["bf"]
        w = np.asarray(spec["w_arr"], dtype=np.float64)
        gs = grad_s.detach().to("cpu", torch.float64).numpy()  # (n_k, n_sets, n_w)
        g_c = (gs * (16.0 * math.pi) * w ** 3)[:, None, :, :] * ctx.c
        coef, pows = ctx.params
        bf.set_gradient_scaling_params(coef.tolist(), pows.tolist(), spec.get("damping"))
        dc, dp = bf.spectrum_vjp(spec["w_arr"], spec["k_hat"], g_c, t_begin=spec.get("t_begin"),
                                 t_end=spec["t_end"], n_t=spec["n_t"],
                                 time_cutoff=spec.get("time_cutoff"))
        return (torch.from_numpy(dc).to(ctx.device), torch.from_numpy(dp).to(ctx.device), None)


def spectrum(coefficients: torch.Tensor, powers: torch.Tensor, **spec) -> torch.Tensor:
    """Differentiable S (n_k, n_sets, n_w); see :class:`BulkFlowSpectrum`."""
    return BulkFlowSpectrum.apply(coefficients, powers, spec)


def theta_spectrum(theta: torch.Tensor, **spec) -> torch.Tensor:
    """S as a function of theta = (xi, A): coefficients [[A]], powers [[xi + 3]]."""
    coef = theta[1].reshape(1, 1)
    pows = (theta[0] + 3.0).reshape(1, 1)
    return spectrum(coef, pows, **spec)


def log_spectrum_loss(S: torch.Tensor, S_target: torch.Tensor) -> torch.Tensor:
    """Mean squared log-spectrum difference over every (k-hat, omega) bin."""
    return torch.mean((torch.log(S) - torch.log(S_target)) ** 2)
