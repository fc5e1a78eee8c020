"""An independent numpy restatement of the bulk-flow integrals, for pinning the oracle.

Written from the mathematics in docs/generalized_bulkflow.md:8-73 (the
reference's physics page), not from the oracle's C++ and not in its operation
order: everything is vectorised, sums run in numpy's pairwise order, the first
collision is taken as the unique minimum of the collision delay (its exact-
arithmetic meaning), and the phases of A and C are multiplied as arrays.  It
is meant for tiny inputs only (it materialises (n_t, n_c, n_phi) arrays).

    C_pm(w)  = 1/(6 pi) sum_n  int dt  e^{i w (t - z_n)} A_n,pm(w, t)
    A_n,pm   = int_{-1}^{1} dzeta e^{-i w (t - t_n) zeta} B_n,pm(zeta, t)
    B_n,pm   = (1 - zeta^2)/2 int_0^{2 pi} dphi g_pm(phi) (t - t_n)^3 f
    f        = 1 before the first collision of the wall element, and
               sum_k a_k (dt_c / (t - t_n))^{p_k} after it,

with the quadratures of the reference: trapezoid weights on the
cos(theta) grid linspace(-1, 1, n_c) (step 2/(n_c-1)), on the phi grid
linspace(0, 2 pi, n_phi) with step 2 pi / n_phi, and on the time grid
linspace(t_begin, t_end, n_t).
"""
from __future__ import annotations

import numpy as np


def minkowski(u, v):
    return -u[..., 0] * v[..., 0] + np.sum(u[..., 1:] * v[..., 1:], axis=-1)


def directions(n_c, n_phi):
    cos_t = -1.0 + (2.0 / (n_c - 1)) * np.arange(n_c)
    phi = (2.0 * np.pi / (n_phi - 1)) * np.arange(n_phi)
    sin_t = np.sqrt(1.0 - cos_t**2)
    x = np.empty((n_c, n_phi, 4))
    x[..., 0] = 1.0
    x[..., 1] = sin_t[:, None] * np.cos(phi)[None, :]
    x[..., 2] = sin_t[:, None] * np.sin(phi)[None, :]
    x[..., 3] = cos_t[:, None]
    return cos_t, phi, x


def first_collision(bubbles, a, n_int, x, tol=1e-10):
    """Unique-minimum first collider per direction: (index or -1, delay)."""
    others = [b for b in range(len(bubbles)) if not (b < n_int and b == a)]
    d = bubbles[others] - bubbles[a]              # (m, 4)
    dsq = minkowski(d, d)                          # (m,)
    dot = np.einsum("ijk,mk->ijm", x * np.array([-1.0, 1, 1, 1]), d)  # (n_c, n_phi, m)
    with np.errstate(divide="ignore", invalid="ignore"):
        delay = np.where(dot > tol, dsq[None, None, :] / (2.0 * dot), np.inf)
    delay = np.where(delay > 0, delay, np.inf)
    idx = np.argmin(delay, axis=-1)
    best = np.take_along_axis(delay, idx[..., None], -1)[..., 0]
    # blockers: any other bubble whose wall reaches the point no later
    with np.errstate(divide="ignore", invalid="ignore"):
        delay_all = np.where(dot > 0, dsq[None, None, :] / (2.0 * dot), np.inf)
    n_tie = np.sum(delay_all <= best[..., None], axis=-1)
    found = np.isfinite(best) & (n_tie == 1)
    first = np.where(found, np.asarray(others)[idx], -1)
    return first, np.where(found, best, 0.0)


def c_integrand(interior, exterior, n_c, n_phi, coefficients, powers, w, t_begin, t_end, n_t):
    """(2, n_sets, n_w, n_t) complex: the terms of C at each time step (their quadrature
    weights included, zero for t <= t_n), summed over the bubbles."""
    interior = np.asarray(interior, float).reshape(-1, 4)
    exterior = np.zeros((0, 4)) if exterior is None else np.asarray(exterior, float).reshape(-1, 4)
    bubbles = np.vstack([interior, exterior])
    n_int = len(interior)
    coefficients = np.asarray(coefficients, float)
    powers = np.asarray(powers, float)
    w = np.asarray(w, float)
    cos_t, phi, x = directions(n_c, n_phi)
    t = t_begin + (t_end - t_begin) / (n_t - 1) * np.arange(n_t)
    dt = t[1] - t[0]
    wt = np.ones(n_t)
    wt[[0, -1]] = 0.5
    wc = np.full(n_c, 2.0 / (n_c - 1))
    wc[[0, -1]] *= 0.5
    wp = np.full(n_phi, 2.0 * np.pi / n_phi)
    wp[[0, -1]] *= 0.5
    g = np.stack([np.cos(2 * phi), np.sin(2 * phi)])          # (2, n_phi)
    out = np.zeros((2, len(coefficients), len(w), n_t), dtype=complex)
    for a in range(n_int):
        first, delay = first_collision(bubbles, a, n_int, x)
        ta, za = interior[a, 0], interior[a, 3]
        dta = t - ta                                           # (n_t,)
        live = dta > 0
        collided = (first[None] >= 0) & (delay[None] > 0) & (dta[:, None, None] >= delay[None])
        with np.errstate(divide="ignore", invalid="ignore"):
            ratio = np.where(collided, delay[None] / np.where(live, dta, 1.0)[:, None, None], 0.0)
        for s in range(len(coefficients)):
            f_post = sum(cf * np.where(collided, ratio, 1.0) ** pw
                         for cf, pw in zip(coefficients[s], powers[s]))
            f = np.where(collided, f_post, 1.0)                # (n_t, n_c, n_phi)
            B = 0.5 * (1 - cos_t**2)[None, None, :] * np.einsum(
                "gj,tij->gti", g * wp[None, :], f) * (np.where(live, dta, 0.0) ** 3)[None, :, None]
            ang = np.exp(-1j * w[:, None, None] * dta[None, :, None] * cos_t[None, None, :])
            A = np.einsum("gti,wti->gwt", B * wc[None, None, :], ang)   # (2, n_w, n_t)
            ph = np.exp(1j * w[:, None] * (t[None, :] - za)) * (wt * dt * (t >= ta))[None, :]
            out[:, s] += A * ph[None] / (6 * np.pi)
    return out


def c_integral(interior, exterior, n_c, n_phi, coefficients, powers, w, t_begin, t_end, n_t,
               cutoff=None):
    """(2, n_sets, n_w) complex; bubbles (t, x, y, z); EmptyLattice semantics.  ``cutoff``:
    None, or a function of the time grid giving the weight of each step."""
    I = c_integrand(interior, exterior, n_c, n_phi, coefficients, powers, w, t_begin, t_end, n_t)
    if cutoff is not None:
        t = t_begin + (t_end - t_begin) / (n_t - 1) * np.arange(n_t)
        I = I * np.asarray(cutoff(t), float)
    return np.sum(I, axis=-1)
