"""Adjoint of the spectrum with respect to the scaling parameters, on the GPU.

dL/d(coefficients, powers) from the hand-written adjoint kernel (through
bgw_generalized_bulk_flow_spectrum_vjp) is checked against
  * the FP64 reference gradient sum_k Re(conj(G_k) dC_k/dtheta), with the
    forward-mode tangents dC/dtheta of the CPU oracle (which follows
    generalized_bulk_flow.rs:481-586 with f replaced by its derivative);
  * central finite differences (h = 1e-6) of the GPU FP64 forward;
and drives the torch.autograd fit of theta = (xi, A).
"""
import numpy as np
import pytest
import torch

import oracle
from gw_fopt_amd import configs
from gw_fopt_amd import many_bubbles as mb
from gw_fopt_amd import nucleation
from gw_fopt_amd.autograd import log_spectrum_loss, spectrum, theta_spectrum

pytestmark = pytest.mark.gpu

REL_ORACLE = 1e-9   # adjoint vs FP64 reference gradient (both FP64)
REL_FD = 1e-5       # adjoint vs central differences, h = 1e-6


def flow(interior, exterior, n_c, n_phi, coef, pows, damping=None):
    lb = mb.LatticeBubbles(mb.EmptyLattice(), interior, exterior)
    bf = mb.GeneralizedBulkFlow(lb)
    bf.set_resolution(n_c, n_phi)
    bf.set_gradient_scaling_params(coef, pows, damping)
    return bf


def reference_gradient(inter, ext, n_c, n_phi, coef, pows, w, k_hat, G, **kw):
    """sum_k Re(conj(G_k) dC_k/dtheta) from the oracle's tangents."""
    n_s, n_terms = np.shape(coef)
    dc = np.zeros((n_s, n_terms))
    dp = np.zeros((n_s, n_terms))
    for ki, k in enumerate(k_hat):
        for j in range(n_terms):
            for name, out in (("coef", dc), ("pow", dp)):
                T = oracle.compute_c_integral(inter, ext, n_c, n_phi, coef, pows, w, k_hat=k,
                                              derivative=(name, j), **kw)
                out[:, j] += np.sum(np.real(np.conj(G[ki]) * T), axis=(0, 2))
    return dc, dp


def random_cotangent(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def small_async(seed, n):
    b = nucleation.exponential_nucleation(12.0, n_max=n, seed=seed)
    return b, nucleation.face_images(b, 1.0)


def assert_rel(a, b, tol):
    a, b = np.asarray(a), np.asarray(b)
    err = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
    assert err <= tol, (err, a, b)


def test_vjp_matches_oracle_two_terms():
    wl = configs.plain_8()
    n_c, n_phi, n_t = 16, 32, 60
    w = wl.w_arr[::8]
    k_hat = wl.k_hat[:3]
    coef, pows = [[0.9, 0.35]], [[5.2, 3.5]]
    bf = flow(wl.interior, wl.exterior, n_c, n_phi, coef, pows)
    G = random_cotangent(np.random.default_rng(11), (len(k_hat), 2, 1, len(w)))
    dc, dp = bf.spectrum_vjp(w, k_hat, G, t_end=wl.t_end, n_t=n_t)
    rc, rp = reference_gradient(wl.interior, wl.exterior, n_c, n_phi, coef, pows, w, k_hat, G,
                                t_end=wl.t_end, n_t=n_t)
    assert_rel(dc, rc, REL_ORACLE)
    assert_rel(dp, rp, REL_ORACLE)


def test_vjp_async_damping_cutoff_three_sets():
    """Staggered nucleation, a damping width, an exponential time cutoff, three
    parameter sets (two adjoint launches: sets 0-1, then set 2)."""
    inter, ext = small_async(5, 7)
    n_c, n_phi, n_t = 12, 24, 50
    w = np.geomspace(0.5, 80.0, 12)
    k_hat = mb.fibonacci_directions(2)
    coef = [[1.0], [0.6], [1.3]]
    pows = [[5.0], [4.1], [3.3]]
    cut = mb.ExponentialTimeCutoff(0.9, 0.7, 0.5)
    bf = flow(inter, ext, n_c, n_phi, coef, pows, damping=0.4)
    G = random_cotangent(np.random.default_rng(2), (len(k_hat), 2, 3, len(w)))
    dc, dp = bf.spectrum_vjp(w, k_hat, G, t_end=1.0, n_t=n_t, time_cutoff=cut)
    rc, rp = reference_gradient(inter, ext, n_c, n_phi, coef, pows, w, k_hat, G, t_end=1.0,
                                n_t=n_t, damping=0.4,
                                time_cutoff=(1, cut.smax, cut.ratio_t_cut, cut.ratio_t_0))
    assert_rel(dc, rc, REL_ORACLE)
    assert_rel(dp, rp, REL_ORACLE)


def test_vjp_zero_cotangent_and_no_collisions():
    inter = np.array([[0.0, 0.0, 0.0, 0.0]])  # a single bubble never collides
    bf = flow(inter, None, 8, 16, [[1.0]], [[5.0]])
    k = mb.fibonacci_directions(1)
    G = random_cotangent(np.random.default_rng(0), (1, 2, 1, 4))
    dc, dp = bf.spectrum_vjp([1.0, 2.0, 3.0, 4.0], k, G, t_end=1.0, n_t=20)
    assert np.all(dc == 0.0) and np.all(dp == 0.0)


def test_autograd_matches_central_differences():
    """Loss = mean (log S - log S*)^2 on a reduced grad_128 workload; the
    adjoint gradient against central differences of the FP64 forward."""
    wl = configs.grad_128(n_k=4)
    lb = mb.LatticeBubbles(mb.EmptyLattice(), wl.interior, wl.exterior)
    bf = mb.GeneralizedBulkFlow(lb)
    bf.set_resolution(16, 32)
    spec = dict(bf=bf, w_arr=wl.w_arr[::4], k_hat=wl.k_hat, t_end=wl.t_end, n_t=200)
    with torch.no_grad():
        target = theta_spectrum(torch.tensor([2.2, 0.9], dtype=torch.float64), **spec)

    def loss(th):
        with torch.no_grad():
            return float(log_spectrum_loss(theta_spectrum(torch.tensor(th, dtype=torch.float64),
                                                          **spec), target))

    th = torch.tensor([1.7, 1.2], dtype=torch.float64, requires_grad=True)
    L = log_spectrum_loss(theta_spectrum(th, **spec), target)
    L.backward()
    h = 1e-6
    fd = []
    for j in range(2):
        e = np.zeros(2)
        e[j] = h
        fd.append((loss(np.array([1.7, 1.2]) + e) - loss(np.array([1.7, 1.2]) - e)) / (2 * h))
    assert_rel(th.grad.numpy(), np.array(fd), REL_FD)


def test_spectrum_function_gradient_of_raw_parameters():
    """BulkFlowSpectrum with a coefficients/powers tensor of two terms: the
    autograd gradient equals the explicit vjp with G = dL/dS 16 pi w^3 C."""
    wl = configs.plain_8()
    bf = flow(wl.interior, wl.exterior, 12, 24, [[1.0]], [[5.0]])
    w, k = wl.w_arr[::16], wl.k_hat[:2]
    coef = torch.tensor([[0.8, 0.2]], dtype=torch.float64, requires_grad=True)
    pows = torch.tensor([[5.0, 4.0]], dtype=torch.float64, requires_grad=True)
    S = spectrum(coef, pows, bf=bf, w_arr=w, k_hat=k, t_end=1.5, n_t=40)
    wts = torch.linspace(0.5, 2.0, S.numel(), dtype=torch.float64).reshape(S.shape)
    (S * wts).sum().backward()
    bf.set_gradient_scaling_params([[0.8, 0.2]], [[5.0, 4.0]])
    _, _, C = bf.compute_spectrum(w, k, t_end=1.5, n_t=40, return_c=True)
    G = (wts.numpy() * (16.0 * np.pi) * w ** 3)[:, None, :, :] * C
    dc, dp = bf.spectrum_vjp(w, k, G, t_end=1.5, n_t=40)
    np.testing.assert_array_equal(coef.grad.numpy(), dc)
    np.testing.assert_array_equal(pows.grad.numpy(), dp)
    rc, rp = reference_gradient(wl.interior, wl.exterior, 12, 24, [[0.8, 0.2]], [[5.0, 4.0]], w,
                                k, G, t_end=1.5, n_t=40)
    assert_rel(dc, rc, REL_ORACLE)
    assert_rel(dp, rp, REL_ORACLE)


def test_vjp_deterministic():
    wl = configs.plain_8()
    bf = flow(wl.interior, wl.exterior, 16, 32, [[0.9]], [[5.2]])
    G = random_cotangent(np.random.default_rng(4), (2, 2, 1, 8))
    a = bf.spectrum_vjp(wl.w_arr[::8], wl.k_hat[:2], G, t_end=1.5, n_t=80)
    b = bf.spectrum_vjp(wl.w_arr[::8], wl.k_hat[:2], G, t_end=1.5, n_t=80)
    assert a[0].tobytes() == b[0].tobytes() and a[1].tobytes() == b[1].tobytes()


def test_grad_128_fd_and_fit_recover_theta_star():
    """The differentiation workload at full size: 128 bubbles (beta L = 15),
    32 x 64, 500 t, 128 omegas, 32 k-hat.  The adjoint gradient agrees with
    central differences (h = 1e-6) to 1e-5, and a 200-step Adam fit from
    theta0 = (1.5, 1.3) recovers theta* = (2.2, 0.9)."""
    from gw_fopt_amd.gradbench import THETA_STAR, GradProblem
    P = GradProblem()
    _, g = P.loss_and_grad((1.8, 1.1))
    fd = P.central_fd((1.8, 1.1), 1e-6)
    assert_rel(g, fd, REL_FD)
    theta, hist = P.fit(theta0=(1.5, 1.3), steps=200)
    assert abs(theta[0] - THETA_STAR[0]) < 1e-3 and abs(theta[1] - THETA_STAR[1]) < 1e-3, theta
    assert hist[-1][0] < 1e-6 * hist[0][0]


def test_vjp_random_configurations():
    """Seeded random configurations and parameters against the FP64 reference gradient,
    including a run of 1100 steps (two adjoint chunks of time steps) and integer and
    non-integer powers."""
    rng = np.random.default_rng(77)
    for case in range(6):
        inter, ext = small_async(int(rng.integers(0, 10**6)), int(rng.integers(3, 8)))
        n_c, n_phi = int(rng.integers(4, 16)), int(rng.integers(4, 33))
        n_t = 1100 if case == 0 else int(rng.integers(30, 120))
        w = np.geomspace(0.5, float(rng.uniform(20.0, 120.0)), 6)
        k_hat = [rng.normal(size=3)]
        k_hat[0] /= np.linalg.norm(k_hat[0])
        p0 = float(rng.integers(3, 7)) if case % 2 else float(rng.uniform(3.0, 6.0))
        coef, pows = [[float(rng.uniform(0.5, 1.5))]], [[p0]]
        print(f"random adjoint case {case}: {len(inter)} bubbles, {n_c} x {n_phi}, n_t {n_t}, "
              f"power {p0:.3f}", flush=True)
        bf = flow(inter, ext, n_c, n_phi, coef, pows)
        G = random_cotangent(rng, (1, 2, 1, len(w)))
        dc, dp = bf.spectrum_vjp(w, k_hat, G, t_end=1.0, n_t=n_t)
        rc, rp = reference_gradient(inter, ext, n_c, n_phi, coef, pows, w, k_hat, G, t_end=1.0,
                                    n_t=n_t)
        assert_rel(dc, rc, REL_ORACLE)
        assert_rel(dp, rp, REL_ORACLE)
