"""The oracle's tensor form (new_generalized_bulk_flow.rs) pinned without the reference:
the projections of the plain form, an isolated bubble in closed form, and covariance
under a quarter turn about z."""
import math

import numpy as np
import pytest

import oracle
from gw_fopt_amd import configs
from gw_fopt_amd import nucleation

XX, XY, XZ, YY, YZ, ZZ = range(6)


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("case", ["plain_8", "async"])
def test_projections_equal_plain_form(case):
    """C+ = (C_xx - C_yy) / 2 and Cx = C_xy against the independently restated plain
    form (generalized_bulk_flow.rs), two coefficient sets."""
    if case == "plain_8":
        wl = configs.plain_8()
        inter, ext = wl.interior, wl.exterior
    else:
        inter = nucleation.exponential_nucleation(12.0, n_max=6, seed=5)
        ext = nucleation.face_images(inter, 1.0)
    coef, pows = [[1.0], [0.5]], [[5.0], [3.5]]
    w = np.geomspace(0.5, 60.0, 9)
    T = oracle.compute_c_tensor(inter, ext, 10, 20, coef, pows, w, t_end=1.6, n_t=90)
    C = oracle.compute_c_integral(inter, ext, 10, 20, coef, pows, w, t_end=1.6, n_t=90)
    assert _rel((T[XX] - T[YY]) / 2, C[0]) < 1e-13
    assert _rel(T[XY], C[1]) < 1e-13


def test_isolated_bubble_closed_form():
    """No collisions (f = 1): the off-diagonal components vanish, C_xx = C_yy, and the
    trace is (1/6pi) sum_t dt w_t e^{iw(t-z)} sum_i dcos w_i e^{-iw dta c_i} (n_phi - 1) dphi dta^3."""
    inter = np.array([[0.1, 0.3, -0.2, 0.4]])
    n_c, n_phi, n_t, t_end = 9, 13, 50, 2.0
    w = np.array([0.7, 2.0, 9.0])
    T = oracle.compute_c_tensor(inter, None, n_c, n_phi, [[1.0]], [[5.0]], w, t_end=t_end, n_t=n_t)
    scale = np.max(np.abs(T))
    for c in (XY, XZ, YZ):
        assert np.max(np.abs(T[c])) < 1e-14 * scale
    assert _rel(T[XX], T[YY]) < 1e-14
    cos = np.linspace(-1.0, 1.0, n_c)
    wc = np.ones(n_c); wc[[0, -1]] = 0.5
    t = np.linspace(0.0, t_end, n_t)
    wt = np.ones(n_t); wt[[0, -1]] = 0.5
    dcos, dphi, dt = 2.0 / (n_c - 1), 2 * math.pi / n_phi, t[1] - t[0]
    trace = np.zeros(len(w), dtype=complex)
    for k in range(n_t):
        dta = t[k] - inter[0, 0]
        if dta <= 0:
            continue
        A = np.sum(dcos * wc[:, None] * np.exp(-1j * w[None, :] * dta * cos[:, None]), axis=0)
        A *= (n_phi - 1) * dphi * dta**3
        trace += dt * wt[k] * np.exp(1j * w * (t[k] - inter[0, 3])) * A
    trace /= 6 * math.pi
    assert _rel(T[XX, 0] + T[YY, 0] + T[ZZ, 0], trace) < 1e-13


def test_quarter_turn_covariance():
    """Bubbles turned by 90 degrees about z ((x, y) -> (-y, x), exact in floating point)
    with n_phi - 1 divisible by 4, so the surface grid maps onto itself: C' = R C R^T."""
    inter = nucleation.exponential_nucleation(10.0, n_max=5, seed=8)
    ext = nucleation.face_images(inter, 1.0)

    def turn(b):
        out = b.copy()
        out[:, 1], out[:, 2] = -b[:, 2], b[:, 1]
        return out

    w = np.geomspace(0.5, 30.0, 7)
    args = dict(t_end=1.4, n_t=80)
    T = oracle.compute_c_tensor(inter, ext, 12, 21, [[1.0]], [[5.0]], w, **args)
    U = oracle.compute_c_tensor(turn(inter), turn(ext), 12, 21, [[1.0]], [[5.0]], w, **args)
    scale = np.max(np.abs(T))
    for got, want in ((U[XX], T[YY]), (U[YY], T[XX]), (U[XY], -T[XY]), (U[XZ], -T[YZ]),
                      (U[YZ], T[XZ]), (U[ZZ], T[ZZ])):
        assert np.max(np.abs(got - want)) < 1e-12 * scale
