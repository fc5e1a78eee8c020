"""The tensor form (NewGeneralizedBulkFlow) on the GPU against the oracle."""
import numpy as np
import pytest

import oracle
from gw_fopt_amd import configs
from gw_fopt_amd import many_bubbles as mb
from gw_fopt_amd import nucleation

pytestmark = pytest.mark.gpu


def _flow(inter, ext, n_c, n_phi, coef, pows):
    bf = mb.NewGeneralizedBulkFlow(mb.LatticeBubbles(mb.EmptyLattice(), inter, ext))
    bf.set_resolution(n_c, n_phi)
    bf.set_gradient_scaling_params(coef, pows)
    return bf


def _rel_per_component(a, b):
    return max(np.max(np.abs(a[c] - b[c])) / np.max(np.abs(b[c])) for c in range(6))


def test_tensor_plain_8_matches_oracle():
    wl = configs.plain_8()
    bf = _flow(wl.interior, wl.exterior, wl.n_cos_thetax, wl.n_phix, [[1.0], [0.5]], [[5.0], [3.5]])
    w = np.geomspace(0.5, 80.0, 16)
    T = bf.compute_c_tensor(w, t_end=wl.t_end, n_t=wl.n_t)
    To = oracle.compute_c_tensor(wl.interior, wl.exterior, wl.n_cos_thetax, wl.n_phix,
                                 [[1.0], [0.5]], [[5.0], [3.5]], w, t_end=wl.t_end, n_t=wl.n_t)
    assert T.shape == (6, 2, 16)
    assert _rel_per_component(T, To) <= 1e-10
    Tb = bf.compute_c_tensor(w, t_end=wl.t_end, n_t=wl.n_t, precision="fp64_blocked")
    assert _rel_per_component(Tb, To) <= 1e-10
    Tm = bf.compute_c_tensor(w, t_end=wl.t_end, n_t=wl.n_t, precision="fp64_matrix")
    assert _rel_per_component(Tm, To) <= 1e-10
    # the plain form's projections from the same handle
    C = bf.compute_c_integral(w, t_end=wl.t_end, n_t=wl.n_t)
    assert np.max(np.abs((T[0] - T[3]) / 2 - C[0])) <= 1e-10 * np.max(np.abs(C[0]))


def test_tensor_selected_and_integrand():
    inter = nucleation.exponential_nucleation(12.0, n_max=6, seed=11)
    ext = nucleation.face_images(inter, 1.0)
    bf = _flow(inter, ext, 12, 24, [[1.0]], [[5.0]])
    w = np.array([0.6, 3.0, 17.0])
    sel = [0, 3, 5]
    T = bf.compute_c_tensor(w, t_end=1.5, n_t=70, selected_bubbles=sel)
    To = oracle.compute_c_tensor(inter, ext, 12, 24, [[1.0]], [[5.0]], w, t_end=1.5, n_t=70,
                                 selected_bubbles=sel)
    assert _rel_per_component(T, To) <= 1e-10
    one = bf.compute_c_tensor_fixed_bubble(3, w, None, 1.5, 70)
    one_o = oracle.compute_c_tensor(inter, ext, 12, 24, [[1.0]], [[5.0]], w, t_end=1.5, n_t=70,
                                    selected_bubbles=[3])
    assert _rel_per_component(one, one_o) <= 1e-10
    I = bf.compute_c_tensor_integrand(w, t_end=1.5, n_t=70, selected_bubbles=sel)
    assert I.shape == (6, 1, 3, 70)
    assert _rel_per_component(I.sum(axis=-1), To) <= 1e-10
    assert not np.any(bf.compute_c_tensor(w, t_end=1.5, n_t=70, selected_bubbles=[]))
    with pytest.raises(IndexError):
        bf.compute_c_tensor(w, t_end=1.5, n_t=70, selected_bubbles=[7])
    with pytest.raises(ValueError, match="n_t must be >= 2"):
        bf.compute_c_tensor(w, t_end=1.5, n_t=1)


def test_tensor_random_configurations():
    """Seeded random configurations: 3-13 bubbles with staggered nucleation (beta 6-20),
    n_cos_thetax 2-39, n_phix 3-139, two parameter sets (an integer and a non-integer
    power), with and without damping.  Each component is judged against its largest entry,
    or 1e-3 of the largest entry of all six where a component is at rounding level (a
    coarse azimuth grid): 1e-10 for both FP64 forms."""
    rng = np.random.default_rng(77)
    for case in range(8):
        n = int(rng.integers(3, 14))
        beta = float(rng.uniform(6, 20))
        inter = nucleation.exponential_nucleation(beta, n_max=n, seed=int(rng.integers(0, 10**6)))
        ext = nucleation.face_images(inter, 1.0)
        n_c, n_phi = int(rng.integers(2, 40)), int(rng.integers(3, 140))
        coef = [[float(rng.uniform(0.5, 1.5))], [float(rng.uniform(0.5, 1.5))]]
        pows = [[5.0], [float(rng.uniform(2.5, 6.0))]]
        damping = None if case % 2 == 0 else float(rng.uniform(0.2, 1.0))
        print(f"tensor random case {case}: {len(inter)} bubbles, grid {n_c} x {n_phi}, "
              f"damping {damping}", flush=True)
        bf = _flow(inter, ext, n_c, n_phi, coef, pows)
        if damping is not None:
            bf.set_gradient_scaling_params(coef, pows, damping)
        w = np.geomspace(0.5, 40.0, 5)
        t_end = float(inter[:, 0].max()) + 0.5
        To = oracle.compute_c_tensor(inter, ext, n_c, n_phi, coef, pows, w, t_end=t_end, n_t=60,
                                     damping=damping)
        floor = 1e-3 * np.max(np.abs(To))
        for precision, bar in (("fp64", 1e-10), ("fp64_blocked", 1e-10)):
            T = bf.compute_c_tensor(w, t_end=t_end, n_t=60, precision=precision)
            for c in range(6):
                err = np.max(np.abs(T[c] - To[c])) / max(np.max(np.abs(To[c])), floor)
                assert err <= bar, (case, precision, c, err)
