"""Host-side logic of the Python surface (CPU only)."""
import numpy as np
import pytest

import oracle
from gw_fopt_amd import configs, nucleation
from gw_fopt_amd import many_bubbles as mb


def test_periodic_images_follow_reference_order():
    lat = mb.ParallelepipedLattice.axis_aligned((0.0, 0.0, 0.0), 1.0, 1.0, 1.0)
    inter = np.array([[0.1, 0.2, 0.3, 0.4], [0.0, 0.5, 0.6, 0.7]])
    lb = mb.LatticeBubbles(lat, inter)
    lb.with_boundary_condition("periodic")
    ext = lb.bubbles_exterior
    # +e1, -e1, +e2, -e2, +e3, -e3 per interior bubble (parallelepiped_lattice.rs:202-211)
    expect = [[0.1, 1.2, 0.3, 0.4], [0.1, -0.8, 0.3, 0.4], [0.1, 0.2, 1.3, 0.4],
              [0.1, 0.2, -0.7, 0.4], [0.1, 0.2, 0.3, 1.4], [0.1, 0.2, 0.3, -0.6]]
    np.testing.assert_array_equal(ext[:6], np.array(expect))
    np.testing.assert_array_equal(ext, nucleation.face_images(inter, 1.0))


def test_lattice_containment_errors():
    lat = mb.ParallelepipedLattice.axis_aligned((0.0, 0.0, 0.0), 1.0, 1.0, 1.0)
    with pytest.raises(ValueError, match="Interior bubbles formed outside lattice"):
        mb.LatticeBubbles(lat, np.array([[0.0, 1.5, 0.5, 0.5]]))
    with pytest.raises(ValueError, match="Exterior bubbles formed inside lattice"):
        mb.LatticeBubbles(lat, np.array([[0.0, 0.5, 0.5, 0.5]]), np.array([[0.0, 0.2, 0.2, 0.2]]))
    with pytest.raises(ValueError, match="Expected 4 columns"):
        mb.LatticeBubbles(lat, np.zeros((2, 3)))
    with pytest.raises(ValueError, match="formed inside bubble"):
        mb.LatticeBubbles(mb.EmptyLattice(), np.array([[0.0, 0.5, 0.5, 0.5], [0.2, 0.55, 0.5, 0.5]]))


def test_cutoff_defaults():
    c = mb.ExponentialTimeCutoff(2.0)
    assert c.t_cut == 0.999999999 * 2.0
    assert c.t_0 == 0.25 * (2.0 - c.t_cut)


def test_set_gradient_scaling_params_validation():
    lb = mb.LatticeBubbles(mb.EmptyLattice(), np.array([[0.0, 0.0, 10.0, 0.0], [0.0, 0.0, 0.0, 0.0]]))
    bf = mb.GeneralizedBulkFlow(lb)
    with pytest.raises(ValueError, match="All coefficient sets must have the same length"):
        bf.set_gradient_scaling_params([[1.0], [1.0, 2.0]], [[3.0], [3.0]])
    with pytest.raises(ValueError, match="Number of coefficient sets and power sets must be equal"):
        bf.set_gradient_scaling_params([[1.0]], [[3.0], [4.0]])
    with pytest.raises(IndexError, match="out of bounds"):
        bf.set_resolution(4, 5)
        bf.compute_first_colliding_bubble(7)
    with pytest.raises(ValueError, match="precision must be"):
        bf.compute_c_integral([1.0], t_end=1.0, n_t=4, precision="fp16")


def test_grid_getters_match_oracle_grids():
    lb = mb.LatticeBubbles(mb.EmptyLattice(), np.array([[0.0, 0.0, 0.0, 0.0]]))
    bf = mb.GeneralizedBulkFlow(lb)
    bf.set_resolution(50, 7)
    cos_t, phi, _ = oracle.grids(50, 7)
    assert np.array_equal(bf.cos_thetax, cos_t)
    assert np.array_equal(bf.phix, phi)


def test_fibonacci_directions():
    k = mb.fibonacci_directions(64)
    np.testing.assert_allclose(np.linalg.norm(k, axis=1), 1.0, atol=1e-15)
    assert abs(k[:, 2].mean()) < 1e-12
    assert np.allclose(k[:, 2], 1 - (2 * np.arange(64) + 1) / 64)


def test_nucleation_rejects_nuclei_inside_bubbles():
    b = nucleation.exponential_nucleation(20.0, n_max=120, seed=7)
    assert b[0, 0] == 0.0 and np.all(np.diff(b[:, 0]) >= 0)
    assert np.all((b[:, 1:] >= 0) & (b[:, 1:] < 1))
    ext = nucleation.face_images(b, 1.0)
    allb = np.vstack([b, ext])
    for i in range(len(b)):
        d = allb - b[i]
        dsq = -d[:, 0] ** 2 + np.sum(d[:, 1:] ** 2, axis=1)
        dsq[i] = 1.0
        assert np.all(dsq >= 0.0)
    # seeded and deterministic
    assert np.array_equal(b, nucleation.exponential_nucleation(20.0, n_max=120, seed=7))


def test_percolation_time_covers_box():
    b = nucleation.exponential_nucleation(20.0, n_max=64, seed=3)
    tp = nucleation.percolation_time(b)
    assert tp > b[:, 0].max() * 0.5
    probe = np.random.default_rng(99).random((2000, 3))
    d = probe[:, None, :] - b[None, :, 1:]
    d -= np.round(d)
    first = np.min(b[None, :, 0] + np.sqrt(np.sum(d * d, axis=-1)), axis=1)
    assert first.max() <= tp + 1e-12


def test_configs_shapes():
    wl = configs.plain_8()
    assert wl.interior.shape == (8, 4) and wl.exterior.shape == (48, 4)
    assert len(wl.w_arr) == 64 and wl.k_hat.shape == (16, 3) and wl.n_t == 200
    assert wl.w_arr[0] == 0.5 and abs(wl.w_arr[-1] - 200.0) < 1e-12
    c, p = configs.theta_to_params(2.2, 0.9)
    assert c == [[0.9]] and p == [[5.2]]
