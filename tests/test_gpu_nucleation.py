"""Nucleation on the GPU against the oracle (same Philox streams).

Sample points, nucleation sites, false-vacuum volumes and images must be bit
exact.  Nucleation times of the time-stepping method are compared at 1e-12
relative: each step multiplies by Gamma(t) = gamma0 exp(beta (t - t0)), and the
device's exp may differ from glibc's by one unit in the last place, which the
running sum of steps carries along; the decisions (sites, counts) do not move.
Times of the distribution method come from one host-side solve and are exact.
"""
import numpy as np
import pytest

import oracle
import oracle.nucleation as on
from gw_fopt_amd import many_bubbles as mb

pytestmark = pytest.mark.gpu

UNIT = ("box", [0.0, 0.0, 0.0], np.eye(3))
SKEW = ("box", [-0.5, 0.25, 1.0], np.array([[1.2, 0.1, 0.0], [0.2, 0.9, 0.1], [0.0, -0.3, 1.1]]))
SPHERE = ("sphere", [0.1, -0.2, 0.3], 1.3)


def _py_lattice(lat):
    if lat[0] == "box":
        return mb.ParallelepipedLattice(lat[1], lat[2])
    return mb.SphericalLattice(lat[1], lat[2])


def _compare(lat, bc, method, *, beta, gamma0, d_p0, n_points, seeds, t0=0.0, max_it=1_000_000,
             cutoff=0.01, lattices=None, betas=None):
    lattices = lattices or [lat] * len(seeds)
    betas = betas or [beta] * len(seeds)
    specs = [(_py_lattice(L), b, gamma0 if gamma0 is not None else b**4 * 0.02, t0, s, 10 + r)
             for r, (L, b, s) in enumerate(zip(lattices, betas, seeds))]
    print(f"nucleation case: {lat[0]} bc={bc} method={method} n_points={n_points} seeds={seeds}",
          flush=True)
    interior, exterior, th, vh, counts = mb.nucleate_fixed_rate_batch(
        specs, boundary_condition=["periodic", "reflection", "none"][bc],
        method=["fixed_probability_time_stepping", "fixed_probability_distribution"][method],
        d_p0=d_p0, n_points=n_points, max_time_iterations=max_it, cutoff_fraction_false_vacuum=cutoff)
    n_img = on.images_per_bubble(lat[0], bc)
    for r, (L, b, g, tt0, s, stream) in enumerate(specs):
        Lo = lattices[r]
        times = None
        if method == 1:
            times, _ = mb.FixedRateNucleation(b, g, tt0, d_p0).solve_bubbles_distribution(
                40.0, on.volume(Lo))
        oi, oe, ot, ov = on.fixed_rate(Lo, bc, method, beta=b, gamma0=g, t0=tt0, d_p0=d_p0, seed=s,
                                       stream=stream, n_points=n_points, max_time_iterations=max_it,
                                       cutoff_fraction=cutoff, times=times)
        n = int(counts[r])
        assert n == len(oi) > 0, (r, n, len(oi))
        np.testing.assert_array_equal(interior[r, :n, 1:], oi[:, 1:])
        np.testing.assert_array_equal(vh[r, :n], ov)
        if method == 1:
            np.testing.assert_array_equal(interior[r, :n, 0], oi[:, 0])
            np.testing.assert_array_equal(th[r, :n], ot)
        else:
            np.testing.assert_allclose(interior[r, :n, 0], oi[:, 0], rtol=1e-12, atol=0)
            np.testing.assert_allclose(th[r, :n], ot, rtol=1e-12, atol=0)
        ge = exterior[r, :n * n_img]
        np.testing.assert_allclose(ge[:, 0], oe[:, 0], rtol=1e-12, atol=0)
        np.testing.assert_array_equal(ge[:, 1:], oe[:, 1:])
    return counts


def test_time_stepping_unit_box_periodic():
    c = _compare(UNIT, 0, 0, beta=10.0, gamma0=None, d_p0=0.05, n_points=2000, seeds=[1, 2, 3])
    assert c.min() > 10


def test_time_stepping_register_limit():
    """10000 points: the largest per-lane register layout (40 points per lane)."""
    _compare(UNIT, 0, 0, beta=8.0, gamma0=None, d_p0=0.1, n_points=10000, seeds=[5])


def test_time_stepping_global_workspace():
    """More points than the registers hold: reach times kept in device memory."""
    _compare(UNIT, 0, 0, beta=8.0, gamma0=None, d_p0=0.1, n_points=11000, seeds=[6])


def test_distribution_unit_box_periodic():
    _compare(UNIT, 0, 1, beta=10.0, gamma0=None, d_p0=0.01, n_points=3000, seeds=[7, 8])


def test_skewed_box_reflection():
    _compare(SKEW, 1, 0, beta=9.0, gamma0=None, d_p0=0.05, n_points=1500, seeds=[9, 10])


def test_sphere_reflection_and_none():
    _compare(SPHERE, 1, 0, beta=6.0, gamma0=None, d_p0=0.05, n_points=1500, seeds=[11])
    _compare(SPHERE, 1, 1, beta=6.0, gamma0=None, d_p0=0.05, n_points=1500, seeds=[12])
    _compare(SPHERE, 2, 0, beta=6.0, gamma0=None, d_p0=0.05, n_points=700, seeds=[13])


def test_uneven_batch():
    """Realisations with their own box sizes and rates in one launch."""
    rng = np.random.default_rng(2)
    sides = rng.uniform(0.7, 1.4, 5)
    lats = [("box", [0.0, 0.0, 0.0], np.eye(3) * L) for L in sides]
    betas = list(rng.uniform(6.0, 12.0, 5))
    _compare(lats[0], 0, 0, beta=None, gamma0=50.0, d_p0=0.05, n_points=1200,
             seeds=[21, 22, 23, 24, 25], lattices=lats, betas=betas)


def test_capacity_error():
    lat = mb.CartesianLattice((0, 0, 0), np.eye(3))
    with pytest.raises(ValueError, match="more than max_bubbles = 3"):
        mb.nucleate_fixed_rate_batch([(lat, 10.0, 200.0, 0.0, 1, 0)], method="fixed_probability_time_stepping",
                                     d_p0=0.05, n_points=500, max_bubbles=3)


@pytest.mark.parametrize("lat,bc", [(UNIT, 0), (SPHERE, 1), (SKEW, 2)])
def test_spontaneous_matches_oracle(lat, bc):
    sp = mb.SpontaneousNucleation(17, t0=0.25, seed=33)
    interior, exterior = sp.nucleate_many(_py_lattice(lat), 3, ["periodic", "reflection", "none"][bc])
    for r in range(3):
        oi, oe = on.spontaneous(lat, bc, 17, 0.25, 33, r)
        np.testing.assert_array_equal(interior[r], oi)
        np.testing.assert_array_equal(exterior[r], oe)


def test_python_surface_streams():
    """nucleate() uses the next stream on every call, like the reference's advancing generator."""
    lat = mb.CartesianLattice((0.0, 0.0, 0.0), np.eye(3))
    fr = mb.FixedRateNucleation(beta=8.0, gamma0=8.0**4 * 0.02, t0=0.0, d_p0=0.05, seed=4,
                                n_points=1000)
    first = fr.nucleate(lat, method="fixed_probability_time_stepping", d_p0=0.05)
    second = fr.nucleate(lat, method="fixed_probability_time_stepping", d_p0=0.05)
    for k, res in enumerate((first, second)):
        oi, oe, ot, ov = on.fixed_rate(UNIT, 0, 0, beta=8.0, gamma0=8.0**4 * 0.02, t0=0.0, d_p0=0.05,
                                       seed=4, stream=k, n_points=1000)
        np.testing.assert_array_equal(res.lattice_bubbles.bubbles_interior[:, 1:], oi[:, 1:])
        np.testing.assert_array_equal(res.volume_false_vacuum_history, ov)
        assert res.lattice_bubbles.bubbles_exterior.shape == (6 * len(oi), 4)
    assert not np.array_equal(first.time_history, second.time_history)
    sp = mb.SpontaneousNucleation(5, seed=1).nucleate(lat)
    assert sp.bubbles_interior.shape == (5, 4) and sp.lattice_type == "Cartesian"


def test_nucleated_bubbles_feed_the_spectrum():
    """A device-nucleated configuration runs through compute_c_integral (oracle parity)."""
    lat = mb.CartesianLattice((0.0, 0.0, 0.0), np.eye(3))
    fr = mb.FixedRateNucleation(beta=6.0, gamma0=6.0**4 * 0.01, t0=0.0, d_p0=0.05, seed=8,
                                n_points=800)
    res = fr.nucleate(lat, method="fixed_probability_time_stepping", d_p0=0.05)
    lb = res.lattice_bubbles
    lb.normalize_time_mut()
    interior, exterior = lb.bubbles_interior[:8], lb.bubbles_exterior[:48]
    bf = mb.GeneralizedBulkFlow(mb.LatticeBubbles(mb.EmptyLattice(), interior, exterior))
    bf.set_resolution(8, 16)
    bf.set_gradient_scaling_params([[1.0]], [[5.0]])
    w = np.array([0.7, 3.0, 11.0])
    got = bf.compute_c_integral(w, t_end=1.0, n_t=40)
    want = oracle.compute_c_integral(interior, exterior, 8, 16, [[1.0]], [[5.0]], w, t_end=1.0, n_t=40)
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-10 * np.abs(want).max())


def test_device_nucleation_feeds_the_ensemble():
    """Realisations nucleated on the device go straight into BubbleEnsemble; each
    realisation's spectrum equals GeneralizedBulkFlow on its bubbles bit for bit, and
    the oracle within 1e-10."""
    from gw_fopt_amd.ensemble import BubbleEnsemble
    lat = mb.CartesianLattice((0.0, 0.0, 0.0), np.eye(3))
    fr = mb.FixedRateNucleation(beta=7.0, gamma0=7.0**4 * 0.01, t0=0.0, d_p0=0.05, seed=21,
                                n_points=600, max_time_iterations=400)
    interior, exterior, _, _, counts = fr.nucleate_many(
        lat, 6, method="fixed_probability_time_stepping", d_p0=0.05, max_bubbles=64)
    assert counts.min() >= 2
    E = BubbleEnsemble.from_nucleation(interior, exterior, counts)
    E.set_resolution(8, 16)
    E.set_gradient_scaling_params([[1.0]], [[5.0]])
    w = np.geomspace(0.5, 20.0, 6)
    k = mb.fibonacci_directions(3)
    t_end = np.array([interior[r, counts[r] - 1, 0] + 0.5 for r in range(6)])
    out = E.compute_spectrum(w, k, t_end=t_end, n_t=40, precision="fp64")
    for r in (0, 5):
        n = int(counts[r])
        inter, ext = interior[r, :n], exterior[r, :6 * n]
        bf = mb.GeneralizedBulkFlow(mb.LatticeBubbles(mb.EmptyLattice(), inter, ext))
        bf.set_resolution(8, 16)
        bf.set_gradient_scaling_params([[1.0]], [[5.0]])
        S, _ = bf.compute_spectrum(w, k, t_end=t_end[r], n_t=40)
        assert out["S"][r].tobytes() == S.tobytes()
        So = oracle.spectrum(inter, ext, k, 8, 16, [[1.0]], [[5.0]], w, t_end=t_end[r], n_t=40)
        assert np.max(np.abs(S - So) / np.abs(So)) <= 1e-10
