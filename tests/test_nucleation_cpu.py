"""Nucleation without a GPU: the oracle pinned against numpy's Philox, brute force and an
independent restatement; the library's host-side ODE against scipy's Radau; the Python
lattice classes against the oracle's lattice geometry; C-ABI argument validation."""
import ctypes
import math

import numpy as np
import pytest

import oracle.nucleation as on
from gw_fopt_amd import _lib
from gw_fopt_amd import many_bubbles as mb

BOX = ("box", [0.0, 0.0, 0.0], np.eye(3))
SKEW = ("box", [-0.5, 0.25, 1.0], np.array([[1.2, 0.1, 0.0], [0.2, 0.9, 0.1], [0.0, -0.3, 1.1]]))
SPHERE = ("sphere", [0.1, -0.2, 0.3], 1.3)


def _uniform(x):
    return (x >> 11) * 2.0 ** -53


# --------------------------------------------------------------------------- Philox
def test_philox_known_answers():
    """Random123's published Philox4x64-10 known-answer vectors."""
    assert on.philox4x64([0, 0, 0, 0], [0, 0]) == [
        0x16554D9ECA36314C, 0xDB20FE9D672D0FDC, 0xD7E772CEE186176B, 0x7E68B68AEC7BA23B]
    m = 2**64 - 1
    assert on.philox4x64([m, m, m, m], [m, m]) == [
        0x87B092C3013FE90B, 0x438C3C67BE8D0224, 0x9CC7D7C69CD777B6, 0xA09CAEBF594F0BA0]


@pytest.mark.parametrize("ctr,key", [((0, 0, 0, 0), (0, 0)), ((41, 3, 0, 0), (12345, 7)),
                                     ((2**40 + 5, 2, 0, 0), (2**63 + 11, 2**32 + 1))])
def test_philox_matches_numpy(ctr, key):
    """numpy's Philox bit generator increments its counter before each block."""
    c = ctr[0] - 1 + (ctr[1] << 64) + (ctr[2] << 128) + (ctr[3] << 192)
    if ctr[0] == 0:
        c = (ctr[1] << 64) - 1 + (ctr[2] << 128) + (ctr[3] << 192)
    if c < 0:
        c += 2**256
    bg = np.random.Philox(counter=c, key=key[0] + (key[1] << 64))
    assert [int(v) for v in bg.random_raw(4)] == on.philox4x64(ctr, key)


# --------------------------------------------------------------------------- lattices
@pytest.mark.parametrize("lat", [BOX, SKEW, SPHERE])
def test_sample_points_from_the_draws(lat):
    """Points follow the draws of stream 0 in attempt order; the sphere keeps the in-ball ones."""
    pts = on.sample_points(lat, 9, 4, 200)
    expect = []
    a = 0
    while len(expect) < 200:
        x = on.philox4x64([a, 0, 0, 0], [9, 4])
        u, v, w = (_uniform(x[i]) for i in range(3))
        a += 1
        if lat[0] == "box":
            o, e = np.asarray(lat[1]), np.asarray(lat[2])
            expect.append([o[i] + u * e[0][i] + v * e[1][i] + w * e[2][i] for i in range(3)])
        else:
            s = [u * 2.0 - 1.0, v * 2.0 - 1.0, w * 2.0 - 1.0]
            if s[0] * s[0] + s[1] * s[1] + s[2] * s[2] > 1.0:
                continue
            expect.append([lat[1][i] + lat[2] * s[i] for i in range(3)])
    np.testing.assert_array_equal(pts, np.array(expect))
    assert on.contains(lat, pts).all()


def test_lattice_volumes():
    assert on.volume(BOX) == 1.0
    assert on.volume(SKEW) == pytest.approx(abs(np.linalg.det(SKEW[2])), rel=1e-14)
    assert on.volume(SPHERE) == pytest.approx(4.0 / 3.0 * math.pi * 1.3**3, rel=1e-15)


def _py_lattice(lat):
    if lat[0] == "box":
        return mb.ParallelepipedLattice(lat[1], lat[2])
    return mb.SphericalLattice(lat[1], lat[2])


@pytest.mark.parametrize("lat,bc", [(BOX, 0), (SKEW, 0), (SKEW, 1), (BOX, 2), (SPHERE, 1), (SPHERE, 0)])
def test_images_python_equals_oracle(lat, bc):
    """Two independent codings of generate_bubbles_exterior agree bit for bit."""
    rng = np.random.default_rng(3)
    pts = on.sample_points(lat, 5, 0, 40)
    ev = np.column_stack([rng.random(40), pts])
    ev[0, 1:] = lat[1]  # a bubble at the origin / center (the sphere's degenerate branch)
    py = _py_lattice(lat).generate_bubbles_exterior(ev, ["periodic", "reflection", "none"][bc])
    np.testing.assert_array_equal(py, on.images(lat, bc, ev))


def test_reflection_images_are_mirror_images():
    """Each reflection image lies as far outside a face as the bubble lies inside it."""
    ev = np.array([[0.3, 0.2, 0.7, 0.9]])
    img = on.images(BOX, 1, ev)
    expect = [[-0.2, 0.7, 0.9], [1.8, 0.7, 0.9], [0.2, -0.7, 0.9], [0.2, 1.3, 0.9],
              [0.2, 0.7, -0.9], [0.2, 0.7, 1.1]]
    np.testing.assert_allclose(img[:, 1:], expect, atol=1e-15)
    assert (img[:, 0] == 0.3).all()
    s = on.images(SPHERE, 1, np.array([[0.0, 0.1, -0.2, 1.3]]))  # 1.0 from the center along z
    np.testing.assert_allclose(s[:, 1:], [[0.1, -0.2, 1.9], [0.1, -0.2, -3.3]], atol=1e-14)


def test_python_lattice_classes():
    c = mb.CartesianLattice.with_origin_and_sizes((0.0, 0.0, 0.0), (2.0, 3.0, 4.0))
    assert c.volume == 24.0 and c.side_lengths == (2.0, 3.0, 4.0) and c.name() == "CartesianLattice"
    with pytest.raises(ValueError, match="not pairwise orthogonal"):
        mb.CartesianLattice((0, 0, 0), [[1, 0, 0], [1, 1, 0], [0, 0, 1]])
    with pytest.raises(ValueError, match="linearly dependent"):
        mb.ParallelepipedLattice((0, 0, 0), [[1, 0, 0], [2, 0, 0], [0, 0, 1]])
    s = mb.SphericalLattice((0, 0, 0), 2.0)
    assert s.volume == pytest.approx(4.0 / 3.0 * math.pi * 8.0, rel=1e-15)
    assert s.contains([[0, 0, 2.0], [0, 0, 2.0000001]]) == [True, False]
    assert c.contains([[1.0, 1.0, 1.0], [2.0, 3.0, 4.0], [2.1, 0, 0]]) == [True, True, False]
    iso = mb.Isometry3((1.0, 0.0, 0.0), euler_angles=(0.0, 0.0, math.pi / 2))
    moved = c.transform(iso)
    np.testing.assert_allclose(moved.origin, (1.0, 0.0, 0.0), atol=1e-15)
    np.testing.assert_allclose(moved.basis, [[0.0, 2.0, 0.0], [-3.0, 0.0, 0.0], [0.0, 0.0, 4.0]],
                               atol=1e-15)
    both = iso.transform(mb.Isometry3.from_translation((0.0, 1.0, 0.0)))
    np.testing.assert_allclose(both.translation, (0.0, 0.0, 0.0), atol=1e-15)
    with pytest.raises(ValueError, match="either euler_angles or rotation_matrix"):
        mb.Isometry3(euler_angles=(0, 0, 0), rotation_matrix=np.eye(3))
    r = mb.Isometry3.from_axis_angle((0, 0, 2.0), math.pi / 2).rotation_matrix
    np.testing.assert_allclose(r, [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-15)


def test_lattice_bubbles_surface():
    lat = mb.CartesianLattice((0, 0, 0), np.eye(3) * 2.0)
    lb = mb.LatticeBubbles(lat, np.array([[0.5, 1.0, 1.0, 1.0], [0.7, 0.2, 0.2, 1.8]]))
    assert lb.lattice_type == "Cartesian" and "n_interior=2" in repr(lb)
    lb.with_boundary_condition("reflection")
    assert lb.bubbles_exterior.shape == (12, 4)
    lb.with_boundary_condition("none")
    assert lb.bubbles_exterior.shape == (0, 4)
    lb.with_boundary_condition("periodic")
    np.testing.assert_array_equal(lb.bubbles_exterior[0], [0.5, 3.0, 1.0, 1.0])
    with pytest.raises(ValueError, match="Invalid boundary condition"):
        lb.with_boundary_condition("mirror")
    shifted = lb.translate_time(-0.5)
    assert shifted.bubbles_interior[0, 0] == 0.0 and lb.bubbles_interior[0, 0] == 0.5
    lb.normalize_time_mut()
    assert lb.bubbles_interior[:, 0].min() == 0.0 and lb.bubbles_exterior[:, 0].min() == 0.0
    assert lb.parameters() == [0.0, 0.0, 0.0, 2.0, 0.0, 0.0, 0.0, 2.0, 0.0, 0.0, 0.0, 2.0]
    moved = lb.transform(mb.Isometry3.from_translation((1.0, 0.0, 0.0)))
    np.testing.assert_array_equal(moved.bubbles_interior[:, 1], lb.bubbles_interior[:, 1] + 1.0)
    assert moved.reference_point == (1.0, 0.0, 0.0)
    with pytest.raises(ValueError, match="formed outside lattice"):
        mb.LatticeBubbles(lat, np.array([[0.0, 3.0, 1.0, 1.0]]))
    sph = mb.LatticeBubbles(mb.SphericalLattice((0, 0, 0), 1.0), np.array([[0.0, 0.5, 0.0, 0.0]]))
    assert sph.lattice_type == "Spherical"


# --------------------------------------------------------------------------- fixed rate
def _covered_by(points, bubbles, t):
    """Brute force: True where some bubble (t_b, x_b) has reached the point by time t."""
    if len(bubbles) == 0:
        return np.zeros(len(points), dtype=bool)
    d = np.sqrt(((points[:, None, :] - bubbles[None, :, 1:]) ** 2).sum(-1))
    return (bubbles[None, :, 0] + d <= t).any(axis=1)


@pytest.mark.parametrize("lat,bc,method", [(BOX, 0, 0), (SKEW, 1, 0), (SPHERE, 1, 0), (BOX, 0, 1),
                                           (SPHERE, 2, 1)])
def test_fixed_rate_decisions_by_brute_force(lat, bc, method):
    """Every nucleation site is the first sample point no earlier bubble or image has
    reached, the recorded volume is the uncovered fraction, images are the lattice's."""
    n_points = 600
    times = None
    beta, gamma0 = 8.0, 8.0**4 * 0.02
    if method == 1:
        times = np.linspace(0.05, 1.5, 200)
    interior, exterior, th, vh = on.fixed_rate(lat, bc, method, beta=beta, gamma0=gamma0, t0=0.0,
                                               d_p0=0.05, seed=17, stream=2, n_points=n_points,
                                               times=times)
    assert len(interior) > 5
    pts = on.sample_points(lat, 17, 2, n_points)
    n_img = on.images_per_bubble(lat[0], bc)
    np.testing.assert_array_equal(exterior, on.images(lat, bc, interior))
    vol = on.volume(lat)
    for k in range(len(interior)):
        t = interior[k, 0]
        prev = np.vstack([interior[:k], exterior[:k * n_img]]) if k else np.zeros((0, 4))
        cov = _covered_by(pts, prev, t)
        site = int(np.argmin(cov))  # first uncovered point
        assert not cov[site]
        np.testing.assert_array_equal(interior[k, 1:], pts[site])
        assert vh[k] == max((n_points - cov.sum()) / n_points * vol, 0.0)
        assert th[k] == (t if method == 0 else times[k])
    assert np.all(np.diff(interior[:, 0]) >= 0)


def test_time_stepping_recurrence_before_the_first_bubble():
    """Before the first bubble the step is d_p0 / Gamma(t) (full false vacuum), and the
    first bubble comes at the first iteration whose draw is <= d_p0."""
    beta, gamma0, d_p0, t0 = 5.0, 3.0, 0.02, 0.25
    interior, _, th, _ = on.fixed_rate(BOX, 0, 0, beta=beta, gamma0=gamma0, t0=t0, d_p0=d_p0,
                                       seed=99, stream=5, n_points=300)
    t = t0
    it = 0
    while True:
        x = _uniform(on.philox4x64([it >> 2, 1, 0, 0], [99, 5])[it & 3])
        if x <= d_p0:
            break
        t += (d_p0 / (gamma0 * math.exp(beta * (t - t0)))) * (1.0 / 1.0)
        it += 1
    assert th[0] == t and interior[0, 0] == t


def test_fixed_rate_independent_restatement():
    """A second, vectorised restatement (coverage recomputed from all bubbles at every
    iteration instead of kept per point) gives the same run."""
    n_points, beta, gamma0, d_p0 = 400, 6.0, 6.0**4 * 0.05, 0.1
    interior, exterior, th, vh = on.fixed_rate(BOX, 0, 0, beta=beta, gamma0=gamma0, t0=0.0,
                                               d_p0=d_p0, seed=3, stream=1, n_points=n_points,
                                               max_time_iterations=3000, cutoff_fraction=0.02)
    pts = on.sample_points(BOX, 3, 1, n_points)
    shifts = [np.eye(3)[j] * s for j in range(3) for s in (1.0, -1.0)]
    bubbles = np.zeros((0, 4))
    got = []
    t = 0.0
    for it in range(3000):
        vrem = max((~_covered_by(pts, bubbles, t)).sum() / n_points * 1.0, 0.0)
        if vrem < 0.02:
            break
        gamma_t = gamma0 * math.exp(beta * (t - 0.0))
        x = _uniform(on.philox4x64([it >> 2, 1, 0, 0], [3, 1])[it & 3])
        if x <= d_p0:
            free = np.nonzero(~_covered_by(pts, bubbles, t))[0]
            if len(free):
                b = np.array([t, *pts[free[0]]])
                imgs = [np.array([t, *(b[1:] + sh)]) for sh in shifts]
                bubbles = np.vstack([bubbles, b, *imgs])
                got.append((b, vrem))
        t += (d_p0 / gamma_t) * (1.0 / vrem)
    assert len(got) == len(interior) > 10
    np.testing.assert_array_equal(np.array([g[0] for g in got]), interior)
    np.testing.assert_array_equal(np.array([g[1] for g in got]), vh)


def test_spontaneous_oracle():
    interior, exterior = on.spontaneous(BOX, 0, 12, 0.5, 8, 3)
    assert (interior[:, 0] == 0.5).all() and on.contains(BOX, interior[:, 1:]).all()
    np.testing.assert_array_equal(exterior, on.images(BOX, 0, interior))
    x = on.philox4x64([0, 2, 0, 0], [8, 3])
    np.testing.assert_array_equal(interior[0, 1:], [_uniform(v) for v in x[:3]])
    s_int, s_ext = on.spontaneous(SPHERE, 1, 30, 0.0, 8, 3)
    assert on.contains(SPHERE, s_int[:, 1:]).all() and s_ext.shape == (60, 4)


# --------------------------------------------------------------------------- ODE (library host code)
def _solve(beta, gamma0, t0, taumax, vol):
    L = _lib.lib()
    n = ctypes.c_size_t(0)
    _lib.check(L.bgw_nucleation_solve_distribution(beta, gamma0, t0, taumax, vol, None, None, 0,
                                                   ctypes.byref(n)))
    t = np.zeros(n.value)
    nb = np.zeros(n.value)
    _lib.check(L.bgw_nucleation_solve_distribution(beta, gamma0, t0, taumax, vol, _lib.dptr(t),
                                                   _lib.dptr(nb), n.value, ctypes.byref(n)))
    return t, nb


@pytest.mark.parametrize("beta,gamma0,vol", [(1.0, 1.0, 30.0), (20.0, 20.0**4 * 0.01, 1.0),
                                             (12.0, 0.3 * 12.0**4, 2.5)])
def test_distribution_times_match_radau(beta, gamma0, vol):
    """The crossing times equal those of scipy's Radau IIA (order 5) at the reference's
    tolerances (rtol 1e-9, atol 1e-12) on the same grid of step 0.001 in tau."""
    from scipy.integrate import solve_ivp
    t, nb = _solve(beta, gamma0, 0.0, 12.0, vol)
    g = math.log(gamma0 * beta**-4)

    def f(tau, m):
        return [math.exp(g + tau - 4.0 * math.pi / 3.0 * m[3]), m[0], 2.0 * m[1], 3.0 * m[2]]

    grid = np.arange(0, 12001) * 0.001
    sol = solve_ivp(f, (0.0, 12.0), [0.0] * 4, method="Radau", rtol=1e-9, atol=1e-12, t_eval=grid)
    n = sol.y[0] * vol * beta**3
    fl = np.floor(n)
    idx = np.nonzero(fl[1:] > fl[:-1])[0] + 1
    assert len(idx) == len(t) > 3
    np.testing.assert_allclose(t, grid[idx] / beta, rtol=1e-12)
    np.testing.assert_allclose(nb, n[idx], rtol=1e-6)


def test_distribution_python_surface():
    fr = mb.FixedRateNucleation(beta=2.0, gamma0=1.0, t0=0.0, d_p0=0.01, seed=0)
    t, n = fr.solve_bubbles_distribution(20.0, 8.0)
    assert len(t) == len(n) > 0 and np.all(np.diff(t) > 0) and np.all(np.diff(np.floor(n)) >= 1)


# --------------------------------------------------------------------------- C ABI and Python validation
def test_nucleation_argument_errors():
    with pytest.raises(ValueError, match=r"d_p0 must be in \(0, 1\]"):
        mb.FixedRateNucleation(1.0, 1.0, 0.0, d_p0=0.0)
    with pytest.raises(ValueError, match="gamma0 must be non-negative"):
        mb.FixedRateNucleation(1.0, -1.0, 0.0, d_p0=0.1)
    with pytest.raises(ValueError, match="n_points must be > 0"):
        mb.FixedRateNucleation(1.0, 1.0, 0.0, d_p0=0.1, n_points=0)
    with pytest.raises(ValueError, match="Parameters must be finite"):
        mb.FixedRateNucleation(float("nan"), 1.0, 0.0, d_p0=0.1)
    fr = mb.FixedRateNucleation(1.0, 1.0, 0.0, d_p0=0.1, seed=4)
    assert repr(fr) == "FixedRateNucleation(beta=1, gamma0=1, t0=0, seed=Some(4), n_points=10000)"
    assert repr(mb.SpontaneousNucleation(4, seed=None)).endswith("seed=None)")
    with pytest.raises(ValueError, match="Invalid boundary condition"):
        fr.nucleate(mb.CartesianLattice((0, 0, 0), np.eye(3)), boundary_condition="mirror")
    with pytest.raises(ValueError, match="method must be"):
        fr.nucleate(mb.CartesianLattice((0, 0, 0), np.eye(3)), method="approximation")
    with pytest.raises(ValueError, match="EmptyLattice"):
        fr.nucleate(mb.EmptyLattice())


def test_c_abi_nucleation_validation():
    L = _lib.lib()
    r = (_lib.BgwFixedRateRealisation * 1)()
    r[0].lattice.kind = 0
    r[0].lattice.basis[:] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    r[0].beta, r[0].gamma0 = 1.0, 1.0
    buf = np.zeros(64)
    cnt = np.zeros(1, dtype=np.uintp)
    args = lambda d_p0, bc, npts: (r, 1, bc, 0, d_p0, npts, 10, 0.01, 4, _lib.dptr(buf),  # noqa: E731
                                   _lib.dptr(buf), _lib.dptr(buf), _lib.dptr(buf), _lib.sptr(cnt))
    assert L.bgw_nucleation_fixed_rate(*args(2.0, 0, 10)) == _lib.CALCULATION_ERROR
    assert "d_p0 must be in (0, 1]" in _lib.last_error()
    assert L.bgw_nucleation_fixed_rate(*args(0.5, 7, 10)) == _lib.CALCULATION_ERROR
    assert "Invalid boundary condition" in _lib.last_error()
    assert L.bgw_nucleation_fixed_rate(*args(0.5, 0, 0)) == _lib.CALCULATION_ERROR
    assert "n_points must be > 0" in _lib.last_error()
    r[0].lattice.basis[:] = [1, 0, 0, 2, 0, 0, 0, 0, 1]
    assert L.bgw_nucleation_fixed_rate(*args(0.5, 0, 10)) == _lib.CALCULATION_ERROR
    assert "degenerate basis" in _lib.last_error()
    assert L.bgw_nucleation_images_per_bubble(0, 0) == 6
    assert L.bgw_nucleation_images_per_bubble(1, 1) == 2
    assert L.bgw_nucleation_images_per_bubble(1, 0) == 0
    assert L.bgw_nucleation_images_per_bubble(0, 2) == 0
