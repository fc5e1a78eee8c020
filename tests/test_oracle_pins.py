"""Pins for the CPU oracle (CPU only).

The reference (Rust) cannot be built here and ships no golden vectors for this
path (its test CSV rust/bubble-gw/tests/envelope_2bubbles_scan=n_phix.csv is
not in the tree), so the oracle is pinned by what the mathematics fixes:

* an independent numpy restatement of the integrals (tests/independent_numpy.py),
* the closed-form collision delay of two simultaneous bubbles,
* a brute-force time-stepping check of every first collision,
* invariants: isolated bubble, linearity in the coefficients, additivity over
  bubbles, the phase of a z-translation, the identity rotation,
* and it is shown that planted mistakes (phase sign, quadrature end weights,
  tie handling) fail these pins.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import independent_numpy as ind
import oracle
from gw_fopt_amd import configs, nucleation

TWO = np.array([[0.0, 0.0, 10.0, 0.0], [0.0, 0.0, 0.0, 0.0]])


def async_bubbles(seed, n=5, beta=12.0):
    b = nucleation.exponential_nucleation(beta, n_max=n, seed=seed)
    return b, nucleation.face_images(b, 1.0)


def rel_max(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


# ---------------------------------------------------------------------------
# independent restatement
# ---------------------------------------------------------------------------
def test_reference_capi_case_matches_numpy():
    args = (TWO, None, 4, 5, [[0.0], [1.0]], [[3.0], [3.0]], [0.5, 1.0])
    C = oracle.compute_c_integral(*args, t_begin=0.0, t_end=8.0, n_t=40)
    Cn = ind.c_integral(*args, 0.0, 8.0, 40)
    assert rel_max(C, Cn) < 1e-13


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_async_periodic_matches_numpy(seed):
    inter, ext = async_bubbles(seed)
    w = np.geomspace(0.5, 60.0, 7)
    coef, pows = [[1.0, 0.4], [0.0, 0.0]], [[5.0, 3.5], [3.0, 3.0]]
    C = oracle.compute_c_integral(inter, ext, 9, 12, coef, pows, w, t_begin=0.0, t_end=0.8,
                                  n_t=35)
    Cn = ind.c_integral(inter, ext, 9, 12, coef, pows, w, 0.0, 0.8, 35)
    assert rel_max(C, Cn) < 1e-12


# ---------------------------------------------------------------------------
# masks
# ---------------------------------------------------------------------------
def test_two_bubble_closed_form_delay():
    """a at y = 10, b at the origin, both at t = 0: along x-hat the wall of a meets b's
    wall when |x_a + dt x - x_b| = dt, i.e. dt = d / (2 x.d_hat) for x.d_hat > 0."""
    n_c, n_phi = 9, 16
    first, dtab = oracle.first_colliding(TWO, None, n_c, n_phi, 0)
    _, _, dirs = ind.directions(n_c, n_phi)
    d_hat = np.array([0.0, -1.0, 0.0])  # from a to b
    proj = dirs[..., 1:] @ d_hat
    hit = 10.0 * proj > 1e-10
    assert np.array_equal(first >= 0, hit)
    np.testing.assert_allclose(dtab[hit], 10.0 / (2.0 * proj[hit]), rtol=1e-14)
    assert np.all(dtab[~hit] == 0.0)


def _covered(p, t, bubbles, exclude):
    for c, bc in enumerate(bubbles):
        if c in exclude:
            continue
        if np.linalg.norm(p - bc[1:]) < t - bc[0]:
            return True
    return False


@pytest.mark.parametrize("seed", [6, 8])
def test_first_collision_bruteforce(seed):
    """Just before the delay the wall point is outside every other bubble; just after it
    is inside the first colliding one."""
    inter, ext = async_bubbles(seed, n=6)
    allb = np.vstack([inter, ext])
    n_c, n_phi = 7, 10
    _, _, dirs = ind.directions(n_c, n_phi)
    for a in range(len(inter)):
        first, dtab = oracle.first_colliding(inter, ext, n_c, n_phi, a)
        for i in range(n_c):
            for j in range(n_phi):
                b = first[i, j]
                if b == -1:
                    continue
                bt = b if b >= 0 else len(inter) + (-b - 2)
                x = dirs[i, j, 1:]
                dt = dtab[i, j]
                ta = inter[a, 0]
                before = inter[a, 1:] + dt * (1 - 1e-7) * x
                after = inter[a, 1:] + dt * (1 + 1e-7) * x
                assert not _covered(before, ta + dt * (1 - 1e-7), allb, {a})
                assert np.linalg.norm(after - allb[bt, 1:]) < ta + dt * (1 + 1e-7) - allb[bt, 0]


def test_exact_tie_is_none():
    inter = np.array([[0.0, 0.0, 0.0, 0.0], [0.0, 0.5, 0.5, 0.0], [0.0, 0.5, -0.5, 0.0]])
    first, dtab = oracle.first_colliding(inter, None, 3, 5, 0)
    assert first[1, 0] == -1 and dtab[1, 0] == 0.0


def test_numpy_first_collision_agrees():
    inter, ext = async_bubbles(9, n=6)
    allb = np.vstack([inter, ext])
    _, _, dirs = ind.directions(8, 12)
    for a in range(len(inter)):
        first, dtab = oracle.first_colliding(inter, ext, 8, 12, a)
        fn, dn = ind.first_collision(allb, a, len(inter), dirs)
        code = np.where(fn < 0, -1, np.where(fn < len(inter), fn, -(fn - len(inter)) - 2))
        assert np.array_equal(first, code)
        np.testing.assert_allclose(dtab, dn, rtol=1e-13)


# ---------------------------------------------------------------------------
# invariants
# ---------------------------------------------------------------------------
def test_isolated_bubble_vanishes():
    one = np.array([[0.0, 0.3, 0.4, 0.5]])
    C = oracle.compute_c_integral(one, None, 12, 16, [[1.0]], [[5.0]], [0.5, 3.0, 20.0],
                                  t_end=1.0, n_t=50)
    assert np.max(np.abs(C)) < 1e-14


def test_linear_in_coefficients():
    inter, ext = async_bubbles(10)
    w = np.geomspace(0.5, 50, 5)
    kw = dict(t_end=0.8, n_t=30)
    c = [oracle.compute_c_integral(inter, ext, 8, 12, [[a]], [[5.0]], w, **kw) for a in (0.0, 1.0, 0.37)]
    np.testing.assert_allclose(c[2], c[0] + 0.37 * (c[1] - c[0]), rtol=0, atol=1e-13 * np.abs(c[1]).max())


def test_tangent_coef_is_the_linear_difference():
    """C is linear in each coefficient: dC/da_k = C(a_k = 1) - C(a_k = 0), other terms fixed."""
    inter, ext = async_bubbles(13)
    w = np.geomspace(0.5, 50, 5)
    kw = dict(t_end=0.8, n_t=30, k_hat=[0.6, 0.0, 0.8])
    pows = [[5.0, 3.5]]
    T = oracle.compute_c_integral(inter, ext, 8, 12, [[0.7, 0.2]], pows, w, derivative=("pow", 0),
                                  **kw)
    d1 = oracle.compute_c_integral(inter, ext, 8, 12, [[0.7, 0.2]], pows, w,
                                   derivative=("coef", 1), **kw)
    hi = oracle.compute_c_integral(inter, ext, 8, 12, [[0.7, 1.0]], pows, w, **kw)
    lo = oracle.compute_c_integral(inter, ext, 8, 12, [[0.7, 0.0]], pows, w, **kw)
    np.testing.assert_allclose(d1, hi - lo, rtol=0, atol=1e-12 * np.abs(hi - lo).max())
    assert np.abs(T).max() > 0 and np.abs(d1).max() > 0


@pytest.mark.parametrize("damping", [None, 0.3])
def test_tangent_pow_matches_central_differences(damping):
    """dC/dp_k (forward mode, f -> a ratio^p ln ratio) against (C(p+h) - C(p-h)) / 2h."""
    inter, ext = async_bubbles(14)
    w = np.geomspace(0.5, 50, 5)
    kw = dict(t_end=0.8, n_t=30, damping=damping, k_hat=[0.0, 0.6, 0.8])
    coef = [[0.7, 0.4], [1.1, 0.0]]
    T = oracle.compute_c_integral(inter, ext, 8, 12, coef, [[5.0, 3.5], [4.2, 3.0]], w,
                                  derivative=("pow", 1), **kw)
    h = 1e-5
    up = oracle.compute_c_integral(inter, ext, 8, 12, coef, [[5.0, 3.5 + h], [4.2, 3.0 + h]], w, **kw)
    dn = oracle.compute_c_integral(inter, ext, 8, 12, coef, [[5.0, 3.5 - h], [4.2, 3.0 - h]], w, **kw)
    fd = (up - dn) / (2 * h)
    np.testing.assert_allclose(T[:, 0], fd[:, 0], rtol=0, atol=1e-7 * np.abs(fd[:, 0]).max())
    assert np.abs(T[:, 1]).max() == 0.0  # coefficient 0: the set does not depend on p


def test_sets_equal_single_runs():
    inter, ext = async_bubbles(11)
    w = np.geomspace(0.5, 50, 5)
    both = oracle.compute_c_integral(inter, ext, 8, 12, [[1.0], [0.5]], [[5.0], [4.0]], w, t_end=0.8, n_t=30)
    one = oracle.compute_c_integral(inter, ext, 8, 12, [[0.5]], [[4.0]], w, t_end=0.8, n_t=30)
    assert np.array_equal(both[:, 1], one[:, 0])


def test_selected_bubbles_add_up():
    inter, ext = async_bubbles(12)
    w = np.geomspace(0.5, 50, 5)
    kw = dict(t_end=0.8, n_t=30)
    total = oracle.compute_c_integral(inter, ext, 8, 12, [[1.0]], [[5.0]], w, **kw)
    parts = sum(oracle.compute_c_integral(inter, ext, 8, 12, [[1.0]], [[5.0]], w,
                                          selected_bubbles=[a], **kw) for a in range(len(inter)))
    np.testing.assert_allclose(parts, total, rtol=0, atol=1e-13 * np.abs(total).max())


def test_z_translation_multiplies_by_phase():
    """C ~ sum e^{i w (t - z_a)}: shifting every z by Z multiplies C by e^{-i w Z}."""
    inter, ext = async_bubbles(13)
    Z = 0.375
    sh = lambda b: b + np.array([0.0, 0.0, 0.0, Z])
    w = np.geomspace(0.5, 50, 6)
    kw = dict(t_end=0.8, n_t=30)
    C = oracle.compute_c_integral(inter, ext, 8, 12, [[1.0]], [[5.0]], w, **kw)
    Cz = oracle.compute_c_integral(sh(inter), sh(ext), 8, 12, [[1.0]], [[5.0]], w, **kw)
    np.testing.assert_allclose(Cz, C * np.exp(-1j * w * Z), rtol=0, atol=1e-9 * np.abs(C).max())


def test_identity_rotation_is_bitexact():
    inter, ext = async_bubbles(14)
    w = np.geomspace(0.5, 50, 5)
    kw = dict(t_end=0.8, n_t=30)
    C = oracle.compute_c_integral(inter, ext, 8, 12, [[1.0]], [[5.0]], w, **kw)
    Ck = oracle.compute_c_integral(inter, ext, 8, 12, [[1.0]], [[5.0]], w, k_hat=[0, 0, 1.0], **kw)
    assert np.array_equal(C, Ck)


def test_rotation_maps_k_to_z():
    for k in np.random.default_rng(0).normal(size=(5, 3)):
        k = k / np.linalg.norm(k)
        R = oracle.rotation_to_z(k)
        np.testing.assert_allclose(R @ k, [0, 0, 1], atol=1e-15)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-15)
        assert abs(np.linalg.det(R) - 1) < 1e-14


def rotate(b, R):
    """Bubbles (t, x, y, z) with positions rotated by R."""
    b = np.asarray(b, float).reshape(-1, 4)
    return np.column_stack([b[:, 0], b[:, 1:] @ R.T])


@pytest.mark.parametrize("seed", [4, 16])
def test_k_hat_frame_matches_numpy_on_rotated_bubbles(seed):
    """The oracle's k-hat frame (positions and deltas rotated, delta^2 kept,
    lattice_bubbles.rs:144-197) against the restatement run on bubbles rotated
    beforehand (delta^2 recomputed): masks equal, delays and C to rounding."""
    inter, ext = async_bubbles(seed, n=6)
    k = np.random.default_rng(seed).normal(size=3)
    k /= np.linalg.norm(k)
    R = oracle.rotation_to_z(k)
    ri, re = rotate(inter, R), rotate(ext, R)
    allb = np.vstack([ri, re])
    _, _, dirs = ind.directions(8, 12)
    for a in range(len(inter)):
        first, dtab = oracle.first_colliding(inter, ext, 8, 12, a, k_hat=k)
        fn, dn = ind.first_collision(allb, a, len(inter), dirs)
        code = np.where(fn < 0, -1, np.where(fn < len(inter), fn, -(fn - len(inter)) - 2))
        assert np.array_equal(first, code), a
        np.testing.assert_allclose(dtab, dn, rtol=1e-12)
    w = np.geomspace(0.5, 60.0, 7)
    coef, pows = [[1.0, 0.4], [0.7, 0.0]], [[5.0, 3.5], [4.0, 3.0]]
    C = oracle.compute_c_integral(inter, ext, 8, 12, coef, pows, w, t_end=0.8, n_t=35, k_hat=k)
    Cn = ind.c_integral(ri, re, 8, 12, coef, pows, w, 0.0, 0.8, 35)
    assert rel_max(C, Cn) < 1e-12
    # the spectrum of that frame: 8 pi w^3 (|C+|^2 + |Cx|^2)
    S = oracle.spectrum(inter, ext, [k], 8, 12, coef, pows, w, t_end=0.8, n_t=35)
    Sn = 8 * np.pi * w**3 * np.sum(np.abs(Cn) ** 2, axis=0)
    np.testing.assert_allclose(S[0], Sn, rtol=1e-11)


@pytest.mark.parametrize("seed", [3, 15])
def test_integrand_matches_numpy_and_sums_to_integral(seed):
    """compute_c_integrand (generalized_bulk_flow.rs:666-767, 875-920): each time step's
    term against the restatement's per-step terms, and their sum equals
    compute_c_integral with the unit cutoff."""
    inter, ext = async_bubbles(seed)
    w = np.geomspace(0.5, 60.0, 6)
    coef, pows = [[1.0, 0.4], [0.7, 0.0]], [[5.0, 3.5], [4.0, 3.0]]
    I = oracle.compute_c_integrand(inter, ext, 9, 12, coef, pows, w, t_begin=0.05, t_end=0.8,
                                   n_t=33)
    In = ind.c_integrand(inter, ext, 9, 12, coef, pows, w, 0.05, 0.8, 33)
    assert I.shape == (2, 2, 6, 33)
    scale = np.max(np.abs(In), axis=-1, keepdims=True)
    assert np.max(np.abs(I - In) / scale) < 1e-12
    C = oracle.compute_c_integral(inter, ext, 9, 12, coef, pows, w, t_begin=0.05, t_end=0.8,
                                  n_t=33)
    np.testing.assert_allclose(I.sum(axis=-1), C, rtol=0, atol=1e-13 * np.abs(C).max())
    # a bubble contributes nothing at or before its nucleation time
    late = int(np.argmax(inter[:, 0]))
    t = np.linspace(0.0, 0.8, 33)
    Il = oracle.compute_c_integrand(inter, ext, 9, 12, coef, pows, w, t_end=0.8, n_t=33,
                                    selected_bubbles=[late])
    assert np.all(Il[..., t <= inter[late, 0]] == 0)
    assert np.any(Il[..., t > inter[late, 0]] != 0)


def exponential_cutoff(t, smax, ratio_t_cut=None, ratio_t_0=None):
    """ExponentialTimeCutoff as documented (time_cutoff.rs:14-43): weight 1 before
    t_cut = r_cut smax (r_cut = 0.999999999 by default), a Gaussian of width
    t_0 = r_0 (smax - t_cut) (r_0 = 0.25 by default) centred on t_cut after it."""
    t_cut = (0.999999999 if ratio_t_cut is None else ratio_t_cut) * smax
    t_0 = (0.25 if ratio_t_0 is None else ratio_t_0) * (smax - t_cut)
    t = np.asarray(t, float)
    return np.where(t < t_cut, 1.0, np.exp(-(((t - t_cut) / t_0) ** 2)))


@pytest.mark.parametrize("cut", [(0.7, 0.6, 0.8), (0.8, None, None), (0.5, 0.3, None)])
def test_exponential_cutoff_weights_each_step(cut):
    """C with an exponential cutoff equals the sum over steps of the cutoff weight at
    t_k times the step's integrand term, and the restatement with that weight."""
    smax, rc, r0 = cut
    inter, ext = async_bubbles(7)
    w = np.geomspace(0.5, 60.0, 6)
    coef, pows = [[1.0], [0.6]], [[5.0], [3.5]]
    kw = dict(t_begin=0.0, t_end=0.8, n_t=41)
    I = oracle.compute_c_integrand(inter, ext, 9, 12, coef, pows, w, **kw)
    C = oracle.compute_c_integral(inter, ext, 9, 12, coef, pows, w, time_cutoff=(1, smax, rc, r0),
                                  **kw)
    t = np.linspace(0.0, 0.8, 41)
    wt = exponential_cutoff(t, smax, rc, r0)
    assert wt.min() < 0.99  # the cutoff acts inside the time range
    np.testing.assert_allclose(C, np.sum(I * wt, axis=-1), rtol=0, atol=1e-13 * np.abs(C).max())
    Cn = ind.c_integral(inter, ext, 9, 12, coef, pows, w, 0.0, 0.8, 41,
                        cutoff=lambda tt: exponential_cutoff(tt, smax, rc, r0))
    assert rel_max(C, Cn) < 1e-12
    # the weight's shape: one at t_cut, 1/e one width later
    t_cut = (0.999999999 if rc is None else rc) * smax
    t_0 = (0.25 if r0 is None else r0) * (smax - t_cut)
    np.testing.assert_allclose(exponential_cutoff([t_cut, t_cut + t_0], smax, rc, r0),
                               [1.0, np.exp(-1.0)], rtol=1e-6)


def test_errors_follow_reference():
    with pytest.raises(oracle.OracleError, match="formed inside bubble"):
        oracle.lattice(np.array([[0.0, 0, 0, 0], [1.0, 0.5, 0, 0]]))
    with pytest.raises(oracle.OracleError, match="n_t must be >= 2"):
        oracle.compute_c_integral(TWO, None, 4, 5, [[1.0]], [[3.0]], [1.0], t_end=1.0, n_t=1)
    with pytest.raises(oracle.OracleError, match="out of bounds"):
        oracle.compute_c_integral(TWO, None, 4, 5, [[1.0]], [[3.0]], [1.0], t_end=1.0, n_t=5,
                                  selected_bubbles=[2])
    with pytest.raises(oracle.OracleError, match="Invalid time range"):
        oracle.compute_c_integral(TWO, None, 4, 5, [[1.0]], [[3.0]], [1.0], t_begin=2.0,
                                  t_end=1.0, n_t=5)
    with pytest.raises(oracle.OracleError, match="Angular resolutions"):
        oracle.grids(1, 5)
    # an empty selection short-circuits before the time-range check (:955-957)
    z = oracle.compute_c_integral(TWO, None, 4, 5, [[1.0]], [[3.0]], [1.0], t_begin=2.0,
                                  t_end=1.0, n_t=5, selected_bubbles=[])
    assert np.all(z == 0)


# ---------------------------------------------------------------------------
# planted mistakes fail the pins
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = os.path.join(os.path.dirname(oracle.__file__), "bulkflow_oracle.cpp")
    out = {}
    d = tmp_path_factory.mktemp("planted")
    for plant in (1, 2, 3, 4, 5):
        so = str(d / f"planted{plant}.so")
        subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off",
                        f"-DORACLE_PLANT={plant}", "-shared", "-o", so, src], check=True)
        out[plant] = so
    return out


def _with_lib(path, fn):
    saved = oracle._lib
    try:
        oracle._lib = None
        old = oracle._LIB_PATH
        oracle._LIB_PATH = path
        return fn()
    finally:
        oracle._LIB_PATH = old
        oracle._lib = saved


@pytest.mark.parametrize("plant", [1, 2])
def test_planted_quadrature_mistakes_are_caught(planted, plant):
    inter, ext = async_bubbles(3)
    w = np.geomspace(0.5, 60.0, 7)
    args = (inter, ext, 9, 12, [[1.0]], [[5.0]], w)
    C = _with_lib(planted[plant], lambda: oracle.compute_c_integral(*args, t_begin=0.0,
                                                                     t_end=0.8, n_t=35))
    Cn = ind.c_integral(*args, 0.0, 0.8, 35)
    assert rel_max(C, Cn) > 1e-6


def test_planted_tie_mistake_is_caught(planted):
    inter = np.array([[0.0, 0.0, 0.0, 0.0], [0.0, 0.5, 0.5, 0.0], [0.0, 0.5, -0.5, 0.0]])
    first, _ = _with_lib(planted[3], lambda: oracle.first_colliding(inter, None, 3, 5, 0))
    assert first[1, 0] != -1


@pytest.mark.parametrize("plant", [4, 5])
def test_planted_frame_mistakes_are_caught(planted, plant):
    """The k-hat frame rotated the wrong way, or positions rotated without the deltas:
    the comparison with the restatement on pre-rotated bubbles fails."""
    inter, ext = async_bubbles(4)
    k = np.array([0.48, -0.6, 0.64])
    w = np.geomspace(0.5, 60.0, 7)
    R = oracle.rotation_to_z(k)
    C = _with_lib(planted[plant], lambda: oracle.compute_c_integral(
        inter, ext, 9, 12, [[1.0]], [[5.0]], w, t_end=0.8, n_t=35, k_hat=k))
    Cn = ind.c_integral(rotate(inter, R), rotate(ext, R), 9, 12, [[1.0]], [[5.0]], w, 0.0, 0.8, 35)
    assert rel_max(C, Cn) > 1e-6


# ---------------------------------------------------------------------------
# golden fixtures (written by tests/golden/make_golden.py from the oracle)
# ---------------------------------------------------------------------------
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_golden_reference_capi_case():
    g = load("capi_two_bubbles.npz")
    C = oracle.compute_c_integral(g["interior"], None, 4, 5, g["coef"], g["pows"], g["w"],
                                  t_begin=0.0, t_end=8.0, n_t=40)
    assert np.array_equal(C, g["C"])
    assert rel_max(g["C"], ind.c_integral(g["interior"], None, 4, 5, g["coef"], g["pows"],
                                          g["w"], 0.0, 8.0, 40)) < 1e-13


def test_golden_plain_8_spectrum_and_masks():
    g = load("plain_8.npz")
    wl = configs.plain_8()
    assert np.array_equal(wl.interior, g["interior"])
    S = oracle.spectrum(g["interior"], g["exterior"], g["k_hat"], 32, 64, [[1.0]], [[5.0]],
                        g["w"], t_end=1.5, n_t=200)
    assert np.array_equal(S, g["S"])
    first, dtab = oracle.first_colliding(g["interior"], g["exterior"], 32, 64, 0,
                                         k_hat=g["k_hat"][3])
    assert np.array_equal(first, g["first_k3_a0"])
    assert np.array_equal(dtab, g["dtab_k3_a0"])
