"""Batched ensembles on the GPU: one persistent, graph-captured launch.

Each realisation's spectrum from BubbleEnsemble equals, bit for bit, the
spectrum of GeneralizedBulkFlow.compute_spectrum on that realisation alone
(same additions in the same order); FP64 results also match the oracle.
"""
import numpy as np
import pytest

import oracle
from gw_fopt_amd import many_bubbles as mb
from gw_fopt_amd import nucleation
from gw_fopt_amd.ensemble import BubbleEnsemble

pytestmark = pytest.mark.gpu

N_C, N_PHI, N_T = 12, 24, 60


def small_ensemble(counts, seed=21):
    inter, ext, t_end, boxes = [], [], [], []
    rng = np.random.default_rng(seed)
    for r, n in enumerate(counts):
        beta = rng.uniform(10.0, 30.0)
        box = float((max(n, 2) * 1.5 * 8.0 * np.pi / (0.68 * beta ** 3)) ** (1.0 / 3.0))
        b = nucleation.exponential_nucleation(beta, box=box, n_max=n, seed=seed * 1000 + r)
        inter.append(b)
        ext.append(nucleation.face_images(b, box))
        t_end.append(nucleation.percolation_time(b, box=box, n_probe=2048))
        boxes.append(box)
    return inter, ext, np.array(t_end), np.array(boxes)


def single(interior, exterior, coef, pows):
    bf = mb.GeneralizedBulkFlow(mb.LatticeBubbles(mb.EmptyLattice(), interior, exterior))
    bf.set_resolution(N_C, N_PHI)
    bf.set_gradient_scaling_params(coef, pows)
    return bf


@pytest.fixture(scope="module")
def ens():
    counts = [9, 1, 17, 4, 26, 12, 2, 7]
    inter, ext, t_end, boxes = small_ensemble(counts)
    w = np.geomspace(0.5, 120.0, 40)[None, :] / boxes[:, None]
    k = mb.fibonacci_directions(3)
    return inter, ext, t_end, w, k


@pytest.mark.parametrize("precision,tiles", [
    ("fp64", None),
    ("fp64_blocked", {"block_steps": 8}),
    ("mixed", {"block_steps": 16, "threads": 64}),
    # pipelined chunks (geometry and phase sums on two streams, double-buffered
    # delay rows): several chunks, some units larger than a chunk
    ("fp64", {"items_per_chunk": 25}),
    ("mixed", {"block_steps": 8, "threads": 64, "items_per_chunk": 40}),
])
def test_batched_equals_per_realisation(ens, precision, tiles):
    inter, ext, t_end, w, k = ens
    coef, pows = [[0.9]], [[5.2]]
    E = BubbleEnsemble(inter, ext)
    E.set_resolution(N_C, N_PHI)
    E.set_gradient_scaling_params(coef, pows)
    out = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision=precision, tiles=tiles)
    assert out["S"].shape == (len(inter), len(k), 1, w.shape[1])
    for r in range(len(inter)):
        bf = single(inter[r], ext[r], coef, pows)
        if tiles:
            bf.set_tiles(tiles.get("block_steps", 0), tiles.get("threads", 0))
        S, _ = bf.compute_spectrum(w[r], k, t_end=t_end[r], n_t=N_T, precision=precision)
        assert out["S"][r].tobytes() == S.tobytes(), r


def test_batched_fp64_matches_oracle(ens):
    inter, ext, t_end, w, k = ens
    E = BubbleEnsemble(inter, ext)
    E.set_resolution(N_C, N_PHI)
    E.set_gradient_scaling_params([[1.0]], [[5.0]])
    out = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="fp64")
    for r in (0, 4):
        So = oracle.spectrum(inter[r], ext[r], k, N_C, N_PHI, [[1.0]], [[5.0]], w[r],
                             t_end=t_end[r], n_t=N_T)
        assert np.max(np.abs(out["S"][r] - So) / np.abs(So)) <= 1e-10


def test_statistics_and_graph_replays(ens):
    inter, ext, t_end, w, k = ens
    E = BubbleEnsemble(inter, ext)
    E.set_resolution(N_C, N_PHI)
    E.set_gradient_scaling_params([[1.0]], [[5.0]])
    tl = {"items_per_chunk": 33}
    a = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="mixed", tiles=tl)
    # other GPU work between the replays of the captured graph
    bf = single(inter[2], ext[2], [[1.0]], [[5.0]])
    bf.compute_spectrum(w[2], k, t_end=t_end[2], n_t=N_T, precision="fp64")
    b = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="mixed", tiles=tl)
    bf.compute_c_integral(w[2], t_end=t_end[2], n_t=N_T)
    c = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="mixed", tiles=tl)
    d = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="mixed")  # fused form
    assert d["S"].tobytes() == a["S"].tobytes()
    for key in ("S", "kmean", "mean", "std"):
        assert a[key].tobytes() == b[key].tobytes() == c[key].tobytes(), key
    S = a["S"]
    km = np.zeros(S.shape[:1] + S.shape[2:])
    for j in range(S.shape[1]):
        km += S[:, j]
    km /= S.shape[1]
    np.testing.assert_array_equal(a["kmean"], km)
    m = np.zeros(km.shape[1:])
    for r in range(len(km)):
        m += km[r]
    m /= len(km)
    v = np.zeros_like(m)
    for r in range(len(km)):
        v += (km[r] - m) ** 2
    np.testing.assert_array_equal(a["mean"], m)
    np.testing.assert_array_equal(a["std"], np.sqrt(v / len(km)))


@pytest.mark.parametrize("chunk", [0, 30])
def test_three_sets(ens, chunk):
    """Three parameter sets: two fused launches (sets 0-1, then 2), or one
    phase-sum launch per pipelined chunk."""
    inter, ext, t_end, w, k = ens
    coef, pows = [[1.0], [0.5], [1.2]], [[5.0], [4.0], [3.5]]
    E = BubbleEnsemble(inter[:4], ext[:4])
    E.set_resolution(N_C, N_PHI)
    E.set_gradient_scaling_params(coef, pows)
    out = E.compute_spectrum(w[:4], k[:2], t_end=t_end[:4], n_t=N_T, precision="fp64",
                             tiles={"items_per_chunk": chunk})
    for r in range(4):
        bf = single(inter[r], ext[r], coef, pows)
        S, _ = bf.compute_spectrum(w[r], k[:2], t_end=t_end[r], n_t=N_T, precision="fp64")
        assert out["S"][r].tobytes() == S.tobytes()


def test_autotune_picks_a_valid_shape(ens):
    inter, ext, t_end, w, k = ens
    E = BubbleEnsemble(inter, ext)
    E.set_resolution(N_C, N_PHI)
    E.set_gradient_scaling_params([[1.0]], [[5.0]])
    best = E.autotune(w, k, t_end=t_end, n_t=N_T, precision="mixed")
    assert best["block_steps"] in (8, 16) and best["threads"] in (64, 128, 256)
    assert best["items_per_chunk"] in (0, 16384, 65536)
    tuned = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="mixed")
    explicit = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="mixed",
                                  tiles={"block_steps": best["block_steps"], "threads": best["threads"],
                                         "items_per_chunk": best["items_per_chunk"]})
    assert tuned["S"].tobytes() == explicit["S"].tobytes()
    t = E.last_timings()
    assert t["n_items"] == int(sum(len(b) for b in inter)) * len(k)


@pytest.mark.parametrize("chunk", [0, 30])
def test_setters_and_new_directions_between_calls(ens, chunk):
    """A captured graph is replayed only while every launch argument is
    unchanged: a smaller grid (no buffer grows), a damping width, more terms
    and another k-hat list of the same length each give the spectrum of a
    fresh ensemble with those settings, bit for bit."""
    inter, ext, t_end, w, k = ens
    tl = {"items_per_chunk": chunk}

    def fresh(n_c, n_phi, coef, pows, damping, kk):
        F = BubbleEnsemble(inter, ext)
        F.set_resolution(n_c, n_phi)
        F.set_gradient_scaling_params(coef, pows, damping)
        return F.compute_spectrum(w, kk, t_end=t_end, n_t=N_T, precision="fp64", tiles=tl)["S"]

    E = BubbleEnsemble(inter, ext)
    E.set_resolution(N_C, N_PHI)
    E.set_gradient_scaling_params([[1.0]], [[5.0]])
    E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="fp64", tiles=tl)
    steps = [
        (N_C - 4, N_PHI - 8, [[1.0]], [[5.0]], None, k),
        (N_C - 4, N_PHI - 8, [[1.0]], [[5.0]], 0.3, k),
        (N_C - 4, N_PHI - 8, [[1.0, 0.5]], [[5.0, 4.0]], 0.3, k),
        (N_C - 4, N_PHI - 8, [[1.0, 0.5]], [[5.0, 4.0]], 0.3, k[::-1].copy()),
    ]
    for n_c, n_phi, coef, pows, damping, kk in steps:
        E.set_resolution(n_c, n_phi)
        E.set_gradient_scaling_params(coef, pows, damping)
        got = E.compute_spectrum(w, kk, t_end=t_end, n_t=N_T, precision="fp64", tiles=tl)["S"]
        want = fresh(n_c, n_phi, coef, pows, damping, kk)
        assert got.tobytes() == want.tobytes(), (n_c, n_phi, coef, damping)


@pytest.mark.parametrize("tiles,coef,pows", [
    (None, [[0.9]], [[5.2]]),
    ({"block_steps": 1}, [[0.9]], [[5.2]]),
    ({"block_steps": 4}, [[0.9]], [[5.2]]),
    # two sets of two terms (one tile column per (item, set, +/x)), integer and
    # non-integer powers
    (None, [[0.5, 0.4], [1.0, 0.2]], [[5.0, 4.0], [5.0, 6.5]]),
])
def test_matrix_form_batched_equals_per_realisation(ens, tiles, coef, pows):
    """The FP64 matrix form over the whole ensemble (the single-configuration launches
    over every realisation at once): each realisation's spectrum bit-equal to its own
    configuration's, the k-hat means and the ensemble mean and spread as the other forms
    compute them, and S within 1e-10 of the oracle."""
    inter, ext, t_end, w, k = ens
    E = BubbleEnsemble(inter, ext)
    E.set_resolution(N_C, N_PHI)
    E.set_gradient_scaling_params(coef, pows)
    out = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="fp64_matrix", tiles=tiles)
    for r in range(len(inter)):
        bf = single(inter[r], ext[r], coef, pows)
        if tiles:
            bf.set_tiles(tiles["block_steps"], 0)
        S, _ = bf.compute_spectrum(w[r], k, t_end=t_end[r], n_t=N_T, precision="fp64_matrix")
        assert out["S"][r].tobytes() == S.tobytes(), r
    S = out["S"]
    km = np.zeros(S.shape[:1] + S.shape[2:])
    for j in range(S.shape[1]):
        km += S[:, j]
    km /= S.shape[1]
    np.testing.assert_array_equal(out["kmean"], km)
    m = np.zeros(km.shape[1:])
    for r in range(len(km)):
        m += km[r]
    m /= len(km)
    v = np.zeros_like(m)
    for r in range(len(km)):
        v += (km[r] - m) ** 2
    np.testing.assert_array_equal(out["mean"], m)
    np.testing.assert_array_equal(out["std"], np.sqrt(v / len(km)))
    for r in (0, 4):
        So = oracle.spectrum(inter[r], ext[r], k, N_C, N_PHI, coef, pows, w[r],
                             t_end=t_end[r], n_t=N_T)
        assert np.max(np.abs(S[r] - So) / np.abs(So)) <= 1e-10, r


def test_matrix_form_autotune(ens):
    inter, ext, t_end, w, k = ens
    E = BubbleEnsemble(inter, ext)
    E.set_resolution(N_C, N_PHI)
    best = E.autotune(w, k, t_end=t_end, n_t=N_T, precision="fp64_matrix")
    assert best["block_steps"] in (1, 2, 4)
    a = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="fp64_matrix")
    b = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="fp64_matrix",
                           tiles={"block_steps": best["block_steps"]})
    assert a["S"].tobytes() == b["S"].tobytes()
    # the other forms keep their own launch shape
    c = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="mixed")
    assert np.all(np.isfinite(c["S"]))

def test_matrix_form_graph_replays(ens):
    """The batched matrix form is captured once and replayed while its launches are
    unchanged: replays with other GPU work in between are bit-equal, and replays after
    new scaling parameters or percolation times (same launch shapes, new contents of the
    uploaded tables) equal a fresh ensemble's first, captured, call."""
    inter, ext, t_end, w, k = ens
    coef, pows = [[0.9]], [[5.2]]
    E = BubbleEnsemble(inter, ext)
    E.set_resolution(N_C, N_PHI)
    E.set_gradient_scaling_params(coef, pows)
    a = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="fp64_matrix")
    bf = single(inter[4], ext[4], coef, pows)
    bf.compute_spectrum(w[4], k, t_end=t_end[4], n_t=N_T, precision="fp64_matrix")
    b = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="fp64_matrix")
    E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="mixed")
    bf.compute_c_integral(w[4], t_end=t_end[4], n_t=N_T)
    c = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="fp64_matrix")
    for key in ("S", "kmean", "mean", "std"):
        assert a[key].tobytes() == b[key].tobytes() == c[key].tobytes(), key

    coef2, pows2 = [[1.3]], [[4.6]]
    t2 = t_end * 1.05
    E.set_gradient_scaling_params(coef2, pows2)
    d = E.compute_spectrum(w, k, t_end=t2, n_t=N_T, precision="fp64_matrix")
    F = BubbleEnsemble(inter, ext)
    F.set_resolution(N_C, N_PHI)
    F.set_gradient_scaling_params(coef2, pows2)
    f = F.compute_spectrum(w, k, t_end=t2, n_t=N_T, precision="fp64_matrix")
    for key in ("S", "kmean", "mean", "std"):
        assert d[key].tobytes() == f[key].tobytes(), key
    assert d["S"].tobytes() != a["S"].tobytes()


def test_matrix_form_many_terms_uncaptured(ens):
    """Two sets of seven terms take two matrix launches per call (at most twelve
    (set, term) pairs each): that request runs uncaptured and still equals each
    realisation's own configuration bit for bit."""
    inter, ext, t_end, w, k = ens
    pw = [5.0, 4.5, 4.0, 3.5, 3.0, 2.5, 2.0]
    coef = [[0.3, 0.1, 0.2, 0.05, 0.1, 0.02, 0.01], [0.2, 0.2, 0.1, 0.1, 0.05, 0.05, 0.02]]
    pows = [pw, pw[::-1]]
    E = BubbleEnsemble(inter, ext)
    E.set_resolution(N_C, N_PHI)
    E.set_gradient_scaling_params(coef, pows)
    out = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="fp64_matrix")
    again = E.compute_spectrum(w, k, t_end=t_end, n_t=N_T, precision="fp64_matrix")
    assert out["S"].tobytes() == again["S"].tobytes()
    for r in (0, 2, 7):
        bf = single(inter[r], ext[r], coef, pows)
        S, _ = bf.compute_spectrum(w[r], k, t_end=t_end[r], n_t=N_T, precision="fp64_matrix")
        assert out["S"][r].tobytes() == S.tobytes(), r
