"""GPU parity at the workloads' own sizes, against oracle fixtures.

* The 256-bubble drop-in (64 x 128 directions, N_tot = 1792 candidates per
  bubble, 1000 steps to percolation, 256 omegas in [0.5, 500]) in two k-hat
  frames: every bubble's first colliders and delays bit-exact (digests of the
  oracle's, and the oracle run live on a subset for a readable diff), S within
  1e-10 per bin and C within 1e-10 of each row's largest entry, for the FP64,
  the blocked FP64 and the FP64 matrix forms.  This is where the candidate walks of the
  geometry kernel are longest and where the uncollided-ring shortcut meets
  the highest omegas.
* Four realisations of the 512-member ensemble at its grids (32 x 64, 500
  steps, 128 omegas / L), the largest (395 bubbles) among them, through the
  batched path in its fused and pipelined forms and in the FP64 matrix form: each
  realisation bit-equal to the single-configuration path and within 1e-10 per bin
  of the oracle.

The fixtures come from tests/golden/make_golden.py.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle
from gw_fopt_amd import many_bubbles as mb
from gw_fopt_amd.ensemble import BubbleEnsemble

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FP64_REL = 1e-10


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def digest(first, dtab):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(first, dtype=np.int32).tobytes())
    h.update(np.ascontiguousarray(dtab, dtype=np.float64).tobytes())
    return h.hexdigest()


def rel_per_bin(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


@pytest.fixture(scope="module")
def p256():
    g = load("plain_256_2k.npz")
    lb = mb.LatticeBubbles(mb.EmptyLattice(), g["interior"], g["exterior"])
    bf = mb.GeneralizedBulkFlow(lb)
    bf.set_resolution(int(g["n_c"]), int(g["n_phi"]))
    bf.set_gradient_scaling_params(g["coef"].tolist(), g["pows"].tolist())
    return g, bf


def test_plain_256_masks_bitexact(p256):
    g, bf = p256
    inter, ext = g["interior"], g["exterior"]
    n = len(inter)
    n_c, n_phi = int(g["n_c"]), int(g["n_phi"])
    # live oracle on the earliest and latest nucleated bubbles and every 16th
    subset = sorted({int(np.argmin(inter[:, 0])), int(np.argmax(inter[:, 0])),
                     *range(0, n, 16)})
    bad = []
    for ki, k in enumerate(g["k_hat"]):
        for a in range(n):
            f, d = bf.first_colliding_in_frame(a, k)
            if a in subset:
                fo, do = oracle.first_colliding(inter, ext, n_c, n_phi, a, k_hat=k)
                np.testing.assert_array_equal(f, fo, err_msg=f"frame {ki} bubble {a}")
                assert d.tobytes() == do.tobytes(), (ki, a)
            if digest(f, d) != str(g["mask_sha256"][ki, a]):
                bad.append((ki, a))
    assert not bad, f"first colliders / delays differ from the oracle's for {bad[:10]}"


@pytest.mark.parametrize("precision", ["fp64", "fp64_blocked", "fp64_matrix"])
def test_plain_256_spectrum(p256, precision):
    g, bf = p256
    S, M, C = bf.compute_spectrum(g["w"], g["k_hat"], t_end=float(g["t_end"]), n_t=int(g["n_t"]),
                                  precision=precision, return_c=True)
    err = rel_per_bin(S, g["S"])
    crow = np.max(np.abs(C - g["C"]), axis=-1) / np.max(np.abs(g["C"]), axis=-1)
    print(f"plain_256 {precision}: S max per-bin rel err {err:.2e}, C worst row {crow.max():.2e}")
    assert err <= FP64_REL
    assert crow.max() <= FP64_REL
    np.testing.assert_allclose(M, S.mean(axis=0), rtol=1e-14)


def test_plain_256_mixed_reported(p256):
    """The mixed form at the drop-in's size: its errors against the oracle, stated."""
    g, bf = p256
    S, _ = bf.compute_spectrum(g["w"], g["k_hat"], t_end=float(g["t_end"]), n_t=int(g["n_t"]),
                               precision="mixed")
    rel = np.abs(S - g["S"]) / np.abs(g["S"])
    print(f"plain_256 mixed: max {rel.max():.2e} median {np.median(rel):.2e} "
          f"bins <= 1e-6: {np.mean(rel <= 1e-6):.3f}")
    assert np.all(np.isfinite(S))
    assert np.median(rel) <= 1e-6


@pytest.fixture(scope="module")
def members():
    return load("ensemble_members.npz")


@pytest.mark.parametrize("precision,tiles", [
    ("fp64", {"block_steps": 8, "threads": 128}),
    ("fp64_blocked", {"block_steps": 8, "threads": 128}),
    # pipelined chunks of about 512 items: several chunks, geometry and phase sums
    # on two streams
    ("fp64_blocked", {"block_steps": 8, "threads": 64, "items_per_chunk": 512}),
    # the FP64 matrix form: the single-configuration launches over all four at once
    ("fp64_matrix", {"block_steps": 2, "threads": 0}),
])
def test_ensemble_members_full_size(members, precision, tiles):
    g = members
    n_r = len(g["members"])
    inter = [g[f"interior_{r}"] for r in range(n_r)]
    ext = [g[f"exterior_{r}"] for r in range(n_r)]
    coef, pows = g["coef"].tolist(), g["pows"].tolist()
    n_c, n_phi, n_t = int(g["n_c"]), int(g["n_phi"]), int(g["n_t"])
    E = BubbleEnsemble(inter, ext)
    E.set_resolution(n_c, n_phi)
    E.set_gradient_scaling_params(coef, pows)
    out = E.compute_spectrum(g["w"], g["k_hat"], t_end=g["t_end"], n_t=n_t, precision=precision,
                             tiles=tiles)
    for r in range(n_r):
        print(f"member {int(g['members'][r])} ({len(inter[r])} bubbles)", flush=True)
        err = rel_per_bin(out["S"][r], g[f"S_{r}"])
        print(f"  {precision} {tiles}: max per-bin rel err vs oracle {err:.2e}")
        assert err <= FP64_REL, r
        bf = mb.GeneralizedBulkFlow(mb.LatticeBubbles(mb.EmptyLattice(), inter[r], ext[r]))
        bf.set_resolution(n_c, n_phi)
        bf.set_gradient_scaling_params(coef, pows)
        bf.set_tiles(tiles["block_steps"], tiles["threads"])
        S, _ = bf.compute_spectrum(g["w"][r], g["k_hat"], t_end=float(g["t_end"][r]), n_t=n_t,
                                   precision=precision)
        assert out["S"][r].tobytes() == S.tobytes(), r


def test_ensemble_members_matrix_form(members):
    """The FP64 matrix form on each of the four realisations alone, against the oracle."""
    g = members
    for r in range(len(g["members"])):
        bf = mb.GeneralizedBulkFlow(mb.LatticeBubbles(mb.EmptyLattice(), g[f"interior_{r}"],
                                                      g[f"exterior_{r}"]))
        bf.set_resolution(int(g["n_c"]), int(g["n_phi"]))
        bf.set_gradient_scaling_params(g["coef"].tolist(), g["pows"].tolist())
        S, _ = bf.compute_spectrum(g["w"][r], g["k_hat"], t_end=float(g["t_end"][r]),
                                   n_t=int(g["n_t"]), precision="fp64_matrix")
        err = rel_per_bin(S, g[f"S_{r}"])
        print(f"member {int(g['members'][r])}: fp64_matrix max per-bin rel err {err:.2e}")
        assert err <= FP64_REL, r
