"""Writes the golden fixtures of tests/golden from the CPU oracle.

Run from the repository root:  python tests/golden/make_golden.py [name ...]
(names: small, plain_256, ensemble; all of them by default).  The fixtures
are data only (inputs and oracle outputs, numpy .npz without pickles).  The
small ones are cross-checked against the independent numpy restatement before
they are written; the large ones (the 256-bubble drop-in at its full grids and
realisations of the 512-member ensemble at theirs) take the oracle's results
as they are, the oracle being pinned at small sizes.
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.dirname(HERE)]

import independent_numpy as ind  # noqa: E402
import oracle  # noqa: E402
from gw_fopt_amd import configs  # noqa: E402


def mask_digest(first, dtab):
    """SHA-256 of the first-collider codes (int32) followed by the delays (float64)."""
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(first, dtype=np.int32).tobytes())
    h.update(np.ascontiguousarray(dtab, dtype=np.float64).tobytes())
    return h.hexdigest()


def plain_256_fixture():
    """The 256-bubble drop-in (64 x 128 directions, 1000 steps to percolation, 256
    omegas in [0.5, 500]) in two k-hat frames: S and C of both frames, and a digest
    of every bubble's first colliders and delays in each frame."""
    wl = configs.plain_256(n_k=2)
    t0 = time.time()
    S, C = oracle.spectrum(wl.interior, wl.exterior, wl.k_hat, wl.n_cos_thetax, wl.n_phix,
                           wl.coefficients_sets, wl.powers_sets, wl.w_arr, t_end=wl.t_end,
                           n_t=wl.n_t, return_c=True)
    print(f"plain_256 spectrum, 2 k-hat: {time.time() - t0:.1f} s", flush=True)
    digests = []
    for k in wl.k_hat:
        for a in range(wl.n_bubbles):
            f, d = oracle.first_colliding(wl.interior, wl.exterior, wl.n_cos_thetax, wl.n_phix, a,
                                          k_hat=k)
            digests.append(mask_digest(f, d))
    np.savez_compressed(os.path.join(HERE, "plain_256_2k.npz"), interior=wl.interior,
                        exterior=wl.exterior, k_hat=wl.k_hat, w=wl.w_arr, t_end=wl.t_end,
                        n_t=wl.n_t, n_c=wl.n_cos_thetax, n_phi=wl.n_phix,
                        coef=np.asarray(wl.coefficients_sets), pows=np.asarray(wl.powers_sets),
                        S=S, C=C, mask_sha256=np.array(digests).reshape(2, wl.n_bubbles))


# realisations of the 512-member ensemble: the largest (395 bubbles), one at the
# 20-bubble floor, and two between
ENSEMBLE_MEMBERS = (284, 296, 17, 181)


def ensemble_fixture():
    """Four realisations of the 512-member ensemble at its grids (32 x 64, 500
    steps to each realisation's percolation, 128 omegas geomspace(0.5, 300) / L)
    in two of its 32 k-hat frames: S and C per realisation."""
    e = configs.ensemble_512_members(ENSEMBLE_MEMBERS)
    k = e.k_hat[[3, 20]]
    w = e.w_arr()
    out = {"k_hat": k, "w": w, "t_end": e.t_end, "boxes": e.boxes, "n_t": e.n_t,
           "n_c": e.n_cos_thetax, "n_phi": e.n_phix, "members": np.array(ENSEMBLE_MEMBERS),
           "coef": np.asarray(e.coefficients_sets), "pows": np.asarray(e.powers_sets)}
    for r in range(len(ENSEMBLE_MEMBERS)):
        t0 = time.time()
        S, C = oracle.spectrum(e.interiors[r], e.exteriors[r], k, e.n_cos_thetax, e.n_phix,
                               e.coefficients_sets, e.powers_sets, w[r], t_end=e.t_end[r],
                               n_t=e.n_t, return_c=True)
        print(f"ensemble member {ENSEMBLE_MEMBERS[r]} ({len(e.interiors[r])} bubbles): "
              f"{time.time() - t0:.1f} s", flush=True)
        out[f"interior_{r}"] = e.interiors[r]
        out[f"exterior_{r}"] = e.exteriors[r]
        out[f"S_{r}"] = S
        out[f"C_{r}"] = C
    np.savez_compressed(os.path.join(HERE, "ensemble_members.npz"), **out)


def small_fixtures():
    # the reference's C-ABI test case (rust/c-bubble-gw/tests/c_api.rs:135-218)
    inter = np.array([[0.0, 0.0, 10.0, 0.0], [0.0, 0.0, 0.0, 0.0]])
    coef = np.array([[0.0], [1.0]])
    pows = np.array([[3.0], [3.0]])
    w = np.array([0.5, 1.0])
    C = oracle.compute_c_integral(inter, None, 4, 5, coef, pows, w, t_begin=0.0, t_end=8.0, n_t=40)
    Cn = ind.c_integral(inter, None, 4, 5, coef, pows, w, 0.0, 8.0, 40)
    assert np.max(np.abs(C - Cn)) / np.max(np.abs(C)) < 1e-13
    np.savez(os.path.join(HERE, "capi_two_bubbles.npz"), interior=inter, coef=coef, pows=pows,
             w=w, C=C)

    wl = configs.plain_8()
    S = oracle.spectrum(wl.interior, wl.exterior, wl.k_hat, 32, 64, [[1.0]], [[5.0]], wl.w_arr,
                        t_end=1.5, n_t=200)
    first, dtab = oracle.first_colliding(wl.interior, wl.exterior, 32, 64, 0, k_hat=wl.k_hat[3])
    np.savez(os.path.join(HERE, "plain_8.npz"), interior=wl.interior, exterior=wl.exterior,
             k_hat=wl.k_hat, w=wl.w_arr, S=S, first_k3_a0=first, dtab_k3_a0=dtab)


def main(names):
    makers = {"small": small_fixtures, "plain_256": plain_256_fixture,
              "ensemble": ensemble_fixture}
    for n in names or list(makers):
        makers[n]()
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main(sys.argv[1:])
