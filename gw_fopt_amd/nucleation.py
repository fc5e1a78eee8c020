"""Seeded bubble nucleation in a periodic cube (input generation for the spectrum path).

Bubbles nucleate with rate Gamma(t) = Gamma0 exp(beta t) per unit volume of
false vacuum and grow at the speed of light.  The process is sampled by
thinning: candidate events come from the Poisson process of rate
Gamma0 V exp(beta t) over the whole box, at uniform positions, and a
candidate is kept only if its point is still in false vacuum, i.e. outside
every earlier bubble and every periodic image of one (this is the rejection
of nuclei inside existing bubbles of the reference's fixed-rate nucleation,
rust/bubble-gw/src/many_bubbles/bubbles_nucleation/fixed_rate_nucleation.rs:
113-142, 210-260).  Times are shifted so that the first bubble nucleates at
t = 0.

The exterior bubbles of a periodic lattice are the six face images of every
interior bubble, in the order +e1, -e1, +e2, -e2, +e3, -e3
(lattice/parallelepiped_lattice.rs:197-215).
"""
from __future__ import annotations

import math

import numpy as np

def face_images(interior: np.ndarray, box: float) -> np.ndarray:
    e = np.eye(3) * box
    shifts = [e[0], -e[0], e[1], -e[1], e[2], -e[2]]
    out = np.empty((len(interior) * 6, 4))
    for i, ev in enumerate(np.asarray(interior, dtype=np.float64)):
        for k, sh in enumerate(shifts):
            out[6 * i + k, 0] = ev[0]
            out[6 * i + k, 1] = ev[1] + sh[0]
            out[6 * i + k, 2] = ev[2] + sh[1]
            out[6 * i + k, 3] = ev[3] + sh[2]
    return out


def _min_image_dist(points: np.ndarray, centres: np.ndarray, box: float) -> np.ndarray:
    """(P, N) periodic distances in a cube (minimum-image convention)."""
    d = points[:, None, :] - centres[None, :, :]
    d -= box * np.round(d / box)
    return np.sqrt(np.sum(d * d, axis=-1))


def _covered(points: np.ndarray, t: float, bubbles: np.ndarray, box: float) -> np.ndarray:
    """True where a point is inside a bubble (or a periodic image) at time t."""
    if len(bubbles) == 0:
        return np.zeros(len(points), dtype=bool)
    rad = t - bubbles[:, 0]
    dist = _min_image_dist(points, bubbles[:, 1:], box)
    return np.any((dist < rad[None, :]) & (rad > 0)[None, :], axis=1)


def exponential_nucleation(beta: float, *, box: float = 1.0, n_max: int | None = None,
                           seed: int = 0, fv_cutoff: float = 0.01, n_probe: int = 4096,
                           max_candidates: int = 2_000_000) -> np.ndarray:
    """Interior bubbles (n, 4) rows (t, x, y, z) in [0, box)^3 from exponential nucleation.

    Stops after ``n_max`` bubbles, or when the false-vacuum fraction (estimated on
    ``n_probe`` fixed random points) drops below ``fv_cutoff``.
    """
    rng = np.random.default_rng(seed)
    V = box ** 3
    rate0 = beta / V  # Gamma0 V / beta = 1: first nucleation expected near t = 0
    probe = rng.random((n_probe, 3)) * box
    probe_first = np.full(n_probe, np.inf)  # earliest time each probe point is covered
    bubbles = np.zeros((0, 4))
    t = -3.0 / beta
    for _ in range(max_candidates):
        e = rng.exponential()
        t = math.log(math.exp(beta * t) + beta * e / (rate0 * V)) / beta
        if np.mean(probe_first > t) < fv_cutoff:
            break
        x = rng.random(3) * box
        if len(bubbles) and _covered(x[None, :], t, bubbles, box)[0]:
            continue
        nb = np.array([[t, *x]])
        bubbles = np.vstack([bubbles, nb])
        probe_first = np.minimum(probe_first, t + _min_image_dist(probe, x[None, :], box)[:, 0])
        if n_max is not None and len(bubbles) >= n_max:
            break
    bubbles[:, 0] -= bubbles[0, 0]
    return bubbles


def simultaneous_nucleation(n: int, *, box: float = 1.0, seed: int = 1) -> np.ndarray:
    """n bubbles at t = 0 at uniform random positions in [0, box)^3."""
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)) * box
    return np.column_stack([np.zeros(n), pos])


def percolation_time(bubbles: np.ndarray, box: float = 1.0, n_probe: int = 32768,
                     seed: int = 12345) -> float:
    """Time at which every probe point of the periodic box is inside some bubble."""
    rng = np.random.default_rng(seed)
    probe = rng.random((n_probe, 3)) * box
    first = np.full(n_probe, np.inf)
    for c0 in range(0, len(bubbles), 64):
        b = bubbles[c0:c0 + 64]
        first = np.minimum(first, np.min(b[None, :, 0] + _min_image_dist(probe, b[:, 1:], box), axis=1))
    return float(np.max(first))
