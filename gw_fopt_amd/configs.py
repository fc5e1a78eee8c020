"""The workloads of the bulk-flow path (seeded, reproducible).

Each returns a ``Workload``: interior bubbles, their six periodic face
images, angular grid, time grid, omegas and k-hat list.

  plain_8       8 bubbles at t = 0 in a periodic unit cube; 32 x 64 surface
                grid; 200 time points on [0, 1.5]; 64 omegas geomspace(0.5,
                200); 16 Fibonacci k-hat; scaling f = 1 * ratio^5 (xi = 2).
  plain_256     256 bubbles from exponential nucleation beta L = 20 (first 256
                nuclei, rejection inside existing bubbles); 64 x 128; 1000 time
                points on [0, t_percolation]; 256 omegas geomspace(0.5, 500);
                64 Fibonacci k-hat; xi = 2.
  mixed_8       plain_8 in the mixed-precision form.
  ensemble_512  512 realisations, beta L ~ U[10, 30], box sides giving ~20-400
                bubbles (log-uniform target counts, mean ~128); 32 x 64; 500
                time points to each realisation's percolation; 128 omegas
                geomspace(0.5, 300) / L; 32 k-hat.
  grad_128      128 bubbles, beta L = 15; 32 x 64; 500 time points; 128
                omegas geomspace(0.5, 300); 32 k-hat; theta = (xi, A) with
                coefficients [[A]], powers [[xi + 3]]; target theta* = (2.2, 0.9).

The scaling function follows the reference's convention f = sum_k a_k
ratio^{p_k} with p = xi + 3 (src/gw_fopt/bubble_dynamics/utils/
data_gw_bulkflow.py:86, docs/generalized_bulkflow.md:45-46).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import nucleation
from .many_bubbles import fibonacci_directions


@dataclass
class Workload:
    name: str
    interior: np.ndarray
    exterior: np.ndarray
    n_cos_thetax: int
    n_phix: int
    t_begin: float
    t_end: float
    n_t: int
    w_arr: np.ndarray
    k_hat: np.ndarray
    coefficients_sets: list = field(default_factory=lambda: [[1.0]])
    powers_sets: list = field(default_factory=lambda: [[5.0]])
    box: float = 1.0

    @property
    def n_bubbles(self) -> int:
        return len(self.interior)

    def points(self) -> int:
        return len(self.w_arr) * len(self.k_hat)


def theta_to_params(xi: float, amplitude: float):
    """(xi, A) -> (coefficients_sets, powers_sets) = ([[A]], [[xi + 3]])."""
    return [[float(amplitude)]], [[float(xi) + 3.0]]


def plain_8(seed: int = 1) -> Workload:
    interior = nucleation.simultaneous_nucleation(8, box=1.0, seed=seed)
    return Workload("plain_8", interior, nucleation.face_images(interior, 1.0), 32, 64, 0.0, 1.5,
                    200, np.geomspace(0.5, 200.0, 64), fibonacci_directions(16))


def plain_256(seed: int = 2, n_k: int = 64) -> Workload:
    interior = nucleation.exponential_nucleation(20.0, n_max=256, seed=seed)
    t_end = nucleation.percolation_time(interior)
    return Workload("plain_256", interior, nucleation.face_images(interior, 1.0), 64, 128, 0.0,
                    t_end, 1000, np.geomspace(0.5, 500.0, 256), fibonacci_directions(n_k))


def grad_128(seed: int = 5, n_k: int = 32) -> Workload:
    interior = nucleation.exponential_nucleation(15.0, n_max=128, seed=seed)
    t_end = nucleation.percolation_time(interior)
    c, p = theta_to_params(2.2, 0.9)
    return Workload("grad_128", interior, nucleation.face_images(interior, 1.0), 32, 64, 0.0,
                    t_end, 500, np.geomspace(0.5, 300.0, 128), fibonacci_directions(n_k), c, p)


@dataclass
class Ensemble:
    name: str
    interiors: list
    exteriors: list
    boxes: np.ndarray
    t_end: np.ndarray
    n_cos_thetax: int
    n_phix: int
    n_t: int
    x_arr: np.ndarray          # omega * L (shared); realisation r uses x_arr / L_r
    k_hat: np.ndarray
    coefficients_sets: list = field(default_factory=lambda: [[1.0]])
    powers_sets: list = field(default_factory=lambda: [[5.0]])

    def w_arr(self) -> np.ndarray:
        return self.x_arr[None, :] / self.boxes[:, None]


def _ensemble_member(args):
    """One realisation: nucleation in a box of side ``box`` stopped at ``n_target``
    bubbles.  A draw that percolates with fewer than ``n_floor`` bubbles is redrawn
    in a box grown for the missing count (next seed of the realisation's stream),
    so every realisation holds at least ``n_floor`` bubbles."""
    beta, n_target, box, seed, n_floor = args
    for attempt in range(16):
        b = nucleation.exponential_nucleation(beta, box=box, n_max=n_target,
                                              seed=seed + 7919 * attempt)
        if len(b) >= n_floor:
            break
        box *= float((1.25 * n_floor / max(len(b), 1)) ** (1.0 / 3.0))
    else:
        raise RuntimeError(f"realisation seed {seed}: fewer than {n_floor} bubbles")
    return b, nucleation.face_images(b, box), nucleation.percolation_time(b, box=box, n_probe=8192), box


def _ensemble_jobs(n_real: int, seed: int, n_min: float, n_max: float):
    """(beta, target count, box side, seed, floor) of each realisation, drawn in order."""
    rng = np.random.default_rng(seed)
    jobs = []
    for r in range(n_real):
        beta = rng.uniform(10.0, 30.0)
        n_target = int(round(float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))))
        box = float((n_target * 8.0 * np.pi / (0.68 * beta ** 3)) ** (1.0 / 3.0))
        jobs.append((beta, n_target, box, seed * 100003 + r, int(round(n_min))))
    return jobs


def _ensemble(name, jobs, n_k, workers):
    import multiprocessing
    import os
    from concurrent.futures import ProcessPoolExecutor

    if workers is None:
        workers = min(16, os.cpu_count() or 1)
    if workers > 1 and len(jobs) > 1:
        ctx = multiprocessing.get_context("fork")
        with ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as ex:
            res = list(ex.map(_ensemble_member, jobs, chunksize=4))
    else:
        res = [_ensemble_member(j) for j in jobs]
    interiors = [r[0] for r in res]
    exteriors = [r[1] for r in res]
    t_end = np.array([r[2] for r in res])
    boxes = np.array([r[3] for r in res])
    return Ensemble(name, interiors, exteriors, boxes, t_end, 32, 64, 500,
                    np.geomspace(0.5, 300.0, 128), fibonacci_directions(n_k))


def ensemble_512(n_real: int = 512, seed: int = 4, n_k: int = 32, workers: int | None = None,
                 n_min: float = 20.0, n_max: float = 400.0) -> Ensemble:
    """Realisation r: beta L_r ~ U[10, 30]; a target count n_r log-uniform in
    [n_min, n_max] (mean ~128 for 20..400); box side for ~1.5 n_r bubbles at
    percolation (density ~0.68 beta^3 / (8 pi)), nucleation stopped at n_r
    bubbles, and redrawn in a larger box if it percolates below n_min bubbles;
    t_end = the realisation's percolation time.  Realisations are generated in
    parallel worker processes (fork); call this before any GPU work in the
    process."""
    return _ensemble("ensemble_512", _ensemble_jobs(n_real, seed, n_min, n_max), n_k, workers)


def ensemble_512_members(indices, seed: int = 4, n_k: int = 32, n_real: int = 512,
                         n_min: float = 20.0, n_max: float = 400.0) -> Ensemble:
    """Realisations ``indices`` of ``ensemble_512`` (the same bubbles, boxes and time
    grids), generated alone."""
    jobs = _ensemble_jobs(n_real, seed, n_min, n_max)
    return _ensemble("ensemble_512", [jobs[i] for i in indices], n_k, workers=1)
