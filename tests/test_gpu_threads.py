"""Separate handles driven from several host threads at once.

ctypes releases the GIL for each library call, so the calls overlap in the
library: they share the kernels' dynamic-LDS limits (process state, set per
launch under a lock) and the device.  Every result must equal the same call
made alone, bit for bit.
"""
import threading

import numpy as np
import pytest

from gw_fopt_amd import many_bubbles as mb
from gw_fopt_amd import nucleation
from gw_fopt_amd.ensemble import BubbleEnsemble

pytestmark = pytest.mark.gpu

W = np.geomspace(0.5, 60.0, 96)
K = mb.fibonacci_directions(4)
# (grid, time steps, form): the two blocked jobs share one kernel instance
# with different LDS sizes
JOBS = [(24, 48, 700, "fp64_blocked"), (16, 40, 300, "mixed"), (32, 64, 520, "fp64"),
        (12, 24, 900, "fp64_blocked")]


def _bubbles(seed):
    inter = nucleation.exponential_nucleation(12.0, n_max=8, seed=seed)
    return inter, nucleation.face_images(inter, 1.0)


def _plain(i):
    n_c, n_phi, n_t, precision = JOBS[i]
    inter, ext = _bubbles(100 + i)
    bf = mb.GeneralizedBulkFlow(mb.LatticeBubbles(mb.EmptyLattice(), inter, ext))
    bf.set_resolution(n_c, n_phi)
    bf.set_gradient_scaling_params([[1.0]], [[5.0]])
    t_end = float(inter[:, 0].max()) + 0.6
    return bf.compute_spectrum(W, K, t_end=t_end, n_t=n_t, precision=precision)[0]


def _ensemble():
    pairs = [_bubbles(200 + r) for r in range(3)]
    E = BubbleEnsemble([p[0] for p in pairs], [p[1] for p in pairs])
    E.set_resolution(20, 36)
    E.set_gradient_scaling_params([[1.0]], [[5.0]])
    t_end = np.array([float(p[0][:, 0].max()) + 0.6 for p in pairs])
    w = np.tile(W, (3, 1))
    return E.compute_spectrum(w, K, t_end=t_end, n_t=400, precision="mixed",
                              tiles={"items_per_chunk": 40})["S"]


def test_separate_handles_on_threads():
    runs = [lambda i=i: _plain(i) for i in range(len(JOBS))] + [_ensemble]
    alone = [f() for f in runs]
    reps = 3
    got = [[None] * reps for _ in runs]
    errors = []

    def worker(j):
        try:
            for r in range(reps):
                got[j][r] = runs[j]()
        except Exception as e:  # reported below, with the job
            errors.append((j, repr(e)))

    threads = [threading.Thread(target=worker, args=(j,)) for j in range(len(runs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    for j in range(len(runs)):
        for r in range(reps):
            assert got[j][r].tobytes() == alone[j].tobytes(), (j, r)


def test_set_num_threads_sizes_the_preparation_only():
    """set_num_threads(n) sets the host threads of the per-configuration candidate
    build (run at the first compute call); the results do not depend on it."""
    inter, ext = _bubbles(7)
    out = []
    for n in (None, 1, 3):
        bf = mb.GeneralizedBulkFlow(mb.LatticeBubbles(mb.EmptyLattice(), inter, ext))
        if n is not None:
            bf.set_num_threads(n)
        bf.set_resolution(24, 48)
        bf.set_gradient_scaling_params([[1.0]], [[5.0]])
        out.append(bf.compute_spectrum(W, K, t_end=float(inter[:, 0].max()) + 0.6, n_t=400,
                                       precision="fp64")[0])
    assert out[0].tobytes() == out[1].tobytes() == out[2].tobytes()
