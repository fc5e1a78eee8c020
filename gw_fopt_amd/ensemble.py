"""Batched ensembles of nucleation realisations (uneven bubble counts) on the GPU.

``BubbleEnsemble`` holds many independent realisations and computes, for a
shared k-hat list, every realisation's spectrum S(k, s, w) = 8 pi w^3
(|C+|^2 + |Cx|^2), the k-hat mean of each, and the mean and population
standard deviation of those over the ensemble.  The batch runs as one
persistent launch over (realisation, k-hat) units, captured in a hipGraph;
``autotune`` times the candidate launch shapes on the ensemble and keeps the
fastest.  Each realisation's numbers equal ``GeneralizedBulkFlow.
compute_spectrum`` on that realisation alone with the same launch shape.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .many_bubbles import _precision


class BubbleEnsemble:
    def __init__(self, interiors, exteriors=None):
        inter = [np.ascontiguousarray(np.asarray(b, dtype=np.float64).reshape(-1, 4)) for b in interiors]
        if exteriors is None:
            exteriors = [np.zeros((0, 4)) for _ in inter]
        exter = [np.ascontiguousarray(np.asarray(b, dtype=np.float64).reshape(-1, 4)) for b in exteriors]
        if len(exter) != len(inter):
            raise ValueError("interiors and exteriors must have the same number of realisations")
        self.n_real = len(inter)
        self.n_interior = np.array([len(b) for b in inter], dtype=np.int64)
        io = np.zeros(self.n_real + 1, dtype=np.uintp)
        eo = np.zeros(self.n_real + 1, dtype=np.uintp)
        io[1:] = np.cumsum([len(b) for b in inter])
        eo[1:] = np.cumsum([len(b) for b in exter])
        ib = np.ascontiguousarray(np.vstack(inter) if self.n_real else np.zeros((0, 4)))
        eb = np.ascontiguousarray(np.vstack(exter) if self.n_real else np.zeros((0, 4)))
        h = ctypes.c_void_p()
        st = _lib.lib().bgw_ensemble_new(ctypes.byref(h), _lib.dptr(ib), _lib.sptr(io),
                                         _lib.dptr(eb), _lib.sptr(eo), self.n_real)
        if st != _lib.OK:
            raise ValueError(_lib.last_error())
        self._h = h
        self._n_sets = 1
        self.set_gradient_scaling_params([[1.0]], [[3.0]])

    @classmethod
    def from_nucleation(cls, interior, exterior, counts) -> "BubbleEnsemble":
        """An ensemble from a device nucleation batch (``nucleate_fixed_rate_batch`` or
        ``FixedRateNucleation.nucleate_many``): realisation r is rows [0, counts[r]) of
        interior[r] and the matching counts[r] * n_img rows of exterior[r]."""
        interior = np.asarray(interior, dtype=np.float64)
        exterior = np.asarray(exterior, dtype=np.float64)
        counts = np.asarray(counts, dtype=np.int64).reshape(-1)
        if interior.ndim != 3 or interior.shape[0] != len(counts) or interior.shape[2] != 4:
            raise ValueError("interior must be (n_realisations, max_bubbles, 4)")
        cap = interior.shape[1]
        n_img = exterior.shape[1] // cap if cap else 0
        if np.any(counts < 0) or np.any(counts > cap):
            raise ValueError("counts must lie in [0, max_bubbles]")
        return cls([interior[r, :n] for r, n in enumerate(counts)],
                   [exterior[r, :n * n_img] for r, n in enumerate(counts)])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().bgw_ensemble_free(h)
            except Exception:
                pass
            self._h = None

    def _call(self, fn, *args):
        st = fn(self._h, *args)
        if st != _lib.OK:
            raise ValueError(_lib.last_error())

    def set_resolution(self, n_cos_thetax: int, n_phix: int) -> None:
        self._call(_lib.lib().bgw_ensemble_set_resolution, int(n_cos_thetax), int(n_phix))

    def set_gradient_scaling_params(self, coefficients_sets, powers_sets, damping_width=None):
        coef = np.ascontiguousarray(np.asarray(coefficients_sets, dtype=np.float64))
        pows = np.ascontiguousarray(np.asarray(powers_sets, dtype=np.float64))
        if coef.ndim != 2 or coef.shape != pows.shape:
            raise ValueError("coefficients_sets and powers_sets must have the same 2-D shape")
        damping = float("nan") if damping_width is None else float(damping_width)
        self._call(_lib.lib().bgw_ensemble_set_gradient_scaling_params, _lib.dptr(coef),
                   _lib.dptr(pows), coef.shape[0], coef.shape[1], damping)
        self._n_sets = coef.shape[0]

    def _args(self, w_arr, k_hat, t_begin, t_end):
        w = np.asarray(w_arr, dtype=np.float64)
        if w.ndim == 1:
            w = np.broadcast_to(w, (self.n_real, len(w)))
        w = np.ascontiguousarray(w)
        if w.shape[0] != self.n_real:
            raise ValueError(f"w_arr must be (n_w,) or ({self.n_real}, n_w)")
        k = np.ascontiguousarray(np.asarray(k_hat, dtype=np.float64).reshape(-1, 3))
        tb = np.ascontiguousarray(np.broadcast_to(np.asarray(0.0 if t_begin is None else t_begin,
                                                             dtype=np.float64), (self.n_real,)))
        te = np.ascontiguousarray(np.broadcast_to(np.asarray(t_end, dtype=np.float64), (self.n_real,)))
        return w, k, tb, te

    def compute_spectrum(self, w_arr, k_hat, *, t_end, n_t, t_begin=None, precision="mixed",
                         tiles=None) -> dict:
        """Returns {"S": (n_real, n_k, n_sets, n_w), "kmean": (n_real, n_sets, n_w),
        "mean": (n_sets, n_w), "std": (n_sets, n_w)}.  ``w_arr``: (n_w,) shared or
        (n_real, n_w) per realisation; ``t_begin`` / ``t_end``: scalars or (n_real,);
        ``tiles``: None (autotuned or automatic) or dict(block_steps, threads,
        items_per_chunk, persistent_blocks)."""
        w, k, tb, te = self._args(w_arr, k_hat, t_begin, t_end)
        n_w = w.shape[1]
        S = np.empty((self.n_real, len(k), self._n_sets, n_w))
        km = np.empty((self.n_real, self._n_sets, n_w))
        mean = np.empty((self._n_sets, n_w))
        std = np.empty((self._n_sets, n_w))
        t = None
        if tiles is not None:
            t = _lib.BgwTileConfig(int(tiles.get("block_steps", 0)), int(tiles.get("threads", 0)),
                                   int(tiles.get("items_per_chunk", 0)),
                                   int(tiles.get("persistent_blocks", 0)))
        self._call(_lib.lib().bgw_ensemble_compute_spectrum, _lib.dptr(w), n_w, _lib.dptr(k),
                   len(k), _lib.dptr(tb), _lib.dptr(te), int(n_t), _precision(precision),
                   None if t is None else ctypes.byref(t), _lib.dptr(S), S.size, _lib.dptr(km),
                   km.size, _lib.dptr(mean), _lib.dptr(std), mean.size)
        return {"S": S, "kmean": km, "mean": mean, "std": std}

    def autotune(self, w_arr, k_hat, *, t_end, n_t, t_begin=None, precision="mixed") -> dict:
        """Times the candidate launch shapes on a sample of this ensemble's work
        units and keeps the fastest for later calls; returns it."""
        w, k, tb, te = self._args(w_arr, k_hat, t_begin, t_end)
        best = _lib.BgwTileConfig()
        self._call(_lib.lib().bgw_ensemble_autotune, _lib.dptr(w), w.shape[1], _lib.dptr(k),
                   len(k), _lib.dptr(tb), _lib.dptr(te), int(n_t), _precision(precision),
                   ctypes.byref(best))
        return {f: int(getattr(best, f)) for f, _ in best._fields_}

    def last_timings(self) -> dict:
        t = _lib.BgwTimings()
        self._call(_lib.lib().bgw_ensemble_last_timings, ctypes.byref(t))
        return {f: getattr(t, f) for f, _ in t._fields_}
