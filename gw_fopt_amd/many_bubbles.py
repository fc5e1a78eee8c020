"""Python surface of the bulk-flow path, mirroring ``gw_fopt.bubble_gw.many_bubbles``.

Same class names, constructor and method signatures, dtypes, shapes and error
types as the reference's PyO3 module for this path
(rust/py-bubble-gw/src/py_many_bubbles/py_generalized_bulk_flow.rs:130-441,
py_lattice_bubbles.rs:18-140; stub src/gw_fopt/bubble_gw/many_bubbles.pyi):

* ``EmptyLattice``, ``ParallelepipedLattice`` (the two lattices the path uses:
  the C ABI's lattice and the periodic box of the configurations),
* ``LatticeBubbles(lattice, bubbles_interior, bubbles_exterior=None)`` with
  ``with_boundary_condition("periodic")``,
* ``ExponentialTimeCutoff(smax, ratio_t_cut=None, ratio_t_0=None)``,
* ``GeneralizedBulkFlow(lattice_bubbles)``.

Every computation runs in libbubble_gw_c.so on the GPU; the Python code only
validates arguments, reshapes arrays and maps errors (InvalidIndex ->
IndexError, everything else -> ValueError, as py_generalized_bulk_flow.rs:117-126).

Additions beyond the reference surface are keyword-only or new methods:
``compute_c_integral(..., precision="fp64"|"mixed")``, ``compute_spectrum``,
``last_timings``.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _lib

__all__ = [
    "EmptyLattice", "ParallelepipedLattice", "LatticeBubbles", "ExponentialTimeCutoff",
    "GeneralizedBulkFlow", "fibonacci_directions",
]

_PRECISIONS = {"fp64": _lib.FP64, "mixed": _lib.MIXED, "fp64_blocked": _lib.FP64_BLOCKED}


def _precision(p) -> int:
    if isinstance(p, str):
        try:
            return _PRECISIONS[p.lower()]
        except KeyError:
            raise ValueError(f"precision must be 'fp64', 'fp64_blocked' or 'mixed', got {p!r}") from None
    if p in (_lib.FP64, _lib.MIXED, _lib.FP64_BLOCKED):
        return int(p)
    raise ValueError(f"precision must be 'fp64', 'fp64_blocked' or 'mixed', got {p!r}")


# ---------------------------------------------------------------------------
# Lattices (rust/bubble-gw/src/many_bubbles/lattice/)
# ---------------------------------------------------------------------------
class EmptyLattice:
    """No geometry: no containment test, no exterior bubbles (empty_lattice.rs)."""

    def name(self) -> str:
        return "EmptyLattice"

    @property
    def volume(self) -> float:
        return 0.0

    def contains(self, points):
        return []

    def generate_bubbles_exterior(self, interior, boundary_condition):
        return np.zeros((0, 4))


class ParallelepipedLattice:
    """origin + u e1 + v e2 + w e3, 0 <= u, v, w <= 1 (parallelepiped_lattice.rs)."""

    def __init__(self, origin, basis):
        self.origin = np.asarray(origin, dtype=np.float64).reshape(3)
        self.basis = np.asarray(basis, dtype=np.float64).reshape(3, 3)
        if abs(np.linalg.det(self.basis.T)) <= np.finfo(float).eps:
            raise ValueError("Invalid ParallelepipedLattice: degenerate basis")

    @staticmethod
    def axis_aligned(origin, lx, ly, lz) -> "ParallelepipedLattice":
        return ParallelepipedLattice(origin, [[lx, 0, 0], [0, ly, 0], [0, 0, lz]])

    @staticmethod
    def cube_centered(center, side) -> "ParallelepipedLattice":
        h = side / 2.0
        origin = np.asarray(center, dtype=np.float64) - h
        return ParallelepipedLattice(origin, [[side, 0, 0], [0, side, 0], [0, 0, side]])

    def name(self) -> str:
        return "ParallelepipedLattice"

    @property
    def volume(self) -> float:
        return float(abs(np.linalg.det(self.basis.T)))

    def contains(self, points):
        pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
        inv = np.linalg.inv(self.basis.T)
        uvw = (pts - self.origin) @ inv.T
        return list(np.all((uvw >= 0.0) & (uvw <= 1.0), axis=1))

    def generate_bubbles_exterior(self, interior, boundary_condition):
        """Periodic: six face images per interior bubble in the order
        +e1, -e1, +e2, -e2, +e3, -e3 (parallelepiped_lattice.rs:197-215)."""
        if boundary_condition != "periodic":
            raise ValueError(f"boundary condition {boundary_condition!r} is not supported")
        e1, e2, e3 = self.basis
        shifts = [e1, -e1, e2, -e2, e3, -e3]
        out = []
        for ev in np.asarray(interior, dtype=np.float64).reshape(-1, 4):
            for sh in shifts:
                out.append([ev[0], ev[1] + sh[0], ev[2] + sh[1], ev[3] + sh[2]])
        return np.asarray(out, dtype=np.float64).reshape(-1, 4)


def _bubble_name(i, n_int):
    return f"Interior({i})" if i < n_int else f"Exterior({i - n_int})"


class LatticeBubbles:
    """Interior (and exterior) bubbles (t, x, y, z) in a lattice (lattice_bubbles.rs:199-365)."""

    def __init__(self, lattice, bubbles_interior, bubbles_exterior=None):
        if not isinstance(lattice, (EmptyLattice, ParallelepipedLattice)):
            raise ValueError("Expected a lattice instance: ParallelepipedLattice or EmptyLattice")
        interior = np.ascontiguousarray(np.asarray(bubbles_interior, dtype=np.float64))
        exterior = (np.zeros((0, 4)) if bubbles_exterior is None
                    else np.ascontiguousarray(np.asarray(bubbles_exterior, dtype=np.float64)))
        if interior.ndim != 2 or exterior.ndim != 2 or interior.shape[1] != 4 or exterior.shape[1] != 4:
            raise ValueError(
                f"Array shape mismatch: Expected 4 columns, got {interior.shape[-1]} for interior, "
                f"{exterior.shape[-1]} for exterior")
        if isinstance(lattice, ParallelepipedLattice):
            inside = lattice.contains(interior[:, 1:])
            bad = [f"Interior({i})" for i, ok in enumerate(inside) if not ok]
            if bad:
                raise ValueError(f"Interior bubbles formed outside lattice: [{', '.join(bad)}]")
            inside = lattice.contains(exterior[:, 1:]) if len(exterior) else []
            bad = [f"Exterior({i})" for i, ok in enumerate(inside) if ok]
            if bad:
                raise ValueError(f"Exterior bubbles formed inside lattice: [{', '.join(bad)}]")
        self.lattice = lattice
        self._interior = interior
        self._exterior = exterior
        self._check_causality()

    def _check_causality(self):
        """check_bubble_formed_inside_bubble (lattice_bubbles.rs:71-119)."""
        n_int = len(self._interior)
        allb = np.vstack([self._interior, self._exterior])
        for a in range(n_int):
            d = allb[a + 1:] - allb[a]
            dsq = -d[:, 0] ** 2 + np.sum(d[:, 1:] ** 2, axis=1)
            bad = np.nonzero(dsq < 0.0)[0]
            if len(bad):
                b = a + 1 + int(bad[0])
                raise ValueError(
                    f"Bubble {_bubble_name(a, n_int)} is formed inside bubble {_bubble_name(b, n_int)}"
                    " at initial time (overlapping light cones)")

    def with_boundary_condition(self, boundary_condition="periodic") -> None:
        self._exterior = self.lattice.generate_bubbles_exterior(self._interior, boundary_condition)

    @property
    def bubbles_interior(self):
        return self._interior.copy()

    @property
    def bubbles_exterior(self):
        return self._exterior.copy()

    @property
    def volume(self) -> float:
        return self.lattice.volume

    def __repr__(self):
        return (f"LatticeBubbles(lattice_type={type(self.lattice).__name__}, "
                f"n_interior={len(self._interior)}, n_exterior={len(self._exterior)})")


@dataclass
class ExponentialTimeCutoff:
    """exp(-(t - t_cut)^2 / t_0^2) for t >= t_cut (time_cutoff.rs:13-43)."""
    smax: float
    ratio_t_cut: float | None = None
    ratio_t_0: float | None = None

    @property
    def t_cut(self) -> float:
        rc = 0.999999999 if self.ratio_t_cut is None else self.ratio_t_cut
        return rc * self.smax

    @property
    def t_0(self) -> float:
        r0 = 0.25 if self.ratio_t_0 is None else self.ratio_t_0
        return r0 * (self.smax - self.t_cut)


def fibonacci_directions(n: int) -> np.ndarray:
    """n unit vectors on a Fibonacci sphere: z_i = 1 - (2 i + 1) / n, golden-angle azimuths."""
    i = np.arange(n, dtype=np.float64)
    z = 1.0 - (2.0 * i + 1.0) / n
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = i * (math.pi * (3.0 - math.sqrt(5.0)))
    k = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    return k / np.linalg.norm(k, axis=1, keepdims=True)


# ---------------------------------------------------------------------------
# GeneralizedBulkFlow
# ---------------------------------------------------------------------------
class GeneralizedBulkFlow:
    """GPU drop-in for ``many_bubbles.GeneralizedBulkFlow`` (py_generalized_bulk_flow.rs:156-441)."""

    def __init__(self, lattice_bubbles: LatticeBubbles):
        if not isinstance(lattice_bubbles, LatticeBubbles):
            raise ValueError("lattice_bubbles must be a LatticeBubbles")
        L = _lib.lib()
        self._interior = np.ascontiguousarray(lattice_bubbles._interior)
        self._exterior = np.ascontiguousarray(lattice_bubbles._exterior)
        h = ctypes.c_void_p()
        st = L.bgw_generalized_bulk_flow_empty_lattice_new(
            ctypes.byref(h), _lib.dptr(self._interior), len(self._interior),
            _lib.dptr(self._exterior), len(self._exterior))
        if st != _lib.OK:
            raise ValueError(_lib.last_error())
        self._h = h
        self._coef = np.ones((1, 1))
        self._pows = np.full((1, 1), 3.0)
        self._damping = None
        self._n_c = None
        self._n_phi = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().bgw_generalized_bulk_flow_free(h)
            except Exception:
                pass
            self._h = None

    # -- plumbing ----------------------------------------------------------
    def _call(self, fn, *args):
        st = fn(self._h, *args)
        if st != _lib.OK:
            raise ValueError(_lib.last_error())

    def _push_params(self):
        coef = np.ascontiguousarray(self._coef, dtype=np.float64)
        pows = np.ascontiguousarray(self._pows, dtype=np.float64)
        if coef.ndim != 2 or coef.shape != pows.shape:
            raise ValueError("Array shape mismatch: coefficients_sets and powers_sets must have the "
                             f"same shape, got {coef.shape} and {pows.shape}")
        damping = float("nan") if self._damping is None else float(self._damping)
        self._call(_lib.lib().bgw_generalized_bulk_flow_set_gradient_scaling_params,
                   _lib.dptr(coef), _lib.dptr(pows), coef.shape[0], coef.shape[1], damping)

    def _check_index(self, a_idx):
        n = len(self._interior)
        if a_idx < 0 or a_idx >= n:
            raise IndexError(f"Index {a_idx} out of bounds (max: {n})")

    def _need_resolution(self):
        if self._n_c is None:
            raise ValueError("Field 'n_cos_thetax' is not initialized")

    # -- configuration -------------------------------------------------------
    def set_num_threads(self, num_threads: int) -> None:
        self._call(_lib.lib().bgw_generalized_bulk_flow_set_num_threads, int(num_threads))

    def set_coefficients_sets(self, coefficients_sets) -> None:
        self._coef = np.array(coefficients_sets, dtype=np.float64, ndmin=2)

    def set_powers_sets(self, powers_sets) -> None:
        self._pows = np.array(powers_sets, dtype=np.float64, ndmin=2)

    def set_gradient_scaling_params(self, coefficients_sets, powers_sets, damping_factor=None) -> None:
        """Validation of generalized_bulk_flow.rs:315-363."""
        coefficients_sets = [list(map(float, v)) for v in coefficients_sets]
        powers_sets = [list(map(float, v)) for v in powers_sets]
        n_c = len(coefficients_sets[0]) if coefficients_sets else 0
        if any(len(v) != n_c for v in coefficients_sets):
            raise ValueError("Array shape mismatch: All coefficient sets must have the same length")
        n_p = len(powers_sets[0]) if powers_sets else 0
        if any(len(v) != n_p for v in powers_sets):
            raise ValueError("Array shape mismatch: All power sets must have the same length")
        if len(coefficients_sets) != len(powers_sets):
            raise ValueError(
                "Array shape mismatch: Number of coefficient sets and power sets must be equal")
        self._coef = np.array(coefficients_sets, dtype=np.float64).reshape(len(coefficients_sets), n_c)
        self._pows = np.array(powers_sets, dtype=np.float64).reshape(len(powers_sets), n_p)
        self._damping = damping_factor

    def set_resolution(self, n_cos_thetax: int, n_phix: int, precompute_first_bubbles: bool = True) -> None:
        self._call(_lib.lib().bgw_generalized_bulk_flow_set_resolution, int(n_cos_thetax),
                   int(n_phix), bool(precompute_first_bubbles))
        self._n_c, self._n_phi = int(n_cos_thetax), int(n_phix)

    # -- getters ------------------------------------------------------------
    @property
    def bubbles_interior(self):
        return self._interior.copy()

    @property
    def bubbles_exterior(self):
        return self._exterior.copy()

    @property
    def coefficients_sets(self):
        return self._coef.copy()

    @property
    def get_powers_sets(self):
        return self._pows.copy()

    @property
    def cos_thetax(self):
        self._need_resolution()
        n = self._n_c
        return -1.0 + (2.0 / (n - 1)) * np.arange(n, dtype=np.float64)

    @property
    def phix(self):
        self._need_resolution()
        n = self._n_phi
        return ((2.0 * math.pi) / (n - 1)) * np.arange(n, dtype=np.float64)

    # -- masks ------------------------------------------------------------------
    def compute_first_colliding_bubble(self, a_idx: int):
        self._need_resolution()
        self._check_index(a_idx)
        out = np.empty((self._n_c, self._n_phi), dtype=np.int32)
        self._call(_lib.lib().bgw_generalized_bulk_flow_first_colliding_bubble, int(a_idx),
                   _lib.iptr(out), out.size)
        return out

    def compute_delta_tab(self, a_idx: int, first_bubble):
        self._need_resolution()
        self._check_index(a_idx)
        fb = np.ascontiguousarray(first_bubble, dtype=np.int32)
        if fb.shape != (self._n_c, self._n_phi):
            raise ValueError(
                "Array shape mismatch: First bubble array must match resolution: expected "
                f"[{self._n_c}, {self._n_phi}], got {list(fb.shape)}")
        out = np.empty(fb.shape, dtype=np.float64)
        self._call(_lib.lib().bgw_generalized_bulk_flow_delta_tab, int(a_idx), _lib.iptr(fb),
                   fb.size, _lib.dptr(out), out.size)
        return out

    def compute_collision_status(self, a_idx: int, t: float, first_bubble, delta_tab_grid):
        self._need_resolution()
        self._check_index(a_idx)
        fb = np.ascontiguousarray(first_bubble, dtype=np.int32)
        dt = np.ascontiguousarray(delta_tab_grid, dtype=np.float64)
        if fb.shape != (self._n_c, self._n_phi) or dt.shape != fb.shape:
            raise ValueError(
                "Array shape mismatch: Input arrays must match resolution: expected "
                f"[{self._n_c}, {self._n_phi}], got first_bubble: {list(fb.shape)}, "
                f"delta_tab_grid: {list(dt.shape)}")
        out = np.empty(fb.shape, dtype=np.int32)
        self._call(_lib.lib().bgw_generalized_bulk_flow_collision_status, int(a_idx), float(t),
                   _lib.iptr(fb), _lib.dptr(dt), fb.size, _lib.iptr(out), out.size)
        return out

    # -- integrals ----------------------------------------------------------------
    def compute_c_integral(self, w_arr, *, t_begin=None, t_end, n_t, selected_bubbles=None,
                           time_cutoff=None, precision="fp64"):
        """C+ and Cx, complex128 (2, n_sets, n_w) (generalized_bulk_flow.rs:928-984)."""
        w = np.ascontiguousarray(np.asarray(w_arr, dtype=np.float64).reshape(-1))
        sel = None
        if selected_bubbles is not None:
            sel_in = np.asarray(selected_bubbles)
            if sel_in.ndim != 1:
                raise ValueError("selected_bubbles must be a contiguous 1-D array in C order")
            sel = np.ascontiguousarray(sel_in.astype(np.uintp))
            for a in sel_in.tolist():
                if a < 0 or a >= len(self._interior):
                    raise IndexError(f"Index {a} out of bounds (max: {len(self._interior)})")
        self._push_params()
        n_sets = self._coef.shape[0]
        out = np.zeros((2, n_sets, len(w)), dtype=np.complex128)
        L = _lib.lib()
        self._call(L.bgw_generalized_bulk_flow_compute_c_integral_ex, _lib.dptr(w), len(w),
                   t_begin is not None, 0.0 if t_begin is None else float(t_begin), float(t_end),
                   int(n_t), _lib.sptr(sel) if sel is not None else None,
                   0 if sel is None else len(sel), _lib.cutoff_config(time_cutoff),
                   _precision(precision), out.ctypes.data_as(_lib._dp), out.size)
        return out

    def compute_c_integral_fixed_bubble(self, a_idx, w_arr, t_begin, t_end, n_t, time_cutoff=None):
        """One bubble's C (generalized_bulk_flow.rs:775-873)."""
        self._check_index(a_idx)
        w = np.asarray(w_arr, dtype=np.float64).reshape(-1)
        if n_t < 2:  # dt = 0: every time weight vanishes (:798-799)
            tb = 0.0 if t_begin is None else t_begin
            if tb > t_end:
                raise ValueError(f"Invalid time range: t_begin={tb} > t_end={t_end}")
            return np.zeros((2, self._coef.shape[0], len(w)), dtype=np.complex128)
        return self.compute_c_integral(w, t_begin=t_begin, t_end=t_end, n_t=n_t,
                                       selected_bubbles=np.array([a_idx]), time_cutoff=time_cutoff)

    def compute_c_integrand(self, w_arr, *, t_begin=None, t_end, n_t, selected_bubbles=None):
        """Time-resolved integrand, complex128 (2, n_sets, n_w, n_t) (:875-920)."""
        w = np.ascontiguousarray(np.asarray(w_arr, dtype=np.float64).reshape(-1))
        sel = None
        if selected_bubbles is not None:
            sel_in = np.asarray(selected_bubbles)
            sel = np.ascontiguousarray(sel_in.astype(np.uintp))
            for a in sel_in.tolist():
                if a < 0 or a >= len(self._interior):
                    raise IndexError(f"Index {a} out of bounds (max: {len(self._interior)})")
        self._push_params()
        n_sets = self._coef.shape[0]
        out = np.zeros((2, n_sets, len(w), int(n_t)), dtype=np.complex128)
        self._call(_lib.lib().bgw_generalized_bulk_flow_compute_c_integrand, _lib.dptr(w), len(w),
                   t_begin is not None, 0.0 if t_begin is None else float(t_begin), float(t_end),
                   int(n_t), _lib.sptr(sel) if sel is not None else None,
                   0 if sel is None else len(sel), out.ctypes.data_as(_lib._dp), out.size)
        return out

    def compute_c_integrand_fixed_bubble(self, a_idx, w_arr, t_begin, t_end, n_t):
        self._check_index(a_idx)
        return self.compute_c_integrand(w_arr, t_begin=t_begin, t_end=t_end, n_t=n_t,
                                        selected_bubbles=np.array([a_idx]))

    # -- additions -----------------------------------------------------------------
    def compute_spectrum(self, w_arr, k_hat, *, t_begin=None, t_end, n_t, time_cutoff=None,
                         precision="fp64", return_c=False):
        """S(k, s, w) = 8 pi w^3 (|C+|^2 + |Cx|^2) in the frame k-hat -> z, and its k-hat mean.

        Returns ``(S (n_k, n_sets, n_w), mean (n_sets, n_w))`` and, with
        ``return_c``, also C (n_k, 2, n_sets, n_w).
        """
        w = np.ascontiguousarray(np.asarray(w_arr, dtype=np.float64).reshape(-1))
        k = np.ascontiguousarray(np.asarray(k_hat, dtype=np.float64).reshape(-1, 3))
        self._push_params()
        n_sets = self._coef.shape[0]
        S = np.empty((len(k), n_sets, len(w)))
        M = np.empty((n_sets, len(w)))
        C = np.empty((len(k), 2, n_sets, len(w)), dtype=np.complex128) if return_c else None
        self._call(_lib.lib().bgw_generalized_bulk_flow_compute_spectrum, _lib.dptr(w), len(w),
                   _lib.dptr(k), len(k), t_begin is not None,
                   0.0 if t_begin is None else float(t_begin), float(t_end), int(n_t),
                   _lib.cutoff_config(time_cutoff), _precision(precision), _lib.dptr(S), S.size,
                   _lib.dptr(M), M.size, None if C is None else C.ctypes.data_as(_lib._dp),
                   0 if C is None else C.size)
        return (S, M, C) if return_c else (S, M)

    def spectrum_vjp(self, w_arr, k_hat, grad_c, *, t_begin=None, t_end, n_t, time_cutoff=None):
        """Gradient of a real loss L with respect to the scaling parameters.

        ``grad_c`` is dL/dRe C + i dL/dIm C for the C of :meth:`compute_spectrum`
        (n_k, 2, n_sets, n_w) complex.  Returns ``(dL/dcoefficients,
        dL/dpowers)``, each (n_sets, n_terms).  The hand-written FP64 adjoint
        kernel recomputes the collision masks.
        """
        w = np.ascontiguousarray(np.asarray(w_arr, dtype=np.float64).reshape(-1))
        k = np.ascontiguousarray(np.asarray(k_hat, dtype=np.float64).reshape(-1, 3))
        self._push_params()
        n_sets, n_terms = self._coef.shape
        g = np.ascontiguousarray(np.asarray(grad_c, dtype=np.complex128))
        if g.shape != (len(k), 2, n_sets, len(w)):
            raise ValueError(f"grad_c has shape {g.shape}, expected {(len(k), 2, n_sets, len(w))}")
        dc = np.empty((n_sets, n_terms))
        dp = np.empty((n_sets, n_terms))
        self._call(_lib.lib().bgw_generalized_bulk_flow_spectrum_vjp, _lib.dptr(w), len(w),
                   _lib.dptr(k), len(k), t_begin is not None,
                   0.0 if t_begin is None else float(t_begin), float(t_end), int(n_t),
                   _lib.cutoff_config(time_cutoff), g.ctypes.data_as(_lib._dp), g.size,
                   _lib.dptr(dc), _lib.dptr(dp), dc.size)
        return dc, dp

    def first_colliding_in_frame(self, a_idx: int, k_hat=None):
        """(first_bubble int32, delta_tab float64), both (n_c, n_phi), as computed by the
        GPU geometry kernel in the frame where ``k_hat`` -> z (None: reference frame)."""
        self._need_resolution()
        self._check_index(a_idx)
        first = np.empty((self._n_c, self._n_phi), dtype=np.int32)
        dtab = np.empty((self._n_c, self._n_phi), dtype=np.float64)
        k = None if k_hat is None else np.ascontiguousarray(np.asarray(k_hat, dtype=np.float64).reshape(3))
        self._call(_lib.lib().bgw_generalized_bulk_flow_first_colliding_bubble_ex, int(a_idx),
                   _lib.dptr(k), _lib.iptr(first), _lib.dptr(dtab), first.size)
        return first, dtab

    def set_tiles(self, block_steps: int = 0, threads: int = 0) -> None:
        """Launch shape of the phase-sum kernel (0: automatic)."""
        t = _lib.BgwTileConfig(int(block_steps), int(threads), 0, 0)
        self._call(_lib.lib().bgw_generalized_bulk_flow_set_tiles, ctypes.byref(t))

    def last_timings(self) -> dict:
        t = _lib.BgwTimings()
        self._call(_lib.lib().bgw_generalized_bulk_flow_last_timings, ctypes.byref(t))
        return {f: getattr(t, f) for f, _ in t._fields_}
