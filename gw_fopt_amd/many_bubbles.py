"""Python surface of the bulk-flow path, mirroring ``gw_fopt.bubble_gw.many_bubbles``.

Same class names, constructor and method signatures, dtypes, shapes and error
types as the reference's PyO3 module (rust/py-bubble-gw/src/py_many_bubbles/:
py_generalized_bulk_flow.rs:130-441, py_lattice_bubbles.rs, py_lattice.rs,
py_isometry.rs, py_bubbles_nucleation/; stub src/gw_fopt/bubble_gw/many_bubbles.pyi):

* ``Isometry3``; lattices ``EmptyLattice``, ``ParallelepipedLattice``,
  ``CartesianLattice``, ``SphericalLattice`` with their exterior images for the
  periodic, reflection and none boundary conditions,
* ``LatticeBubbles(lattice, bubbles_interior, bubbles_exterior=None)`` with
  ``with_boundary_condition``, ``transform``, ``translate_time``, ``normalize_time``,
* ``SpontaneousNucleation``, ``FixedRateNucleation`` (sampled on the device),
* ``ExponentialTimeCutoff(smax, ratio_t_cut=None, ratio_t_0=None)``,
* ``GeneralizedBulkFlow(lattice_bubbles)``.

Every computation runs in libbubble_gw_c.so on the GPU; the Python code only
validates arguments, reshapes arrays and maps errors (InvalidIndex ->
IndexError, everything else -> ValueError, as py_generalized_bulk_flow.rs:117-126).
Lattice geometry (containment, images) is evaluated here in the reference's
operation order.

Additions beyond the reference surface are keyword-only or new methods:
``compute_c_integral(..., precision="fp64"|"mixed")``, ``compute_spectrum``,
``last_timings``, ``nucleate_many`` and ``nucleate_fixed_rate_batch``.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

from . import _lib

__all__ = [
    "Isometry3", "EmptyLattice", "ParallelepipedLattice", "CartesianLattice", "SphericalLattice",
    "LatticeBubbles", "ExponentialTimeCutoff", "SpontaneousNucleation", "FixedRateNucleation",
    "FixedRateNucleationResult", "nucleate_fixed_rate_batch", "GeneralizedBulkFlow",
    "NewGeneralizedBulkFlow", "fibonacci_directions",
]

_PRECISIONS = {"fp64": _lib.FP64, "mixed": _lib.MIXED, "fp64_blocked": _lib.FP64_BLOCKED,
               "fp64_matrix": _lib.FP64_MATRIX}
_PRECISION_NAMES = "'fp64', 'fp64_blocked', 'fp64_matrix' or 'mixed'"


def _precision(p) -> int:
    if isinstance(p, str):
        try:
            return _PRECISIONS[p.lower()]
        except KeyError:
            raise ValueError(f"precision must be {_PRECISION_NAMES}, got {p!r}") from None
    if p in tuple(_PRECISIONS.values()):
        return int(p)
    raise ValueError(f"precision must be {_PRECISION_NAMES}, got {p!r}")


# ---------------------------------------------------------------------------
# Isometries and lattices (rust/bubble-gw/src/many_bubbles/lattice/,
# rust/py-bubble-gw/src/py_many_bubbles/py_lattice.rs, py_isometry.rs)
# ---------------------------------------------------------------------------
_EPS = float(np.finfo(np.float64).eps)


def _vec3(v, name="vector") -> np.ndarray:
    a = np.asarray(v, dtype=np.float64)
    if a.shape != (3,):
        raise ValueError(f"{name} must have 3 components")
    return a.copy()


def _mat3(m, name="matrix") -> np.ndarray:
    a = np.asarray(m, dtype=np.float64)
    if a.shape != (3, 3):
        raise ValueError(f"{name} must be 3x3")
    return a.copy()


def _tuple3(v):
    return tuple(float(x) for x in v)


def _euler_matrix(roll, pitch, yaw) -> np.ndarray:
    """nalgebra Rotation3::from_euler_angles: Rz(yaw) Ry(pitch) Rx(roll)."""
    sr, cr = math.sin(roll), math.cos(roll)
    sp, cp = math.sin(pitch), math.cos(pitch)
    sy, cy = math.sin(yaw), math.cos(yaw)
    return np.array([[cy * cp, cy * sp * sr - sy * cr, cy * sp * cr + sy * sr],
                     [sy * cp, sy * sp * sr + cy * cr, sy * sp * cr - cy * sr],
                     [-sp, cp * sr, cp * cr]])


def _nearest_rotation(m) -> np.ndarray:
    """The rotation closest to m (the reference goes through UnitQuaternion::from_matrix)."""
    u, _, vt = np.linalg.svd(_mat3(m, "rotation_matrix"))
    r = u @ vt
    if np.linalg.det(r) < 0.0:
        u[:, -1] = -u[:, -1]
        r = u @ vt
    return r


def _rotation_from_input(euler_angles, rotation_matrix) -> np.ndarray:
    if euler_angles is not None and rotation_matrix is not None:
        raise ValueError("Specify either euler_angles or rotation_matrix, not both.")
    if euler_angles is not None:
        return _euler_matrix(*_vec3(euler_angles, "euler_angles"))
    if rotation_matrix is not None:
        return _nearest_rotation(rotation_matrix)
    return np.eye(3)


class Isometry3:
    """Rotation followed by translation, x -> R x + t (py_isometry.rs).

    The rotation is held as a 3x3 matrix; the reference holds a unit
    quaternion, so composed rotations agree to rounding.
    """

    def __init__(self, translation=(0.0, 0.0, 0.0), *, euler_angles=None, rotation_matrix=None):
        self._t = _vec3(translation, "translation")
        self._r = _rotation_from_input(euler_angles, rotation_matrix)

    @classmethod
    def _from(cls, r, t) -> "Isometry3":
        iso = cls.__new__(cls)
        iso._r = np.asarray(r, dtype=np.float64)
        iso._t = np.asarray(t, dtype=np.float64)
        return iso

    @staticmethod
    def identity() -> "Isometry3":
        return Isometry3()

    @staticmethod
    def from_translation(translation) -> "Isometry3":
        return Isometry3(translation)

    @staticmethod
    def from_rotation(*, euler_angles=None, rotation_matrix=None) -> "Isometry3":
        return Isometry3(euler_angles=euler_angles, rotation_matrix=rotation_matrix)

    @staticmethod
    def from_euler_angles(roll: float, pitch: float, yaw: float) -> "Isometry3":
        return Isometry3(euler_angles=(roll, pitch, yaw))

    @staticmethod
    def from_axis_angle(axis, angle: float) -> "Isometry3":
        a = _vec3(axis, "axis")
        n = float(np.linalg.norm(a))
        if n < _EPS:
            raise ValueError("Axis must be a non-zero vector.")
        k = a / n
        K = np.array([[0.0, -k[2], k[1]], [k[2], 0.0, -k[0]], [-k[1], k[0], 0.0]])
        r = np.eye(3) + math.sin(angle) * K + (1.0 - math.cos(angle)) * (K @ K)
        return Isometry3._from(r, np.zeros(3))

    def transform(self, other: "Isometry3") -> "Isometry3":
        """self * other: apply other first."""
        return Isometry3._from(self._r @ other._r, self._r @ other._t + self._t)

    def rotate(self, *, euler_angles=None, rotation_matrix=None) -> "Isometry3":
        return Isometry3._from(_rotation_from_input(euler_angles, rotation_matrix) @ self._r, self._t)

    def translate(self, translation) -> "Isometry3":
        return Isometry3._from(self._r, self._t + _vec3(translation, "translation"))

    @property
    def translation(self):
        return _tuple3(self._t)

    @property
    def rotation_matrix(self):
        return tuple(_tuple3(row) for row in self._r)

    def apply_points(self, points) -> np.ndarray:
        p = np.asarray(points, dtype=np.float64).reshape(-1, 3)
        return p @ self._r.T + self._t

    def apply_vectors(self, vectors) -> np.ndarray:
        return np.asarray(vectors, dtype=np.float64).reshape(-1, 3) @ self._r.T

    def __repr__(self):
        t = self._t
        return f"Isometry3(translation=[{t[0]:.3f}, {t[1]:.3f}, {t[2]:.3f}], rotation=...)"


def _det3_columns(e) -> float:
    """nalgebra's 3x3 determinant of Matrix3::from_columns(e) (cofactor form)."""
    m = lambda r, c: float(e[c][r])  # noqa: E731
    minor_m12_m23 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2)
    minor_m11_m23 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2)
    minor_m11_m22 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1)
    return m(0, 0) * minor_m12_m23 - m(0, 1) * minor_m11_m23 + m(0, 2) * minor_m11_m22


def _inverse3_columns(e):
    """nalgebra's 3x3 try_inverse of Matrix3::from_columns(e); None if singular."""
    m = lambda r, c: float(e[c][r])  # noqa: E731
    minor_m12_m23 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2)
    minor_m11_m23 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2)
    minor_m11_m22 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1)
    det = m(0, 0) * minor_m12_m23 - m(0, 1) * minor_m11_m23 + m(0, 2) * minor_m11_m22
    if det == 0.0:
        return None
    return np.array([
        [minor_m12_m23 / det, (m(0, 2) * m(2, 1) - m(2, 2) * m(0, 1)) / det,
         (m(0, 1) * m(1, 2) - m(1, 1) * m(0, 2)) / det],
        [-minor_m11_m23 / det, (m(0, 0) * m(2, 2) - m(2, 0) * m(0, 2)) / det,
         (m(0, 2) * m(1, 0) - m(1, 2) * m(0, 0)) / det],
        [minor_m11_m22 / det, (m(0, 1) * m(2, 0) - m(2, 1) * m(0, 0)) / det,
         (m(0, 0) * m(1, 1) - m(1, 0) * m(0, 1)) / det]])


def _check_bc(boundary_condition) -> int:
    bc = str(boundary_condition).lower()
    if bc == "periodic":
        return _lib.BC_PERIODIC
    if bc == "reflection":
        return _lib.BC_REFLECTION
    if bc == "none":
        return _lib.BC_NONE
    raise ValueError("Invalid boundary condition. Expected 'periodic', 'reflection', or 'none'.")


def _dot3(a, b) -> float:
    return float(a[0]) * float(b[0]) + float(a[1]) * float(b[1]) + float(a[2]) * float(b[2])


class EmptyLattice:
    """No geometry: no containment test, no exterior bubbles (empty_lattice.rs)."""

    def name(self) -> str:
        return "EmptyLattice"

    def transform(self, iso: Isometry3) -> "EmptyLattice":
        return EmptyLattice()

    @property
    def volume(self) -> float:
        return 0.0

    def reference_point(self):
        return (0.0, 0.0, 0.0)

    def parameters(self):
        return []

    def contains(self, points):
        return []

    def generate_bubbles_exterior(self, interior, boundary_condition):
        _check_bc(boundary_condition)
        return np.zeros((0, 4))

    def __repr__(self):
        return "EmptyLattice"


class ParallelepipedLattice:
    """origin + u e1 + v e2 + w e3, 0 <= u, v, w <= 1 (parallelepiped_lattice.rs)."""

    _kind = _lib.LATTICE_PARALLELEPIPED

    def __init__(self, origin, basis):
        self._origin = _vec3(origin, "origin")
        self._basis = _mat3(basis, "basis")
        if abs(_det3_columns(self._basis)) <= _EPS:
            raise ValueError("Basis vectors are linearly dependent (zero volume)")

    @classmethod
    def _unchecked(cls, origin, basis):
        obj = cls.__new__(cls)
        obj._origin = np.asarray(origin, dtype=np.float64)
        obj._basis = np.asarray(basis, dtype=np.float64)
        return obj

    @staticmethod
    def axis_aligned(origin, lx, ly, lz) -> "ParallelepipedLattice":
        return ParallelepipedLattice(origin, [[lx, 0, 0], [0, ly, 0], [0, 0, lz]])

    @staticmethod
    def cube_centered(center, side) -> "ParallelepipedLattice":
        h = side / 2.0
        origin = _vec3(center, "center") - h
        return ParallelepipedLattice(origin, [[side, 0, 0], [0, side, 0], [0, 0, side]])

    def name(self) -> str:
        return "ParallelepipedLattice"

    def transform(self, iso: Isometry3):
        return type(self)._unchecked(iso.apply_points(self._origin)[0], iso.apply_vectors(self._basis))

    @property
    def origin(self):
        return _tuple3(self._origin)

    @property
    def basis(self):
        return tuple(_tuple3(e) for e in self._basis)

    @property
    def normalized_basis(self):
        out = []
        for e in self._basis:
            n2 = _dot3(e, e)
            out.append(_tuple3(e / math.sqrt(n2)) if n2 > _EPS else (0.0, 0.0, 0.0))
        return tuple(out)

    @property
    def volume(self) -> float:
        return abs(_det3_columns(self._basis))

    def reference_point(self):
        return _tuple3(self._origin)

    def parameters(self):
        return [float(x) for x in (*self._origin, *self._basis.reshape(-1))]

    def contains(self, points):
        pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
        inv = _inverse3_columns(self._basis)
        if inv is None:
            return [False] * len(pts)
        out = []
        for p in pts:
            d = p - self._origin
            ok = True
            for i in range(3):
                u = inv[i, 0] * d[0] + inv[i, 1] * d[1] + inv[i, 2] * d[2]
                ok = ok and 0.0 <= u <= 1.0
            out.append(bool(ok))
        return out

    def _lattice_struct(self) -> "_lib.BgwNucleationLattice":
        st = _lib.BgwNucleationLattice()
        st.kind = self._kind
        st.origin[:] = list(self._origin)
        st.basis[:] = list(self._basis.reshape(-1))
        st.radius = 0.0
        return st

    def generate_bubbles_exterior(self, interior, boundary_condition):
        """Periodic: six face translates per interior bubble in the order +e1, -e1, +e2,
        -e2, +e3, -e3 (parallelepiped_lattice.rs:197-216); reflection: mirror images
        in the faces -e1, +e1, -e2, +e2, -e3, +e3 (:217-295); none: no images."""
        bc = _check_bc(boundary_condition)
        ev = np.asarray(interior, dtype=np.float64).reshape(-1, 4)
        if bc == _lib.BC_NONE:
            return np.zeros((0, 4))
        out = np.empty((len(ev), 6, 4))
        out[:, :, 0] = ev[:, None, 0]
        if bc == _lib.BC_PERIODIC:
            for j in range(3):
                out[:, 2 * j, 1:] = ev[:, 1:] + self._basis[j]
                out[:, 2 * j + 1, 1:] = ev[:, 1:] + (-self._basis[j])
            return out.reshape(-1, 4)
        axis = np.eye(3)
        for j in range(3):
            e = self._basis[j]
            nrm = math.sqrt(_dot3(e, e))
            n = e / nrm if nrm > _EPS else axis[j]
            q = self._origin + e
            for b in range(len(ev)):
                p = ev[b, 1:]
                nn = -n
                d0 = _dot3(p - self._origin, nn)
                out[b, 2 * j, 1:] = p - 2.0 * d0 * nn
                d1 = _dot3(p - q, n)
                out[b, 2 * j + 1, 1:] = p - 2.0 * d1 * n
        return out.reshape(-1, 4)

    def __repr__(self):
        return f"{self.name()}(origin={self.origin}, basis={self.basis})"


class CartesianLattice(ParallelepipedLattice):
    """A parallelepiped with pairwise orthogonal basis vectors (cartesian_lattice.rs)."""

    def __init__(self, origin, basis):
        super().__init__(origin, basis)
        e1, e2, e3 = self._basis
        eps = 1e-10
        if abs(_dot3(e1, e2)) > eps or abs(_dot3(e1, e3)) > eps or abs(_dot3(e2, e3)) > eps:
            raise ValueError("Basis vectors are not pairwise orthogonal")

    @staticmethod
    def with_origin_and_sizes(origin, sizes) -> "CartesianLattice":
        lx, ly, lz = (float(x) for x in sizes)
        return CartesianLattice(origin, [[lx, 0, 0], [0, ly, 0], [0, 0, lz]])

    def name(self) -> str:
        return "CartesianLattice"

    @property
    def side_lengths(self):
        return tuple(math.sqrt(_dot3(e, e)) for e in self._basis)


class SphericalLattice:
    """|x - center| <= radius (spherical_lattice.rs)."""

    _kind = _lib.LATTICE_SPHERICAL

    def __init__(self, center, radius: float):
        self._center = _vec3(center, "center")
        self._radius = float(radius)

    def name(self) -> str:
        return "SphericalLattice"

    def transform(self, iso: Isometry3) -> "SphericalLattice":
        return SphericalLattice(iso.apply_points(self._center)[0], self._radius)

    @property
    def center(self):
        return _tuple3(self._center)

    @property
    def radius(self) -> float:
        return self._radius

    @property
    def volume(self) -> float:
        r = self._radius
        return (4.0 / 3.0) * math.pi * (r * r * r)

    def reference_point(self):
        return _tuple3(self._center)

    def parameters(self):
        return [*map(float, self._center), self._radius]

    def contains(self, points):
        pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
        r2 = self._radius * self._radius
        return [bool(_dot3(p - self._center, p - self._center) <= r2) for p in pts]

    def _lattice_struct(self) -> "_lib.BgwNucleationLattice":
        st = _lib.BgwNucleationLattice()
        st.kind = self._kind
        st.origin[:] = list(self._center)
        st.basis[:] = [0.0] * 9
        st.radius = self._radius
        return st

    def generate_bubbles_exterior(self, interior, boundary_condition):
        """Reflection: the mirror images in the tangent planes where the line through
        the center and the bubble meets the sphere (spherical_lattice.rs:72-131);
        periodic and none: no images."""
        bc = _check_bc(boundary_condition)
        ev = np.asarray(interior, dtype=np.float64).reshape(-1, 4)
        if bc != _lib.BC_REFLECTION:
            return np.zeros((0, 4))
        c, r = self._center, self._radius
        out = np.empty((len(ev), 2, 4))
        for b, e in enumerate(ev):
            p = e[1:]
            v = p - c
            d = math.sqrt(_dot3(v, v))
            out[b, :, 0] = e[0]
            if d < _EPS:
                out[b, 0, 1:] = (c[0] + r, c[1], c[2])
                out[b, 1, 1:] = (c[0] - r, c[1], c[2])
                continue
            u = v / d
            for side, (q, n) in enumerate(((c + r * u, u), (c - r * u, -u))):
                dd = _dot3(p - q, n)
                out[b, side, 1:] = p - 2.0 * dd * n
        return out.reshape(-1, 4)

    def __repr__(self):
        return f"SphericalLattice(center={self.center}, radius={self._radius})"


_LATTICE_TYPES = (ParallelepipedLattice, CartesianLattice, SphericalLattice, EmptyLattice)


def _lattice_type_name(lattice) -> str:
    if isinstance(lattice, CartesianLattice):
        return "Cartesian"
    if isinstance(lattice, ParallelepipedLattice):
        return "Parallelepiped"
    if isinstance(lattice, SphericalLattice):
        return "Spherical"
    return "Empty"


def _bubble_name(i, n_int):
    return f"Interior({i})" if i < n_int else f"Exterior({i - n_int})"


class LatticeBubbles:
    """Interior (and exterior) bubbles (t, x, y, z) in a lattice (lattice_bubbles.rs:199-460,
    py_lattice_bubbles.rs)."""

    def __init__(self, lattice, bubbles_interior, bubbles_exterior=None):
        if not isinstance(lattice, _LATTICE_TYPES):
            raise ValueError("Expected a lattice instance: ParallelepipedLattice, CartesianLattice, "
                             "SphericalLattice, or EmptyLattice")
        interior = np.ascontiguousarray(np.asarray(bubbles_interior, dtype=np.float64))
        exterior = (np.zeros((0, 4)) if bubbles_exterior is None
                    else np.ascontiguousarray(np.asarray(bubbles_exterior, dtype=np.float64)))
        if interior.ndim != 2 or exterior.ndim != 2 or interior.shape[1] != 4 or exterior.shape[1] != 4:
            raise ValueError(
                f"Array shape mismatch: Expected 4 columns, got {interior.shape[-1]} for interior, "
                f"{exterior.shape[-1]} for exterior")
        if not isinstance(lattice, EmptyLattice):
            inside = lattice.contains(interior[:, 1:])
            bad = [f"Interior({i})" for i, ok in enumerate(inside) if not ok]
            if bad:
                raise ValueError(f"Interior bubbles formed outside lattice: [{', '.join(bad)}]")
            inside = lattice.contains(exterior[:, 1:]) if len(exterior) else []
            bad = [f"Exterior({i})" for i, ok in enumerate(inside) if ok]
            if bad:
                raise ValueError(f"Exterior bubbles formed inside lattice: [{', '.join(bad)}]")
        self._lattice = lattice
        self._interior = interior
        self._exterior = exterior
        self._check_causality()

    @classmethod
    def _unchecked(cls, lattice, interior, exterior) -> "LatticeBubbles":
        obj = cls.__new__(cls)
        obj._lattice = lattice
        obj._interior = np.ascontiguousarray(interior, dtype=np.float64)
        obj._exterior = np.ascontiguousarray(exterior, dtype=np.float64).reshape(-1, 4)
        return obj

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
        """Replace the exterior by the lattice's images of the interior (lattice_bubbles.rs:323-365)."""
        self._exterior = np.ascontiguousarray(
            self._lattice.generate_bubbles_exterior(self._interior, boundary_condition))

    def transform(self, iso: Isometry3) -> "LatticeBubbles":
        """Positions x -> R x + t, times unchanged, lattice transformed (lattice_bubbles.rs:144-197)."""
        def move(b):
            out = b.copy()
            if len(b):
                out[:, 1:] = iso.apply_points(b[:, 1:])
            return out
        return LatticeBubbles._unchecked(self._lattice.transform(iso), move(self._interior),
                                         move(self._exterior))

    def translate_time_mut(self, t_shift: float) -> None:
        """Add t_shift to every bubble time (lattice_bubbles.rs:392-402)."""
        self._interior = self._interior.copy()
        self._exterior = self._exterior.copy()
        self._interior[:, 0] += t_shift
        self._exterior[:, 0] += t_shift

    def translate_time(self, t_shift: float) -> "LatticeBubbles":
        out = LatticeBubbles._unchecked(self._lattice, self._interior, self._exterior)
        out.translate_time_mut(t_shift)
        return out

    def normalize_time_mut(self) -> None:
        """Shift times so the earliest bubble (interior or exterior) is at t = 0 (:439-460)."""
        times = np.concatenate([self._interior[:, 0], self._exterior[:, 0]])
        if len(times) and np.isfinite(times.min()):
            self.translate_time_mut(-float(times.min()))

    def normalize_time(self) -> "LatticeBubbles":
        out = LatticeBubbles._unchecked(self._lattice, self._interior, self._exterior)
        out.normalize_time_mut()
        return out

    @property
    def lattice(self):
        return self._lattice

    @property
    def lattice_type(self) -> str:
        return _lattice_type_name(self._lattice)

    @property
    def bubbles_interior(self):
        return self._interior.copy()

    @property
    def bubbles_exterior(self):
        return self._exterior.copy()

    @property
    def volume(self) -> float:
        return self._lattice.volume

    @property
    def reference_point(self):
        return self._lattice.reference_point()

    def parameters(self):
        return self._lattice.parameters()

    def contains(self, points):
        return self._lattice.contains(points)

    def __repr__(self):
        return (f"LatticeBubbles(lattice_type={self.lattice_type}, "
                f"n_interior={len(self._interior)}, n_exterior={len(self._exterior)})")

    # -- CSV files: rows t,x,y,z (lattice_bubbles.rs:468-542) --------------------------
    @staticmethod
    def load_bubbles_from_csv(path, has_headers: bool = False) -> np.ndarray:
        """(n, 4) float64 from a comma-separated file; every row must have 4 fields."""
        import csv
        with open(path, newline="") as fh:
            rows = list(csv.reader(fh))
        if has_headers and rows:
            rows = rows[1:]
        rows = [r for r in rows if r]
        width = {len(r) for r in rows}
        if len(width) > 1:
            raise ValueError("deserialize_array2 error")
        if rows and width != {4}:
            raise ValueError("Invalid number of columns")
        try:
            arr = np.array([[float(v) for v in r] for r in rows], dtype=np.float64)
        except ValueError:
            raise ValueError("deserialize_array2 error") from None
        return arr.reshape(-1, 4)

    @classmethod
    def from_csv_files(cls, interior_path, exterior_path, lattice,
                       has_headers: bool = False) -> "LatticeBubbles":
        return cls(lattice, cls.load_bubbles_from_csv(interior_path, has_headers),
                   cls.load_bubbles_from_csv(exterior_path, has_headers))

    @staticmethod
    def _write_csv(bubbles, path, has_headers: bool) -> None:
        def fmt(x: float) -> str:
            r = repr(float(x))  # shortest representation that reads back exactly
            if "e" in r:
                mant, exp = r.split("e")
                r = f"{mant}e{int(exp)}"
            return r
        with open(path, "w", newline="") as fh:
            if has_headers:
                fh.write("t,x,y,z\n")
            for row in np.asarray(bubbles, dtype=np.float64).reshape(-1, 4):
                fh.write(",".join(fmt(v) for v in row) + "\n")

    def write_bubbles_interior_to_csv(self, path, has_headers: bool = False) -> None:
        self._write_csv(self._interior, path, has_headers)

    def save_bubbles_exterior_to_csv(self, path, has_headers: bool = False) -> None:
        self._write_csv(self._exterior, path, has_headers)

    def save_to_csv_files(self, interior_path, exterior_path, has_headers: bool = False) -> None:
        self.write_bubbles_interior_to_csv(interior_path, has_headers)
        self.save_bubbles_exterior_to_csv(exterior_path, has_headers)


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


# ---------------------------------------------------------------------------
# Nucleation on the device (bubbles_nucleation/*.rs, py_bubbles_nucleation/*.rs)
# ---------------------------------------------------------------------------
_METHODS = {"fixed_probability_time_stepping": _lib.FIXED_RATE_TIME_STEPPING,
            "fixed_probability_distribution": _lib.FIXED_RATE_DISTRIBUTION}


def _nucleation_lattice(lattice):
    if isinstance(lattice, (ParallelepipedLattice, SphericalLattice)):
        return lattice._lattice_struct()
    if isinstance(lattice, EmptyLattice):
        raise ValueError("nucleation needs a lattice with a volume: EmptyLattice has no sample points")
    raise ValueError("Expected a lattice instance: ParallelepipedLattice, CartesianLattice, "
                     "SphericalLattice, or EmptyLattice")


def _random_seed() -> int:
    return int.from_bytes(os.urandom(8), "little")


def _rust_float(x: float) -> str:
    return str(int(x)) if float(x).is_integer() and abs(x) < 1e16 else repr(float(x))


def _rust_option(x) -> str:
    return "None" if x is None else f"Some({x})"


def _images_per_bubble(lattice, bc: int) -> int:
    return int(_lib.lib().bgw_nucleation_images_per_bubble(lattice._kind, bc))


class SpontaneousNucleation:
    """n_bubbles bubbles at t0, uniform in the lattice (spontaneous_nucleation.rs,
    py_spontaneous_nucleation.rs), sampled on the device.

    Random numbers come from Philox4x64-10 with key (seed, stream); each call
    to ``nucleate`` uses the next stream, as the reference's generator advances
    between calls.
    """

    def __init__(self, n_bubbles: int, t0: float = 0.0, seed: int | None = None):
        self._n_bubbles = int(n_bubbles)
        self._t0 = float(t0)
        self._seed = seed
        self._key = int(seed) if seed is not None else _random_seed()
        self._stream = 0

    @property
    def n_bubbles(self) -> int:
        return self._n_bubbles

    @property
    def t0(self) -> float:
        return self._t0

    @property
    def seed(self):
        return self._seed

    def __repr__(self):
        return (f"SpontaneousNucleation(n_bubbles={self._n_bubbles}, t0={_rust_float(self._t0)}, "
                f"seed={_rust_option(self._seed)})")

    def nucleate_many(self, lattice, n_realisations: int, boundary_condition="periodic"):
        """Interior (n_real, n_bubbles, 4) and exterior (n_real, n_bubbles * n_img, 4) of
        n_realisations independent realisations, in one launch."""
        bc = _check_bc(boundary_condition)
        st = _nucleation_lattice(lattice)
        n_img = _images_per_bubble(lattice, bc)
        n_real = int(n_realisations)
        interior = np.zeros((n_real, self._n_bubbles, 4))
        exterior = np.zeros((n_real, self._n_bubbles * n_img, 4))
        status = _lib.lib().bgw_nucleation_spontaneous(
            ctypes.byref(st), bc, self._n_bubbles, self._t0, ctypes.c_uint64(self._key & (2**64 - 1)),
            ctypes.c_uint64(self._stream), n_real, _lib.dptr(interior),
            _lib.dptr(exterior) if n_img else None)
        if status != _lib.OK:
            raise ValueError(_lib.last_error())
        self._stream += n_real
        return interior, exterior

    def nucleate(self, lattice, boundary_condition="periodic") -> LatticeBubbles:
        interior, exterior = self.nucleate_many(lattice, 1, boundary_condition)
        return LatticeBubbles(lattice, interior[0], exterior[0])


@dataclass
class FixedRateNucleationResult:
    """lattice_bubbles, time_history and volume_false_vacuum_history of one run."""
    lattice_bubbles: LatticeBubbles
    time_history: np.ndarray
    volume_false_vacuum_history: np.ndarray


def nucleate_fixed_rate_batch(specs, *, boundary_condition="periodic",
                              method="fixed_probability_distribution", d_p0: float = 0.01,
                              n_points: int = 10000, max_time_iterations: int | None = None,
                              cutoff_fraction_false_vacuum: float | None = None,
                              max_bubbles: int | None = None):
    """Fixed-rate nucleation of many realisations in one device launch.

    ``specs``: a sequence of (lattice, beta, gamma0, t0, seed, stream); every
    lattice must be of one kind.  Returns (interior, exterior, time_history,
    volume_history, counts) with interior (n_real, max_bubbles, 4), exterior
    (n_real, max_bubbles * n_img, 4), histories (n_real, max_bubbles); row
    counts[r] and beyond are unused.  ``max_bubbles`` defaults to n_points,
    which no realisation can exceed.
    """
    bc = _check_bc(boundary_condition)
    m = _METHODS.get(str(method).lower())
    if m is None:
        raise ValueError("method must be 'fixed_probability_time_stepping' or "
                         "'fixed_probability_distribution'")
    specs = list(specs)
    n_real = len(specs)
    reals = (_lib.BgwFixedRateRealisation * max(n_real, 1))()
    n_img = None
    for r, (lattice, beta, gamma0, t0, seed, stream) in enumerate(specs):
        reals[r].lattice = _nucleation_lattice(lattice)
        reals[r].beta, reals[r].gamma0, reals[r].t0 = float(beta), float(gamma0), float(t0)
        reals[r].seed = int(seed) & (2**64 - 1)
        reals[r].stream = int(stream) & (2**64 - 1)
        k = _images_per_bubble(lattice, bc)
        if n_img is not None and k != n_img:
            raise ValueError("every realisation must use the same lattice kind")
        n_img = k
    n_img = n_img or 0
    cap = int(n_points if max_bubbles is None else max_bubbles)
    max_it = 1_000_000 if max_time_iterations is None else int(max_time_iterations)
    cutoff = 0.01 if cutoff_fraction_false_vacuum is None else float(cutoff_fraction_false_vacuum)
    interior = np.zeros((n_real, cap, 4))
    exterior = np.zeros((n_real, cap * n_img, 4))
    t_hist = np.zeros((n_real, cap))
    v_hist = np.zeros((n_real, cap))
    counts = np.zeros(n_real, dtype=np.uintp)
    status = _lib.lib().bgw_nucleation_fixed_rate(
        reals, n_real, bc, m, float(d_p0), int(n_points), max_it, cutoff, cap,
        _lib.dptr(interior), _lib.dptr(exterior) if n_img else None, _lib.dptr(t_hist),
        _lib.dptr(v_hist), _lib.sptr(counts))
    if status != _lib.OK:
        raise ValueError(_lib.last_error())
    return interior, exterior, t_hist, v_hist, counts.astype(np.int64)


def last_nucleation_kernel_ms() -> float:
    """Device time of the most recent fixed-rate nucleation kernel on this thread."""
    out = ctypes.c_double(0.0)
    _lib.check(_lib.lib().bgw_nucleation_last_kernel_ms(ctypes.byref(out)))
    return out.value


class FixedRateNucleation:
    """Monte Carlo nucleation at rate gamma0 exp(beta (t - t0)) per unit volume of false
    vacuum (fixed_rate_nucleation.rs, py_fixed_rate_nucleation.rs), on the device.

    The false-vacuum volume is the fraction of n_points sample points that no
    bubble (or exterior image) has reached; a new bubble nucleates at the first
    sample point still in the false vacuum.  Random numbers come from
    Philox4x64-10 with key (seed, stream); each realisation uses the next stream.
    """

    def __init__(self, beta: float, gamma0: float, t0: float, d_p0: float, seed: int | None = None,
                 n_points: int = 10000, max_time_iterations: int | None = None,
                 cutoff_fraction_false_vacuum: float | None = None):
        if any(math.isnan(float(x)) for x in (beta, gamma0, t0, d_p0)):
            raise ValueError("Parameters must be finite")
        if d_p0 <= 0.0 or d_p0 > 1.0:
            raise ValueError("d_p0 must be in (0, 1]")
        if gamma0 < 0.0:
            raise ValueError("gamma0 must be non-negative")
        if int(n_points) == 0:
            raise ValueError("n_points must be > 0")
        self._beta, self._gamma0, self._t0 = float(beta), float(gamma0), float(t0)
        self._seed = seed
        self._key = int(seed) if seed is not None else _random_seed()
        self._stream = 0
        self._n_points = int(n_points)
        self._max_it = 1_000_000 if max_time_iterations is None else int(max_time_iterations)
        self._cutoff = 0.01 if cutoff_fraction_false_vacuum is None else float(cutoff_fraction_false_vacuum)

    @property
    def beta(self) -> float:
        return self._beta

    @property
    def gamma0(self) -> float:
        return self._gamma0

    @property
    def t0(self) -> float:
        return self._t0

    @property
    def seed(self):
        return self._seed

    @property
    def n_points(self) -> int:
        return self._n_points

    def __repr__(self):
        return (f"FixedRateNucleation(beta={_rust_float(self._beta)}, gamma0={_rust_float(self._gamma0)}, "
                f"t0={_rust_float(self._t0)}, seed={_rust_option(self._seed)}, n_points={self._n_points})")

    def solve_bubbles_distribution(self, taumax: float, volume_lattice: float, rtol: float = 1e-9,
                                   atol: float = 1e-12):
        """(t, n): the points of the even grid of step 0.001 in tau = beta t where the
        expected bubble count n crosses an integer (fixed_rate_nucleation.rs:144-186).
        The integrator is fourth-order Runge-Kutta on that grid (error far below
        rtol and atol, which are accepted for the reference's signature)."""
        L = _lib.lib()
        n = ctypes.c_size_t(0)
        args = (float(self._beta), float(self._gamma0), float(self._t0), float(taumax),
                float(volume_lattice))
        _lib.check(L.bgw_nucleation_solve_distribution(*args, None, None, 0, ctypes.byref(n)))
        t = np.zeros(n.value)
        nb = np.zeros(n.value)
        _lib.check(L.bgw_nucleation_solve_distribution(*args, _lib.dptr(t), _lib.dptr(nb), n.value,
                                                       ctypes.byref(n)))
        return t, nb

    def nucleate_many(self, lattice, n_realisations: int, boundary_condition="periodic",
                      method="fixed_probability_distribution", d_p0: float = 0.01,
                      max_bubbles: int | None = None):
        """n_realisations independent runs in one device launch: (interior, exterior,
        time_history, volume_history, counts) as ``nucleate_fixed_rate_batch``."""
        n_real = int(n_realisations)
        specs = [(lattice, self._beta, self._gamma0, self._t0, self._key, self._stream + r)
                 for r in range(n_real)]
        out = nucleate_fixed_rate_batch(
            specs, boundary_condition=boundary_condition, method=method, d_p0=d_p0,
            n_points=self._n_points, max_time_iterations=self._max_it,
            cutoff_fraction_false_vacuum=self._cutoff, max_bubbles=max_bubbles)
        self._stream += n_real
        return out

    def nucleate(self, lattice, boundary_condition="periodic", method="fixed_probability_distribution",
                 d_p0: float = 0.01) -> FixedRateNucleationResult:
        interior, exterior, t_hist, v_hist, counts = self.nucleate_many(
            lattice, 1, boundary_condition, method, d_p0)
        n = int(counts[0])
        n_img = exterior.shape[1] // max(interior.shape[1], 1)
        lb = LatticeBubbles(lattice, interior[0, :n], exterior[0, :n * n_img])
        return FixedRateNucleationResult(lb, t_hist[0, :n].copy(), v_hist[0, :n].copy())


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


class NewGeneralizedBulkFlow(GeneralizedBulkFlow):
    """The reference's experimental tensor form (new_generalized_bulk_flow.rs,
    py_new_generalized_bulk_flow.rs): C_ij of the bubbles in the reference frame,
    complex128 (6, n_sets, n_w) in the order xx, xy, xz, yy, yz, zz, without a time
    cutoff.  The projections of the plain form are C+ = (C_xx - C_yy) / 2 and
    Cx = C_xy.  Everything else (resolution, scaling parameters, masks) is that of
    ``GeneralizedBulkFlow``."""

    def _selection(self, selected_bubbles):
        if selected_bubbles is None:
            return None
        sel_in = np.asarray(selected_bubbles)
        if sel_in.ndim != 1:
            raise ValueError("selected_bubbles must be a contiguous 1-D array in C order")
        for a in sel_in.tolist():
            if a < 0 or a >= len(self._interior):
                raise IndexError(f"Index {a} out of bounds (max: {len(self._interior)})")
        return np.ascontiguousarray(sel_in.astype(np.uintp))

    def compute_c_tensor(self, w_arr, *, t_begin=None, t_end, n_t, selected_bubbles=None,
                         precision="fp64"):
        """C_ij, complex128 (6, n_sets, n_w) (new_generalized_bulk_flow.rs:1205-1275)."""
        w = np.ascontiguousarray(np.asarray(w_arr, dtype=np.float64).reshape(-1))
        sel = self._selection(selected_bubbles)
        self._push_params()
        out = np.zeros((6, self._coef.shape[0], len(w)), dtype=np.complex128)
        self._call(_lib.lib().bgw_generalized_bulk_flow_compute_c_tensor, _lib.dptr(w), len(w),
                   t_begin is not None, 0.0 if t_begin is None else float(t_begin), float(t_end),
                   int(n_t), _lib.sptr(sel) if sel is not None else None,
                   0 if sel is None else len(sel), _precision(precision),
                   out.ctypes.data_as(_lib._dp), out.size)
        return out

    def compute_c_tensor_fixed_bubble(self, a_idx, w_arr, t_begin, t_end, n_t):
        """One bubble's C_ij (new_generalized_bulk_flow.rs:1067-1203)."""
        self._check_index(a_idx)
        return self.compute_c_tensor(w_arr, t_begin=t_begin, t_end=t_end, n_t=n_t,
                                     selected_bubbles=np.array([a_idx]))

    def compute_c_tensor_integrand(self, w_arr, *, t_begin=None, t_end, n_t, selected_bubbles=None):
        """Time-resolved integrand of C_ij, complex128 (6, n_sets, n_w, n_t) (:993-1065)."""
        w = np.ascontiguousarray(np.asarray(w_arr, dtype=np.float64).reshape(-1))
        sel = self._selection(selected_bubbles)
        self._push_params()
        out = np.zeros((6, self._coef.shape[0], len(w), int(n_t)), dtype=np.complex128)
        self._call(_lib.lib().bgw_generalized_bulk_flow_compute_c_tensor_integrand, _lib.dptr(w),
                   len(w), t_begin is not None, 0.0 if t_begin is None else float(t_begin),
                   float(t_end), int(n_t), _lib.sptr(sel) if sel is not None else None,
                   0 if sel is None else len(sel), out.ctypes.data_as(_lib._dp), out.size)
        return out

    def compute_c_tensor_integrand_fixed_bubble(self, a_idx, w_arr, t_begin, t_end, n_t):
        self._check_index(a_idx)
        return self.compute_c_tensor_integrand(w_arr, t_begin=t_begin, t_end=t_end, n_t=n_t,
                                               selected_bubbles=np.array([a_idx]))
