"""CPU oracle for bubble nucleation — TEST INFRASTRUCTURE ONLY.

Wraps ``liboracle_nucleation.so`` (built from ``nucleation_oracle.cpp`` by
``oracle/Makefile``), a C++ restatement of the reference's fixed-rate and
spontaneous nucleation (rust/bubble-gw/src/many_bubbles/bubbles_nucleation/)
and lattice geometry, drawing from the same Philox4x64-10 streams as the
device kernels.  Only ``tests/`` and the CPU baseline of ``bench.py`` use it.

Lattice parameters: ("box", origin(3), basis rows (3x3)) or ("sphere",
center(3), radius); boundary conditions 0 periodic, 1 reflection, 2 none;
methods 0 time stepping, 1 distribution.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import build as _build_all

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_nucleation.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_sp = ctypes.POINTER(ctypes.c_size_t)
_u64 = ctypes.c_uint64
_u64p = ctypes.POINTER(ctypes.c_uint64)
_sz = ctypes.c_size_t
_d = ctypes.c_double
_i = ctypes.c_int


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            _build_all()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_nuc_last_error.restype = ctypes.c_char_p
        L.oracle_philox4x64.argtypes = [_u64p, _u64, _u64, _u64p]
        L.oracle_lattice_volume.argtypes = [_dp, _i]
        L.oracle_lattice_volume.restype = ctypes.c_double
        L.oracle_images_per_bubble.argtypes = [_i, _i]
        L.oracle_lattice_images.argtypes = [_dp, _i, _i, _dp, _sz, _dp]
        L.oracle_lattice_contains.argtypes = [_dp, _i, _dp, _sz, _ip]
        L.oracle_sample_points.argtypes = [_dp, _i, _u64, _u64, _sz, _dp]
        L.oracle_fixed_rate.argtypes = [_dp, _i, _i, _i, _d, _d, _d, _d, _u64, _u64, _sz, _sz, _d,
                                        _dp, _sz, _sz, _dp, _dp, _dp, _dp, _sp]
        L.oracle_spontaneous.argtypes = [_dp, _i, _i, _sz, _d, _u64, _u64, _dp, _dp]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise OracleError(lib().oracle_nuc_last_error().decode())


def _params(lattice):
    kind, *rest = lattice
    if kind == "box":
        origin, basis = rest
        p = np.concatenate([np.asarray(origin, float).reshape(3), np.asarray(basis, float).reshape(9)])
        return np.ascontiguousarray(p), 0
    center, radius = rest
    return np.ascontiguousarray(np.concatenate([np.asarray(center, float).reshape(3), [float(radius)]])), 1


def _ptr(a):
    return a.ctypes.data_as(_dp)


def philox4x64(counter, key):
    """Philox4x64-10 block: counter (4 x u64), key (2 x u64) -> 4 x u64."""
    c = (ctypes.c_uint64 * 4)(*[int(x) & (2**64 - 1) for x in counter])
    out = (ctypes.c_uint64 * 4)()
    lib().oracle_philox4x64(c, int(key[0]) & (2**64 - 1), int(key[1]) & (2**64 - 1), out)
    return [int(x) for x in out]


def volume(lattice) -> float:
    p, kind = _params(lattice)
    return float(lib().oracle_lattice_volume(_ptr(p), kind))


def images_per_bubble(kind: str, bc: int) -> int:
    return int(lib().oracle_images_per_bubble(0 if kind == "box" else 1, bc))


def images(lattice, bc, interior):
    p, kind = _params(lattice)
    ev = np.ascontiguousarray(np.asarray(interior, float).reshape(-1, 4))
    n_img = images_per_bubble(lattice[0], bc)
    out = np.zeros((len(ev) * n_img, 4))
    lib().oracle_lattice_images(_ptr(p), kind, bc, _ptr(ev), len(ev), _ptr(out))
    return out


def contains(lattice, points):
    p, kind = _params(lattice)
    pts = np.ascontiguousarray(np.asarray(points, float).reshape(-1, 3))
    out = np.zeros(len(pts), dtype=np.int32)
    lib().oracle_lattice_contains(_ptr(p), kind, _ptr(pts), len(pts), out.ctypes.data_as(_ip))
    return out.astype(bool)


def sample_points(lattice, seed, stream, n_points):
    p, kind = _params(lattice)
    out = np.zeros((n_points, 3))
    _check(lib().oracle_sample_points(_ptr(p), kind, int(seed), int(stream), n_points, _ptr(out)))
    return out


def fixed_rate(lattice, bc, method, *, beta, gamma0, t0, d_p0, seed, stream, n_points,
               max_time_iterations=1_000_000, cutoff_fraction=0.01, times=None, cap=None):
    """One realisation: (interior (n, 4), exterior (n * n_img, 4), time_history, volume_history)."""
    p, kind = _params(lattice)
    n_img = images_per_bubble(lattice[0], bc)
    cap = n_points if cap is None else cap
    tm = np.ascontiguousarray(np.zeros(1) if times is None else np.asarray(times, float))
    interior = np.zeros((cap, 4))
    exterior = np.zeros((cap * max(n_img, 1), 4))
    th = np.zeros(cap)
    vh = np.zeros(cap)
    n = ctypes.c_size_t(0)
    _check(lib().oracle_fixed_rate(
        _ptr(p), kind, bc, method, beta, gamma0, t0, d_p0, int(seed), int(stream), n_points,
        max_time_iterations, cutoff_fraction, _ptr(tm), 0 if times is None else len(tm), cap,
        _ptr(interior), _ptr(exterior), _ptr(th), _ptr(vh), ctypes.byref(n)))
    k = n.value
    return interior[:k], exterior[:k * n_img], th[:k], vh[:k]


def spontaneous(lattice, bc, n_bubbles, t0, seed, stream):
    p, kind = _params(lattice)
    n_img = images_per_bubble(lattice[0], bc)
    interior = np.zeros((n_bubbles, 4))
    exterior = np.zeros((n_bubbles * max(n_img, 1), 4))
    _check(lib().oracle_spontaneous(_ptr(p), kind, bc, n_bubbles, t0, int(seed), int(stream),
                                    _ptr(interior), _ptr(exterior)))
    return interior, exterior[:n_bubbles * n_img]
