"""CPU oracle for the generalized bulk-flow spectrum — TEST INFRASTRUCTURE ONLY.

This package is the checker: ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package
``gw_fopt_amd`` never does.  It wraps ``liboracle_bulkflow.so`` (built from
``bulkflow_oracle.cpp`` by ``oracle/Makefile``), a C++ restatement of
``rust/bubble-gw/src/many_bubbles/generalized_bulk_flow.rs`` that follows the
reference's operation order (see the header of ``bulkflow_oracle.cpp``).

All arrays are numpy float64 / int32, row-major.  Complex outputs are numpy
complex128 with the reference's layouts:
``compute_c_integral`` -> (2, n_sets, n_w), ``compute_c_integrand`` ->
(2, n_sets, n_w, n_t), ``spectrum`` -> (n_k, n_sets, n_w).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_bulkflow.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_sp = ctypes.POINTER(ctypes.c_size_t)
_sz = ctypes.c_size_t
_d = ctypes.c_double
_i = ctypes.c_int


def build(force: bool = False) -> str:
    """Compile the oracle with ``make`` (g++ -O2 -fopenmp -ffp-contract=off)."""
    libs = (_LIB_PATH, os.path.join(_HERE, "liboracle_nucleation.so"))
    if force or not all(os.path.exists(p) for p in libs):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_grids.argtypes = [_sz, _sz, _dp, _dp, _dp]
        L.oracle_rotation_to_z.argtypes = [_dp, _dp]
        L.oracle_lattice.argtypes = [_dp, _sz, _dp, _sz, _dp, _dp]
        L.oracle_first_colliding.argtypes = [_dp, _sz, _dp, _sz, _dp, _sz, _sz, _sz, _i, _ip, _dp]
        L.oracle_collision_status.argtypes = [_d, _d, _sz, _ip, _dp, _ip]
        L.oracle_c_integral.argtypes = [
            _dp, _sz, _dp, _sz, _dp, _sz, _sz, _dp, _dp, _sz, _sz, _d,
            _dp, _sz, _i, _d, _d, _sz, _sp, _sz, _i, _i, _d, _d, _d, _i, _sz, _i, _dp]
        L.oracle_c_integrand.argtypes = [
            _dp, _sz, _dp, _sz, _sz, _sz, _dp, _dp, _sz, _sz, _d,
            _dp, _sz, _i, _d, _d, _sz, _sp, _sz, _i, _i, _dp]
        L.oracle_c_tensor.argtypes = [
            _dp, _sz, _dp, _sz, _sz, _sz, _dp, _dp, _sz, _sz, _d, _dp, _sz, _i, _d, _d, _sz,
            _sp, _sz, _i, _i, _dp]
        L.oracle_spectrum.argtypes = [
            _dp, _sz, _dp, _sz, _dp, _sz, _sz, _sz, _dp, _dp, _sz, _sz, _d,
            _dp, _sz, _i, _d, _d, _sz, _i, _d, _d, _d, _i, _dp, _dp]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _f64(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape is not None:
        a = a.reshape(shape)
    return a


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _check(rc):
    if rc != 0:
        raise OracleError(lib().oracle_last_error().decode())


def _bubbles(interior, exterior):
    interior = _f64(interior).reshape(-1, 4)
    exterior = _f64(np.zeros((0, 4)) if exterior is None else exterior).reshape(-1, 4)
    return interior, exterior


def grids(n_c, n_phi):
    cos_t = np.empty(n_c)
    phi = np.empty(n_phi)
    dirs = np.empty((n_c, n_phi, 4))
    _check(lib().oracle_grids(n_c, n_phi, _ptr(cos_t), _ptr(phi), _ptr(dirs)))
    return cos_t, phi, dirs


def rotation_to_z(k):
    k = _f64(k, (3,))
    R = np.empty((3, 3))
    _check(lib().oracle_rotation_to_z(_ptr(k), _ptr(R)))
    return R


def lattice(interior, exterior=None):
    interior, exterior = _bubbles(interior, exterior)
    n_int, n_tot = len(interior), len(interior) + len(exterior)
    delta = np.empty((n_int, n_tot, 4))
    dsq = np.empty((n_int, n_tot))
    _check(lib().oracle_lattice(_ptr(interior), n_int, _ptr(exterior), len(exterior),
                                _ptr(delta), _ptr(dsq)))
    return delta, dsq


def first_colliding(interior, exterior, n_c, n_phi, a, k_hat=None, n_threads=0):
    """(first_bubble int32 (n_c, n_phi), delta_tab float64 (n_c, n_phi))."""
    interior, exterior = _bubbles(interior, exterior)
    first = np.empty((n_c, n_phi), dtype=np.int32)
    dtab = np.empty((n_c, n_phi))
    k = None if k_hat is None else _f64(k_hat, (3,))
    _check(lib().oracle_first_colliding(
        _ptr(interior), len(interior), _ptr(exterior), len(exterior), _ptr(k), n_c, n_phi, a,
        n_threads, first.ctypes.data_as(_ip), _ptr(dtab)))
    return first, dtab


def collision_status(t_a, t, first, dtab):
    first = np.ascontiguousarray(first, dtype=np.int32)
    dtab = _f64(dtab)
    out = np.empty(first.shape, dtype=np.int32)
    _check(lib().oracle_collision_status(t_a, t, first.size, first.ctypes.data_as(_ip),
                                         _ptr(dtab), out.ctypes.data_as(_ip)))
    return out


def _params(coefficients_sets, powers_sets):
    coef = _f64(coefficients_sets)
    pows = _f64(powers_sets)
    if coef.ndim != 2 or coef.shape != pows.shape:
        raise ValueError("coefficients_sets and powers_sets must have the same 2-D shape")
    return coef, pows


def _cutoff(time_cutoff):
    if time_cutoff is None:
        return 0, 0.0, float("nan"), float("nan")
    kind, smax, rc, r0 = time_cutoff
    return kind, smax, (float("nan") if rc is None else rc), (float("nan") if r0 is None else r0)


def compute_c_integral(interior, exterior, n_c, n_phi, coefficients_sets, powers_sets, w_arr, *,
                       t_end, n_t, t_begin=None, selected_bubbles=None, time_cutoff=None,
                       damping=None, k_hat=None, derivative=None, n_threads=0):
    """compute_c_integral -> complex128 (2, n_sets, n_w).

    ``time_cutoff``: None (unit) or (1, smax, ratio_t_cut|None, ratio_t_0|None).
    ``derivative``: None, ("coef", k) or ("pow", k) for forward-mode dC/dθ.
    """
    interior, exterior = _bubbles(interior, exterior)
    coef, pows = _params(coefficients_sets, powers_sets)
    w = _f64(w_arr)
    n_s, n_terms = coef.shape
    out = np.zeros((2, n_s, len(w)), dtype=np.complex128)
    sel = None
    if selected_bubbles is not None:
        sel = np.ascontiguousarray(np.asarray(selected_bubbles, dtype=np.uint64)).astype(np.uintp)
    kind, smax, rc, r0 = _cutoff(time_cutoff)
    mk, mt = 0, 0
    if derivative is not None:
        mk = {"coef": 1, "pow": 2}[derivative[0]]
        mt = int(derivative[1])
    k = None if k_hat is None else _f64(k_hat, (3,))
    _check(lib().oracle_c_integral(
        _ptr(interior), len(interior), _ptr(exterior), len(exterior), _ptr(k), n_c, n_phi,
        _ptr(coef), _ptr(pows), n_s, n_terms, float("nan") if damping is None else damping,
        _ptr(w), len(w), int(t_begin is not None), 0.0 if t_begin is None else t_begin, t_end, n_t,
        None if sel is None else sel.ctypes.data_as(_sp), 0 if sel is None else len(sel),
        int(sel is not None), kind, smax, rc, r0, mk, mt, n_threads,
        out.ctypes.data_as(_dp)))
    return out


def compute_c_integrand(interior, exterior, n_c, n_phi, coefficients_sets, powers_sets, w_arr, *,
                        t_end, n_t, t_begin=None, selected_bubbles=None, damping=None,
                        n_threads=0):
    interior, exterior = _bubbles(interior, exterior)
    coef, pows = _params(coefficients_sets, powers_sets)
    w = _f64(w_arr)
    n_s, n_terms = coef.shape
    out = np.zeros((2, n_s, len(w), n_t), dtype=np.complex128)
    sel = None
    if selected_bubbles is not None:
        sel = np.ascontiguousarray(np.asarray(selected_bubbles, dtype=np.uint64)).astype(np.uintp)
    _check(lib().oracle_c_integrand(
        _ptr(interior), len(interior), _ptr(exterior), len(exterior), n_c, n_phi,
        _ptr(coef), _ptr(pows), n_s, n_terms, float("nan") if damping is None else damping,
        _ptr(w), len(w), int(t_begin is not None), 0.0 if t_begin is None else t_begin, t_end, n_t,
        None if sel is None else sel.ctypes.data_as(_sp), 0 if sel is None else len(sel),
        int(sel is not None), n_threads, out.ctypes.data_as(_dp)))
    return out


def spectrum(interior, exterior, k_hat, n_c, n_phi, coefficients_sets, powers_sets, w_arr, *,
             t_end, n_t, t_begin=None, time_cutoff=None, damping=None, n_threads=0,
             return_c=False):
    """S(k, s, w) = 8 pi w^3 (|C+|^2 + |Cx|^2) in the frame k_hat -> z."""
    interior, exterior = _bubbles(interior, exterior)
    coef, pows = _params(coefficients_sets, powers_sets)
    w = _f64(w_arr)
    kh = _f64(k_hat).reshape(-1, 3)
    n_s, n_terms = coef.shape
    S = np.empty((len(kh), n_s, len(w)))
    C = np.zeros((len(kh), 2, n_s, len(w)), dtype=np.complex128) if return_c else None
    kind, smax, rc, r0 = _cutoff(time_cutoff)
    _check(lib().oracle_spectrum(
        _ptr(interior), len(interior), _ptr(exterior), len(exterior), _ptr(kh), len(kh), n_c,
        n_phi, _ptr(coef), _ptr(pows), n_s, n_terms,
        float("nan") if damping is None else damping, _ptr(w), len(w), int(t_begin is not None),
        0.0 if t_begin is None else t_begin, t_end, n_t, kind, smax, rc, r0, n_threads,
        _ptr(S), None if C is None else C.ctypes.data_as(_dp)))
    return (S, C) if return_c else S


def compute_c_tensor(interior, exterior, n_c, n_phi, coefficients_sets, powers_sets, w_arr, *,
                     t_end, n_t, t_begin=None, selected_bubbles=None, damping=None, n_threads=0):
    """NewGeneralizedBulkFlow.compute_c_tensor -> complex128 (6, n_sets, n_w), components
    xx, xy, xz, yy, yz, zz (new_generalized_bulk_flow.rs:1205-1275)."""
    interior, exterior = _bubbles(interior, exterior)
    coef, pows = _params(coefficients_sets, powers_sets)
    w = _f64(w_arr)
    n_s, n_terms = coef.shape
    out = np.zeros((6, n_s, len(w)), dtype=np.complex128)
    sel = None
    if selected_bubbles is not None:
        sel = np.ascontiguousarray(np.asarray(selected_bubbles, dtype=np.uint64)).astype(np.uintp)
    _check(lib().oracle_c_tensor(
        _ptr(interior), len(interior), _ptr(exterior), len(exterior), n_c, n_phi, _ptr(coef),
        _ptr(pows), n_s, n_terms, float("nan") if damping is None else damping, _ptr(w), len(w),
        int(t_begin is not None), 0.0 if t_begin is None else t_begin, t_end, n_t,
        None if sel is None else sel.ctypes.data_as(_sp), 0 if sel is None else len(sel),
        int(sel is not None), n_threads, out.ctypes.data_as(_dp)))
    return out
