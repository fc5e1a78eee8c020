"""ctypes binding of libbubble_gw_c.so (the C ABI declared in include/bubble_gw.h).

The library must have been built (``python -m gw_fopt_amd.build``); importing
this module never compiles anything and never substitutes another
implementation: if the library is missing or cannot be loaded, every call
raises ``LibraryNotBuilt``.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbubble_gw_c.so")

# BgwStatus (include/bubble_gw.h)
OK, NULL_POINTER, INVALID_ARGUMENT, CALCULATION_ERROR, PANIC = range(5)
# BgwPrecision
FP64, MIXED, FP64_BLOCKED, FP64_MATRIX = 0, 1, 2, 3
# BgwTimeCutoffKind
CUTOFF_UNIT, CUTOFF_EXPONENTIAL = 0, 1


class LibraryNotBuilt(RuntimeError):
    pass


class BgwError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


class BgwTimeCutoffConfig(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("smax", ctypes.c_double),
                ("ratio_t_cut", ctypes.c_double), ("ratio_t_0", ctypes.c_double)]


class BgwTimings(ctypes.Structure):
    _fields_ = [("geometry_ms", ctypes.c_double), ("spectrum_ms", ctypes.c_double),
                ("reduce_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("n_items", ctypes.c_size_t), ("n_terms", ctypes.c_size_t),
                ("n_directions", ctypes.c_size_t), ("matrix_ms", ctypes.c_double)]


class BgwTileConfig(ctypes.Structure):
    _fields_ = [("block_steps", ctypes.c_size_t), ("threads", ctypes.c_size_t),
                ("items_per_chunk", ctypes.c_size_t), ("persistent_blocks", ctypes.c_size_t)]


# BgwLatticeKind, BgwBoundaryCondition, BgwFixedRateMethod
LATTICE_PARALLELEPIPED, LATTICE_SPHERICAL = 0, 1
BC_PERIODIC, BC_REFLECTION, BC_NONE = 0, 1, 2
FIXED_RATE_TIME_STEPPING, FIXED_RATE_DISTRIBUTION = 0, 1


class BgwNucleationLattice(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("origin", ctypes.c_double * 3),
                ("basis", ctypes.c_double * 9), ("radius", ctypes.c_double)]


class BgwFixedRateRealisation(ctypes.Structure):
    _fields_ = [("lattice", BgwNucleationLattice), ("beta", ctypes.c_double),
                ("gamma0", ctypes.c_double), ("t0", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("stream", ctypes.c_uint64)]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_sp = ctypes.POINTER(ctypes.c_size_t)
_sz = ctypes.c_size_t
_d = ctypes.c_double
_b = ctypes.c_bool
_vp = ctypes.c_void_p
_st = ctypes.c_int

# name -> (restype, argtypes)
SIGNATURES = {
    "bgw_version": (ctypes.c_char_p, []),
    "bgw_last_error_message": (ctypes.c_char_p, []),
    "bgw_generalized_bulk_flow_empty_lattice_new": (_st, [ctypes.POINTER(_vp), _dp, _sz, _dp, _sz]),
    "bgw_generalized_bulk_flow_free": (None, [_vp]),
    "bgw_generalized_bulk_flow_set_num_threads": (_st, [_vp, _sz]),
    "bgw_generalized_bulk_flow_set_resolution": (_st, [_vp, _sz, _sz, _b]),
    "bgw_generalized_bulk_flow_set_gradient_scaling_params": (_st, [_vp, _dp, _dp, _sz, _sz, _d]),
    "bgw_generalized_bulk_flow_compute_c_integral": (
        _st, [_vp, _dp, _sz, _b, _d, _d, _sz, _sp, _sz, BgwTimeCutoffConfig, _dp, _sz]),
    "bgw_generalized_bulk_flow_compute_c_integral_ex": (
        _st, [_vp, _dp, _sz, _b, _d, _d, _sz, _sp, _sz, BgwTimeCutoffConfig, _st, _dp, _sz]),
    "bgw_generalized_bulk_flow_compute_c_integral_no_alloc": (
        _st, [_vp, _dp, _sz, _b, _d, _d, _sz, _sp, _sz, BgwTimeCutoffConfig, _st, _dp, _sz]),
    "bgw_generalized_bulk_flow_first_colliding_bubble": (_st, [_vp, _sz, _ip, _sz]),
    "bgw_generalized_bulk_flow_first_colliding_bubble_ex": (_st, [_vp, _sz, _dp, _ip, _dp, _sz]),
    "bgw_generalized_bulk_flow_delta_tab": (_st, [_vp, _sz, _ip, _sz, _dp, _sz]),
    "bgw_generalized_bulk_flow_collision_status": (_st, [_vp, _sz, _d, _ip, _dp, _sz, _ip, _sz]),
    "bgw_generalized_bulk_flow_compute_c_tensor": (
        _st, [_vp, _dp, _sz, _b, _d, _d, _sz, _sp, _sz, _st, _dp, _sz]),
    "bgw_generalized_bulk_flow_compute_c_tensor_integrand": (
        _st, [_vp, _dp, _sz, _b, _d, _d, _sz, _sp, _sz, _dp, _sz]),
    "bgw_generalized_bulk_flow_compute_c_integrand": (
        _st, [_vp, _dp, _sz, _b, _d, _d, _sz, _sp, _sz, _dp, _sz]),
    "bgw_generalized_bulk_flow_compute_spectrum": (
        _st, [_vp, _dp, _sz, _dp, _sz, _b, _d, _d, _sz, BgwTimeCutoffConfig, _st,
              _dp, _sz, _dp, _sz, _dp, _sz]),
    "bgw_generalized_bulk_flow_spectrum_vjp": (
        _st, [_vp, _dp, _sz, _dp, _sz, _b, _d, _d, _sz, BgwTimeCutoffConfig, _dp, _sz,
              _dp, _dp, _sz]),
    "bgw_generalized_bulk_flow_set_tiles": (_st, [_vp, ctypes.POINTER(BgwTileConfig)]),
    "bgw_generalized_bulk_flow_last_timings": (_st, [_vp, ctypes.POINTER(BgwTimings)]),
    "bgw_ensemble_new": (_st, [ctypes.POINTER(_vp), _dp, _sp, _dp, _sp, _sz]),
    "bgw_ensemble_free": (None, [_vp]),
    "bgw_ensemble_set_resolution": (_st, [_vp, _sz, _sz]),
    "bgw_ensemble_set_gradient_scaling_params": (_st, [_vp, _dp, _dp, _sz, _sz, _d]),
    "bgw_ensemble_compute_spectrum": (
        _st, [_vp, _dp, _sz, _dp, _sz, _dp, _dp, _sz, _st, ctypes.POINTER(BgwTileConfig),
              _dp, _sz, _dp, _sz, _dp, _dp, _sz]),
    "bgw_ensemble_autotune": (
        _st, [_vp, _dp, _sz, _dp, _sz, _dp, _dp, _sz, _st, ctypes.POINTER(BgwTileConfig)]),
    "bgw_ensemble_last_timings": (_st, [_vp, ctypes.POINTER(BgwTimings)]),
    "bgw_nucleation_images_per_bubble": (_sz, [ctypes.c_int32, _st]),
    "bgw_nucleation_fixed_rate": (
        _st, [ctypes.POINTER(BgwFixedRateRealisation), _sz, _st, _st, _d, _sz, _sz, _d, _sz,
              _dp, _dp, _dp, _dp, _sp]),
    "bgw_nucleation_solve_distribution": (_st, [_d, _d, _d, _d, _d, _dp, _dp, _sz, _sp]),
    "bgw_nucleation_spontaneous": (
        _st, [ctypes.POINTER(BgwNucleationLattice), _st, _sz, _d, ctypes.c_uint64,
              ctypes.c_uint64, _sz, _dp, _dp]),
    "bgw_nucleation_last_kernel_ms": (_st, [_dp]),
}

_lib = None


def lib():
    """The loaded library (raises LibraryNotBuilt if it is absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryNotBuilt(
                f"{LIB_PATH} is missing; build it with `python -m gw_fopt_amd.build`")
        try:
            L = ctypes.CDLL(LIB_PATH)
        except OSError as e:  # pragma: no cover - environment dependent
            raise LibraryNotBuilt(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().bgw_last_error_message().decode(errors="replace")


def check(status: int) -> None:
    if status != OK:
        raise BgwError(status, last_error())


def dptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def iptr(a):
    return None if a is None else a.ctypes.data_as(_ip)


def sptr(a):
    return None if a is None else a.ctypes.data_as(_sp)


def cutoff_config(time_cutoff) -> BgwTimeCutoffConfig:
    """None -> unit cutoff; an object with smax/ratio_t_cut/ratio_t_0 -> exponential."""
    nan = float("nan")
    if time_cutoff is None:
        return BgwTimeCutoffConfig(CUTOFF_UNIT, 0.0, nan, nan)
    rc = time_cutoff.ratio_t_cut
    r0 = time_cutoff.ratio_t_0
    return BgwTimeCutoffConfig(CUTOFF_EXPONENTIAL, float(time_cutoff.smax),
                               nan if rc is None else float(rc), nan if r0 is None else float(r0))
