"""C ABI of libbubble_gw_c.so without a GPU: exports, header, host-side validation.

Mirrors the reference's C-ABI tests where they do not need numbers
(rust/c-bubble-gw/tests/c_api.rs:220-262: the header compiles as C11 and
C++17) and the error contract of rust/c-bubble-gw/src/lib.rs:104-134,
386-574.  Nothing here launches a kernel.
"""
import ctypes
import math
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from gw_fopt_amd import _lib
from gw_fopt_amd import many_bubbles as mb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bubble_gw.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bgw_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    syms = declared_symbols()
    assert "bgw_generalized_bulk_flow_compute_c_integral" in syms
    for name in syms:
        assert hasattr(L, name), name
    assert set(syms) == set(_lib.SIGNATURES), "ctypes table and header disagree"


def test_reference_symbols_present():
    # the bulk-flow symbols of the reference's cbindgen export list (cbindgen.toml:16-40)
    ref = ["bgw_version", "bgw_last_error_message", "bgw_generalized_bulk_flow_empty_lattice_new",
           "bgw_generalized_bulk_flow_free", "bgw_generalized_bulk_flow_set_num_threads",
           "bgw_generalized_bulk_flow_set_resolution",
           "bgw_generalized_bulk_flow_set_gradient_scaling_params",
           "bgw_generalized_bulk_flow_compute_c_integral"]
    assert set(ref) <= set(declared_symbols())


@pytest.mark.parametrize("lang,std,compiler", [("c", "c11", "gcc"), ("c++", "c++17", "g++")])
def test_header_compiles(tmp_path, lang, std, compiler):
    if shutil.which(compiler) is None:
        pytest.skip(f"{compiler} not available")
    src = tmp_path / ("smoke.c" if lang == "c" else "smoke.cpp")
    nul = "0" if lang == "c" else "nullptr"
    src.write_text('#include "bubble_gw.h"\nint main(void) { return bgw_version() == %s; }\n' % nul)
    subprocess.run([compiler, f"-std={std}", "-fsyntax-only", "-Wall", "-Werror", "-I",
                    os.path.dirname(HEADER), str(src)], check=True)


def test_version_and_clean_error_state():
    L = _lib.lib()
    assert L.bgw_version().decode() == "0.1.0"
    h = ctypes.c_void_p()
    inter = np.array([[0.0, 0.0, 10.0, 0.0], [0.0, 0.0, 0.0, 0.0]])
    st = L.bgw_generalized_bulk_flow_empty_lattice_new(ctypes.byref(h), _lib.dptr(inter), 2, None, 0)
    assert st == _lib.OK and h.value
    assert L.bgw_last_error_message().decode() == ""
    assert L.bgw_generalized_bulk_flow_set_num_threads(h, 0) == _lib.CALCULATION_ERROR
    assert L.bgw_last_error_message().decode() == "num_threads must be > 0"
    assert L.bgw_generalized_bulk_flow_set_num_threads(h, 2) == _lib.OK
    assert L.bgw_last_error_message().decode() == ""  # cleared on success (lib.rs:121-124)
    L.bgw_generalized_bulk_flow_free(h)
    L.bgw_generalized_bulk_flow_free(None)  # null is allowed (lib.rs:424-430)


def test_null_out_is_null_pointer_status():
    L = _lib.lib()
    inter = np.zeros((1, 4))
    st = L.bgw_generalized_bulk_flow_empty_lattice_new(None, _lib.dptr(inter), 1, None, 0)
    assert st == _lib.NULL_POINTER
    assert L.bgw_last_error_message().decode() == "out must not be null"


def test_null_arrays_and_handles():
    L = _lib.lib()
    h = ctypes.c_void_p(123)
    st = L.bgw_generalized_bulk_flow_empty_lattice_new(ctypes.byref(h), None, 3, None, 0)
    assert st == _lib.CALCULATION_ERROR and h.value is None  # *out set to NULL first (lib.rs:402)
    assert "bubbles_interior must not be null" in _lib.last_error()
    assert L.bgw_generalized_bulk_flow_set_resolution(None, 4, 4, True) == _lib.CALCULATION_ERROR
    assert _lib.last_error() == "bulk_flow must not be null"


def test_bubble_inside_bubble_rejected():
    L = _lib.lib()
    h = ctypes.c_void_p()
    inter = np.array([[0.0, 0.0, 0.0, 0.0], [1.0, 0.5, 0.0, 0.0]])
    st = L.bgw_generalized_bulk_flow_empty_lattice_new(ctypes.byref(h), _lib.dptr(inter), 2, None, 0)
    assert st == _lib.CALCULATION_ERROR
    assert _lib.last_error() == ("Bubble Interior(0) is formed inside bubble Interior(1) at "
                                 "initial time (overlapping light cones)")
    ext = np.array([[0.5, 0.2, 0.0, 0.0]])
    st = L.bgw_generalized_bulk_flow_empty_lattice_new(ctypes.byref(h), _lib.dptr(inter[:1]), 1,
                                                       _lib.dptr(ext), 1)
    assert "Exterior(0)" in _lib.last_error()


def test_validation_before_any_device_work():
    L = _lib.lib()
    h = ctypes.c_void_p()
    inter = np.array([[0.0, 0.0, 10.0, 0.0], [0.0, 0.0, 0.0, 0.0]])
    L.bgw_generalized_bulk_flow_empty_lattice_new(ctypes.byref(h), _lib.dptr(inter), 2, None, 0)
    assert L.bgw_generalized_bulk_flow_set_resolution(h, 1, 5, True) == _lib.CALCULATION_ERROR
    assert _lib.last_error() == "Invalid resolution: Angular resolutions must be greater than one"
    w = np.array([0.5, 1.0])
    out = np.zeros(8)
    cut = _lib.cutoff_config(None)
    st = L.bgw_generalized_bulk_flow_compute_c_integral(h, _lib.dptr(w), 2, True, 0.0, 8.0, 1,
                                                        None, 0, cut, _lib.dptr(out), 4)
    assert st == _lib.CALCULATION_ERROR
    assert _lib.last_error() == "Invalid resolution: n_t must be >= 2 for integration"
    sel = np.array([5], dtype=np.uintp)
    st = L.bgw_generalized_bulk_flow_compute_c_integral(h, _lib.dptr(w), 2, True, 0.0, 8.0, 40,
                                                        _lib.sptr(sel), 1, cut, _lib.dptr(out), 4)
    assert _lib.last_error() == "Index 5 out of bounds for max 2"
    # empty selection: zeros, no device needed (generalized_bulk_flow.rs:955-957)
    sel0 = np.zeros(1, dtype=np.uintp)
    out[:] = 7.0
    st = L.bgw_generalized_bulk_flow_compute_c_integral(h, _lib.dptr(w), 2, True, 0.0, 8.0, 40,
                                                        _lib.sptr(sel0), 0, cut, _lib.dptr(out), 4)
    assert st == _lib.OK and np.all(out[:4] == 0.0), _lib.last_error()
    st = L.bgw_generalized_bulk_flow_compute_c_integral(h, _lib.dptr(w), 2, True, 0.0, 8.0, 40,
                                                        _lib.sptr(sel0), 0, cut, _lib.dptr(out), 3)
    assert _lib.last_error() == "out_len=3 does not match expected length 4"
    coef = np.array([1.0, 2.0])
    st = L.bgw_generalized_bulk_flow_set_gradient_scaling_params(h, _lib.dptr(coef), _lib.dptr(coef),
                                                                 2, 1, math.nan)
    assert st == _lib.OK
    st = L.bgw_generalized_bulk_flow_compute_c_integral_ex(h, _lib.dptr(w), 2, True, 0.0, 8.0,
                                                           40, None, 0, cut, 7, _lib.dptr(out), 8)
    assert st == _lib.INVALID_ARGUMENT
    L.bgw_generalized_bulk_flow_free(h)


def test_no_device_fails_loudly():
    """Without a GPU the compute path raises; there is no CPU fallback."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    lb = mb.LatticeBubbles(mb.EmptyLattice(), np.array([[0.0, 0.0, 10.0, 0.0], [0.0, 0.0, 0.0, 0.0]]))
    bf = mb.GeneralizedBulkFlow(lb)
    bf.set_resolution(4, 5)
    with pytest.raises(ValueError, match="no HIP device"):
        bf.compute_c_integral([0.5, 1.0], t_begin=0.0, t_end=8.0, n_t=40)


def _handle(inter, ext=None):
    L = _lib.lib()
    h = ctypes.c_void_p()
    ext = np.zeros((0, 4)) if ext is None else ext
    assert L.bgw_generalized_bulk_flow_empty_lattice_new(ctypes.byref(h), _lib.dptr(inter), len(inter),
                                                         _lib.dptr(ext), len(ext)) == _lib.OK
    return h


def test_integrand_validation_and_zero_cases():
    """compute_c_integrand (generalized_bulk_flow.rs:875-920): n_t is checked
    first, an empty selection or bubbles nucleated at or after t_end give zeros
    without touching a device, and a wrong out_len is rejected."""
    L = _lib.lib()
    inter = np.array([[0.0, 0.0, 10.0, 0.0], [2.0, 0.0, 0.0, 0.0]])
    h = _handle(inter)
    L.bgw_generalized_bulk_flow_set_resolution(h, 4, 5, True)
    w = np.array([0.5, 1.0])
    out = np.full(2 * 1 * 2 * 6 * 2, 7.0)
    st = L.bgw_generalized_bulk_flow_compute_c_integrand(h, _lib.dptr(w), 2, True, 0.0, 1.0, 1, None,
                                                         0, _lib.dptr(out), 24)
    assert st == _lib.CALCULATION_ERROR and _lib.last_error() == "Invalid resolution: n_t must be >= 2"
    sel = np.array([1], dtype=np.uintp)  # nucleated at t = 2 > t_end = 1: zeros, no device
    st = L.bgw_generalized_bulk_flow_compute_c_integrand(h, _lib.dptr(w), 2, True, 0.0, 1.0, 6,
                                                         _lib.sptr(sel), 1, _lib.dptr(out), 24)
    assert st == _lib.OK and np.all(out == 0.0), _lib.last_error()
    # ... even with t_begin > t_end: that bubble never reaches the range check (:681-692)
    st = L.bgw_generalized_bulk_flow_compute_c_integrand(h, _lib.dptr(w), 2, True, 1.5, 1.0, 6,
                                                         _lib.sptr(sel), 1, _lib.dptr(out), 24)
    assert st == _lib.OK
    st = L.bgw_generalized_bulk_flow_compute_c_integrand(h, _lib.dptr(w), 2, True, 0.0, 1.0, 6,
                                                         _lib.sptr(sel), 1, _lib.dptr(out), 23)
    assert "out_len=23 does not match expected length 24" in _lib.last_error()
    L.bgw_generalized_bulk_flow_free(h)


def test_spectrum_and_vjp_validation():
    L = _lib.lib()
    h = _handle(np.array([[0.0, 0.0, 10.0, 0.0], [0.0, 0.0, 0.0, 0.0]]))
    L.bgw_generalized_bulk_flow_set_resolution(h, 4, 5, True)
    w = np.array([0.5, 1.0])
    k = np.array([[0.0, 0.0, 1.0], [0.6, 0.0, 0.8]])
    cut = _lib.cutoff_config(None)
    S = np.zeros(4)
    st = L.bgw_generalized_bulk_flow_compute_spectrum(h, _lib.dptr(w), 2, _lib.dptr(k), 2, False, 0.0,
                                                      1.0, 10, cut, _lib.FP64, _lib.dptr(S), 3, None,
                                                      0, None, 0)
    assert st == _lib.CALCULATION_ERROR and "spectrum_len=3" in _lib.last_error()
    bad = np.array([[0.0, 0.0, 2.0]])
    st = L.bgw_generalized_bulk_flow_compute_spectrum(h, _lib.dptr(w), 2, _lib.dptr(bad), 1, False, 0.0,
                                                      1.0, 10, cut, _lib.FP64, _lib.dptr(S), 2, None,
                                                      0, None, 0)
    assert _lib.last_error() == "k_hat[0] is not a unit vector"
    g = np.zeros(2 * 2 * 2 * 2)
    dc = np.zeros(1)
    dp = np.zeros(1)
    st = L.bgw_generalized_bulk_flow_spectrum_vjp(h, _lib.dptr(w), 2, _lib.dptr(k), 2, False, 0.0, 1.0,
                                                  10, cut, _lib.dptr(g), 7, _lib.dptr(dc), _lib.dptr(dp), 1)
    assert "grad_c_len=7 does not match expected length 8" in _lib.last_error()
    st = L.bgw_generalized_bulk_flow_spectrum_vjp(h, _lib.dptr(w), 2, _lib.dptr(k), 2, False, 0.0, 1.0,
                                                  10, cut, _lib.dptr(g), 8, _lib.dptr(dc), _lib.dptr(dp), 2)
    assert "grad_len=2 does not match expected length 1" in _lib.last_error()
    st = L.bgw_generalized_bulk_flow_spectrum_vjp(h, _lib.dptr(w), 2, _lib.dptr(k), 2, False, 0.0, 1.0,
                                                  1, cut, _lib.dptr(g), 8, _lib.dptr(dc), _lib.dptr(dp), 1)
    assert _lib.last_error() == "Invalid resolution: n_t must be >= 2 for integration"
    t = _lib.BgwTileConfig(5, 0, 0, 0)
    assert L.bgw_generalized_bulk_flow_set_tiles(h, ctypes.byref(t)) == _lib.CALCULATION_ERROR
    assert _lib.last_error() == ("tiles.block_steps must be 0, 8 or 16 (1, 2 or 4 for the FP64 "
                                 "matrix form)")
    t = _lib.BgwTileConfig(8, 96, 0, 0)
    assert L.bgw_generalized_bulk_flow_set_tiles(h, ctypes.byref(t)) == _lib.CALCULATION_ERROR
    L.bgw_generalized_bulk_flow_free(h)


def test_ensemble_validation():
    L = _lib.lib()
    inter = np.array([[0.0, 0.1, 0.1, 0.1], [0.0, 0.6, 0.6, 0.6], [0.0, 0.3, 0.3, 0.3]])
    io = np.array([0, 2, 3], dtype=np.uintp)
    eo = np.array([0, 0, 0], dtype=np.uintp)
    h = ctypes.c_void_p()
    assert L.bgw_ensemble_new(None, _lib.dptr(inter), _lib.sptr(io), None, _lib.sptr(eo), 2) == _lib.NULL_POINTER
    bad = np.array([0, 2, 1], dtype=np.uintp)
    st = L.bgw_ensemble_new(ctypes.byref(h), _lib.dptr(inter), _lib.sptr(bad), None, _lib.sptr(eo), 2)
    assert st == _lib.CALCULATION_ERROR and _lib.last_error() == "offsets must be non-decreasing"
    st = L.bgw_ensemble_new(ctypes.byref(h), _lib.dptr(inter), _lib.sptr(io), None, _lib.sptr(eo), 2)
    assert st == _lib.OK
    assert L.bgw_ensemble_set_resolution(h, 1, 8) == _lib.CALCULATION_ERROR
    assert L.bgw_ensemble_set_resolution(h, 4, 8) == _lib.OK
    w = np.array([[0.5, 1.0], [0.5, 1.0]])
    k = np.array([[0.0, 0.0, 1.0]])
    te = np.array([1.0, 1.0])
    S = np.zeros(4)
    st = L.bgw_ensemble_compute_spectrum(h, _lib.dptr(w), 2, _lib.dptr(k), 1, None, _lib.dptr(te), 10,
                                         _lib.MIXED, None, _lib.dptr(S), 4, None, 0, None, None, 0)
    assert st == _lib.CALCULATION_ERROR and _lib.last_error() == "t_begin must not be null"
    tb = np.zeros(2)
    st = L.bgw_ensemble_compute_spectrum(h, _lib.dptr(w), 2, _lib.dptr(k), 1, _lib.dptr(tb), _lib.dptr(te),
                                         10, _lib.MIXED, None, _lib.dptr(S), 3, None, 0, None, None, 0)
    assert "spectrum_len=3 does not match expected length 4" in _lib.last_error()
    st = L.bgw_ensemble_compute_spectrum(h, _lib.dptr(w), 2, _lib.dptr(k), 1, _lib.dptr(tb), _lib.dptr(te),
                                         10, 9, None, _lib.dptr(S), 4, None, 0, None, None, 0)
    assert st == _lib.INVALID_ARGUMENT
    t = _lib.BgwTileConfig(0, 0, 0, 70000)
    st = L.bgw_ensemble_compute_spectrum(h, _lib.dptr(w), 2, _lib.dptr(k), 1, _lib.dptr(tb), _lib.dptr(te),
                                         10, _lib.MIXED, ctypes.byref(t), _lib.dptr(S), 4, None, 0, None,
                                         None, 0)
    assert _lib.last_error() == "tiles.persistent_blocks must be <= 65536"
    L.bgw_ensemble_free(h)
    L.bgw_ensemble_free(None)


def test_tensor_validation_and_zero_cases():
    """compute_c_tensor / compute_c_tensor_integrand (new_generalized_bulk_flow.rs:993-1275):
    n_t first, out_len, an empty selection and late bubbles give zeros without a device."""
    L = _lib.lib()
    inter = np.array([[0.0, 0.0, 10.0, 0.0], [2.0, 0.0, 0.0, 0.0]])
    h = _handle(inter)
    L.bgw_generalized_bulk_flow_set_resolution(h, 4, 5, True)
    w = np.array([0.5, 1.0])
    out = np.full(6 * 1 * 2 * 2, 7.0)
    st = L.bgw_generalized_bulk_flow_compute_c_tensor(h, _lib.dptr(w), 2, True, 0.0, 1.0, 1, None, 0,
                                                      _lib.FP64, _lib.dptr(out), 12)
    assert st == _lib.CALCULATION_ERROR
    assert _lib.last_error() == "Invalid resolution: n_t must be >= 2 for integration"
    st = L.bgw_generalized_bulk_flow_compute_c_tensor(h, _lib.dptr(w), 2, True, 0.0, 1.0, 6, None, 0,
                                                      _lib.FP64, _lib.dptr(out), 11)
    assert "out_len=11 does not match expected length 12" in _lib.last_error()
    empty = np.zeros(0, dtype=np.uintp)
    st = L.bgw_generalized_bulk_flow_compute_c_tensor(h, _lib.dptr(w), 2, True, 0.0, 1.0, 6,
                                                      _lib.sptr(empty), 0, _lib.FP64, _lib.dptr(out), 12)
    assert st == _lib.OK and np.all(out == 0.0)
    bad = np.array([5], dtype=np.uintp)
    st = L.bgw_generalized_bulk_flow_compute_c_tensor(h, _lib.dptr(w), 2, True, 0.0, 1.0, 6,
                                                      _lib.sptr(bad), 1, _lib.FP64, _lib.dptr(out), 12)
    assert st == _lib.CALCULATION_ERROR and "out of bounds" in _lib.last_error()
    assert L.bgw_generalized_bulk_flow_compute_c_tensor(h, _lib.dptr(w), 2, True, 0.0, 1.0, 6, None, 0,
                                                        9, _lib.dptr(out), 12) == _lib.INVALID_ARGUMENT
    outi = np.full(6 * 2 * 2 * 6, 7.0)
    sel = np.array([1], dtype=np.uintp)  # nucleated at t = 2 > t_end: zeros, no device
    st = L.bgw_generalized_bulk_flow_compute_c_tensor_integrand(h, _lib.dptr(w), 2, True, 0.0, 1.0, 6,
                                                                _lib.sptr(sel), 1, _lib.dptr(outi), 72)
    assert st == _lib.OK and np.all(outi == 0.0), _lib.last_error()
    L.bgw_generalized_bulk_flow_free(h)
