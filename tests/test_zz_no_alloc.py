"""The no-allocation entry point against compute_c_integral_ex, through the C ABI.

bgw_generalized_bulk_flow_compute_c_integral_no_alloc runs the same kernels as
bgw_generalized_bulk_flow_compute_c_integral_ex with every device buffer carved
from the library's 64 MiB scratch, so its C must equal the allocating entry
point's bit for bit.  Two calls in a row, the second smaller than the first (the
scratch keeps the first call's contents), in every precision; the allocating
entry point still works afterwards on the same handle; an oversized request
fails with the scratch limit in its message.
"""
import numpy as np
import pytest

from gw_fopt_amd import _lib
from gw_fopt_amd import configs
from gw_fopt_amd import many_bubbles as mb

pytestmark = pytest.mark.gpu

PRECISIONS = {"fp64": _lib.FP64, "fp64_blocked": _lib.FP64_BLOCKED, "mixed": _lib.MIXED,
              "fp64_matrix": _lib.FP64_MATRIX}


def flow(interior, exterior, n_c, n_phi, coef, pows):
    bf = mb.GeneralizedBulkFlow(mb.LatticeBubbles(mb.EmptyLattice(), interior, exterior))
    bf.set_resolution(n_c, n_phi)
    bf.set_gradient_scaling_params(coef, pows)
    bf._push_params()
    return bf


def call(bf, entry, w, t_end, n_t, precision, selected=None):
    """C (2, n_sets, n_w) complex from one C-ABI entry point."""
    L = _lib.lib()
    fn = getattr(L, entry)
    w = np.ascontiguousarray(w, dtype=np.float64)
    sel = None if selected is None else np.ascontiguousarray(selected, dtype=np.uintp)
    n_sets = bf._coef.shape[0]
    out = np.full((2, n_sets, len(w)), np.nan, dtype=np.complex128)
    st = fn(bf._h, _lib.dptr(w), len(w), True, 0.0, float(t_end), int(n_t),
            None if sel is None else _lib.sptr(sel), 0 if sel is None else len(sel),
            _lib.cutoff_config(None), PRECISIONS[precision], out.ctypes.data_as(_lib._dp),
            out.size)
    if st != _lib.OK:
        raise ValueError(_lib.last_error())
    return out


def bench_case():
    """The reference's criterion-bench bubbles (rust/benches/bulk_flow_bench.rs:8-44)."""
    inter = np.array([[0.0, 0.0, 10.0, 0.0], [0.0, 0.0, 0.0, 0.0], [0.0, 4.0, 0.0, 4.0],
                      [0.0, 4.0, 2.0, 4.0], [0.0, 0.0, 2.0, 1.0]])
    return flow(inter, None, 50, 100, [[0.0], [1.0]], [[3.0], [3.0]])


@pytest.mark.parametrize("precision", list(PRECISIONS))
def test_no_alloc_equals_allocating_entry(precision):
    bf = bench_case()
    big = (np.geomspace(1e-2, 1e2, 120), 15.0, 800, None)
    small = (np.geomspace(1e-1, 1e1, 40), 12.0, 300, [4, 1])
    for w, t_end, n_t, sel in (big, small):
        print(f"no_alloc {precision}: n_w {len(w)}, n_t {n_t}, selection {sel}", flush=True)
        got = call(bf, "bgw_generalized_bulk_flow_compute_c_integral_no_alloc", w, t_end, n_t,
                   precision, sel)
        ref = call(bf, "bgw_generalized_bulk_flow_compute_c_integral_ex", w, t_end, n_t,
                   precision, sel)
        assert np.all(np.isfinite(got))
        assert got.tobytes() == ref.tobytes()
    # the allocating entry point is unchanged by the calls in between
    again = call(bf, "bgw_generalized_bulk_flow_compute_c_integral_ex", *small[:3], precision,
                 small[3])
    assert again.tobytes() == ref.tobytes()


def test_no_alloc_async_bubbles_with_images():
    """Staggered nucleation in the periodic box (face images as exterior bubbles), the
    plain_8 workload's bubbles and grid, then a smaller selection: both FP64 forms."""
    wl = configs.plain_8()
    bf = flow(wl.interior, wl.exterior, wl.n_cos_thetax, wl.n_phix, wl.coefficients_sets,
              wl.powers_sets)
    for precision in ("fp64", "fp64_blocked"):
        for w, sel in ((wl.w_arr, None), (wl.w_arr[::3], [0, 2, 5])):
            got = call(bf, "bgw_generalized_bulk_flow_compute_c_integral_no_alloc", w, wl.t_end,
                       wl.n_t, precision, sel)
            ref = call(bf, "bgw_generalized_bulk_flow_compute_c_integral_ex", w, wl.t_end,
                       wl.n_t, precision, sel)
            assert got.tobytes() == ref.tobytes(), (precision, sel)


def test_no_alloc_rejects_requests_beyond_the_scratch():
    """256 bubbles on a 128 x 256 grid: the collision delays alone exceed 64 MiB."""
    wl = configs.plain_256(n_k=1)
    bf = flow(wl.interior, wl.exterior, 128, 256, [[1.0]], [[3.0]])
    w = np.geomspace(0.5, 500.0, 16)
    with pytest.raises(ValueError, match="64 MiB scratch of the no-allocation entry point"):
        call(bf, "bgw_generalized_bulk_flow_compute_c_integral_no_alloc", w, wl.t_end, 50, "fp64")
    # the handle is still usable through both entry points
    sel = [0, 1]
    got = call(bf, "bgw_generalized_bulk_flow_compute_c_integral_no_alloc", w, wl.t_end, 50,
               "fp64", sel)
    ref = call(bf, "bgw_generalized_bulk_flow_compute_c_integral_ex", w, wl.t_end, 50, "fp64",
               sel)
    assert got.tobytes() == ref.tobytes()
