"""Post-processing of bulk-flow spectra: the step after the device path.

What the reference's Python layer does with the C integrals, over the arrays
the device path returns (n_w x n_k numbers, so plain numpy / scipy on the host):

* ``DataGwBulkFlow`` — src/gw_fopt/bubble_dynamics/utils/data_gw_bulkflow.py:9-190:
  a linear combination of scaling powers xi (power = xi + 3, None = the envelope
  C), S = 8 pi w^3 (|C+|^2 + |Cx|^2) per cos(theta_k), the angular integral
  2 * trapz over cos(theta_k), and the rescaling (d / d0)^5 rho^2 with a cubic
  spline in w / d.  ``DataGwBulkFlow.from_bulk_flow`` fills it from the device
  (one coefficient set per power, k-hat = (sin theta_k, 0, cos theta_k)).
* ``general_broken_power_law`` / ``asymptotic_broken_power_law`` —
  src/gw_fopt/bubble_dynamics/visualizer/general_broken_power_law.py:19-131: the
  smoothly broken power law used to fit the spectra and its asymptote in one
  region; ``fit_broken_power_law`` fits it in log space (scipy least squares; the
  reference fits with lmfit, which this image does not have).
"""
from __future__ import annotations

import numpy as np
from scipy.interpolate import CubicSpline
from scipy.optimize import least_squares

__all__ = ["spectrum_from_c", "DataGwBulkFlow", "general_broken_power_law",
           "asymptotic_broken_power_law", "fit_broken_power_law", "sample", "sample_arr",
           "filter_dataframe"]


def sample(start, stop, n_sample, n_grid, n_iter, sample_type, base=10.0) -> np.ndarray:
    """The reference's ``gw_fopt.bubble_gw.utils.sample`` (rust/py-bubble-gw/src/py_utils.rs:42-110,
    rust/bubble-gw/src/utils/sample.rs:174-315): n_sample + 1 points from start to stop,
    evenly spaced in the transformed space of sample_type ("uniform"/"linear": x, "log":
    log_base x, "exp": base^x of an even grid); with n_iter > 0 only the points that
    refinement level n_iter adds (n_grid - 1 per cell of the level n_iter - 1 grid)."""
    kind = str(sample_type).lower()
    if kind not in ("uniform", "linear", "log", "exp"):
        raise ValueError(f"Invalid sample_type: {sample_type}. Expected 'uniform', 'linear', "
                         "'log', or 'exp'.")
    base = float(base)
    if kind in ("log", "exp") and base <= 0.0:
        kname = "logarithmic" if kind == "log" else "exponential"
        raise ValueError(f"`base` must be greater than 0 for {kname} sampling")
    start, stop = float(start), float(stop)
    if start >= stop:
        raise ValueError(f"`start` ({start}) must be less than `stop` ({stop})")
    if np.isnan(start) or np.isnan(stop):
        raise ValueError("Invalid parameter start or stop: NaN")
    if kind in ("log", "exp") and (start <= 0.0 or stop <= 0.0):
        raise ValueError(f"`start` ({start}) must be less than `stop` ({stop})")
    n_sample, n_grid, n_iter = int(n_sample), int(n_grid), int(n_iter)
    if n_sample == 0:
        return np.array([start])
    if n_iter > 0 and n_grid < 2:
        raise ValueError(f"Invalid parameter n_grid: {float(n_grid)}")
    if kind == "log":
        fwd, inv = (lambda x: np.log(x) / np.log(base)), (lambda x: base ** x)
    elif kind == "exp":
        fwd, inv = (lambda x: x), (lambda x: base ** x)
    else:
        fwd, inv = (lambda x: x), (lambda x: x)
    a, b = fwd(start), fwd(stop)
    if n_iter == 0:
        t = np.arange(n_sample + 1, dtype=np.float64) / n_sample
    else:
        factor = float(n_grid) ** n_iter
        sub = float(n_grid) ** (n_iter - 1)
        cells = np.arange(n_sample * int(sub), dtype=np.float64)
        fine = np.arange(1, n_grid, dtype=np.float64)
        t = ((fine[None, :] / factor + cells[:, None] / sub) / n_sample).reshape(-1)
    return np.asarray(inv(a + t * (b - a)), dtype=np.float64)


def sample_arr(start, stop, n_sample, n_grid, n_iter, sample_type, base=10.0) -> np.ndarray:
    """Levels 0..n_iter of ``sample`` concatenated (sample.rs:234-249)."""
    return np.concatenate([sample(start, stop, n_sample, n_grid, it, sample_type, base)
                           for it in range(int(n_iter) + 1)])


def filter_dataframe(df, filter_dict):
    """Rows of df that meet every criterion: numeric columns against a number within
    rtol 1e-5 / atol 1e-8, anything else by equality
    (src/gw_fopt/bubble_dynamics/utils/filter_dataframe.py:69-131, the cases the bulk-flow
    post-processing uses)."""
    import pandas as pd
    mask = pd.Series(True, index=df.index)
    for col, crit in filter_dict.items():
        if col not in df.columns:
            raise ValueError(f"Column '{col}' not found.")
        series = df[col]
        if pd.api.types.is_numeric_dtype(series) and isinstance(crit, (int, float, np.number)):
            mask &= np.isclose(series, crit, rtol=1e-5, atol=1e-8)
        else:
            mask &= series == crit
    return df[mask]


def _is_dataframe(x) -> bool:
    return hasattr(x, "iterrows") and hasattr(x, "columns")


def spectrum_from_c(w, c_plus, c_cross) -> np.ndarray:
    """S = 8 pi w^3 (|C+|^2 + |Cx|^2) (data_gw_bulkflow.py:108-113)."""
    w = np.asarray(w, dtype=np.float64)
    return 8 * np.pi * w**3 * (np.abs(c_plus) ** 2 + np.abs(c_cross) ** 2)


class DataGwBulkFlow:
    """Bulk-flow C integrals on a cos(theta_k) grid, and the spectra built from them.

    Two constructors: the reference's ``DataGwBulkFlow(df, d0=1.0)`` over its
    DataFrame (columns d, cos_theta_k, w_sample, powers_sets_sample, c_bulkflow,
    c_envelope, _pset_id; rows taken in the frame's order, as the reference
    iterates them, against the sorted cos grid), and the array form below.

    ``c_bulkflow``: (n_cos, 2, n_powers, n_w) or (n_cos, 2, n_w) complex, [C+, Cx]
    for each power of ``powers_sets_sample``; ``c_envelope``: (n_cos, 2, n_w) or
    None.  Rows are kept in ascending cos(theta_k) order (the reference sorts the
    cos grid, data_gw_bulkflow.py:22, and integrates over it).
    """

    def __init__(self, w_arr, cos_thetak_arr=None, c_bulkflow=None, powers_sets_sample=None,
                 c_envelope=None, d0: float = 1.0):
        self.df = None
        if _is_dataframe(w_arr):
            # the reference's DataGwBulkFlow(df, d0=1.0) (data_gw_bulkflow.py:10-24)
            if c_bulkflow is not None or powers_sets_sample is not None or c_envelope is not None:
                raise TypeError("DataGwBulkFlow(df, d0) takes no arrays")
            self._init_frame(w_arr, d0 if cos_thetak_arr is None else cos_thetak_arr)
            return
        if cos_thetak_arr is None or c_bulkflow is None:
            raise TypeError("DataGwBulkFlow takes (df, d0) or (w_arr, cos_thetak_arr, c_bulkflow, ...)")
        self.d0 = float(d0)
        self.w_arr = np.asarray(w_arr, dtype=np.float64)
        cos = np.asarray(cos_thetak_arr, dtype=np.float64)
        order = np.argsort(cos, kind="stable")
        self.cos_thetak_arr = cos[order]
        c = np.asarray(c_bulkflow, dtype=np.complex128)
        if c.ndim == 3:
            c = c[:, :, None, :]
        if c.ndim != 4 or c.shape[1] != 2 or c.shape[0] != len(cos) or c.shape[3] != len(self.w_arr):
            raise ValueError("c_bulkflow must be (n_cos, 2, n_powers, n_w) or (n_cos, 2, n_w)")
        self.c_bulkflow = c[order]
        self.powers_sets_sample = (None if powers_sets_sample is None
                                   else [float(p) for p in powers_sets_sample])
        if self.powers_sets_sample is not None and len(self.powers_sets_sample) != c.shape[2]:
            raise ValueError("powers_sets_sample must list one power per set of c_bulkflow")
        self.c_envelope = None
        if c_envelope is not None:
            e = np.asarray(c_envelope, dtype=np.complex128)
            if e.shape != (len(cos), 2, len(self.w_arr)):
                raise ValueError("c_envelope must be (n_cos, 2, n_w)")
            self.c_envelope = e[order]

    def _init_frame(self, df, d0):
        """Rows with column d equal to d0; cos(theta_k) sorted; omegas from the first
        row's w_sample arguments of ``sample`` (label 0, as the reference indexes)."""
        self.d0 = d0
        self.df = filter_dataframe(df, {"d": d0})
        self.cos_thetak_arr = np.sort(self.df.cos_theta_k.values)
        self.w_arr = sample(*self.df["w_sample"][0])
        self.c_bulkflow = None
        self.c_envelope = None
        self.powers_sets_sample = None

    def _frame_spectra(self, xi_powers, xi_coefficients) -> np.ndarray:
        """Per row of the frame, in its order (data_gw_bulkflow.py:65-115): C+ and Cx of
        the combination from c_envelope (None) or the row's c_bulkflow set of power
        xi + 3 among sample(*powers_sets_sample)."""
        w = self.w_arr
        out = []
        for _, row in self.df.iterrows():
            cp = np.zeros_like(w, dtype=complex)
            cx = np.zeros_like(w, dtype=complex)
            available = None
            if any(p is not None for p in xi_powers):
                available = list(sample(*row["powers_sets_sample"]))
            for p, coeff in zip(xi_powers, xi_coefficients):
                if p is None:
                    cm = np.asarray(row["c_envelope"])
                    if cm.ndim != 2 or cm.shape[0] != 2:
                        raise ValueError(f"c_envelope for _pset_id {row['_pset_id']} must be a 2D "
                                         "array with shape[0]=2")
                    cp += coeff * cm[0, :]
                    cx += coeff * cm[1, :]
                    continue
                power = p + 3
                if available is None:
                    raise ValueError("powers_sets_sample required for non-None powers")
                if power not in available:
                    raise ValueError(f"Power {power} not in powers_sets_sample for _pset_id "
                                     f"{row['_pset_id']}")
                k = available.index(power)
                cm = np.asarray(row["c_bulkflow"])
                if cm.ndim == 2 and cm.shape[0] == 2:
                    cm = cm[:, np.newaxis, :]
                elif cm.ndim != 3 or cm.shape[0] != 2:
                    raise ValueError(f"c_bulkflow for _pset_id {row['_pset_id']} must be a 2D or "
                                     "3D array with shape[0]=2")
                cp += coeff * cm[0, k, :]
                cx += coeff * cm[1, k, :]
            out.append(spectrum_from_c(w, cp, cx))
        return np.array(out)

    @classmethod
    def from_bulk_flow(cls, bulk_flow, w_arr, cos_thetak_arr, powers, *, t_end, n_t, t_begin=None,
                       coefficient: float = 1.0, envelope: bool = True, precision="fp64",
                       d0: float = 1.0) -> "DataGwBulkFlow":
        """Run the device path for every power (one coefficient set each) and, with
        ``envelope``, the envelope C (coefficient 0: f = 0 after collision)."""
        powers = [float(p) for p in powers]
        cos = np.asarray(cos_thetak_arr, dtype=np.float64)
        k_hat = np.column_stack([np.sqrt(np.maximum(1.0 - cos**2, 0.0)), np.zeros_like(cos), cos])
        coef = [[float(coefficient)] for _ in powers] + ([[0.0]] if envelope else [])
        pows = [[p] for p in powers] + ([[3.0]] if envelope else [])
        bulk_flow.set_gradient_scaling_params(coef, pows)
        _, _, C = bulk_flow.compute_spectrum(w_arr, k_hat, t_begin=t_begin, t_end=t_end, n_t=n_t,
                                             precision=precision, return_c=True)
        n_p = len(powers)
        return cls(w_arr, cos, C[:, :, :n_p, :], powers,
                   C[:, :, n_p, :] if envelope else None, d0=d0)

    def de_dlogw_dcosthetak(self, xi_powers, xi_coefficients, check_sum: bool = True) -> np.ndarray:
        """(n_cos, n_w): S of the combination sum_i coeff_i C(xi_i) per cos(theta_k)
        (data_gw_bulkflow.py:25-115)."""
        if not isinstance(xi_powers, (list, np.ndarray)) or not isinstance(xi_coefficients,
                                                                            (list, np.ndarray)):
            raise ValueError("xi_powers and xi_coefficients must be lists or numpy arrays")
        xi_powers = list(np.asarray(xi_powers, dtype=object))
        xi_coefficients = np.asarray(xi_coefficients)
        if len(xi_powers) != len(xi_coefficients):
            raise ValueError("xi_powers and xi_coefficients must have the same length")
        if check_sum and not np.isclose(np.sum(xi_coefficients), 1.0):
            raise ValueError("xi_coefficients must sum to 1")
        if not all(p is None or isinstance(p, (int, float, np.integer, np.floating)) for p in xi_powers):
            raise ValueError("xi_powers must contain numbers or None")
        if self.df is not None:
            return self._frame_spectra(xi_powers, xi_coefficients)
        n_cos, n_w = self.c_bulkflow.shape[0], self.c_bulkflow.shape[3]
        c_plus = np.zeros((n_cos, n_w), dtype=np.complex128)
        c_cross = np.zeros((n_cos, n_w), dtype=np.complex128)
        for p, coeff in zip(xi_powers, xi_coefficients):
            if p is None:
                if self.c_envelope is None:
                    raise ValueError("c_envelope required for None powers")
                c_plus += coeff * self.c_envelope[:, 0, :]
                c_cross += coeff * self.c_envelope[:, 1, :]
                continue
            power = float(p) + 3
            if self.powers_sets_sample is None:
                raise ValueError("powers_sets_sample required for non-None powers")
            if power not in self.powers_sets_sample:
                raise ValueError(f"Power {power} not in powers_sets_sample")
            k = self.powers_sets_sample.index(power)
            c_plus += coeff * self.c_bulkflow[:, 0, k, :]
            c_cross += coeff * self.c_bulkflow[:, 1, k, :]
        return spectrum_from_c(self.w_arr, c_plus, c_cross)

    def de_dlogw(self, xi_powers, xi_coefficients, check_sum: bool = True) -> np.ndarray:
        """(n_w): 2 * trapz over cos(theta_k) (data_gw_bulkflow.py:117-146)."""
        s = self.de_dlogw_dcosthetak(xi_powers, xi_coefficients, check_sum)
        return 2.0 * np.trapezoid(s, x=self.cos_thetak_arr, axis=0)

    def de_dlogw_interp(self, xi_powers, xi_coefficients, check_sum: bool = True, d=None,
                        rho_vacuum: float = 1.0) -> CubicSpline:
        """Cubic spline in w / d of (d / d0)^5 rho^2 de_dlogw (data_gw_bulkflow.py:148-190)."""
        if d is None:
            d = self.d0
        y = (d / self.d0) ** 5 * rho_vacuum**2 * self.de_dlogw(xi_powers, xi_coefficients, check_sum)
        return CubicSpline(self.w_arr / d, y)


def _log1pexp(x):
    return np.logaddexp(0.0, x)


def general_broken_power_law(w, amplitude, wscales, widths, powers, windex_amplitude):
    """amplitude * (w / w_A)^p_0 * prod_k [(1 + (w/w_k)^(1/e_k)) / (1 + (w_A/w_k)^(1/e_k))]^(e_k (p_{k+1} - p_k)),
    w_A = wscales[windex_amplitude - 1] (general_broken_power_law.py:19-66), in log space."""
    w = np.asarray(w, dtype=np.float64)
    wscales = np.asarray(wscales, dtype=np.float64)
    widths = np.asarray(widths, dtype=np.float64)
    powers = np.asarray(powers, dtype=np.float64)
    if len(powers) != len(wscales) + 1 or len(widths) != len(wscales):
        raise ValueError("need n scales, n widths and n + 1 powers")
    w_amp = wscales[windex_amplitude - 1]
    out = powers[0] * np.log(w / w_amp)
    for k in range(len(wscales)):
        e = 1.0 / widths[k]
        dp = powers[k + 1] - powers[k]
        log_den = _log1pexp(e * np.log(w_amp / wscales[k]))
        out = out + widths[k] * dp * (_log1pexp(e * np.log(w / wscales[k])) - log_den)
    return amplitude * np.exp(out)


def asymptotic_broken_power_law(w, amplitude, wscales, widths, powers, windex_amplitude, region):
    """The pure power law of ``region`` (0: below the first scale, n: above the last)
    that the general form approaches (general_broken_power_law.py:69-107)."""
    w = np.asarray(w, dtype=np.float64)
    wscales = np.asarray(wscales, dtype=np.float64)
    widths = np.asarray(widths, dtype=np.float64)
    powers = np.asarray(powers, dtype=np.float64)
    w_amp = wscales[windex_amplitude - 1]
    pref = 1.0
    for j in range(len(wscales)):
        dp = powers[j + 1] - powers[j]
        ratio = (w_amp / wscales[j]) ** (1.0 / widths[j])
        pref /= (1.0 + ratio) ** (widths[j] * dp)
        if j < region:
            pref *= (w_amp / wscales[j]) ** dp
    return amplitude * pref * (w / w_amp) ** powers[region]


def fit_broken_power_law(w, y, *, wscales, widths, powers, windex_amplitude, amplitude=None,
                         sigma=None) -> dict:
    """Least-squares fit of ``general_broken_power_law`` to y(w) in log space.

    Starting values: the given scales, widths and powers (and amplitude, default
    y interpolated at the amplitude scale).  Scales, widths and the amplitude are
    fitted as logarithms (kept positive).  Returns the fitted parameters, the
    residual sum of squares of log y and the fit's success flag.
    """
    w = np.asarray(w, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = len(wscales)
    if amplitude is None:
        amplitude = float(np.exp(np.interp(np.log(wscales[windex_amplitude - 1]), np.log(w), np.log(y))))
    x0 = np.concatenate([[np.log(amplitude)], np.log(wscales), np.log(widths), powers])
    wt = 1.0 if sigma is None else 1.0 / np.asarray(sigma, dtype=np.float64)

    def unpack(x):
        return np.exp(x[0]), np.exp(x[1:1 + n]), np.exp(x[1 + n:1 + 2 * n]), x[1 + 2 * n:]

    def resid(x):
        a, s, e, p = unpack(x)
        return wt * (np.log(general_broken_power_law(w, a, s, e, p, windex_amplitude)) - np.log(y))

    r = least_squares(resid, x0, method="lm" if len(w) >= len(x0) else "trf", xtol=1e-15,
                      ftol=1e-15, gtol=1e-15, max_nfev=20000)
    a, s, e, p = unpack(r.x)
    return {"amplitude": float(a), "wscales": s.tolist(), "widths": e.tolist(), "powers": p.tolist(),
            "rss_log": float(np.sum(r.fun**2)), "success": bool(r.success)}
