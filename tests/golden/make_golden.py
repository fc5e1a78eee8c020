"""Writes the golden fixtures of tests/golden from the CPU oracle.

Run from the repository root:  python tests/golden/make_golden.py
The fixtures are data only (inputs and oracle outputs, numpy .npz without
pickles).  Each is cross-checked against the independent numpy restatement
before it is written.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.dirname(HERE)]

import independent_numpy as ind  # noqa: E402
import oracle  # noqa: E402
from gw_fopt_amd import configs  # noqa: E402


def main():
    # the reference's C-ABI test case (rust/c-bubble-gw/tests/c_api.rs:135-218)
    inter = np.array([[0.0, 0.0, 10.0, 0.0], [0.0, 0.0, 0.0, 0.0]])
    coef = np.array([[0.0], [1.0]])
    pows = np.array([[3.0], [3.0]])
    w = np.array([0.5, 1.0])
    C = oracle.compute_c_integral(inter, None, 4, 5, coef, pows, w, t_begin=0.0, t_end=8.0, n_t=40)
    Cn = ind.c_integral(inter, None, 4, 5, coef, pows, w, 0.0, 8.0, 40)
    assert np.max(np.abs(C - Cn)) / np.max(np.abs(C)) < 1e-13
    np.savez(os.path.join(HERE, "capi_two_bubbles.npz"), interior=inter, coef=coef, pows=pows,
             w=w, C=C)

    wl = configs.plain_8()
    S = oracle.spectrum(wl.interior, wl.exterior, wl.k_hat, 32, 64, [[1.0]], [[5.0]], wl.w_arr,
                        t_end=1.5, n_t=200)
    first, dtab = oracle.first_colliding(wl.interior, wl.exterior, 32, 64, 0, k_hat=wl.k_hat[3])
    np.savez(os.path.join(HERE, "plain_8.npz"), interior=wl.interior, exterior=wl.exterior,
             k_hat=wl.k_hat, w=wl.w_arr, S=S, first_k3_a0=first, dtab_k3_a0=dtab)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
