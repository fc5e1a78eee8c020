#!/usr/bin/env python3
"""Benchmark of the generalized bulk-flow spectrum on one MI355X.

Workload (plain_256): 256 bubbles from seeded exponential nucleation
(beta L = 20) in a periodic unit cube, six face images each; 64 x 128 surface
grid; 1000 time steps to full percolation; 256 omegas geomspace(0.5, 500);
64 Fibonacci k-hat; scaling f = ratio^5.  One step = the whole spectrum
S(k-hat, omega) over 64 x 256 points: first-collision geometry, phase sums,
per-k-hat reduction.

Four forms run on the same inputs: the plain FP64 form (one FP64 sincos per
term), the blocked FP64 form (phases factorised over blocks of time steps,
FP64 throughout), the mixed form (FP64 masks, argument reduction, anchors and
accumulation, FP32 phase tables, weights and block sums) and the FP64 matrix
form (the phase sums on the FP64 matrix cores).  Each form's per-bin error
against the plain FP64 form is reported.  `value` is the (omega, k-hat)
points/s of the fastest form whose largest per-bin relative error is
<= 1e-6, from the device time of the step (HIP events on the library's
stream, bubbles resident in HBM).

Also reported: the batched ensemble (512 realisations, uneven bubble counts,
one graph-captured launch with autotuned tiles), device nucleation of 512
fixed-rate realisations in one launch, the differentiation workload
(adjoint gradients per second, finite-difference check, 200-step fit), the
roofline of the phase-sum kernel, its HBM traffic from rocprofv3 counters,
and the CPU oracle on a bounded sample.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6    # MI355X vector FP64 (published)
HBM_PEAK_GBS = 8000.0      # MI355X HBM3E (published)
REL_BAR = 1e-6


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n-k", type=int, default=64, help="k-hat directions (workload: 64)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-grad", action="store_true")
    p.add_argument("--no-ensemble", action="store_true")
    p.add_argument("--no-nucleation", action="store_true")
    p.add_argument("--profile-nucleation", action="store_true",
                   help="only run the 512-realisation nucleation batch --steps times (for rocprofv3)")
    p.add_argument("--profile-ensemble", default=None, metavar="FORM",
                   help="only run ensemble_512 in this form --steps times with the tiles the "
                        "autotuner picks for it (for rocprofv3)")
    p.add_argument("--profile-grad", action="store_true",
                   help="only run --steps grad_128 gradients (forward + adjoint, for rocprofv3)")
    p.add_argument("--profile-tensor", default=None, metavar="FORM",
                   help="only run --steps compute_c_tensor calls (the tensor form C_ij) in this "
                        "form at plain_256's bubbles, grid, times and omegas (for rocprofv3)")
    p.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 traffic counters")
    p.add_argument("--modes", default="mixed,fp64_blocked,fp64_matrix,fp64")
    p.add_argument("--profile-mode", default=None,
                   help="run only this form for --steps steps (the program rocprofv3 profiles)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="one device copy of 1 GiB (reads 1 GiB, writes 1 GiB): the reference "
                        "launch for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    return p.parse_args()


# ---------------------------------------------------------------------------
# rocprofv3 counters (child processes, started before this process touches the GPU)
# ---------------------------------------------------------------------------
def calibrate_copy():
    import torch
    n = (1 << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes_read": n * 4, "bytes_written": n * 4}), flush=True)


def _pmc_run(counters: list[str], args: list[str], timeout: int = 240) -> dict:
    """{counter: {kernel name: [value per dispatch]}} from one rocprofv3 run."""
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    out = tempfile.mkdtemp(prefix="bgw_pmc_", dir=os.environ.get("TMPDIR", "/tmp"))
    cmd = [prof, "--pmc", *counters, "--output-format", "csv", "-d", out, "-o", "run", "--",
           sys.executable, os.path.abspath(__file__), *args]
    try:
        subprocess.run(cmd, check=True, timeout=timeout, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, cwd=os.environ.get("TMPDIR", "/tmp"))
        res: dict = {c: {} for c in counters}
        for path in glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as f:
                for row in csv.DictReader(f):
                    c = row.get("Counter_Name")
                    if c in res:
                        res[c].setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
        return res
    finally:
        shutil.rmtree(out, ignore_errors=True)


def _per_call(res: dict, needles, n_calls: int):
    """Counter summed over the dispatches of the kernels named by `needles`, per call."""
    vals = [v for k, vs in res.items() if any(n in k for n in needles) for v in vs]
    return sum(vals) / n_calls if vals else None


def pmc_copy_calibration():
    """FETCH_SIZE and WRITE_SIZE read on a device copy of 1 GiB (the largest
    dispatch of the calibration run)."""
    try:
        cf = _pmc_run(["FETCH_SIZE"], ["--calibrate-copy"])["FETCH_SIZE"]
        cw = _pmc_run(["WRITE_SIZE"], ["--calibrate-copy"])["WRITE_SIZE"]
        copy_f = max((v for vs in cf.values() for v in vs), default=None)
        copy_w = max((v for vs in cw.values() for v in vs), default=None)
        if copy_f is None or copy_w is None or copy_f <= 0 or copy_w <= 0:
            return None
        return copy_f, copy_w
    except Exception:
        return None


# kernels of each part of a step, by name
KERNELS = {
    "geometry": ("geometry_kernel",),
    "spectrum": ("spectrum_kernel", "mx_time_kernel", "mx_weights_kernel", "mx_gemm_kernel"),
    "matrix": ("mx_gemm_kernel",),
    "weights": ("mx_weights_kernel",),
}
PMC_CALLS = 3  # --warmup 1 --steps 2


def pmc_flops(mode: str, n_k: int):
    """Executed FP64 flops per call of each part of the step: vector FP64
    (SQ_INSTS_VALU_FLOPS_FP64 counts flops per wave instruction, an FMA as 2,
    times 64 lanes) plus matrix-core FP64 (SQ_INSTS_VALU_MFMA_MOPS_F64 x 512).
    Returns {part: {"valu": F, "mfma": F, "total": F}} (None where missing)."""
    empty = {k: None for k in KERNELS}
    try:
        prog = ["--profile-mode", mode, "--steps", "2", "--warmup", "1", "--n-k", str(n_k)]
        res = _pmc_run(["SQ_INSTS_VALU_FLOPS_FP64", "SQ_INSTS_VALU_MFMA_MOPS_F64"], prog)
        out = {}
        for part, needles in KERNELS.items():
            v = _per_call(res["SQ_INSTS_VALU_FLOPS_FP64"], needles, PMC_CALLS)
            mf = _per_call(res["SQ_INSTS_VALU_MFMA_MOPS_F64"], needles, PMC_CALLS)
            if v is None:
                out[part] = None
                continue
            valu, mfma = 64.0 * v, 512.0 * (mf or 0.0)
            out[part] = {"valu": valu, "mfma": mfma, "total": valu + mfma}
        return out
    except Exception:
        return empty


def pmc_traffic(mode: str, n_k: int, calibration, parts=("spectrum",)):
    """HBM bytes per call of the named parts: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch), each scaled by bytes per unit measured the same way on a 1 GiB
    device copy."""
    if calibration is None:
        return None
    copy_f, copy_w = calibration
    try:
        prog = ["--profile-mode", mode, "--steps", "2", "--warmup", "1", "--n-k", str(n_k)]
        fetch = _pmc_run(["FETCH_SIZE"], prog)["FETCH_SIZE"]
        write = _pmc_run(["WRITE_SIZE"], prog)["WRITE_SIZE"]
        f_scale = (1 << 30) / copy_f   # bytes per counted unit, reads
        w_scale = (1 << 30) / copy_w   # writes
        out = {"fetch_scale_bytes_per_kib": f_scale / 1024.0,
               "write_scale_bytes_per_kib": w_scale / 1024.0}
        for part in parts:
            f = _per_call(fetch, KERNELS[part], PMC_CALLS)
            w = _per_call(write, KERNELS[part], PMC_CALLS)
            if f is None or w is None:
                out[part] = None
            else:
                out[part] = {"fetch_bytes": f * f_scale, "write_bytes": w * w_scale,
                             "bytes": f * f_scale + w * w_scale}
        return out
    except Exception:
        return None


# ---------------------------------------------------------------------------
# plain_256
# ---------------------------------------------------------------------------
def algorithmic_flops(wl):
    """F = n_k (10 S_A + 10 S_C + 8 E_B + 10 G) with S_A = N n_t n_c n_w, S_C = N n_t n_w,
    E_B = N n_t n_c n_phi, G = N n_c n_phi N_tot (FP64-equivalent flops)."""
    N = wl.n_bubbles
    n_tot = N + len(wl.exterior)
    n_k, n_t, n_w = len(wl.k_hat), wl.n_t, len(wl.w_arr)
    n_c, n_p = wl.n_cos_thetax, wl.n_phix
    S_A = N * n_t * n_c * n_w
    S_C = N * n_t * n_w
    E_B = N * n_t * n_c * n_p
    G = N * n_c * n_p * n_tot
    return {"spectrum": n_k * (10 * S_A + 10 * S_C + 8 * E_B), "geometry": n_k * 10 * G}


def _sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def run_mode(bf, wl, mode, steps, warmup):
    import numpy as np
    for _ in range(warmup):
        bf.compute_spectrum(wl.w_arr, wl.k_hat, t_end=wl.t_end, n_t=wl.n_t, precision=mode)
    dev, geo, spec, red, mat = [], [], [], [], []
    _sync()
    t0 = time.perf_counter()
    S = None
    for _ in range(steps):
        S, _ = bf.compute_spectrum(wl.w_arr, wl.k_hat, t_end=wl.t_end, n_t=wl.n_t, precision=mode)
        tm = bf.last_timings()
        dev.append(tm["total_ms"])
        geo.append(tm["geometry_ms"])
        spec.append(tm["spectrum_ms"])
        red.append(tm["reduce_ms"])
        mat.append(tm["matrix_ms"])
    _sync()
    wall = (time.perf_counter() - t0) / steps
    return S, {"device_ms": float(np.mean(dev)), "geometry_ms": float(np.mean(geo)),
               "spectrum_ms": float(np.mean(spec)), "reduce_ms": float(np.mean(red)),
               "matrix_ms": float(np.mean(mat)), "wall_ms": wall * 1e3}


def cpu_baseline(wl, n_k_sample=2):
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    k = wl.k_hat[:n_k_sample]
    t0 = time.perf_counter()
    oracle.spectrum(wl.interior, wl.exterior, k, wl.n_cos_thetax, wl.n_phix,
                    wl.coefficients_sets, wl.powers_sets, wl.w_arr, t_end=wl.t_end, n_t=wl.n_t,
                    n_threads=threads)
    dt = time.perf_counter() - t0
    pts = len(wl.w_arr) * len(k)
    return {"value": pts / dt, "unit": "points/s", "cores": threads, "kind": "port",
            "sample": f"{len(k)} of {len(wl.k_hat)} k-hat directions of plain_256 (all 256 "
                      f"bubbles, 64x128 grid, 1000 t, 256 omegas) on the CPU oracle with "
                      f"{threads} OpenMP threads, {dt:.1f} s; k-hat directions are independent "
                      "equal-cost units, so the rate holds for the whole workload"}


# ---------------------------------------------------------------------------
# ensemble_512
# ---------------------------------------------------------------------------
def ensemble_bench(ens, steps):
    import numpy as np
    from gw_fopt_amd.ensemble import BubbleEnsemble
    t0 = time.perf_counter()
    E = BubbleEnsemble(ens.interiors, ens.exteriors)
    E.set_resolution(ens.n_cos_thetax, ens.n_phix)
    E.set_gradient_scaling_params(ens.coefficients_sets, ens.powers_sets)
    build_s = time.perf_counter() - t0
    w = ens.w_arr()
    n_real = len(ens.interiors)
    pts = n_real * len(ens.k_hat) * w.shape[1]
    out, res = {}, {}
    tuned = {}
    for m in ("mixed", "fp64_blocked", "fp64_matrix", "fp64"):
        if m == "fp64":
            tl = dict(tuned["fp64_blocked"])
            tl["block_steps"] = 16
        else:
            t1 = time.perf_counter()
            tl = E.autotune(w, ens.k_hat, t_end=ens.t_end, n_t=ens.n_t, precision=m)
            tl["autotune_s"] = time.perf_counter() - t1
            tuned[m] = {k: v for k, v in tl.items() if k != "autotune_s"}
        tiles = {k: v for k, v in tl.items() if k != "autotune_s"}
        r = E.compute_spectrum(w, ens.k_hat, t_end=ens.t_end, n_t=ens.n_t, precision=m, tiles=tiles)
        ms, wall = [], []
        for _ in range(max(1, steps)):
            t1 = time.perf_counter()
            r = E.compute_spectrum(w, ens.k_hat, t_end=ens.t_end, n_t=ens.n_t, precision=m,
                                   tiles=tiles)
            wall.append(time.perf_counter() - t1)
            ms.append(E.last_timings()["total_ms"])
        res[m] = r
        out[m] = {"device_ms": float(np.mean(ms)), "points_per_s": pts / (np.mean(ms) * 1e-3),
                  "wall_ms": float(np.mean(wall)) * 1e3, "tiles": tl}
        print(f"# ensemble {m}: {out[m]}", file=sys.stderr, flush=True)
    ref = res["fp64"]
    for m in ("mixed", "fp64_blocked", "fp64_matrix"):
        rel = np.abs(res[m]["S"] - ref["S"]) / np.abs(ref["S"])
        out[m]["err_vs_fp64"] = {
            "max": float(rel.max()), "median": float(np.median(rel)),
            "frac_le_1e-6": float(np.mean(rel <= REL_BAR)),
            "ensemble_mean_max": float(np.max(np.abs(res[m]["mean"] - ref["mean"]) / np.abs(ref["mean"])))}
        out[m]["speedup_vs_fp64"] = out[m]["points_per_s"] / out["fp64"]["points_per_s"]
    # the batched target: the fastest form whose largest per-bin error is within the bar
    ok = [m for m in ("mixed", "fp64_blocked", "fp64_matrix") if out[m]["err_vs_fp64"]["max"] <= REL_BAR]
    best = max(ok, key=lambda m: out[m]["points_per_s"]) if ok else None
    counts = np.array([len(b) for b in ens.interiors])
    del E  # release the batch's device buffers before the rest of the run
    return {"workload": "ensemble_512: 512 realisations, beta L ~ U[10,30], bubble counts "
                        f"{int(counts.min())}-{int(counts.max())} (mean {counts.mean():.1f}), "
                        "32x64, 500 t to each percolation time, 128 omegas geomspace(0.5,300)/L, "
                        "32 k-hat; mixed / blocked / FP64: one graph-captured launch sequence "
                        "per call; FP64 matrix form: the single-configuration launches over "
                        "every realisation at once",
            "points": pts, "setup_s": build_s, "modes": out,
            "best_within_1e-6": None if best is None else {
                "form": best, "points_per_s": out[best]["points_per_s"],
                "speedup_vs_fp64": out[best]["speedup_vs_fp64"],
                "max_rel_err_vs_fp64": out[best]["err_vs_fp64"]["max"]},
            "ensemble_mean_spread": {"mean_S_first_bin": float(ref["mean"][0, 0]),
                                     "std_S_first_bin": float(ref["std"][0, 0])}}


NUC_KW = dict(boundary_condition="periodic", method="fixed_probability_time_stepping", d_p0=0.01,
              n_points=10000, max_bubbles=4096)


def nucleation_specs(n_real=512, seed=11):
    """512 fixed-rate realisations in the unit Cartesian box, beta ~ U[10, 30],
    Gamma(t) = 0.01 beta^4 exp(beta t) (gamma0 beta^-4 = 0.01), seed 11, stream r."""
    import numpy as np
    from gw_fopt_amd import many_bubbles as mb
    rng = np.random.default_rng(seed)
    lat = mb.CartesianLattice((0.0, 0.0, 0.0), np.eye(3))
    return [(lat, float(b), 0.01 * float(b) ** 4, 0.0, seed, r)
            for r, b in enumerate(rng.uniform(10.0, 30.0, n_real))]


def nucleation_bench(steps=2, cpu_sample=4):
    """Device nucleation of 512 realisations (one launch) and the oracle on a sample."""
    import numpy as np
    from gw_fopt_amd import many_bubbles as mb
    specs = nucleation_specs()
    mb.nucleate_fixed_rate_batch(specs[:8], **NUC_KW)  # load and warm up
    ms, walls = [], []
    for _ in range(max(1, steps)):
        t0 = time.perf_counter()
        interior, exterior, th, vh, counts = mb.nucleate_fixed_rate_batch(specs, **NUC_KW)
        walls.append(time.perf_counter() - t0)
        ms.append(mb.last_nucleation_kernel_ms())
    kernel_ms = float(np.median(ms))
    out = {"workload": "nucleation_512: 512 fixed-rate realisations (time stepping, d_p0 0.01, "
                       "10000 sample points, cutoff 1 %), unit Cartesian box, periodic, "
                       "beta ~ U[10,30], gamma0 = 0.01 beta^4; one launch",
           "kernel_ms": kernel_ms, "wall_s": float(np.median(walls)),
           "realisations_per_s": len(specs) / (kernel_ms * 1e-3),
           "bubbles": int(counts.sum()), "bubbles_per_s": float(counts.sum()) / (kernel_ms * 1e-3),
           "bubble_counts": [int(counts.min()), int(counts.max())]}
    try:
        import oracle.nucleation as on
        rng = np.random.default_rng(5)
        sample = sorted(rng.choice(len(specs), cpu_sample, replace=False).tolist())
        t0 = time.perf_counter()
        same = True
        for r in sample:
            _, beta, gamma0, t0_, seed, stream = specs[r]
            oi, _, _, ov = on.fixed_rate(("box", [0.0, 0.0, 0.0], np.eye(3)), 0, 0, beta=beta,
                                         gamma0=gamma0, t0=t0_, d_p0=NUC_KW["d_p0"], seed=seed,
                                         stream=stream, n_points=NUC_KW["n_points"])
            n = int(counts[r])
            same = same and n == len(oi) and np.array_equal(interior[r, :n, 1:], oi[:, 1:]) \
                and np.array_equal(vh[r, :n], ov)
        cpu_s = time.perf_counter() - t0
        share = float(counts[sample].sum()) / float(counts.sum())
        out["cpu_baseline"] = {
            "kind": "port", "cores": 1, "sample": f"realisations {sample} (beta "
            f"{[round(specs[r][1], 2) for r in sample]}) on the nucleation oracle, one thread, "
            f"{cpu_s:.1f} s; scaled to the whole batch by bubble count (the oracle's work is "
            "proportional to the iterations, about bubbles / d_p0)",
            "sample_s": cpu_s, "batch_s_extrapolated": cpu_s / share,
            "realisations_per_s": len(specs) * share / cpu_s}
        out["speedup_vs_cpu_baseline"] = out["realisations_per_s"] / out["cpu_baseline"]["realisations_per_s"]
        out["sample_matches_oracle"] = bool(same)
    except Exception as e:  # reported, not hidden
        out["cpu_baseline"] = {"error": f"{type(e).__name__}: {e}"}
    return out


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the path shards by k-hat "
                          "direction (no exchange) but multi-GPU launch is not part of it",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)
    import numpy as np
    from gw_fopt_amd import configs

    if args.profile_nucleation:
        from gw_fopt_amd import many_bubbles as mb
        specs = nucleation_specs()
        for _ in range(args.steps):
            _, _, _, _, counts = mb.nucleate_fixed_rate_batch(specs, **NUC_KW)
        print(json.dumps({"profile": "nucleation_512", "kernel_ms": mb.last_nucleation_kernel_ms(),
                          "bubbles": int(counts.sum())}), flush=True)
        return
    if args.profile_ensemble:
        from gw_fopt_amd.ensemble import BubbleEnsemble
        ens = configs.ensemble_512()
        E = BubbleEnsemble(ens.interiors, ens.exteriors)
        E.set_resolution(ens.n_cos_thetax, ens.n_phix)
        E.set_gradient_scaling_params(ens.coefficients_sets, ens.powers_sets)
        w = ens.w_arr()
        tiles = E.autotune(w, ens.k_hat, t_end=ens.t_end, n_t=ens.n_t,
                           precision=args.profile_ensemble)
        ms = []
        for _ in range(args.steps):
            E.compute_spectrum(w, ens.k_hat, t_end=ens.t_end, n_t=ens.n_t,
                               precision=args.profile_ensemble, tiles=tiles)
            ms.append(E.last_timings()["total_ms"])
        print(json.dumps({"profile": "ensemble_512", "form": args.profile_ensemble, "tiles": tiles,
                          "device_ms": ms}), flush=True)
        return
    if args.profile_grad:
        from gw_fopt_amd.gradbench import GradProblem
        P = GradProblem(precision="fp64_blocked")
        for _ in range(args.steps):
            P.loss_and_grad((1.8, 1.1))
        print(json.dumps({"profile": "grad_128", "gradients": args.steps}), flush=True)
        return
    if args.profile_tensor:
        from gw_fopt_amd import many_bubbles as mb
        tw = configs.plain_256(n_k=1)
        bf = mb.NewGeneralizedBulkFlow(mb.LatticeBubbles(mb.EmptyLattice(), tw.interior, tw.exterior))
        bf.set_resolution(tw.n_cos_thetax, tw.n_phix)
        bf.set_gradient_scaling_params(tw.coefficients_sets, tw.powers_sets)
        form = args.profile_tensor
        for _ in range(args.warmup):
            bf.compute_c_tensor(tw.w_arr, t_end=tw.t_end, n_t=tw.n_t, precision=form)
        walls = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            # returns host arrays, so the call has finished on the device
            bf.compute_c_tensor(tw.w_arr, t_end=tw.t_end, n_t=tw.n_t, precision=form)
            walls.append((time.perf_counter() - t0) * 1e3)
        print(json.dumps({"profile": "tensor plain_256", "form": form, "calls": args.steps,
                          "wall_ms_per_call": walls}), flush=True)
        return
    wl = configs.plain_256(n_k=args.n_k)
    if args.profile_mode:
        from gw_fopt_amd import many_bubbles as mb
        lb = mb.LatticeBubbles(mb.EmptyLattice(), wl.interior, wl.exterior)
        bf = mb.GeneralizedBulkFlow(lb)
        bf.set_resolution(wl.n_cos_thetax, wl.n_phix)
        bf.set_gradient_scaling_params(wl.coefficients_sets, wl.powers_sets)
        S, st = run_mode(bf, wl, args.profile_mode, args.steps, args.warmup)
        print(json.dumps({"profile_mode": args.profile_mode, **st}), flush=True)
        return

    # CPU-side inputs and counter runs before this process touches the GPU
    ens = None if args.no_ensemble else configs.ensemble_512()
    traffic, executed = {}, {}
    if not args.no_pmc:
        cal = pmc_copy_calibration()
        traffic["fp64_matrix"] = pmc_traffic("fp64_matrix", args.n_k, cal,
                                             parts=("matrix", "weights", "spectrum", "geometry"))
        traffic["fp64_blocked"] = pmc_traffic("fp64_blocked", args.n_k, cal)
        for m in ("fp64_matrix", "fp64_blocked", "fp64"):
            executed[m] = pmc_flops(m, args.n_k)

    from gw_fopt_amd import many_bubbles as mb
    lb = mb.LatticeBubbles(mb.EmptyLattice(), wl.interior, wl.exterior)
    bf = mb.GeneralizedBulkFlow(lb)
    bf.set_resolution(wl.n_cos_thetax, wl.n_phix)
    bf.set_gradient_scaling_params(wl.coefficients_sets, wl.powers_sets)
    points = wl.points()
    flops = algorithmic_flops(wl)

    modes = args.modes.split(",")
    results, spectra = {}, {}
    for m in modes:
        spectra[m], results[m] = run_mode(bf, wl, m, args.steps, args.warmup)
        print(f"# {m}: {results[m]}", file=sys.stderr, flush=True)
    ref = spectra.get("fp64")
    errs = {}
    for m in modes:
        if ref is None:
            break
        rel = np.abs(spectra[m] - ref) / np.abs(ref)
        errs[m] = {"max": float(rel.max()), "median": float(np.median(rel)),
                   "frac_le_1e-6": float(np.mean(rel <= REL_BAR))}
    # headline: the fastest form within the error bar (the FP64 form is the reference)
    ok = [m for m in modes if m == "fp64" or (m in errs and errs[m]["max"] <= REL_BAR)]
    main_mode = min(ok or modes, key=lambda m: results[m]["device_ms"])
    r = results[main_mode]
    value = points / (r["device_ms"] * 1e-3)
    spec_flops = flops["spectrum"]
    algorithmic_rate = spec_flops / (r["spectrum_ms"] * 1e-3) / 1e12
    ex = executed.get(main_mode) or {}

    def rate(part, ms):
        e = ex.get(part)
        return None if e is None or not ms else e["total"] / (ms * 1e-3) / 1e12

    dtypes = {"fp64": "f64", "fp64_blocked": "f64", "fp64_matrix": "f64",
              "mixed": "f32 phase tables, weights and block sums; f64 masks, argument reduction, anchors and accumulation"}
    tr = (traffic.get(main_mode) or {})
    if main_mode == "fp64_matrix":
        # the dominant kernel: the matrix-core products (HIP events around its launches)
        kern, part, ms, bound = "mx_gemm_kernel (FP64 matrix form)", "matrix", r["matrix_ms"], "mfma"
        basis = ("executed FP64 flops per launch: matrix cores (rocprofv3 "
                 "SQ_INSTS_VALU_MFMA_MOPS_F64 x 512) + vector (SQ_INSTS_VALU_FLOPS_FP64 x 64 "
                 "lanes), same workload, / the kernel's HIP-event time; peak: MI355X dense "
                 "FP64 matrix.  The products pair conjugate steps about 16-step block "
                 "centres: half the matrix flops of the direct phase sum for the same "
                 "outputs, plus the vector Horner steps that combine the blocks")
    else:
        kern, part, ms, bound = f"spectrum_kernel ({main_mode} form)", "spectrum", r["spectrum_ms"], "valu"
        basis = ("executed FP64 VALU flops per launch (rocprofv3 SQ_INSTS_VALU_FLOPS_FP64 x 64 "
                 "lanes, same workload) / the kernel's HIP-event time; peak: MI355X vector FP64")
    achieved = rate(part, ms)
    ktr = tr.get(part) if isinstance(tr, dict) else None
    line = {
        "metric": "bulk-flow (omega,k-hat) points/s: mixed vs FP64 GPU @ <=1e-6 rel err; grad/s",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["device_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtypes.get(main_mode, main_mode),
        "data": "synthetic (seeded nucleation)",
        "config": {"workload": "plain_256: 256 bubbles (beta L=20, periodic unit cube, 6 face "
                               "images), 64x128 surface grid, 1000 t to percolation, 256 omegas "
                               "geomspace(0.5,500), 64 Fibonacci k-hat, f=ratio^5",
                   "form": main_mode, "n_bubbles": wl.n_bubbles, "n_k": len(wl.k_hat),
                   "n_w": len(wl.w_arr), "n_t": wl.n_t, "grid": [wl.n_cos_thetax, wl.n_phix],
                   "timed_region": "device time of one whole spectrum call (geometry, phase "
                                   "sums, reduction; HIP events on the library's stream) with "
                                   "the bubbles resident in HBM; excludes the host-side call "
                                   "set-up and the copies of omegas in and spectra out "
                                   "(wall_points_per_s includes them)"},
        "roofline": {"bound": bound, "kernel": kern,
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": None if achieved is None else achieved / FP64_PEAK_TFLOPS,
                     "traffic": None if ktr is None else ktr["bytes"],
                     "achieved_basis": basis,
                     "executed_flops_per_launch": ex.get(part), "launch_ms": ms,
                     "hbm_GBs": None if ktr is None or not ms else ktr["bytes"] / (ms * 1e-3) / 1e9,
                     "algorithmic_flops_per_step": spec_flops,
                     "algorithmic_rate_tflops": algorithmic_rate,
                     "algorithmic_note": "the workload's formula count 10 S_A + 10 S_C + 8 E_B "
                                         "over the phase-sum time; it assumes one sincos per "
                                         "phase term, which none of the factorised forms "
                                         "executes, so this rate is not a fraction of any peak"},
        "modes": {m: {"points_per_s": points / (v["device_ms"] * 1e-3),
                      "wall_points_per_s": points / (v["wall_ms"] * 1e-3),
                      "err_vs_fp64": errs.get(m), **v}
                  for m, v in results.items()},
        "wall_points_per_s": points / (r["wall_ms"] * 1e-3),
    }
    if "fp64" in results:
        base = line["modes"]["fp64"]["points_per_s"]
        line["speedup_vs_fp64"] = {m: v["points_per_s"] / base for m, v in line["modes"].items()}
    # executed rates of every measured form's phase sums (all their kernels) and geometry
    line["executed"] = {}
    for m, exm in executed.items():
        if m not in results or not exm:
            continue
        rm = results[m]
        ent = {}
        for p_, ms_ in (("spectrum", rm["spectrum_ms"]), ("geometry", rm["geometry_ms"]),
                        ("matrix", rm["matrix_ms"])):
            e = exm.get(p_)
            if e is None or not ms_:
                continue
            t = e["total"] / (ms_ * 1e-3) / 1e12
            ent[p_] = {"flops_per_call": e, "ms": ms_, "tflops": t, "frac": t / FP64_PEAK_TFLOPS}
        line["executed"][m] = ent
    geo_ex = (ex.get("geometry") or {}).get("total") if ex.get("geometry") else None
    geo_rate = None if geo_ex is None else geo_ex / (r["geometry_ms"] * 1e-3) / 1e12
    line["geometry"] = {"kernel": "geometry_kernel", "launch_ms": r["geometry_ms"],
                        "executed_fp64_flops_per_launch": geo_ex,
                        "executed_tflops": geo_rate,
                        "frac": None if geo_rate is None else geo_rate / FP64_PEAK_TFLOPS,
                        "algorithmic_flops_per_launch": flops["geometry"],
                        "algorithmic_rate_tflops": flops["geometry"] / (r["geometry_ms"] * 1e-3) / 1e12,
                        "algorithmic_note": "10 flops per (bubble, direction, candidate) test; "
                                            "the key-ordered walks test far fewer candidates"}
    if traffic:
        line["pmc_traffic"] = traffic
    if not args.no_grad:
        try:
            from gw_fopt_amd import gradbench
            line["grad"] = gradbench.measure()
        except Exception as e:  # reported, not hidden
            line["grad"] = {"error": f"{type(e).__name__}: {e}"}
    if ens is not None:
        try:
            line["ensemble"] = ensemble_bench(ens, steps=2)
        except Exception as e:  # reported, not hidden
            line["ensemble"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_nucleation:
        try:
            line["nucleation"] = nucleation_bench()
        except Exception as e:  # reported, not hidden
            line["nucleation"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_cpu_baseline:
        cb = cpu_baseline(wl)
        line["cpu_baseline"] = cb
        line["speedup_vs_cpu_baseline"] = value / cb["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
