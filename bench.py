#!/usr/bin/env python3
"""Benchmark of the MI355X LM bundle adjuster.

Default run (one GPU, no flags): one JSON line whose headline is the "mid"
scene — 200 cameras with 9 parameters (angle-axis, translation, f, k1, k2),
50k points, ~1.0M keys, each point seen by 3..40 cameras (truncated geometric,
mean 20) — solved in mixed precision (FP32 Jacobian / Hessian / Schur, FP64
Cholesky, cost and gain ratio).  A "step" is one LM iteration (J + N + g
build, Schur, FP64 Cholesky, point back-substitution, candidate cost, accept /
reject); the timed region runs exactly --steps iterations from x0 with every
input already resident in HBM, bracketed by device synchronisation.

The same line carries the other workloads as sub-objects under "workloads":
  ring     10 cameras, 500 points, ~4.5k keys, FP64, 20 iterations
  windows  512 independent windows (5..40 cameras, 200..5000 points) batched
           in one set of launches per LM iteration vs solved one at a time
  large    2000 cameras x 9 along a trajectory, 600k points, ~5.1M keys,
           block-Jacobi PCG (rtol 1e-6), 30 iterations
  serving  Poisson arrivals, 95 % windows / 5 % 1M-key jobs, p50 / p99
  covariance  the full dense x covariance (hsba_x_covariance, FP64) of a
           30-camera ring with 500 points, gauge fixed (Schur route) and
           free (dense Jacobi route), against the oracle's SVD
Each is also available alone with --workload.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import platform
import statistics
import sys
import time
import traceback

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from hs_sfm_amd import abi, synthetic as syn  # noqa: E402
from hs_sfm_amd.hsba import Batch, Problem  # noqa: E402

METRIC = "BA LM iters/s (1M obs) + batched solves/s, p99 ms, at FP64-matched RMSE"
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
CHECK_THREADS = 16  # host threads for the oracle checks (the box's CPU share)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def pair_flops(scene: syn.Scene, E: int) -> float:
    """Schur products of one trial: every point contributes k(k+1)/2 camera
    pairs, each an E x E block of 3-term dot products (6 E^2 flops), plus the
    right-hand side Z^T q per key (6 E flops)."""
    k = np.bincount(scene.feature_maps[:, 1], minlength=scene.num_points).astype(np.float64)
    return float(np.sum(k * (k + 1) / 2) * 6 * E * E + scene.num_keys * 6 * E)


def linearize_bytes(scene: syn.Scene, E: int, tbytes: int) -> float:
    """K1+K2 algorithmic HBM bytes: per key the observation (16 B), image id
    (4 B), point id (4 B) and the stored W block (3 E T), per point its
    parameters (24 B), CSR offset (4 B) and V / g blocks (9 T)."""
    N, P = scene.num_keys, scene.num_points
    return float(N * (16 + 4 + 4 + 3 * E * tbytes) + P * (24 + 4 + 9 * tbytes))


def chol_flops(n: int) -> float:
    """Dense Cholesky n^3 / 3 plus the two triangular solves 2 n^2."""
    return n ** 3 / 3 + 2 * n ** 2


def load_traffic(name: str):
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def oracle_reference(name: str):
    path = os.path.join(ROOT, "tests", "golden", f"{name}_fp64_oracle.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def roofline_entry(kind, achieved, peak, unit, algorithmic, traffic=None, kernel=None):
    out = {"bound": kind, "achieved": achieved, "peak": peak, "unit": unit, "frac": achieved / peak,
           "algorithmic": algorithmic, "traffic": traffic}
    if kernel:
        out["kernel"] = kernel
    return out


def oracle_x(scene, iters):
    """FP64 oracle's x after `iters` LM iterations (checker only)."""
    from oracle.oracle import Oracle
    _, xo, _, _ = Oracle(scene).lm(max_iterations=iters)
    return xo


def oracle_x_all(scenes, iters):
    """Oracle x of every scene, on host threads (the oracle is a ctypes call,
    which releases the GIL)."""
    with cf.ThreadPoolExecutor(max_workers=CHECK_THREADS) as ex:
        return list(ex.map(lambda s: oracle_x(s, iters), scenes))


def oracle_rmse_all(scenes, iters):
    return [syn.reprojection_rmse(s, xo) for s, xo in zip(scenes, oracle_x_all(scenes, iters))]


def param_deviation(scene, x, xo):
    """Deviation of the parameters x from the oracle's xo (reference x layout:
    points | images | cameras, fixed blocks absent): max |dx| per block kind
    and |dx| / |xo| over all of x.  Both runs start from the same x0 and their
    LM trajectories agree, so they end in the same gauge and no similarity
    alignment is applied."""
    x, xo = np.asarray(x, np.float64), np.asarray(xo, np.float64)
    d = np.abs(x - xo)
    P, I = scene.num_points, scene.num_images
    o = 0
    out = {}
    if not scene.fix_mask & syn.FIX_POINTS:
        out["points_max_abs_m"] = float(d[o:o + 3 * P].max()) if P else 0.0
        o += 3 * P
    if not scene.fix_mask & syn.FIX_IMAGES:
        blk = d[o:o + 6 * I].reshape(-1, 6)
        out["rotation_max_abs_rad"] = float(blk[:, :3].max()) if I else 0.0
        out["translation_max_abs_m"] = float(blk[:, 3:].max()) if I else 0.0
        o += 6 * I
    if o < len(d):
        out["camera_params_max_abs"] = float(d[o:].max())
    out["relative_l2"] = float(np.linalg.norm(x - xo) / max(np.linalg.norm(xo), 1e-300))
    return out


# ------------------------------------------------------------------ configs[1]
def cpu_baseline_mid(scene: syn.Scene, samples: int = 3):
    """The FP64 oracle (single-threaded C++ restatement of the reference path)
    on the same scene: one LM iteration per sample, from two runs (0 and 1
    iterations) so the initial linearization is not counted; median of
    `samples` samples."""
    from oracle.oracle import Oracle  # CPU baseline leg only
    o = Oracle(scene)
    per = []
    for _ in range(samples):
        t0 = time.perf_counter()
        o.lm(max_iterations=0)
        t1 = time.perf_counter()
        o.lm(max_iterations=1)
        t2 = time.perf_counter()
        per.append(max(1e-9, (t2 - t1) - (t1 - t0)))
    med = float(np.median(per))
    return {"value": 1.0 / med, "unit": "iters/s", "cores": 1, "kind": "port",
            "sample": f"median of {samples} samples of one LM iteration of the same {scene.num_keys}-key scene "
                      f"(FP64, point-wise Schur + dense LDL^T, single thread; initial linearization excluded; "
                      f"per-sample s: {', '.join(f'{t:.2f}' for t in per)}), cpu={cpu_model()}, "
                      f"nproc={os.cpu_count()}"}


def run_mid(args):
    scene = syn.mid_scene()
    prec = abi.PRECISION_FP64 if args.precision == "fp64" else abi.PRECISION_MIXED
    E = 6 + scene.camera_params
    tbytes = 8 if prec == abi.PRECISION_FP64 else 4
    p = Problem(scene, precision=prec)
    K, W = args.steps, args.warmup
    for _ in range(W):
        abi.check(p.lm_begin(max_iterations=K), "lm_begin")
        abi.check(p.lm_iterate(K)[0], "lm_iterate")
        p.lm_end()
    # timed region: K iterations from x0, inputs resident, synchronised both sides
    abi.check(p.lm_begin(max_iterations=K), "lm_begin")
    t0 = time.perf_counter()
    rc, _ = p.lm_iterate(K)
    t1 = time.perf_counter()
    abi.check(rc, "lm_iterate")
    rc, x, st = p.lm_end()
    dt = t1 - t0
    rmse = syn.reprojection_rmse(scene, x)
    # stage breakdown: the same K iterations replayed with HIP events recorded
    # between the stage graphs on the library's stream
    abi.check(p.lm_begin(max_iterations=K), "lm_begin")
    rc, stages = p.lm_iterate(K, timed=True)
    abi.check(rc, "lm_iterate(timed)")
    p.lm_end()
    flops = pair_flops(scene, E)
    lin_bytes = linearize_bytes(scene, E, tbytes)
    n = 9 * scene.num_images
    cf_ = chol_flops(n)
    kernels = {
        "k_pairs": roofline_entry("mfma", flops / (stages["pairs_ms"] * 1e-3) / 1e12,
                                  PEAK_FP32_TFLOPS if tbytes == 4 else PEAK_FP64_TFLOPS, "TFLOP/s", flops,
                                  (load_traffic(f"k_pairs_mid_{args.precision}") or {}).get("bytes_per_launch")),
        "cholesky": roofline_entry("mfma", cf_ / (stages["cholesky_ms"] * 1e-3) / 1e12, PEAK_FP64_TFLOPS,
                                   "TFLOP/s", cf_,
                                   (load_traffic(f"cholesky_mid_{args.precision}") or {}).get("bytes_per_launch")),
        "k_linearize": roofline_entry("hbm", lin_bytes / (stages["linearize_ms"] * 1e-3) / 1e9, PEAK_HBM_GBS,
                                      "GB/s", lin_bytes),
    }
    dominant = max(("pairs_ms", "cholesky_ms", "linearize_ms", "image_reduce_ms", "point_factor_ms",
                    "assemble_ms", "backsub_ms", "step_cost_ms"), key=lambda k: stages[k])
    dom_kernel = {"pairs_ms": "k_pairs", "cholesky_ms": "cholesky", "linearize_ms": "k_linearize"}.get(dominant,
                                                                                                       "k_pairs")
    names = {"k_pairs": "k_pairs_mma (camera-pair Schur products on the FP32 matrix cores)",
             "cholesky": f"dense FP64 Cholesky + triangular solves of the {n} x {n} reduced system",
             "k_linearize": "k_linearize (residual + Jacobian + point blocks)"}
    roof = dict(kernels[dom_kernel], kernel=names[dom_kernel])
    ref = oracle_reference("mid")
    out = {
        "metric": METRIC, "value": K / dt, "unit": "iters/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": 1e3 * dt / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+fp64" if prec == abi.PRECISION_MIXED else "fp64", "data": "synthetic",
        "config": {"workload": "mid: 200 cameras x 9 params (r, t, f, k1, k2), 50k points, "
                               f"{scene.num_keys} keys, LM from x0",
                   "cameras": scene.num_images, "points": scene.num_points, "keys": scene.num_keys,
                   "precision": args.precision, "reduced_size": n},
        "roofline": roof, "kernels": kernels,
        "stage_ms": stages, "dominant_stage": dominant, "autotuner": abi.tuning_table(),
        "lm": {"status": st.status, "iterations": st.iterations, "accepted": st.accepted,
               "initial_cost": st.initial_cost, "final_cost": st.final_cost, "rmse_px": rmse,
               "mean_error_px": syn.mean_reprojection_error(scene, x),
               "rmse_x0_px": syn.reprojection_rmse(scene, scene.x0)},
        "fp64_gpu_plain": {"status": "switched off (HSBA_ERR_UNSUPPORTED) for this scene's 9-parameter cameras",
                           "why": "an FP64 LM on this scene faulted the GPU once and no cause in the code has "
                                  "been shown; the FP64 path with camera blocks >= 7 stays off",
                           "calibrated_cameras": fp64_vs_mixed_calibrated(scene, K)},
    }
    if ref and ref.get("iterations") == K:
        out["lm"]["fp64_oracle_rmse_px"] = ref["rmse_px"]
        out["lm"]["rmse_delta_px"] = abs(rmse - ref["rmse_px"])
        out["lm"]["final_cost_rel_err"] = abs(st.final_cost - ref["final_cost"]) / ref["final_cost"]
        xo_path = os.path.join(ROOT, "tests", "golden", "mid_fp64_oracle_x.npy")
        if os.path.exists(xo_path):
            out["lm"]["param_deviation_vs_fp64_oracle"] = param_deviation(scene, x, np.load(xo_path))
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_mid(scene, samples=args.cpu_samples)
    p.close()
    return out


def fp64_vs_mixed_calibrated(scene, K):
    """The same 1M-key problem with every camera held at its true value
    (calibrated cameras, camera block E = 6, which the FP64 path serves):
    K LM iterations from x0 in FP64 end to end and in mixed precision, each
    after one untimed run, with the FP64 result as the accuracy reference."""
    s = syn.hold_cameras(scene)
    out = {"workload": f"the 1M-key scene with cameras held (camera block 6): {s.num_keys} keys, "
                       f"reduced size {6 * s.num_images}, {K} LM iterations"}
    xs = {}
    for name, prec in (("fp64", abi.PRECISION_FP64), ("mixed", abi.PRECISION_MIXED)):
        p = Problem(s, precision=prec)
        abi.check(p.lm_begin(max_iterations=K), "lm_begin")
        abi.check(p.lm_iterate(K)[0], "lm_iterate")
        p.lm_end()
        abi.check(p.lm_begin(max_iterations=K), "lm_begin")
        t0 = time.perf_counter()
        rc, _ = p.lm_iterate(K)
        dt = time.perf_counter() - t0
        abi.check(rc, "lm_iterate")
        _, x, st = p.lm_end()
        p.close()
        xs[name] = x
        out[name] = {"iters_per_s": K / dt, "ms_per_iter": 1e3 * dt / K, "accepted": st.accepted,
                     "final_cost": st.final_cost, "rmse_px": syn.reprojection_rmse(s, x)}
    out["mixed_over_fp64"] = out["mixed"]["iters_per_s"] / out["fp64"]["iters_per_s"]
    out["mixed_rmse_delta_vs_fp64_px"] = abs(out["mixed"]["rmse_px"] - out["fp64"]["rmse_px"])
    out["mixed_final_cost_rel_err_vs_fp64"] = (abs(out["mixed"]["final_cost"] - out["fp64"]["final_cost"]) /
                                               out["fp64"]["final_cost"])
    return out


# ------------------------------------------------------------------ configs[0]
def run_ring(args, K=20):
    """10 cameras on a 10 m ring, 500 points, ~4.5k keys, FP64 (camera block
    6: fixed shared intrinsics), 20 iterations; CPU baseline = the oracle with
    the reference's literal O(I^2 P) Schur loops."""
    scene = syn.ring_scene()
    p = Problem(scene, precision=abi.PRECISION_FP64)
    p.lm(max_iterations=K)  # warm-up: captures the iteration graph
    abi.check(p.lm_begin(max_iterations=K), "lm_begin")
    t0 = time.perf_counter()
    abi.check(p.lm_iterate(K)[0], "lm_iterate")
    t1 = time.perf_counter()
    rc, x, st = p.lm_end()
    p.close()
    from oracle.oracle import Oracle  # checker / CPU baseline
    o = Oracle(scene)
    c0 = time.perf_counter()
    _, xo, sto, _ = o.lm(max_iterations=K, literal=True)
    c1 = time.perf_counter()
    rm, rmo = syn.reprojection_rmse(scene, x), syn.reprojection_rmse(scene, xo)
    return {"metric": "BA LM iters/s (ring, 10 cameras, FP64)", "value": K / (t1 - t0), "unit": "iters/s",
            "steps": K, "ms_per_step": 1e3 * (t1 - t0) / K, "dtype": "fp64",
            "config": {"workload": f"ring: 10 cameras, 500 points, {scene.num_keys} keys, FP64"},
            "rmse_px": rm, "fp64_oracle_rmse_px": rmo, "rmse_delta_px": abs(rm - rmo),
            "final_cost_rel_err": abs(st.final_cost - sto.final_cost) / sto.final_cost,
            "param_deviation_vs_fp64_oracle": param_deviation(scene, x, xo),
            "cpu_baseline": {"value": K / (c1 - c0), "unit": "iters/s", "cores": 1, "kind": "port",
                             "sample": f"{K} LM iterations of the same scene with the reference's literal "
                                       f"O(I^2 P) Schur loops, cpu={cpu_model()}"}}


# ------------------------------------------------------------------ configs[3]
def window_flops(scenes, E=6):
    pf = sum(pair_flops(s, E) for s in scenes)
    cf_ = sum(chol_flops(6 * s.num_images) for s in scenes)
    return pf, cf_


def run_windows(args):
    """512 independent windows (5..40 cameras, 200..5000 points, 6-parameter
    cameras with fixed shared intrinsics): one batch (one set of launches per
    LM iteration for every window) vs the same windows solved one at a time
    through the single-problem path.  Both legs include the host-to-device
    upload of x0 and the observations and the download of x (PCIe-inclusive);
    the batch's graph capture and each single problem's are done in an
    untimed warm-up solve."""
    scenes = syn.windows_batch(count=args.windows)
    K = args.steps
    n = len(scenes)
    b = Batch(scenes, precision=abi.PRECISION_MIXED)
    b.lm(max_iterations=K)  # warm-up
    # median of three timed solves (each one ~0.1 s including its PCIe
    # transfers, so a single sample moves by a few percent with host load)
    samples = []
    for _ in range(3):
        t0 = time.perf_counter()
        rc, xs, sts = b.lm(max_iterations=K)
        t1 = time.perf_counter()
        abi.check(rc, "batch lm")
        samples.append(n / (t1 - t0))
    batched = statistics.median(samples)
    # per-stage device time of the batch (HIP events between the stage graphs)
    abi.check(b.lm_begin(max_iterations=K), "batch lm_begin")
    rc, stages = b.lm_iterate(K, timed=True)
    abi.check(rc, "batch lm_iterate(timed)")
    b.lm_end()
    b.close()
    keys = sum(s.num_keys for s in scenes)
    if args.batch_only:  # kernel profiles of the batched launch alone
        return {"metric": "batched BA solves/s (512 windows)", "value": batched, "unit": "solves/s", "steps": K,
                "config": {"workload": f"windows: {n} windows, {keys} keys, batch only"}, "stage_ms": stages}
    # one at a time: one live problem (its own stream and graph), created,
    # warmed up, solved (timed) and destroyed in turn
    single_s = 0.0
    for s in scenes:
        p = Problem(s, precision=abi.PRECISION_MIXED)
        p.lm(max_iterations=K)
        t = time.perf_counter()
        rc1, _, _ = p.lm(max_iterations=K)
        single_s += time.perf_counter() - t
        abi.check(rc1, "single lm")
        p.close()
    one_at_a_time = n / single_s
    # accuracy: every window against the FP64 oracle
    xos = oracle_x_all(scenes, K)
    deltas = [abs(syn.reprojection_rmse(s, x) - syn.reprojection_rmse(s, xo)) for s, x, xo in zip(scenes, xs, xos)]
    devs = [param_deviation(s, x, xo) for s, x, xo in zip(scenes, xs, xos)]
    worst_dev = {k: max(dv[k] for dv in devs if k in dv) for k in devs[0]}
    # CPU baseline: the oracle solving a sample of windows on one core
    from oracle.oracle import Oracle  # CPU baseline leg only
    sample = scenes[:args.cpu_windows]
    c0 = time.perf_counter()
    for s in sample:
        Oracle(s).lm(max_iterations=K)
    c1 = time.perf_counter()
    pf, cfl = window_flops(scenes)
    lin = sum(linearize_bytes(s, 6, 4) for s in scenes)
    kernels = {
        "k_pairs": roofline_entry("mfma", pf / (stages["pairs_ms"] * 1e-3) / 1e12, PEAK_FP32_TFLOPS, "TFLOP/s", pf),
        "cholesky": roofline_entry("mfma", cfl / (stages["cholesky_ms"] * 1e-3) / 1e12, PEAK_FP64_TFLOPS,
                                   "TFLOP/s", cfl),
        "k_linearize": roofline_entry("hbm", lin / (stages["linearize_ms"] * 1e-3) / 1e9, PEAK_HBM_GBS, "GB/s", lin),
    }
    dominant = max(("pairs_ms", "cholesky_ms", "linearize_ms", "image_reduce_ms", "point_factor_ms",
                    "assemble_ms", "backsub_ms", "step_cost_ms"), key=lambda k: stages[k])
    dom_kernel = {"pairs_ms": "k_pairs", "cholesky_ms": "cholesky", "linearize_ms": "k_linearize"}.get(dominant,
                                                                                                       "k_pairs")
    names = {"k_pairs": "camera-pair Schur products, E = 6, FP32 (k_pairs_mma or k_pairs per window class, "
                        "as the autotuner chose)",
             "cholesky": "k_chol_small (one workgroup per window: FP64 factor + both solves)",
             "k_linearize": "k_linearize (residual + Jacobian + point blocks)"}
    return {"metric": "batched BA solves/s (512 windows)", "value": batched, "unit": "solves/s", "steps": K,
            "ms_per_batch_solve": 1e3 * n / batched, "batched_samples": samples, "dtype": "fp32+fp64",
            "config": {"workload": f"windows: {n} windows, 5..40 cameras x 6, 200..5000 points, "
                                   f"{keys} keys, {K} LM iterations each"},
            "one_at_a_time_solves_per_s": one_at_a_time, "speedup_vs_one_at_a_time": batched / one_at_a_time,
            "rmse_worst_delta_vs_fp64_oracle_px": max(deltas), "oracle_checked_windows": len(deltas),
            "param_deviation_worst_vs_fp64_oracle": worst_dev,
            "statuses": sorted(set(int(s.status) for s in sts)),
            "stage_ms": stages, "dominant_stage": dominant, "autotuner": abi.tuning_table(),
            "roofline": dict(kernels[dom_kernel], kernel=names[dom_kernel]), "kernels": kernels,
            "cpu_baseline": {"value": len(sample) / (c1 - c0), "unit": "solves/s", "cores": 1, "kind": "port",
                             "sample": f"the FP64 oracle solving the first {len(sample)} windows ({K} LM "
                                       f"iterations each, point-wise Schur + dense LDL^T), one thread, "
                                       f"cpu={cpu_model()}"}}


# ------------------------------------------------------------------ configs[2]
def pcg_bytes(scene: syn.Scene, E: int = 9) -> float:
    """Algorithmic HBM bytes of one PCG iteration of the implicit Schur
    complement: both matvec passes read every key's Z block (3 E FP32 values,
    point-major and image-major copies) and a 4-byte index, and the point-space
    intermediate (3 doubles per point) is written once and read once."""
    return float(scene.num_keys * (2 * 3 * E * 4 + 8) + scene.num_points * 48)


def run_large(args, K=30):
    """Large banded strip: 2000 cameras x 9, 600k points, ~5.1M keys; block-Jacobi
    PCG on the implicit Schur complement (FP32 matvec, FP64 vectors), rtol 1e-6."""
    scene = syn.large_scene()
    p = Problem(scene, precision=abi.PRECISION_MIXED, solver=abi.SOLVER_PCG, pcg_rtol=1e-6)
    abi.check(p.lm_begin(max_iterations=K), "lm_begin")
    t0 = time.perf_counter()
    rc, _ = p.lm_iterate(K)
    t1 = time.perf_counter()
    abi.check(rc, "lm_iterate")
    rc, x, st = p.lm_end()
    abi.check(p.lm_begin(max_iterations=K), "lm_begin")
    rc, stages = p.lm_iterate(K, timed=True)
    rc2, _, st2 = p.lm_end()
    p.close()
    # PCG loop time per PCG iteration (the host polls the convergence flags
    # every 8 iterations, so this includes those round trips)
    pcg_it = max(1, st2.pcg_iterations)
    per_it_s = stages["cholesky_ms"] * 1e-3 * K / pcg_it
    b = pcg_bytes(scene)
    return {"metric": "BA LM iters/s (5.1M keys, PCG)", "value": K / (t1 - t0), "unit": "iters/s", "steps": K,
            "ms_per_step": 1e3 * (t1 - t0) / K, "dtype": "fp32+fp64",
            "config": {"workload": "large: 2000 cameras x 9, 600k points, banded, PCG rtol 1e-6",
                       "keys": scene.num_keys, "reduced_size": 9 * scene.num_images},
            "stage_ms": stages,
            "roofline": roofline_entry("hbm", b / per_it_s / 1e9, PEAK_HBM_GBS, "GB/s", b,
                                       kernel="one PCG iteration (k_pcg_points + k_pcg_images + dot / vector "
                                              "kernels), loop time per iteration incl. host polls"),
            "lm": {"status": st.status, "iterations": st.iterations, "accepted": st.accepted,
                   "pcg_iterations": st.pcg_iterations, "initial_cost": st.initial_cost,
                   "final_cost": st.final_cost, "rmse_px": syn.reprojection_rmse(scene, x),
                   "rmse_x0_px": syn.reprojection_rmse(scene, scene.x0),
                   "rmse_true_px": syn.reprojection_rmse(scene, scene.x_true)},
            "cpu_baseline": None,
            "cpu_baseline_note": "not timed: the oracle's dense LDL^T of the 18000 x 18000 reduced system "
                                 "(~2e12 flop per iteration) exceeds the bounded CPU sample"}


# ------------------------------------------------------------------ configs[4]
def run_serving(args):
    """Serving loop: Poisson arrivals at each rate in --rates; 95% of jobs from
    the windows distribution, 5% of the 200-camera x 9, ~1M-key class (the
    mid scene itself).  Jobs are drawn with replacement from a pool of
    pre-built problems; every served result is compared bit for bit with the
    same problem solved alone, and every pool problem's alone result with the
    FP64 oracle (the 1M-key one with its committed 20-iteration oracle run),
    so every served job's RMSE is checked."""
    from hs_sfm_amd.hsba import Server
    rng = np.random.default_rng(syn.SEED_SERVING)
    cams, pts = syn.window_sizes(rng, args.pool)
    pool = [syn.window_scene(rng, int(c), int(p), name=f"pool{w}") for w, (c, p) in enumerate(zip(cams, pts))]
    mid = syn.mid_scene()
    K = args.steps
    srv = Server(streams=args.streams, max_batch_jobs=args.max_batch, max_iterations=K,
                 hbm_budget_bytes=args.hbm_budget_gib << 30)
    alone = []
    for s in pool + [mid]:
        pr = Problem(s, precision=abi.PRECISION_MIXED)
        rc, x, st = pr.lm(max_iterations=K)
        abi.check(rc, "alone")
        alone.append(x)
        pr.close()
    prepared = [srv.prepare(s) for s in pool + [mid]]
    rows = []
    mismatches = 0
    for lam in args.rates:
        n_jobs = int(min(args.max_jobs, max(50, lam * args.seconds)))
        gaps = rng.exponential(1.0 / lam, n_jobs)
        kinds = np.where(rng.random(n_jobs) < 0.05, len(pool), rng.integers(0, len(pool), n_jobs))
        ids, which = [], []
        due = srv.clock()
        for g, k in zip(gaps, kinds):
            due += g
            while srv.clock() < due:
                pass
            d, x0, y, _ = prepared[k]
            ids.append(srv.submit((d, x0, y, np.zeros_like(alone[k]))))
            which.append(int(k))
        srv.drain()
        res = {r.job_id: r for r in srv.poll(len(ids) + 16)}
        lat, lat_w, lat_m, bad = [], [], [], 0
        for jid, k in zip(ids, which):
            r = res[jid]
            if r.status != 0:
                bad += 1
                print(f"job {jid} (pool {k}) failed: {r.status}", file=sys.stderr)
            mismatches += int(not np.array_equal(srv.x_out(jid), alone[k]))
            ms = 1e3 * (r.finish_time_s - r.submit_time_s)
            lat.append(ms)
            (lat_m if k == len(pool) else lat_w).append(ms)
        first = min(res[j].submit_time_s for j in ids)
        last = max(res[j].finish_time_s for j in ids)
        srv._jobs.clear()
        rows.append({"rate_per_s": lam, "jobs": len(ids), "mid_jobs": len(lat_m),
                     "solves_per_s": len(ids) / (last - first),
                     "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                     "window_p50_ms": float(np.percentile(lat_w, 50)) if lat_w else None,
                     "window_p99_ms": float(np.percentile(lat_w, 99)) if lat_w else None,
                     "mid_p50_ms": float(np.percentile(lat_m, 50)) if lat_m else None,
                     "failed": bad})
        print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
    srv.close()
    ro = oracle_rmse_all(pool, K)
    worst = max(abs(syn.reprojection_rmse(s, a) - r) for s, a, r in zip(pool, alone, ro))
    ref = oracle_reference("mid")
    mid_delta = abs(syn.reprojection_rmse(mid, alone[-1]) - ref["rmse_px"]) if ref and ref["iterations"] == K else None
    best = max(rows, key=lambda r: r["solves_per_s"])
    return {"metric": "serving solves/s at Poisson arrivals (95% windows, 5% 1M-key), p50/p99 ms",
            "value": best["solves_per_s"], "unit": "solves/s", "steps": K, "dtype": "fp32+fp64",
            "config": {"workload": f"serving: pool of {len(pool)} windows + the 1M-key scene, {K} LM iterations "
                                   f"per job, {args.streams} streams, batches of <= {args.max_batch}"},
            "rates": rows, "results_identical_to_alone": mismatches == 0, "mismatches": mismatches,
            "rmse_worst_delta_vs_fp64_oracle_px": worst, "oracle_checked_pool_problems": len(pool),
            "mid_rmse_delta_vs_fp64_oracle_px": mid_delta}


# ------------------------------------------------------------------ covariance
def run_covariance(args, reps=3):
    """hsba_x_covariance: the reference's whole dense x covariance (pseudo-
    inverse of N with singular values <= 1e-6 dropped), FP64, on a ring of 30
    cameras (6 parameters) and 500 points.  fixed: three points constrained in
    all coordinates (gauge fixed, Schur route); free: no constraints (seven
    null directions, dense Jacobi route).  The timed call includes the normal
    equations at x, the device work and the copy of x_size^2 doubles to the
    host.  Each result is checked against the oracle's SVD pseudo-inverse; the
    CPU baseline is that oracle (dense N + LAPACK SVD, one BLAS thread)."""
    from oracle.covariance import dense_normal, pseudo_inverse  # checker / CPU baseline
    from oracle.oracle import Oracle

    def scene(fixed):
        s = syn.ring_scene(num_cameras=30, num_points=500)
        if fixed:
            ids = [0, 1, 2]
            s.point_constraints = np.array([[i, 7] for i in ids], np.int64)
            s.y_obs = np.concatenate([s.y_obs[:2 * s.num_keys], s.x_true[:9]])
            s.constraint_weights = np.full(9, 100.0)
        return s

    out = {}
    for name, fixed in (("fixed", True), ("free", False)):
        s = scene(fixed)
        p = Problem(s, precision=abi.PRECISION_FP64)
        rc, G, info = p.x_covariance()  # warm-up (and the checked result)
        abi.check(rc, "hsba_x_covariance")
        ts = []
        for _ in range(reps if fixed else 1):
            t0 = time.perf_counter()
            abi.check(p.x_covariance()[0], "hsba_x_covariance")
            ts.append(time.perf_counter() - t0)
        p.close()
        o = Oracle(s)
        N = dense_normal(o, s.x0)
        try:
            from threadpoolctl import threadpool_limits
            limit = threadpool_limits(1)
        except ImportError:
            limit = None
        c0 = time.perf_counter()
        C = pseudo_inverse(N)
        c1 = time.perf_counter()
        if limit is not None:
            limit.restore_original_limits()
        err = float(np.abs(G - C).max() / np.abs(C).max())
        t = float(np.median(ts))
        out[name] = {"x_size": int(G.shape[0]), "route": {1: "schur", 2: "dense"}[info["method"]],
                     "kept": info["rank"], "jacobi_sweeps": info["sweeps"], "ms": 1e3 * t,
                     "rel_err_vs_fp64_oracle": err,
                     "cpu_baseline": {"value": 1.0 / (c1 - c0), "unit": "covariances/s", "cores": 1,
                                      "kind": "port",
                                      "sample": f"the oracle's LAPACK SVD pseudo-inverse of the same dense N "
                                                f"({G.shape[0]}^2), one BLAS thread, cpu={cpu_model()}"}}
    return {"metric": "full dense x covariances/s (FP64)", "value": 1e3 / out["fixed"]["ms"],
            "unit": "covariances/s", "dtype": "fp64",
            "config": {"workload": "covariance: ring of 30 cameras x 6, 500 points, x_size "
                                   f"{out['fixed']['x_size']}; gauge fixed (Schur) and free (dense Jacobi)"},
            "fixed": out["fixed"], "free": out["free"]}


def standalone(out, args):
    """A sub-workload result printed as its own line (--workload X)."""
    base = {"n_gpus": 1, "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "data": "synthetic"}
    base.update(out)
    return base


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="default",
                    choices=["default", "mid", "large", "windows", "ring", "serving", "covariance"])
    ap.add_argument("--extras", default="ring,windows,large,serving,covariance",
                    help="workloads carried in the default line (comma list, or 'none')")
    ap.add_argument("--rates", type=float, nargs="+", default=[100, 500, 2000, 5000])
    ap.add_argument("--seconds", type=float, default=1.0, help="arrival window per rate (serving)")
    ap.add_argument("--max-jobs", type=int, default=1500, help="job cap per rate (serving)")
    ap.add_argument("--pool", type=int, default=128, help="distinct window problems (serving)")
    ap.add_argument("--streams", type=int, default=4)
    ap.add_argument("--max-batch", type=int, default=64)
    ap.add_argument("--hbm-budget-gib", type=int, default=64)
    ap.add_argument("--windows", type=int, default=512)
    ap.add_argument("--cpu-windows", type=int, default=16)
    ap.add_argument("--batch-only", action="store_true",
                    help="windows: only the batched solve (no one-at-a-time leg, no oracle checks)")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-samples", type=int, default=3)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="MiB copied device-to-device before the run (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build measures one GPU (every GPU box here has one); multi-GPU "
                                   "sharding is design-only"}))
        sys.exit(2)
    if args.copy_calibration > 0:
        import torch  # device buffers for the calibration copy only
        src = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.workload != "default":
        fn = {"mid": run_mid, "large": run_large, "windows": run_windows, "ring": run_ring,
              "serving": run_serving, "covariance": run_covariance}[args.workload]
        out = fn(args)
        print(json.dumps(out if args.workload == "mid" else standalone(out, args)))
        return
    out = run_mid(args)
    extras = [] if args.extras == "none" else [e for e in args.extras.split(",") if e]
    out["workloads"] = {}
    for name in extras:
        fn = {"ring": run_ring, "windows": run_windows, "large": run_large, "serving": run_serving,
              "covariance": run_covariance}[name]
        t = time.perf_counter()
        print(f"bench: {name} ...", file=sys.stderr, flush=True)
        try:
            r = fn(args)
        except Exception as e:  # a failing extra leaves the headline line intact
            traceback.print_exc()
            r = {"error": f"{type(e).__name__}: {e}"}
        r["wall_s"] = time.perf_counter() - t
        out["workloads"][name] = r
        print(f"bench: {name} done in {r['wall_s']:.1f} s", file=sys.stderr, flush=True)
    w = out["workloads"].get("windows", {})
    if "value" in w:
        out["batched_solves_per_s"] = w["value"]
        out["one_at_a_time_solves_per_s"] = w["one_at_a_time_solves_per_s"]
        out["batched_speedup"] = w["speedup_vs_one_at_a_time"]
    s = out["workloads"].get("serving", {})
    if "rates" in s:
        out["serving_p99_ms"] = {str(r["rate_per_s"]): r["p99_ms"] for r in s["rates"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
