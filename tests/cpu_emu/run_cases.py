"""Runs the library's own kernel sources on the CPU HIP emulation (built by
build.sh, AddressSanitizer on) through the C ABI, against the FP64 oracle.

Started by tests/test_cpu_emulation.py in a child process with the ASan
runtime preloaded and HSBA_LIB pointing at libhsba_cpu.so.  The cases are
small but take every code path the GPU parity tests take on large inputs:
the tiled Cholesky (more than four 64-row tiles) with its forward and
backward solves, the one-workgroup factorization, the matrix-core pair
kernel, batched windows, the allocation-free covariance.  Prints one line
per case and exits nonzero on the first mismatch.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from hs_sfm_amd import abi, synthetic as syn  # noqa: E402
from hs_sfm_amd.hsba import Batch, Problem  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def sparse_ring(num_cameras, num_points, seed=5, kmax=4):
    """Cameras on a ring (fixed intrinsics, 6 parameters each), each point
    seen by 2..kmax consecutive cameras: a reduced system of 6 C rows with
    few keys, so the emulation stays fast."""
    rng = np.random.default_rng(seed)
    C = num_cameras
    ang = 2 * np.pi * np.arange(C) / C
    centers = np.stack([10 * np.cos(ang), 10 * np.sin(ang), 2.0 + 0.2 * np.sin(3 * ang)], 1)
    rot, trans = syn.look_at(centers, np.zeros(3))
    pts = rng.uniform([-2, -2, -1], [2, 2, 1], (num_points, 3))
    fm = []
    for p in range(num_points):
        c0, k = int(rng.integers(C)), int(rng.integers(2, kmax + 1))
        for c in sorted({(c0 + j) % C for j in range(k)}):
            fm.append((c, p))
    # every camera observed
    seen = {c for c, _ in fm}
    for c in range(C):
        if c not in seen:
            fm.append((c, 0))
    fm = np.asarray(fm, np.int32)[rng.permutation(len(fm))]
    intr = np.array([[1000.0, 0.0, 500.0, 500.0, 1.0]])
    s = syn.Scene(num_points=num_points, num_images=C, num_cameras=1, feature_maps=fm,
                  image_camera_map=np.zeros(C, np.int32), comp_mask=syn.COMPUTE_INTRINSIC, free_mask=0,
                  intrinsic_defaults=None, fix_mask=syn.FIX_CAMERAS, fix_cameras=intr.copy(), name="sparse_ring")
    x_true = np.concatenate([pts.reshape(-1), np.concatenate([rot, trans], 1).reshape(-1)])
    return syn._finish(s, rng, x_true)


def case_tiled_cholesky():
    """n = 6 x 55 = 330 reduced rows (six 64-row tiles, 10 real rows in the
    last: its factor takes the one-block path; HSBA_EMU_TILED="50,55" adds
    n = 300, five tiles with 44 real rows in the last): FP64 step vs the
    oracle and the device's own S solved in numpy."""
    out = []
    for cams in [int(c) for c in os.environ.get("HSBA_EMU_TILED", "55").split(",")]:
        n = 6 * cams
        s = sparse_ring(cams, 60)
        o, p = Oracle(s), Problem(s, precision=abi.PRECISION_FP64)
        assert p.n_reduced == n
        o.build_normal(s.x0)
        p.build_normal(s.x0)
        mu = 1e-3 * o.max_diagonal()
        rc, S, rhs = p.reduced_system(mu)
        assert rc == 0
        rc_o, So, ro = o.reduced_system(mu)
        assert rel(S, So) < 1e-10 and rel(rhs, ro) < 1e-10
        rc, dg = p.solve(mu)
        assert rc == 0
        rc_o, do = o.solve(mu)
        assert rel(dg, do) < 1e-8, rel(dg, do)
        dr = dg[-n:]
        back = np.abs(S @ dr - rhs).max() / (np.abs(S).max() * np.abs(dr).max())
        assert back < 1e-12, back
        out.append(f"n={n}: step rel err {rel(dg, do):.1e}, backward err {back:.1e}")
    return "tiled factor " + "; ".join(out)


def case_small_mma():
    """E = 9: each of the three pair kernels (forced) against the oracle, then
    the one-workgroup factorization, mixed."""
    s = syn.tiny_scene(comp_mask=syn.COMPUTE_RADIAL | syn.COMPUTE_INTRINSIC,
                       free_mask=(1 << syn.K1) | (1 << syn.K2) | (1 << syn.FOCAL), constraints=True,
                       random_cov=True, num_images=6, num_points=20, num_cameras=2)
    o = Oracle(s)
    o.build_normal(s.x0)
    mu = 1e-3 * o.max_diagonal()
    rc_o, So, ro = o.reduced_system(mu)
    assert rc_o == 0
    for kernel in (abi.PAIRS_VALU, abi.PAIRS_MFMA_WAVE, abi.PAIRS_MFMA):
        p = Problem(s, precision=abi.PRECISION_MIXED, pair_kernel=kernel)
        p.build_normal(s.x0)
        rc, S, rhs = p.reduced_system(mu)
        assert rc == 0
        assert rel(S, So) < 2e-5 and rel(rhs, ro) < 2e-5, (kernel, rel(S, So))
    rc, dg = p.solve(mu)
    assert rc == 0
    back = np.abs(S @ dg[-p.n_reduced:] - rhs).max() / (np.abs(S).max() * np.abs(dg[-p.n_reduced:]).max())
    assert back < 1e-12, back
    return f"E=9 pair kernel + small factor: S rel err {rel(S, So):.1e}, backward err {back:.1e}"


def case_lm_tiled():
    """Three FP64 LM iterations through the graph on the tiled path."""
    s = sparse_ring(50, 60, seed=6)
    o, p = Oracle(s), Problem(s, precision=abi.PRECISION_FP64)
    rc, x, st = p.lm(max_iterations=3)
    rc_o, xo, sto, _ = o.lm(max_iterations=3)
    assert rc == 0 and rc_o == 0
    err = abs(st.final_cost - sto.final_cost) / sto.final_cost
    assert err < 1e-8, err
    return f"LM x3 tiled FP64: final cost rel err {err:.1e}"


def case_batch():
    rng = np.random.default_rng(3)
    scenes = [syn.window_scene(rng, 6, 20, name="w0"), syn.window_scene(rng, 8, 25, name="w1"),
              syn.window_scene(rng, 5, 16, name="w2")]
    rc, xs, sts = Batch(scenes).lm(max_iterations=2)
    assert rc == 0
    for s, x in zip(scenes, xs):
        p = Problem(s)
        rc1, xa, _ = p.lm(max_iterations=2)
        assert rc1 == 0 and np.array_equal(x, xa)
    return f"batch of {len(scenes)} windows bit-identical to the windows alone"


def case_batch_xcd():
    """Sixteen windows in one batch, bit-identical to each window alone."""
    rng = np.random.default_rng(4)
    scenes = [syn.window_scene(rng, 5 + (j % 3), 12 + j, name=f"x{j}") for j in range(16)]
    rc, xs, _ = Batch(scenes).lm(max_iterations=1)
    assert rc == 0
    for s, x in zip(scenes, xs):
        rc1, xa, _ = Problem(s).lm(max_iterations=1)
        assert rc1 == 0 and np.array_equal(x, xa)
    return f"batch of {len(scenes)} windows bit-identical to the windows alone"


def case_covariance_static():
    s = sparse_ring(12, 30, seed=7)
    ids = [0, 1, 2]
    s.point_constraints = np.array([[i, 7] for i in ids], np.int64)
    s.y_obs = np.concatenate([s.y_obs[:2 * s.num_keys], s.x_true[:9]])
    s.constraint_weights = np.full(9, 100.0)
    p = Problem(s, precision=abi.PRECISION_FP64)
    rc_a, a = p.covariance()
    rc_s, b = p.covariance(static=True)
    assert rc_a == 0 and rc_s == 0
    assert all(np.array_equal(a[k], b[k]) for k in a)
    return "covariance_static bit-identical to covariance"


def case_x_covariance():
    """Full dense covariance vs the oracle's SVD pseudo-inverse on a ring
    with a free gauge (dense Jacobi route; the singular values the device
    drops are those numpy's SVD finds at or below 1e-6), then with the gauge
    fixed by three constrained points (Schur route)."""
    from oracle.covariance import dense_normal, x_covariance
    s = sparse_ring(6, 16, seed=3, kmax=4)
    o = Oracle(s)
    sv = np.linalg.svd(dense_normal(o, s.x0), compute_uv=False)
    free = Problem(s, precision=abi.PRECISION_FP64)
    rc, Gf, info_f = free.x_covariance()
    assert rc == 0 and info_f["method"] == abi.COV_DENSE, (rc, info_f)
    assert info_f["rank"] == int((sv > 1e-6).sum()) < free.x_size, (info_f, sv.min())
    e_free = rel(Gf, x_covariance(o, s.x0))
    assert e_free < 1e-8, e_free
    ids = [0, 1, 2]
    s.point_constraints = np.array([[i, 7] for i in ids], np.int64)
    s.y_obs = np.concatenate([s.y_obs[:2 * s.num_keys], s.x_true[:9]])
    s.constraint_weights = np.full(9, 100.0)
    p = Problem(s, precision=abi.PRECISION_FP64)
    C = x_covariance(Oracle(s), s.x0)
    rc, Ga, info_a = p.x_covariance()
    assert rc == 0 and info_a["method"] == abi.COV_SCHUR, (rc, info_a)
    e_s = rel(Ga, C)
    assert e_s < 1e-8, e_s
    return (f"x_size {p.x_size}: free gauge (dense, rank {info_f['rank']}, {info_f['sweeps']} sweeps) "
            f"{e_free:.1e}; fixed gauge (Schur) {e_s:.1e}")


def case_fp64_wide(num_cameras=None, num_points=None, iterations=None):
    """The FP64 LM on 9-parameter cameras (camera block E = 9, free points):
    the path switched off on the GPU.  Default: a 40-camera, 2000-point
    scene of the 1M-key scene's kind for two iterations; HSBA_EMU_WIDE =
    "cameras,points,iterations" sets the size (200,50000,1 is the full
    1M-key scene).  Every device index is bounds-checked by AddressSanitizer;
    the final cost is compared with the FP64 oracle."""
    spec = os.environ.get("HSBA_EMU_WIDE", "40,2000,2").split(",")
    C = num_cameras or int(spec[0])
    P = num_points or int(spec[1])
    K = iterations or int(spec[2])
    s = syn.mid_scene(num_cameras=C, num_points=P, kmax=min(40, C))
    p = Problem(s, precision=abi.PRECISION_FP64)
    rc, x, st = p.lm(max_iterations=K)
    assert rc == 0, rc
    rc_o, xo, sto, _ = Oracle(s).lm(max_iterations=K)
    assert rc_o == 0
    err = abs(st.final_cost - sto.final_cost) / sto.final_cost
    assert err < 1e-9, err
    return (f"FP64 E=9 LM x{K}: {s.num_keys} keys, n = {p.n_reduced}, final cost rel err {err:.1e}, "
            f"rmse {syn.reprojection_rmse(s, x):.6f} px")


CASES = {"tiled_cholesky": case_tiled_cholesky, "small_mma": case_small_mma, "lm_tiled": case_lm_tiled,
         "fp64_wide": case_fp64_wide,
         "batch": case_batch, "batch_xcd": case_batch_xcd, "covariance_static": case_covariance_static,
         "x_covariance": case_x_covariance}

if __name__ == "__main__":
    names = sys.argv[1:] or list(CASES)
    for n in names:
        print(f"{n}: {CASES[n]()}", flush=True)
    print("EMU_OK", flush=True)
