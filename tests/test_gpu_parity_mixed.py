"""Mixed-precision parity of the wide-camera-block Schur path against the
FP64 oracle (camera_shared_normal_equation_solver.hpp:116-538).

The 9-parameter camera of the 1M-key workload gives a camera block
E = 6 + 3 = 9, which runs the matrix-core pair kernel (7 <= E <= 15); the
reduced system of 40 such cameras (n = 360) runs the tiled Cholesky (k_chol_first
/ k_chol_step, forward solve riding along, k_chol_back).  Here S, rhs and the
step are compared element-wise with the oracle for E = 9, 10 (fixed images),
14 and 16 (the VALU pair kernel), and the FP64 factorization and both
triangular solves are checked on their own against numpy on the device's own S.

Tolerances (mixed precision stores W, Z, V, D and the pair blocks in FP32):
  S, rhs      2e-5 of the largest entry (FP32 storage, FP64 assembly)
  step        (N + mu I) delta = g residual within 1e-4 of |g|, and the
              step within 5e-3 of the oracle's (relative to its largest entry;
              the FP32 perturbation of S is amplified by its conditioning)
  solves      S delta_r = rhs on the device's own S: 1e-12 backward error
"""
import json
import os

import numpy as np
import pytest

from hs_sfm_amd import abi, synthetic as syn
from hs_sfm_amd.hsba import Problem
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ALL = syn.COMPUTE_RADIAL | syn.COMPUTE_DECENTERING | syn.COMPUTE_INTRINSIC
RK_F = (1 << syn.K1) | (1 << syn.K2) | (1 << syn.FOCAL)
EIGHT = 0x3FF & ~((1 << syn.SKEW) | (1 << syn.PR))


def scene(kind, fix=0):
    if kind == "e9":
        return syn.tiny_scene(fix_mask=fix, comp_mask=syn.COMPUTE_RADIAL | syn.COMPUTE_INTRINSIC,
                              free_mask=RK_F, constraints=True, random_cov=True, num_images=8,
                              num_points=40, num_cameras=3)
    if kind == "e10":  # images fixed, every intrinsic free
        return syn.tiny_scene(fix_mask=syn.FIX_IMAGES, comp_mask=ALL, constraints=True, random_cov=True,
                              num_images=8, num_points=40, num_cameras=3)
    if kind == "e14":
        return syn.tiny_scene(fix_mask=fix, comp_mask=ALL, free_mask=EIGHT, constraints=True,
                              random_cov=True, num_images=8, num_points=40, num_cameras=3)
    if kind == "e16":
        return syn.tiny_scene(fix_mask=fix, comp_mask=ALL, constraints=True, random_cov=True,
                              num_images=8, num_points=40, num_cameras=3)
    if kind == "mid40":  # n = 360: tiled Cholesky
        return syn.mid_scene(num_cameras=40, num_points=2500, name="mid40")
    raise ValueError(kind)


def e_width(s):
    return (0 if s.fix_mask & syn.FIX_IMAGES else 6) + (0 if s.fix_mask & syn.FIX_CAMERAS else s.camera_params)


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("kind", ["e9", "e10", "e14", "e16", "mid40"])
def test_reduced_system_and_step_mixed(gpu, kind):
    s = scene(kind)
    o, p = Oracle(s), Problem(s, precision=abi.PRECISION_MIXED)
    print(f"mixed reduced system {kind}: E={e_width(s)} n={p.n_reduced} keys={s.num_keys}", flush=True)
    if kind != "e16":
        assert 7 <= e_width(s) <= 15  # the matrix-core pair kernel
    o.build_normal(s.x0)
    p.build_normal(s.x0)
    mu = 1e-3 * o.max_diagonal()
    rc_o, So, ro = o.reduced_system(mu)
    rc_g, Sg, rg = p.reduced_system(mu)
    assert rc_o == 0 and rc_g == 0
    assert rel(Sg, So) < 2e-5
    assert rel(rg, ro) < 2e-5
    rc_o, do = o.solve(mu)
    rc_g, dg = p.solve(mu)
    assert rc_o == 0 and rc_g == 0
    _, N, g = o.dense_normal(s.x0) if s.num_keys < 5000 else (None, None, None)
    if N is not None:
        res = (N + mu * np.eye(o.x_size)) @ dg - g
        assert np.abs(res).max() < 1e-4 * np.abs(g).max()
    assert rel(dg, do) < 5e-3


@pytest.mark.parametrize("kind", ["e9", "mid40"])
def test_factor_and_solves_on_device_system(gpu, kind):
    """K5 alone: the device's FP64 Cholesky and its two triangular solves
    applied to the device's own reduced system, against numpy.  mid40 runs
    the tiled path (7 tile rows: k_chol_first, six k_chol_step launches with
    the forward solve, two k_chol_back launches); e9 the one-workgroup path."""
    s = scene(kind)
    p = Problem(s, precision=abi.PRECISION_MIXED)
    print(f"device factor/solves {kind}: n={p.n_reduced}", flush=True)
    p.build_normal(s.x0)
    mu = 1e-3 * p.max_diagonal()
    rc, S, rhs = p.reduced_system(mu)
    assert rc == 0
    rc, dg = p.solve(mu)
    assert rc == 0
    n = p.n_reduced
    dr = dg[-n:]
    ref = np.linalg.solve(S, rhs)
    back = np.abs(S @ dr - rhs).max() / (np.abs(S).max() * np.abs(dr).max())
    assert back < 1e-12, back
    cond = np.linalg.cond(S)
    assert rel(dr, ref) < max(1e-12, 1e-15 * cond * 10), (rel(dr, ref), cond)


def test_lm_mid_scene_against_fp64_oracle(gpu):
    """configs[1] end to end: the 200-camera x 9-parameter, ~1M-key scene,
    20 LM iterations in mixed precision from x0, against the FP64 oracle's
    20-iteration run committed in tests/golden/mid_fp64_oracle.json (final
    RMSE within 1e-3 px, final cost within 1e-6 relative)."""
    with open(os.path.join(ROOT, "tests", "golden", "mid_fp64_oracle.json")) as f:
        ref = json.load(f)
    s = syn.mid_scene()
    assert ref["scene"] == "mid (seed %d)" % syn.SEED_MID
    p = Problem(s, precision=abi.PRECISION_MIXED)
    print(f"lm mid keys={s.num_keys}", flush=True)
    rc, x, st = p.lm(max_iterations=ref["iterations"])
    assert rc == 0 and st.iterations == ref["iterations"]
    assert abs(st.initial_cost - ref["initial_cost"]) < 1e-9 * ref["initial_cost"]
    rmse = syn.reprojection_rmse(s, x)
    assert abs(rmse - ref["rmse_px"]) < 1e-3, (rmse, ref["rmse_px"])
    assert abs(st.final_cost - ref["final_cost"]) < 1e-6 * ref["final_cost"]
    p.close()
