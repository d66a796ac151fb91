"""GPU parity of every C-ABI entry point against the FP64 oracle.

Tolerances: FP64 path within 1e-9 relative of the oracle (different
summation order only); mixed path (FP32 Jacobian / Hessian / Schur storage)
within the stated looser bounds.  Observation indexing and sparsity are
exercised through scenes whose keys arrive in arbitrary order.
"""
import numpy as np
import pytest

from hs_sfm_amd import abi, synthetic as syn
from hs_sfm_amd.hsba import Batch, Problem
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

FIX_MASKS = [0, syn.FIX_POINTS, syn.FIX_IMAGES, syn.FIX_CAMERAS,
             syn.FIX_POINTS | syn.FIX_IMAGES, syn.FIX_POINTS | syn.FIX_CAMERAS,
             syn.FIX_IMAGES | syn.FIX_CAMERAS]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = max(np.abs(b).max(), 1e-300) if b.size else 1.0
    return float(np.abs(a - b).max() / den) if a.size else 0.0


def tiny(fix, **kw):
    return syn.tiny_scene(fix_mask=fix, constraints=True, random_cov=True, **kw)


def fp64_supported(s):
    """FP64 dense Schur with a camera block of width >= 7 is switched off in the
    library (HSBA_ERR_UNSUPPORTED) until the cause of one GPU fault is shown."""
    e = (0 if s.fix_mask & syn.FIX_IMAGES else 6) + (0 if s.fix_mask & syn.FIX_CAMERAS else s.camera_params)
    return bool(s.fix_mask & syn.FIX_POINTS) or e < 7


def fp64_problem(s):
    if not fp64_supported(s):
        pytest.skip("FP64 dense Schur with camera block >= 7 switched off (GPU fault, cause not shown)")
    return Problem(s, precision=abi.PRECISION_FP64)


@pytest.mark.parametrize("fix", FIX_MASKS)
def test_evaluate_and_jacobian(gpu, fix):
    s = tiny(fix)
    o = Oracle(s)
    p = Problem(s, precision=abi.PRECISION_FP64) if fp64_supported(s) else Problem(s)
    assert p.x_size == o.x_size and p.y_size == o.y_size
    print(f"evaluate/jacobian fix={fix}", flush=True)
    assert rel(p.evaluate(s.x0), o.evaluate(s.x0)) < 1e-13
    Jg, Jo = p.jacobian(s.x0), o.jacobian(s.x0)
    # the reference produces only the blocks of free parameters
    # (camera_shared_analytical_jacobian_matrix_calculator.hpp:209-229)
    free = [not fix & syn.FIX_POINTS, not fix & syn.FIX_IMAGES, not fix & syn.FIX_CAMERAS]
    for a, b, f in zip(Jg, Jo, free):
        if f:
            assert rel(a, b) < 1e-11


@pytest.mark.parametrize("fix", FIX_MASKS)
@pytest.mark.parametrize("prec,tol", [(abi.PRECISION_FP64, 1e-10), (abi.PRECISION_MIXED, 2e-5)])
def test_build_normal(gpu, fix, prec, tol):
    s = tiny(fix)
    o = Oracle(s)
    p = fp64_problem(s) if prec == abi.PRECISION_FP64 else Problem(s, precision=prec)
    print(f"build_normal fix={fix} prec={prec}", flush=True)
    bo = o.build_normal(s.x0)
    bg = p.build_normal(s.x0)
    if not fix & syn.FIX_POINTS:
        assert rel(bg["V"], bo["V"]) < tol
    if not fix & syn.FIX_IMAGES:
        assert rel(bg["Ui"], bo["Ui"]) < tol
    if not fix & syn.FIX_CAMERAS:
        assert rel(bg["Uc"], bo["Uc"]) < tol
    assert rel(bg["g"], bo["g"]) < tol
    assert abs(p.max_diagonal() - o.max_diagonal()) <= tol * abs(o.max_diagonal())
    assert abs(p.gradient_inf_norm() - o.gradient_inf_norm()) <= tol * abs(o.gradient_inf_norm())
    dx = np.linspace(-1, 1, o.x_size)
    assert abs(p.gradient_dot(dx) - o.gradient_dot(dx)) <= tol * np.abs(bo["g"]).sum()


@pytest.mark.parametrize("fix", FIX_MASKS)
def test_reduced_system_and_solve(gpu, fix):
    s = tiny(fix)
    o, p = Oracle(s), fp64_problem(s)
    print(f"solve fix={fix}", flush=True)
    o.build_normal(s.x0)
    p.build_normal(s.x0)
    mu = 1e-3 * o.max_diagonal()
    if not fix & syn.FIX_POINTS and (fix & 6) != 6:
        rc_o, So, ro = o.reduced_system(mu)
        rc_g, Sg, rg = p.reduced_system(mu)
        assert rc_o == 0 and rc_g == 0
        assert rel(Sg, So) < 1e-10
        assert rel(rg, ro) < 1e-10
    rc_o, do = o.solve(mu)
    rc_g, dg = p.solve(mu)
    assert rc_o == 0 and rc_g == 0
    assert rel(dg, do) < 1e-8
    # invariant of the reference solver test: (N + mu I) delta = g
    _, N, g = o.dense_normal(s.x0)
    res = (N + mu * np.eye(o.x_size)) @ dg - g
    assert np.abs(res).max() < 1e-8 * np.abs(g).max()


def test_mahalanobis(gpu):
    s = tiny(syn.FIX_CAMERAS)
    o, p = Oracle(s), Problem(s, precision=abi.PRECISION_FP64)
    assert abs(p.mahalanobis(s.x0) - o.mahalanobis(s.x0)) < 1e-11 * o.mahalanobis(s.x0)


@pytest.mark.parametrize("prec", [abi.PRECISION_FP64, abi.PRECISION_MIXED])
def test_lm_ring(gpu, prec):
    s = syn.ring_scene()
    o, p = Oracle(s), Problem(s, precision=prec)
    print(f"lm ring prec={prec}", flush=True)
    rc_o, xo, st_o, _ = o.lm(max_iterations=20)
    rc_g, xg, st_g = p.lm(max_iterations=20)
    assert rc_o == 0 and rc_g == 0
    assert st_g.iterations == st_o.iterations == 20
    rm_o, rm_g = syn.reprojection_rmse(s, xo), syn.reprojection_rmse(s, xg)
    assert abs(rm_g - rm_o) < 1e-3
    assert abs(st_g.initial_cost - st_o.initial_cost) < 1e-9 * st_o.initial_cost
    if prec == abi.PRECISION_FP64:
        assert abs(st_g.final_cost - st_o.final_cost) < 1e-7 * st_o.final_cost
        # once converged, rho is rounding noise: accept / reject of the tail
        # steps may differ, so accepted counts are not compared
        assert rel(xg, xo) < 1e-5


def test_lm_general_tiny(gpu):
    """Every camera group, constraints and random covariances, in mixed
    precision (the FP64 form of this E = 16 scene is switched off): same
    optimum as the FP64 oracle within 1e-3 px of RMSE."""
    s = tiny(0, num_images=6, num_points=40, num_cameras=3)
    o, p = Oracle(s), Problem(s, precision=abi.PRECISION_MIXED)
    rc_o, xo, st_o, _ = o.lm(max_iterations=10)
    rc_g, xg, st_g = p.lm(max_iterations=10)
    assert rc_o == 0 and rc_g == 0
    # mixed precision: same optimum to the tolerance the benchmark states
    assert abs(syn.reprojection_rmse(s, xg) - syn.reprojection_rmse(s, xo)) < 1e-3


def test_batch_matches_single(gpu):
    scenes = syn.windows_batch(count=6, scale_points=0.1)
    b = Batch(scenes, precision=abi.PRECISION_FP64)
    rc, xs, sts = b.lm(max_iterations=8)
    assert rc == 0
    for s, x, st in zip(scenes, xs, sts):
        o = Oracle(s)
        _, xo, st_o, _ = o.lm(max_iterations=8)
        assert abs(st.final_cost - st_o.final_cost) < 1e-7 * st_o.final_cost
        assert abs(syn.reprojection_rmse(s, x) - syn.reprojection_rmse(s, xo)) < 1e-6


@pytest.mark.parametrize("prec", [abi.PRECISION_MIXED])
def test_pcg_step_matches_direct(gpu, prec):
    """Block-Jacobi PCG on the implicit Schur complement vs the oracle's direct
    solve (banded trajectory scene)."""
    s = syn.large_scene(num_cameras=40, num_points=600)
    o = Oracle(s)
    o.build_normal(s.x0)
    mu = 1e-3 * o.max_diagonal()
    _, do = o.solve(mu)
    p = Problem(s, precision=prec, solver=abi.SOLVER_PCG, pcg_rtol=1e-10)
    p.build_normal(s.x0)
    rc, dg = p.solve(mu)
    assert rc == 0
    assert rel(dg, do) < 1e-4


def test_pcg_lm(gpu):
    s = syn.large_scene(num_cameras=60, num_points=2000)
    o = Oracle(s)
    p = Problem(s, precision=abi.PRECISION_MIXED, solver=abi.SOLVER_PCG, pcg_rtol=1e-6)
    rc, x, st = p.lm(max_iterations=10)
    _, xo, sto, _ = o.lm(max_iterations=10)
    assert rc == 0 and st.pcg_iterations > 0
    assert abs(syn.reprojection_rmse(s, x) - syn.reprojection_rmse(s, xo)) < 1e-3


def test_pcg_fp64_fixed_cameras(gpu):
    """FP64 PCG (camera block width 6) on the ring scene vs the direct path."""
    s = syn.ring_scene()
    o = Oracle(s)
    p = Problem(s, precision=abi.PRECISION_FP64, solver=abi.SOLVER_PCG, pcg_rtol=1e-9)
    rc, x, st = p.lm(max_iterations=10)
    _, xo, sto, _ = o.lm(max_iterations=10)
    assert rc == 0
    assert abs(st.final_cost - sto.final_cost) < 1e-6 * sto.final_cost


def test_auto_solver_keeps_shared_cameras_dense(gpu):
    """Above 4096 reduced unknowns the automatic choice is PCG, whose
    preconditioner needs one image per camera; with one camera group shared by
    700 images (free intrinsics) the dense path is taken instead."""
    s = syn.tiny_scene(num_images=700, num_points=300, num_cameras=1, kmin=2)
    p = Problem(s, precision=abi.PRECISION_MIXED)
    print(f"auto solver n_reduced={p.n_reduced}", flush=True)
    assert p.n_reduced == 6 * 700 + p.cam_d
    p.build_normal(s.x0)
    rc, dx = p.solve(1e-3 * p.max_diagonal())
    assert rc == 0 and np.isfinite(dx).all()


def test_long_tracks(gpu):
    """Points seen by 150..200 images (pair incidences index keys with 32 bits
    each, so track length is not limited); mixed precision LM vs the oracle."""
    s = syn.tiny_scene(num_images=200, num_points=40, num_cameras=2, kmin=150, constraints=True)
    assert np.bincount(s.feature_maps[:, 1]).max() > 128
    o, p = Oracle(s), Problem(s, precision=abi.PRECISION_MIXED)
    print(f"long tracks keys={s.num_keys}", flush=True)
    rc_o, xo, st_o, _ = o.lm(max_iterations=5)
    rc_g, xg, st_g = p.lm(max_iterations=5)
    assert rc_o == 0 and rc_g == 0
    assert abs(syn.reprojection_rmse(s, xg) - syn.reprojection_rmse(s, xo)) < 1e-3
