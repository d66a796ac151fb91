"""Pins the oracle's vector function and analytic Jacobian.

* The single fixed input of the reference's derivative test
  (unit_test/bundle_adjustment/test_camera_shared_analytical_derivative.cpp:749-770),
  checked stage by stage (normalized key, radial, decentering, intrinsic) as
  that test does, but against complex-step derivatives (exact to rounding)
  instead of forward differences, at the reference's 1e-5 plus a 1e-9 check.
* The (C, R) identity of test_camera_shared_vector_function.cpp:190-194:
  x_c = R (X - C) equals the angle-axis form with t = -R C.
* Closed forms: the theta = 0 branch and a pinhole camera at the origin.
* Golden vectors committed under tests/golden/.
"""
import json
import os

import numpy as np
import pytest

from hs_sfm_amd import synthetic as syn
from oracle.oracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))

POINT = [89.06332903711264, -254.7119654139254, 27.36102981725708]
ROT = [0.0461972342123952, -0.01398317345699249, -1.02689361003788]
TRANS = [228.8962690432249, 77.85673263411233, -476.5385356161123]
K123 = [-0.005049046200604883, 0.03129380429860988, -0.03079482096044222]
D12 = [-0.0005537654832018913, -0.0004987771776838148]
INTR = [4835.47665904517, 0.0, -42.4095312016, -31.699212823, 1.0]


def single_key_scene(comp_mask):
    """One point, one image, one camera group; x = [point | image | camera]."""
    cam_full = np.array(K123 + D12 + INTR)
    s = syn.Scene(num_points=1, num_images=1, num_cameras=1,
                  feature_maps=np.array([[0, 0]], np.int32),
                  image_camera_map=np.zeros(1, np.int32), comp_mask=comp_mask, free_mask=0,
                  intrinsic_defaults=cam_full[None, :].copy(), fix_mask=0)
    cam = cam_full[s.free_indices]
    s.x0 = np.concatenate([POINT, ROT, TRANS, cam])
    s.y_obs = np.zeros(2)
    return s


@pytest.mark.parametrize("comp", [0, syn.COMPUTE_RADIAL, syn.COMPUTE_DECENTERING,
                                  syn.COMPUTE_RADIAL | syn.COMPUTE_DECENTERING | syn.COMPUTE_INTRINSIC])
def test_reference_fixed_input_derivatives(comp):
    s = single_key_scene(comp)
    o = Oracle(s)
    Ja = o.jacobian(s.x0, mode=0)
    Jcs = o.jacobian(s.x0, mode=3)
    Jfd = o.jacobian(s.x0, mode=1, delta=1e-6, min_delta=1e-8)  # the reference's FD settings
    for a, c, f in zip(Ja, Jcs, Jfd):
        if a.size == 0:
            continue
        # reference tolerance: isApprox(., 1e-5) against forward differences
        assert np.linalg.norm(a - f) <= 1e-5 * min(np.linalg.norm(a), np.linalg.norm(f))
        # complex step is exact to rounding: a much tighter pin
        assert np.linalg.norm(a - c) <= 1e-9 * np.linalg.norm(c)


def test_fixed_input_golden():
    """Regression against committed vectors (generated by tests/golden/make_golden.py)."""
    with open(os.path.join(HERE, "golden", "fixed_input.json")) as f:
        gold = json.load(f)
    s = single_key_scene(syn.COMPUTE_RADIAL | syn.COMPUTE_DECENTERING | syn.COMPUTE_INTRINSIC)
    o = Oracle(s)
    y = o.evaluate(s.x0)
    Jp, Ji, Jc = o.jacobian(s.x0)
    np.testing.assert_allclose(y, gold["key"], rtol=1e-14, atol=0)
    np.testing.assert_allclose(Jp.reshape(-1), gold["point_block"], rtol=1e-12)
    np.testing.assert_allclose(Ji.reshape(-1), gold["image_block"], rtol=1e-12)
    np.testing.assert_allclose(Jc.reshape(-1), gold["camera_block"], rtol=1e-12, atol=1e-9)


def test_camera_center_identity():
    """x_c = R (X - C) (ProjectiveFunctions::WorldPointProjectToImageKey) equals
    the reference's angle-axis form with t = -R C."""
    rng = np.random.default_rng(3)
    s = syn.tiny_scene(seed=3)
    o = Oracle(s)
    y = o.evaluate(s.x0)[:2 * s.num_keys].reshape(-1, 2)
    pts, imgs, intr = syn.scene_params(s, s.x0)
    im, pt = s.feature_maps[:, 0], s.feature_maps[:, 1]
    R = syn.angle_axis_to_matrix(imgs[:, :3])
    C = -np.einsum("nji,nj->ni", R, imgs[:, 3:])  # C = -R^T t
    xc = np.einsum("nij,nj->ni", R[im], pts[pt] - C[im])
    n = xc[:, :2] / xc[:, 2:3]
    c = intr[s.image_camera_map[im]]
    r2 = (n ** 2).sum(1)
    coeff = c[:, 0] * r2 + c[:, 1] * r2 ** 2 + c[:, 2] * r2 ** 3
    dx = 2 * c[:, 3] * n[:, 0] * n[:, 1] + c[:, 4] * (r2 + 2 * n[:, 0] ** 2)
    dy = 2 * c[:, 4] * n[:, 0] * n[:, 1] + c[:, 3] * (r2 + 2 * n[:, 1] ** 2)
    m = n + np.stack([coeff * n[:, 0] + dx, coeff * n[:, 1] + dy], 1)
    K = np.zeros((len(c), 3, 3))
    K[:, 0, 0] = c[:, 5]; K[:, 0, 1] = c[:, 6]; K[:, 0, 2] = c[:, 7]
    K[:, 1, 1] = c[:, 5] * c[:, 9]; K[:, 1, 2] = c[:, 8]; K[:, 2, 2] = 1
    uv = np.einsum("nij,nj->ni", K, np.concatenate([m, np.ones((len(m), 1))], 1))[:, :2]
    np.testing.assert_allclose(y, uv, rtol=1e-10, atol=1e-8)
    del rng


def test_theta_zero_branch_and_pinhole():
    """theta = 0: x_c = X + t (camera_shared_vector_function.hpp:695-698); with
    no distortion the key is the pinhole f X/Z + p."""
    s = single_key_scene(syn.COMPUTE_INTRINSIC)
    x = s.x0.copy()
    x[3:6] = 0.0
    X = np.array(POINT) + np.array(TRANS)
    f, sk, px, py, pr = INTR
    expect = [f * X[0] / X[2] + sk * X[1] / X[2] + px, f * pr * X[1] / X[2] + py]
    o = Oracle(s)
    np.testing.assert_allclose(o.evaluate(x), expect, rtol=1e-15)
    # the Taylor branch of the Jacobian at theta = 0 equals the limit of the
    # exact formula (complex step away from zero is not defined; compare with
    # a central difference instead)
    Ja = o.jacobian(x, mode=0)
    Jc = o.jacobian(x, mode=2, delta=0.0, min_delta=1e-6)
    for a, c in zip(Ja, Jc):
        np.testing.assert_allclose(a, c, rtol=1e-7, atol=1e-7 * np.abs(c).max())


def test_jacobian_all_masks_complex_step():
    for comp in range(8):
        for free in (0,):
            s = syn.tiny_scene(comp_mask=comp, free_mask=free, seed=11 + comp)
            o = Oracle(s)
            for a, c in zip(o.jacobian(s.x0), o.jacobian(s.x0, mode=3)):
                if a.size:
                    assert np.abs(a - c).max() <= 1e-9 * max(1.0, np.abs(c).max())


def test_free_mask_subset_matches_full_columns():
    """The 9-parameter camera (r, t, f, k1, k2) is the reference's {radial,
    intrinsic} camera with some columns removed: same projection, same columns."""
    full = syn.tiny_scene(comp_mask=syn.COMPUTE_RADIAL | syn.COMPUTE_INTRINSIC, seed=5)
    sub_mask = (1 << syn.K1) | (1 << syn.K2) | (1 << syn.FOCAL)
    sub = syn.tiny_scene(comp_mask=syn.COMPUTE_RADIAL | syn.COMPUTE_INTRINSIC, free_mask=sub_mask, seed=5)
    of, osb = Oracle(full), Oracle(sub)
    np.testing.assert_allclose(of.evaluate(full.x_true), osb.evaluate(sub.x_true), rtol=1e-14)
    Jf = of.jacobian(full.x_true)[2]
    Js = osb.jacobian(sub.x_true)[2]
    cols = [full.free_indices.index(j) for j in sub.free_indices]
    np.testing.assert_allclose(Js, Jf[:, :, cols], rtol=1e-13, atol=1e-12)
