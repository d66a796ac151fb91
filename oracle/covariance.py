"""Posterior covariance restated: CameraSharedAnalyticalXCovarianceCalculator
(include/hs_sfm/bundle_adjustment/camera_shared_analytical_x_covariance_calculator.hpp:45-90).

TEST INFRASTRUCTURE ONLY: the checker of hsba_covariance, never the product.

The reference forms the dense normal matrix N = J^T Sigma^-1 J at x (with the
constraint weights of ComputeNormalMatrix), takes Eigen::JacobiSVD (thin U, V)
and keeps the singular values above the absolute threshold 1e-6:
    C = V diag(1/d_i if |d_i| > 1e-6 else 0) U^T          (:71-86)
numpy's SVD gives the same factors up to the order and signs of singular
pairs, which the product V D U^T does not depend on.
"""
from __future__ import annotations

import numpy as np

THRESHOLD = 1e-6  # camera_shared_analytical_x_covariance_calculator.hpp:77


def pseudo_inverse(N: np.ndarray) -> np.ndarray:
    U, d, Vt = np.linalg.svd(N)
    D = np.where(np.abs(d) > THRESHOLD, 1.0 / np.where(d == 0, 1.0, d), 0.0)
    return (Vt.T * D) @ U.T


def dense_normal(oracle, x) -> np.ndarray:
    """N in the x layout assembled from the oracle builder's blocks (V_p, U_i,
    U_c, per-key W_pi and W_pc, per-image W_ic), as
    CameraSharedNormalMatrix::coeff reads them
    (BA/camera_shared_normal_matrix.hpp:234-443)."""
    s = oracle.scene
    b = oracle.build_normal(x)
    P, I, C, d = s.num_points, s.num_images, s.num_cameras, oracle.cam_d
    pts = not s.fix_mask & 1
    imgs = not s.fix_mask & 2
    cams = not s.fix_mask & 4 and d > 0
    op = 0
    oi = op + (3 * P if pts else 0)
    oc = oi + (6 * I if imgs else 0)
    n = oc + (d * C if cams else 0)
    N = np.zeros((n, n))
    if pts:
        for p in range(P):
            N[op + 3 * p:op + 3 * p + 3, op + 3 * p:op + 3 * p + 3] = b["V"][p]
    if imgs:
        for i in range(I):
            N[oi + 6 * i:oi + 6 * i + 6, oi + 6 * i:oi + 6 * i + 6] = b["Ui"][i]
    if cams:
        for c in range(C):
            N[oc + d * c:oc + d * c + d, oc + d * c:oc + d * c + d] = b["Uc"][c]
    fm, icm = s.feature_maps, s.image_camera_map
    for k in range(s.num_keys):
        i, p = int(fm[k, 0]), int(fm[k, 1])
        c = int(icm[i])
        if pts and imgs:
            blk = b["Wpi"][k]
            N[op + 3 * p:op + 3 * p + 3, oi + 6 * i:oi + 6 * i + 6] += blk
            N[oi + 6 * i:oi + 6 * i + 6, op + 3 * p:op + 3 * p + 3] += blk.T
        if pts and cams:
            blk = b["Wpc"][k]
            N[op + 3 * p:op + 3 * p + 3, oc + d * c:oc + d * c + d] += blk
            N[oc + d * c:oc + d * c + d, op + 3 * p:op + 3 * p + 3] += blk.T
    if imgs and cams:
        for i in range(I):
            c = int(icm[i])
            blk = b["Wic"][i]
            N[oi + 6 * i:oi + 6 * i + 6, oc + d * c:oc + d * c + d] += blk
            N[oc + d * c:oc + d * c + d, oi + 6 * i:oi + 6 * i + 6] += blk.T
    return N


def x_covariance(oracle, x) -> np.ndarray:
    """The reference's x covariance for the oracle's scene at x."""
    return pseudo_inverse(dense_normal(oracle, x))
