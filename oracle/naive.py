"""FP64 CPU restatement of the ba_naive vector function.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the
hsba_naive_* entry points, never by the product path.

ba_naive (include/hs_sfm/bundle_adjustment/ba_naive_*.hpp) keeps 6-DoF
cameras in normalised coordinates with x = [cameras 6C | points 3P]
(ba_naive_vector_function.hpp:117-135).  Each feature is
PointProjectToFeature (ba_naive_vector_function.hpp:170-196):
    theta = |r|;  theta == 0:  h = p + t
    else k = r / theta,  h = cos(theta) p + sin(theta) k x p
                             + (1 - cos(theta)) (p . k) k + t
    feature = h.xy / h.z
Written with arithmetic that stays analytic in complex numbers (theta from
sqrt(r . r), no abs), so the tests can differentiate it by complex step.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def naive_evaluate(x: np.ndarray, num_cameras: int, num_points: int, feature_maps: np.ndarray) -> np.ndarray:
    """BANaiveVectorFunction::operator() (ba_naive_vector_function.hpp:137-168):
    y[2i:2i+2] = PointProjectToFeature(r_c, t_c, p_q) for feature i = (c, q)."""
    C, P = num_cameras, num_points
    cams = x[:6 * C].reshape(C, 6)
    pts = x[6 * C:6 * C + 3 * P].reshape(P, 3)
    c = feature_maps[:, 0]
    q = feature_maps[:, 1]
    r, t, p = cams[c, :3], cams[c, 3:], pts[q]
    th2 = np.sum(r * r, axis=1)
    zero = th2 == 0
    th = np.sqrt(np.where(zero, 1.0, th2))
    k = r / th[:, None]
    ct, st = np.cos(th)[:, None], np.sin(th)[:, None]
    h = ct * p + st * np.cross(k, p) + (1 - ct) * np.sum(p * k, axis=1)[:, None] * k + t
    h = np.where(zero[:, None], p + t, h)
    return (h[:, :2] / h[:, 2:3]).reshape(-1)


def naive_to_shared(x: np.ndarray, num_cameras: int, num_points: int) -> np.ndarray:
    """[cameras 6C | points 3P] -> camera_shared [points 3P | images 6C]."""
    return np.concatenate([x[6 * num_cameras:6 * num_cameras + 3 * num_points], x[:6 * num_cameras]])


def shared_to_naive(xs: np.ndarray, num_cameras: int, num_points: int) -> np.ndarray:
    return np.concatenate([xs[3 * num_points:3 * num_points + 6 * num_cameras], xs[:3 * num_points]])


@dataclass
class NaiveScene:
    num_cameras: int
    num_points: int
    feature_maps: np.ndarray  # [F, 2] int32 (camera, point)
    x0: np.ndarray
    x_true: np.ndarray
    y_obs: np.ndarray         # normalised coordinates, 2F
    name: str = "naive"


def focal_and_principal_point(scene):
    """(f, px, py) of a scene whose cameras share fixed intrinsics: from the
    held intrinsic block [f skew px py pr] or from the defaults [C][10]."""
    if scene.intrinsic_defaults is not None:
        full = scene.intrinsic_defaults[0]
        return full[5], full[7], full[8]
    fc = np.asarray(scene.fix_cameras).reshape(scene.num_cameras, -1)[0]
    return fc[0], fc[2], fc[3]


def naive_from_scene(scene) -> NaiveScene:
    """ba_naive form of a camera_shared scene with fixed, distortion-free shared
    intrinsics: features divided through by the intrinsics, as the ba_naive
    synthetic generator does (ba_naive_synthetic_data_generator.hpp:252-255)."""
    f, px, py = focal_and_principal_point(scene)
    uv = scene.y_obs[:2 * scene.num_keys].reshape(-1, 2)
    y = np.stack([(uv[:, 0] - px) / f, (uv[:, 1] - py) / f], axis=1).reshape(-1)
    P, C = scene.num_points, scene.num_images

    def conv(x):
        return shared_to_naive(x[:3 * P + 6 * C], C, P)

    return NaiveScene(C, P, scene.feature_maps.astype(np.int32), conv(scene.x0), conv(scene.x_true), y,
                      name=scene.name + "-naive")


# ---------------------------------------------------------------- GCP family
@dataclass
class GCPScene:
    """ba_gcp_constrained problem: a NaiveScene whose last num_gcps points are
    ground control points with observed coordinates."""
    naive: NaiveScene
    num_gcps: int
    gcp_blocks: np.ndarray    # [G, 3, 3] symmetric positive definite inverse covariances
    y_obs: np.ndarray         # [2F + 3G]: features, then GCP coordinates


def gcp_evaluate(x: np.ndarray, scene: GCPScene) -> np.ndarray:
    """BAGCPConstrainedVectorFunction::operator()
    (ba_gcp_constrained_vector_function.hpp:73-113): the naive features, then
    the x entries of the last G points in GCP order."""
    n = scene.naive
    y = naive_evaluate(x, n.num_cameras, n.num_points, n.feature_maps)
    G = scene.num_gcps
    gx = x[6 * n.num_cameras + 3 * (n.num_points - G):6 * n.num_cameras + 3 * n.num_points]
    return np.concatenate([y, gx])


def gcp_mahalanobis(x: np.ndarray, scene: GCPScene, feature_cov=None) -> float:
    """BAGCPConstrainedMahalanobisDistanceCalculator
    (ba_gcp_constrained_mahalanobis_distance_calculator.hpp:33-62):
    sum r_f^T W_f r_f over features plus sum r_g^T B_g r_g over GCPs."""
    r = gcp_evaluate(x, scene) - scene.y_obs
    F = scene.naive.feature_maps.shape[0]
    rf = r[:2 * F].reshape(-1, 2)
    W = np.broadcast_to(np.eye(2), (F, 2, 2)) if feature_cov is None else np.asarray(feature_cov).reshape(F, 2, 2)
    rg = r[2 * F:].reshape(-1, 3)
    return float(np.einsum("ki,kij,kj->", rf, W, rf) + np.einsum("gi,gij,gj->", rg, scene.gcp_blocks, rg))


def make_gcp_scene(n: NaiveScene, num_gcps: int, seed: int = 7, gcp_noise: float = 0.0) -> GCPScene:
    """GCPs = the last num_gcps points, observed at their true coordinates
    plus N(0, gcp_noise^2) noise, each with a random SPD 3x3 weight block."""
    rng = np.random.default_rng(seed)
    C, P = n.num_cameras, n.num_points
    pts = n.x_true[6 * C:6 * C + 3 * P].reshape(P, 3)
    g = pts[P - num_gcps:] + gcp_noise * rng.standard_normal((num_gcps, 3))
    A = rng.standard_normal((num_gcps, 3, 3))
    B = np.einsum("gij,gkj->gik", A, A) + 3.0 * np.eye(3)
    B = 0.5 * (B + np.transpose(B, (0, 2, 1)))
    return GCPScene(n, num_gcps, np.ascontiguousarray(B), np.concatenate([n.y_obs, g.reshape(-1)]))


def gcp_shared_scene(gs: GCPScene):
    """The camera_shared form the oracle runs: one image per camera, identity
    intrinsics, and the GCPs as point constraints on all three coordinates
    with full 3x3 weight blocks."""
    from hs_sfm_amd import synthetic as syn
    n = gs.naive
    C, P, G = n.num_cameras, n.num_points, gs.num_gcps
    t = syn.Scene(num_points=P, num_images=C, num_cameras=C, feature_maps=n.feature_maps.astype(np.int32),
                  image_camera_map=np.arange(C, dtype=np.int32), comp_mask=0, free_mask=0,
                  intrinsic_defaults=None, fix_mask=0, name=n.name + "-gcp-shared")
    t.point_constraints = np.array([[P - G + i, 7] for i in range(G)], np.int64).reshape(-1, 2)
    t.point_constraint_blocks = gs.gcp_blocks.copy()
    t.x0 = naive_to_shared(n.x0, C, P)
    t.x_true = naive_to_shared(n.x_true, C, P)
    t.y_obs = gs.y_obs.copy()
    return t
