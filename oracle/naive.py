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
