"""Seeded synthetic bundle-adjustment scenes (numpy).

The reference builds its test scenes with Boost RNG generators
(include/hs_sfm/bundle_adjustment/camera_shared_synthetic_data_generator.hpp:46-841)
that are not reproducible here; these generators keep only the reference's
visibility rule (a key is kept when it projects strictly inside the image,
include/hs_sfm/synthetic/multiple_camera_keyset_generator.hpp:76-96) and its
camera model, and draw the five benchmark workloads from fixed seeds:

  ring      10 cameras on a 10 m circle, 500 points, ~4.5k keys, fixed intrinsics
  mid       200 cameras x 9 params (r, t, f, k1, k2), 50k points, ~1M keys
  large     2000 cameras along a trajectory, 600k points, ~5.1M keys, banded
  windows   512 independent local windows of 5-40 cameras, 200-5000 points
  serving   jobs drawn from the windows distribution plus 5% mid-class jobs

Scenes are plain numpy arrays in the reference's layouts (x = [points | images |
cameras], y = [keys | constraints], feature_maps = (image_id, point_id)).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

# bit layout mirrors include/hsba.h
COMPUTE_RADIAL = 1
COMPUTE_DECENTERING = 2
COMPUTE_INTRINSIC = 4
FIX_POINTS = 1
FIX_IMAGES = 2
FIX_CAMERAS = 4
K1, K2, K3, D1, D2, FOCAL, SKEW, PX, PY, PR = range(10)

SEED_RING = 20261203
SEED_MID = 20261204
SEED_LARGE = 20261205
SEED_WINDOWS = 20261206
SEED_SERVING = 20261207


@dataclasses.dataclass
class Scene:
    """One bundle-adjustment problem in the reference's layouts."""
    num_points: int
    num_images: int
    num_cameras: int
    feature_maps: np.ndarray          # int32 [N, 2] (image_id, point_id)
    image_camera_map: np.ndarray      # int32 [I]
    comp_mask: int
    free_mask: int                    # 0 = every parameter of the enabled groups
    intrinsic_defaults: Optional[np.ndarray]  # float64 [C, 10]
    fix_mask: int
    fix_points: Optional[np.ndarray] = None   # [P, 3]
    fix_images: Optional[np.ndarray] = None   # [I, 6]
    fix_cameras: Optional[np.ndarray] = None  # [C, d]
    point_constraints: Optional[np.ndarray] = None   # int64 [n, 2] (id, mask)
    image_constraints: Optional[np.ndarray] = None
    camera_constraints: Optional[np.ndarray] = None
    x_true: Optional[np.ndarray] = None
    x0: Optional[np.ndarray] = None
    y_obs: Optional[np.ndarray] = None
    key_blocks: Optional[np.ndarray] = None          # [N, 2, 2] or None = identity
    constraint_weights: Optional[np.ndarray] = None
    point_constraint_blocks: Optional[np.ndarray] = None  # [n, 3, 3]: full weights of mask-7 point constraints
    name: str = ""

    @property
    def num_keys(self) -> int:
        return int(self.feature_maps.shape[0])

    @property
    def enabled_bits(self) -> int:
        b = 0
        if self.comp_mask & COMPUTE_RADIAL:
            b |= 0x7
        if self.comp_mask & COMPUTE_DECENTERING:
            b |= 0x18
        if self.comp_mask & COMPUTE_INTRINSIC:
            b |= 0x3E0
        return b

    @property
    def free_bits(self) -> int:
        return self.free_mask if self.free_mask else self.enabled_bits

    @property
    def camera_params(self) -> int:
        return bin(self.free_bits).count("1")

    @property
    def free_indices(self) -> List[int]:
        return [j for j in range(10) if (self.free_bits >> j) & 1]

    def x_layout(self):
        pts = 0 if self.fix_mask & FIX_POINTS else 3 * self.num_points
        imgs = 0 if self.fix_mask & FIX_IMAGES else 6 * self.num_images
        cams = 0 if self.fix_mask & FIX_CAMERAS else self.camera_params * self.num_cameras
        return pts, imgs, cams

    @property
    def x_size(self) -> int:
        return sum(self.x_layout())


# ---------------------------------------------------------------- rotations
def angle_axis_to_matrix(r: np.ndarray) -> np.ndarray:
    """Rodrigues' formula for a batch of angle-axis vectors [..., 3]."""
    r = np.asarray(r, dtype=np.float64)
    th = np.linalg.norm(r, axis=-1, keepdims=True)
    k = np.where(th > 0, r / np.where(th > 0, th, 1.0), 0.0)
    kx, ky, kz = k[..., 0], k[..., 1], k[..., 2]
    K = np.zeros(r.shape[:-1] + (3, 3))
    K[..., 0, 1], K[..., 0, 2] = -kz, ky
    K[..., 1, 0], K[..., 1, 2] = kz, -kx
    K[..., 2, 0], K[..., 2, 1] = -ky, kx
    s = np.sin(th)[..., None]
    c = np.cos(th)[..., None]
    eye = np.broadcast_to(np.eye(3), K.shape)
    return eye + s * K + (1 - c) * (K @ K)


def matrix_to_angle_axis(R: np.ndarray) -> np.ndarray:
    """Logarithm of rotation matrices [n, 3, 3] -> angle-axis [n, 3], including
    angles at or near pi (axis from the symmetric part (R + I) / 2 = k k^T)."""
    R = np.asarray(R, dtype=np.float64)
    shape = R.shape[:-2]
    R = R.reshape(-1, 3, 3)
    cos_t = np.clip((np.trace(R, axis1=-2, axis2=-1) - 1) / 2, -1.0, 1.0)
    th = np.arccos(cos_t)
    w = np.stack([R[:, 2, 1] - R[:, 1, 2], R[:, 0, 2] - R[:, 2, 0], R[:, 1, 0] - R[:, 0, 1]], axis=-1)
    s = np.sin(th)
    out = w * np.where(s > 1e-12, th / (2 * np.where(s > 1e-12, s, 1.0)), 0.5)[:, None]
    near_pi = th > np.pi - 1e-3
    for n in np.nonzero(near_pi)[0]:
        B = (R[n] + np.eye(3)) / 2
        j = int(np.argmax(np.diag(B)))
        k = B[:, j] / np.sqrt(max(B[j, j], 1e-300))
        k /= np.linalg.norm(k)
        if np.dot(k, w[n]) < 0:
            k = -k
        out[n] = th[n] * k
    return out.reshape(shape + (3,))


def look_at(centers: np.ndarray, target: np.ndarray, up=(0.0, 0.0, 1.0)):
    """World-to-camera rotations for cameras at `centers` looking at `target`
    (x right, y down, z forward); returns (angle-axis [n,3], t [n,3])."""
    centers = np.asarray(centers, dtype=np.float64)
    z = np.asarray(target, dtype=np.float64) - centers
    z /= np.linalg.norm(z, axis=1, keepdims=True)
    x = np.cross(z, np.asarray(up, dtype=np.float64))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    y = np.cross(z, x)
    R = np.stack([x, y, z], axis=1)
    t = -np.einsum("nij,nj->ni", R, centers)
    return matrix_to_angle_axis(R), t


# --------------------------------------------------------------- projection
def project(points: np.ndarray, rot: np.ndarray, trans: np.ndarray,
            intr_full: np.ndarray):
    """Reference projection (BA/camera_shared_vector_function.hpp:634-744),
    vectorized over keys.  Returns (uv [N,2], depth [N])."""
    th = np.linalg.norm(rot, axis=1)
    safe = np.where(th > 0, th, 1.0)
    k = rot / safe[:, None]
    ct, st = np.cos(th), np.sin(th)
    kx = np.cross(k, points)
    kd = np.einsum("ij,ij->i", k, points)
    y = ct[:, None] * points + st[:, None] * kx + ((1 - ct) * kd)[:, None] * k + trans
    y = np.where((th > 0)[:, None], y, points + trans)
    n0, n1 = y[:, 0] / y[:, 2], y[:, 1] / y[:, 2]
    r2 = n0 * n0 + n1 * n1
    r4 = r2 * r2
    r6 = r2 * r4
    c = intr_full
    coeff = c[:, K1] * r2 + c[:, K2] * r4 + c[:, K3] * r6
    dd0 = 2 * c[:, D1] * n0 * n1 + c[:, D2] * (r2 + 2 * n0 * n0)
    dd1 = 2 * c[:, D2] * n0 * n1 + c[:, D1] * (r2 + 2 * n1 * n1)
    m0 = n0 + (coeff * n0 + dd0)
    m1 = n1 + (coeff * n1 + dd1)
    u = c[:, FOCAL] * m0 + c[:, SKEW] * m1 + c[:, PX]
    v = c[:, FOCAL] * c[:, PR] * m1 + c[:, PY]
    return np.stack([u, v], axis=1), y[:, 2]


def _identity_intrinsics(n: int) -> np.ndarray:
    a = np.zeros((n, 10))
    a[:, FOCAL] = 1.0
    a[:, PR] = 1.0
    return a


def full_intrinsics(scene: Scene, x: np.ndarray) -> np.ndarray:
    """Per-camera [C, 10] intrinsic vectors at x (held values from defaults)."""
    C = scene.num_cameras
    full = _identity_intrinsics(C)
    en = scene.enabled_bits
    if scene.intrinsic_defaults is not None:
        for j in range(10):
            if (en >> j) & 1:
                full[:, j] = scene.intrinsic_defaults[:, j]
    fi = scene.free_indices
    d = len(fi)
    if scene.fix_mask & FIX_CAMERAS:
        vals = scene.fix_cameras.reshape(C, d)
    else:
        p, i, _ = scene.x_layout()
        vals = x[p + i:p + i + d * C].reshape(C, d)
    for q, j in enumerate(fi):
        full[:, j] = vals[:, q]
    return full


def scene_params(scene: Scene, x: np.ndarray):
    """(points [P,3], images [I,6], intrinsics [C,10]) at x."""
    p, i, _ = scene.x_layout()
    pts = scene.fix_points if scene.fix_mask & FIX_POINTS else x[:p].reshape(-1, 3)
    imgs = scene.fix_images if scene.fix_mask & FIX_IMAGES else x[p:p + i].reshape(-1, 6)
    return pts, imgs, full_intrinsics(scene, x)


def evaluate_keys(scene: Scene, x: np.ndarray) -> np.ndarray:
    pts, imgs, intr = scene_params(scene, x)
    im = scene.feature_maps[:, 0]
    pt = scene.feature_maps[:, 1]
    uv, _ = project(pts[pt], imgs[im, :3], imgs[im, 3:], intr[scene.image_camera_map[im]])
    return uv


def reprojection_rmse(scene: Scene, x: np.ndarray) -> float:
    """sqrt(sum ||f_k(x) - y_k||^2 / N) in pixels."""
    uv = evaluate_keys(scene, x)
    r = uv - scene.y_obs[:2 * scene.num_keys].reshape(-1, 2)
    return float(np.sqrt(np.sum(r * r) / max(1, scene.num_keys)))


def mean_reprojection_error(scene: Scene, x: np.ndarray) -> float:
    """Mean L2 norm of key residuals, the statistic of the reference's
    ReprojectiveErrorCalculator (sfm_pipeline/reprojective_error_calculator.hpp:80-91)."""
    uv = evaluate_keys(scene, x)
    r = uv - scene.y_obs[:2 * scene.num_keys].reshape(-1, 2)
    return float(np.mean(np.linalg.norm(r, axis=1)))


# --------------------------------------------------------------- generators
def _perturb(scene: Scene, rng: np.random.Generator, x_true: np.ndarray,
             rot_sigma=0.01, trans_sigma=0.05, point_sigma=0.05) -> np.ndarray:
    x0 = x_true.copy()
    p, i, _ = scene.x_layout()
    if p:
        x0[:p] += rng.normal(0.0, point_sigma, p)
    if i:
        im = x0[p:p + i].reshape(-1, 6)
        im[:, :3] += rng.normal(0.0, rot_sigma, (im.shape[0], 3))
        im[:, 3:] += rng.normal(0.0, trans_sigma, (im.shape[0], 3))
    return x0


def _finish(scene: Scene, rng: np.random.Generator, x_true: np.ndarray,
            pixel_sigma=1.0, perturb=True) -> Scene:
    scene.x_true = x_true
    uv = evaluate_keys(scene, x_true)
    scene.y_obs = (uv + rng.normal(0.0, pixel_sigma, uv.shape)).reshape(-1)
    scene.x0 = _perturb(scene, rng, x_true) if perturb else x_true.copy()
    return scene


def ring_scene(seed: int = SEED_RING, num_cameras: int = 10, num_points: int = 500,
               radius: float = 10.0, height: float = 2.0, focal: float = 1000.0,
               image_size: float = 1000.0, pixel_sigma: float = 1.0,
               rng: Optional[np.random.Generator] = None, name="ring") -> Scene:
    """Cameras evenly on a circle looking at the origin; points uniform in a
    4 x 4 x 2 m box; a key exists when it projects strictly inside the image
    and in front of the camera.  Shared intrinsics, held fixed (FIX_CAMERAS)."""
    rng = rng if rng is not None else np.random.default_rng(seed)
    ang = 2 * np.pi * np.arange(num_cameras) / num_cameras
    centers = np.stack([radius * np.cos(ang), radius * np.sin(ang),
                        np.full(num_cameras, height)], axis=1)
    rot, trans = look_at(centers, np.zeros(3))
    pts = rng.uniform([-2, -2, -1], [2, 2, 1], size=(num_points, 3))
    intr = np.array([[focal, 0.0, image_size / 2, image_size / 2, 1.0]])
    full = _identity_intrinsics(1)
    full[0, FOCAL:PR + 1] = intr[0]
    fm = []
    for c in range(num_cameras):  # image-major key order
        uv, z = project(pts, np.repeat(rot[c:c + 1], num_points, 0),
                        np.repeat(trans[c:c + 1], num_points, 0),
                        np.repeat(full, num_points, 0))
        vis = (z > 0) & (uv[:, 0] > 0) & (uv[:, 0] < image_size) & \
              (uv[:, 1] > 0) & (uv[:, 1] < image_size)
        for p in np.nonzero(vis)[0]:
            fm.append((c, p))
    fm = np.asarray(fm, dtype=np.int32).reshape(-1, 2)
    scene = Scene(num_points=num_points, num_images=num_cameras, num_cameras=1,
                  feature_maps=fm, image_camera_map=np.zeros(num_cameras, np.int32),
                  comp_mask=COMPUTE_INTRINSIC, free_mask=0, intrinsic_defaults=None,
                  fix_mask=FIX_CAMERAS, fix_cameras=intr.copy(), name=name)
    x_true = np.concatenate([pts.reshape(-1), np.concatenate([rot, trans], 1).reshape(-1)])
    return _finish(scene, rng, x_true, pixel_sigma)


def truncated_geometric(rng, p: float, lo: int, hi: int, size: int) -> np.ndarray:
    """k in [lo, hi] with P(k) proportional to (1-p)^(k-lo)."""
    ks = np.arange(lo, hi + 1)
    w = (1 - p) ** (ks - lo)
    return rng.choice(ks, size=size, p=w / w.sum())


def _choose_without_replacement(rng, n_total: int, counts: np.ndarray) -> List[np.ndarray]:
    """For each count, a sorted random subset of range(n_total) (vectorized
    through argsort of random keys, one row per point)."""
    out = []
    P = counts.shape[0]
    chunk = 20000
    for s in range(0, P, chunk):
        cnt = counts[s:s + chunk]
        keys = rng.random((cnt.shape[0], n_total))
        order = np.argsort(keys, axis=1)
        for row, k in zip(order, cnt):
            out.append(np.sort(row[:k]))
    return out


def mid_scene(seed: int = SEED_MID, num_cameras: int = 200, num_points: int = 50000,
              geometric_p: float = 0.01244, kmin: int = 3, kmax: int = 40,
              focal: float = 1000.0, image_size: float = 1000.0,
              name="mid") -> Scene:
    """200 cameras, 9 parameters each: extrinsics (r, t) per image plus its
    own camera group with free (k1, k2, f); k3, skew, principal point and
    pixel ratio held.  Each point is seen by k ~ truncated geometric cameras
    drawn without replacement, re-drawn until every view is in front."""
    rng = np.random.default_rng(seed)
    C = num_cameras
    ang = 2 * np.pi * np.arange(C) / C
    heights = 2.0 + 1.5 * np.sin(3 * ang)
    centers = np.stack([10 * np.cos(ang), 10 * np.sin(ang), heights], axis=1)
    rot, trans = look_at(centers, np.zeros(3))
    pts = rng.uniform([-2, -2, -1], [2, 2, 1], size=(num_points, 3))
    counts = truncated_geometric(rng, geometric_p, kmin, kmax, num_points)
    subsets = _choose_without_replacement(rng, C, counts)
    pt_ids = np.repeat(np.arange(num_points), counts)
    img_ids = np.concatenate(subsets)
    fm = np.stack([img_ids, pt_ids], axis=1).astype(np.int32)  # track-major
    defaults = _identity_intrinsics(C)
    defaults[:, PX] = image_size / 2
    defaults[:, PY] = image_size / 2
    defaults[:, FOCAL] = focal
    k12 = rng.normal(0.0, 0.01, (C, 2))
    scene = Scene(num_points=num_points, num_images=C, num_cameras=C,
                  feature_maps=fm, image_camera_map=np.arange(C, dtype=np.int32),
                  comp_mask=COMPUTE_RADIAL | COMPUTE_INTRINSIC,
                  free_mask=(1 << K1) | (1 << K2) | (1 << FOCAL),
                  intrinsic_defaults=defaults, fix_mask=0, name=name)
    # all points are inside the 10 m ring, so every view has positive depth
    cams_true = np.stack([k12[:, 0], k12[:, 1], np.full(C, focal)], axis=1)
    x_true = np.concatenate([pts.reshape(-1), np.concatenate([rot, trans], 1).reshape(-1),
                             cams_true.reshape(-1)])
    scene = _finish(scene, rng, x_true)
    _, depth = project(pts[pt_ids], rot[img_ids], trans[img_ids],
                       full_intrinsics(scene, x_true)[img_ids])
    assert np.all(depth > 0)
    # unknown distortion starts at zero
    p3, i6, _ = scene.x_layout()
    scene.x0[p3 + i6:] = np.stack([np.zeros(C), np.zeros(C), np.full(C, focal)], 1).reshape(-1)
    return scene


def hold_cameras(scene: Scene, name: Optional[str] = None) -> Scene:
    """The same problem with every camera block held at its true value
    (FIX_CAMERAS, calibrated cameras): same keys and observations, camera
    block E = 6 (the image extrinsics)."""
    p, i, _ = scene.x_layout()
    return dataclasses.replace(scene, fix_mask=scene.fix_mask | FIX_CAMERAS,
                               fix_cameras=scene.x_true[p + i:].reshape(scene.num_cameras, -1).copy(),
                               x_true=scene.x_true[:p + i].copy(), x0=scene.x0[:p + i].copy(),
                               name=name or scene.name + "_calibrated")


def large_scene(seed: int = SEED_LARGE, num_cameras: int = 2000, num_points: int = 600000,
                kmin: int = 2, kmax: int = 15, altitude: float = 40.0, step: float = 2.0,
                focal: float = 1000.0, image_size: float = 1000.0, name="large") -> Scene:
    """Aerial strip: cameras along a seeded random-walk trajectory looking
    down; each point is seen by k ~ U{kmin..kmax} consecutive cameras, so the
    reduced camera system is banded.  9 parameters per camera as in mid_scene."""
    rng = np.random.default_rng(seed)
    C = num_cameras
    heading = np.cumsum(rng.normal(0.0, 0.05, C))
    pos = np.cumsum(np.stack([step * np.cos(heading), step * np.sin(heading)], 1), axis=0)
    centers = np.concatenate([pos, np.full((C, 1), altitude)], axis=1)
    # look straight down, image x along the heading
    fwd = np.stack([np.cos(heading), np.sin(heading), np.zeros(C)], 1)
    z = np.tile([0.0, 0.0, -1.0], (C, 1))
    y = np.cross(z, fwd)
    R = np.stack([fwd, y, z], axis=1)
    rot = matrix_to_angle_axis(R)
    trans = -np.einsum("nij,nj->ni", R, centers)
    counts = rng.integers(kmin, kmax + 1, num_points)
    starts = (rng.random(num_points) * (C - counts + 1)).astype(np.int64)
    mid = starts + (counts - 1) / 2.0
    lo = np.floor(mid).astype(np.int64)
    frac = mid - lo
    hi = np.minimum(lo + 1, C - 1)
    base = pos[lo] * (1 - frac)[:, None] + pos[hi] * frac[:, None]
    pts = np.concatenate([base + rng.normal(0.0, 3.0, (num_points, 2)),
                          rng.uniform(-2.0, 2.0, (num_points, 1))], axis=1)
    pt_ids = np.repeat(np.arange(num_points), counts)
    offs = np.arange(pt_ids.shape[0]) - np.repeat(np.cumsum(counts) - counts, counts)
    img_ids = np.repeat(starts, counts) + offs
    fm = np.stack([img_ids, pt_ids], axis=1).astype(np.int32)
    defaults = _identity_intrinsics(C)
    defaults[:, PX] = image_size / 2
    defaults[:, PY] = image_size / 2
    defaults[:, FOCAL] = focal
    k12 = rng.normal(0.0, 0.01, (C, 2))
    scene = Scene(num_points=num_points, num_images=C, num_cameras=C, feature_maps=fm,
                  image_camera_map=np.arange(C, dtype=np.int32),
                  comp_mask=COMPUTE_RADIAL | COMPUTE_INTRINSIC,
                  free_mask=(1 << K1) | (1 << K2) | (1 << FOCAL),
                  intrinsic_defaults=defaults, fix_mask=0, name=name)
    cams_true = np.stack([k12[:, 0], k12[:, 1], np.full(C, focal)], axis=1)
    x_true = np.concatenate([pts.reshape(-1), np.concatenate([rot, trans], 1).reshape(-1),
                             cams_true.reshape(-1)])
    scene = _finish(scene, rng, x_true)
    p3, i6, _ = scene.x_layout()
    # The strip reaches kilometres from the origin, where an angle-axis
    # perturbation of t = -R C would swing cameras by metres; x0 perturbs the
    # rotation and the camera centre instead (a GPS-style initialisation).
    im0 = scene.x0[p3:p3 + i6].reshape(C, 6)
    c0 = centers + rng.normal(0.0, 0.05, (C, 3))
    im0[:, 3:] = -np.einsum("nij,nj->ni", angle_axis_to_matrix(im0[:, :3]), c0)
    scene.x0[p3 + i6:] = np.stack([np.zeros(C), np.zeros(C), np.full(C, focal)], 1).reshape(-1)
    return scene


def window_scene(rng: np.random.Generator, num_cameras: int, num_points: int,
                 kmax_cap: int = 10, focal: float = 1000.0, image_size: float = 1000.0,
                 name="window") -> Scene:
    """One local-BA window: cameras on an arc looking at the scene centre,
    each point seen by k ~ U{2..min(kmax_cap, C)} cameras drawn without
    replacement; 6-parameter cameras with fixed shared intrinsics."""
    C = num_cameras
    ang = np.linspace(0.0, 1.2 * np.pi, C, endpoint=False) + rng.uniform(0, 2 * np.pi)
    centers = np.stack([10 * np.cos(ang), 10 * np.sin(ang),
                        2.0 + rng.normal(0.0, 0.3, C)], axis=1)
    rot, trans = look_at(centers, np.zeros(3))
    pts = rng.uniform([-2, -2, -1], [2, 2, 1], size=(num_points, 3))
    counts = rng.integers(2, min(kmax_cap, C) + 1, num_points)
    subsets = _choose_without_replacement(rng, C, counts)
    # every camera observes at least one point (an unobserved image leaves
    # the reduced system singular, as in the reference)
    seen = np.zeros(C, bool)
    for sub in subsets:
        seen[sub] = True
    for c in np.nonzero(~seen)[0]:
        p = int(rng.integers(num_points))
        subsets[p] = np.sort(np.append(subsets[p], c))
    counts = np.array([len(sub) for sub in subsets])
    pt_ids = np.repeat(np.arange(num_points), counts)
    img_ids = np.concatenate(subsets)
    fm = np.stack([img_ids, pt_ids], axis=1).astype(np.int32)
    intr = np.array([[focal, 0.0, image_size / 2, image_size / 2, 1.0]])
    scene = Scene(num_points=num_points, num_images=C, num_cameras=1, feature_maps=fm,
                  image_camera_map=np.zeros(C, np.int32), comp_mask=COMPUTE_INTRINSIC,
                  free_mask=0, intrinsic_defaults=None, fix_mask=FIX_CAMERAS,
                  fix_cameras=intr.copy(), name=name)
    x_true = np.concatenate([pts.reshape(-1), np.concatenate([rot, trans], 1).reshape(-1)])
    return _finish(scene, rng, x_true)


def window_sizes(rng: np.random.Generator, count: int, cmin=5, cmax=40, pmin=200, pmax=5000):
    cams = rng.integers(cmin, cmax + 1, count)
    pts = np.exp(rng.uniform(math.log(pmin), math.log(pmax), count)).astype(np.int64)
    return cams, pts


def windows_batch(seed: int = SEED_WINDOWS, count: int = 512, scale_points: float = 1.0,
                  name="windows") -> List[Scene]:
    """`count` independent windows: C ~ U{5..40}, P ~ log-U[200, 5000]."""
    rng = np.random.default_rng(seed)
    cams, pts = window_sizes(rng, count)
    out = []
    for w in range(count):
        P = max(8, int(pts[w] * scale_points))
        out.append(window_scene(rng, int(cams[w]), P, name=f"{name}[{w}]"))
    return out


def serving_jobs(seed: int = SEED_SERVING, count: int = 256, mid_fraction: float = 0.05,
                 mid_points: int = 50000, mid_cameras: int = 200) -> List[Scene]:
    """Job mix for the serving loop: windows-distribution jobs plus a
    `mid_fraction` share of mid-class jobs (200 cameras x 9, ~1M keys)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(count):
        if rng.random() < mid_fraction:
            jobs.append(mid_scene(seed=int(rng.integers(1 << 30)), num_cameras=mid_cameras,
                                  num_points=mid_points, name=f"job{j}-mid"))
        else:
            c, p = window_sizes(rng, 1)
            jobs.append(window_scene(rng, int(c[0]), int(p[0]), name=f"job{j}-window"))
    return jobs


def tiny_scene(seed: int = 7, num_images: int = 4, num_points: int = 12, num_cameras: int = 2,
               comp_mask: int = COMPUTE_RADIAL | COMPUTE_DECENTERING | COMPUTE_INTRINSIC,
               free_mask: int = 0, fix_mask: int = 0, constraints: bool = False,
               random_cov: bool = False, kmin: int = 2, name="tiny") -> Scene:
    """Small fully general problem for brute-force checks: several images per
    camera group, every distortion group, optional constraints and a random
    SPD inverse covariance per key."""
    rng = np.random.default_rng(seed)
    I, P, C = num_images, num_points, num_cameras
    ang = 2 * np.pi * np.arange(I) / I
    centers = np.stack([8 * np.cos(ang), 8 * np.sin(ang), 1.5 + 0.3 * np.sin(ang)], 1)
    rot, trans = look_at(centers, np.zeros(3))
    pts = rng.uniform([-1.5, -1.5, -0.7], [1.5, 1.5, 0.7], (P, 3))
    icm = (np.arange(I) % C).astype(np.int32)
    fm = []
    for p in range(P):
        k = int(rng.integers(kmin, I + 1))
        for i in np.sort(rng.choice(I, k, replace=False)):
            fm.append((i, p))
    fm = np.asarray(fm, np.int32)
    perm = rng.permutation(fm.shape[0])  # arbitrary key order
    fm = fm[perm]
    full = _identity_intrinsics(C)
    full[:, K1] = rng.normal(0, 0.02, C)
    full[:, K2] = rng.normal(0, 0.01, C)
    full[:, K3] = rng.normal(0, 0.005, C)
    full[:, D1] = rng.normal(0, 0.001, C)
    full[:, D2] = rng.normal(0, 0.001, C)
    full[:, FOCAL] = 800 + 50 * rng.random(C)
    full[:, SKEW] = rng.normal(0, 0.5, C)
    full[:, PX] = 400 + rng.normal(0, 5, C)
    full[:, PY] = 300 + rng.normal(0, 5, C)
    full[:, PR] = 1 + rng.normal(0, 0.01, C)
    scene = Scene(num_points=P, num_images=I, num_cameras=C, feature_maps=fm,
                  image_camera_map=icm, comp_mask=comp_mask, free_mask=free_mask,
                  intrinsic_defaults=full.copy(), fix_mask=fix_mask, name=name)
    fi = scene.free_indices
    cams = full[:, fi]
    imgs = np.concatenate([rot, trans], 1)
    if fix_mask & FIX_POINTS:
        scene.fix_points = pts.copy()
    if fix_mask & FIX_IMAGES:
        scene.fix_images = imgs.copy()
    if fix_mask & FIX_CAMERAS:
        scene.fix_cameras = cams.copy()
    parts = []
    if not fix_mask & FIX_POINTS:
        parts.append(pts.reshape(-1))
    if not fix_mask & FIX_IMAGES:
        parts.append(imgs.reshape(-1))
    if not fix_mask & FIX_CAMERAS:
        parts.append(cams.reshape(-1))
    x_true = np.concatenate(parts) if parts else np.zeros(0)
    if constraints:
        scene.point_constraints = np.array([[0, 0b101], [min(3, P - 1), 0b010]], np.int64)
        scene.image_constraints = np.array([[0, 0b1111], [min(2, I - 1), 0b0110]], np.int64)
        en = scene.free_bits
        m1 = en & 0b1111100001
        scene.camera_constraints = np.array([[0, m1 if m1 else en]], np.int64)
    if random_cov:
        A = rng.normal(0, 0.3, (fm.shape[0], 2, 2))
        scene.key_blocks = np.einsum("nij,nkj->nik", A, A) + np.eye(2)[None] * 0.5
    uv = evaluate_keys(scene, x_true)
    scene.x_true = x_true
    y = (uv + rng.normal(0, 0.5, uv.shape)).reshape(-1)
    ncons = count_constraint_rows(scene)
    if ncons:
        scene.constraint_weights = 0.5 + rng.random(ncons)
    x0 = x_true.copy()
    if x0.size:
        x0 = x0 + rng.normal(0, 1e-3, x0.shape) * np.maximum(1.0, np.abs(x0)) * 0.5
    scene.x0 = x0
    scene.y_obs = np.concatenate([y, constraint_targets(scene, x_true, rng)])
    return scene


def count_constraint_rows(scene: Scene) -> int:
    n = 0
    if scene.point_constraints is not None and not scene.fix_mask & FIX_POINTS:
        n += sum(bin(int(m) & 7).count("1") for m in scene.point_constraints[:, 1])
    if scene.image_constraints is not None and not scene.fix_mask & FIX_IMAGES:
        for m in scene.image_constraints[:, 1]:
            m = int(m)
            n += (3 if m & 1 else 0) + bin(m & 0b1110).count("1")
    if scene.camera_constraints is not None and not scene.fix_mask & FIX_CAMERAS:
        n += sum(bin(int(m) & 0x3FF).count("1") for m in scene.camera_constraints[:, 1])
    return n


def constraint_targets(scene: Scene, x: np.ndarray, rng) -> np.ndarray:
    """Observed values of the constraint rows: the true parameter plus noise."""
    n = count_constraint_rows(scene)
    if n == 0:
        return np.zeros(0)
    idx = constraint_x_indices(scene)
    return x[idx] + rng.normal(0, 1e-3, n)


def constraint_x_indices(scene: Scene) -> np.ndarray:
    """x index of every constraint row, in y order
    (BA/camera_shared_jacobian_matrix.hpp:391-511)."""
    p, i, _ = scene.x_layout()
    out = []
    if scene.point_constraints is not None and not scene.fix_mask & FIX_POINTS:
        for pid, m in scene.point_constraints:
            out += [3 * int(pid) + b for b in range(3) if int(m) >> b & 1]
    if scene.image_constraints is not None and not scene.fix_mask & FIX_IMAGES:
        for iid, m in scene.image_constraints:
            base = p + 6 * int(iid)
            m = int(m)
            if m & 1:
                out += [base, base + 1, base + 2]
            out += [base + 2 + b for b in range(1, 4) if m >> b & 1]
    if scene.camera_constraints is not None and not scene.fix_mask & FIX_CAMERAS:
        fi = scene.free_indices
        d = len(fi)
        for cid, m in scene.camera_constraints:
            base = p + i + d * int(cid)
            out += [base + fi.index(j) for j in range(10) if int(m) >> j & 1]
    return np.asarray(out, np.int64)
