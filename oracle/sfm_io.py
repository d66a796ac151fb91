"""Restatement of the problem assembly of the pipeline's bundle adjuster and
writers of the hs_sfm text formats and of BAL files.

TEST INFRASTRUCTURE ONLY: used by tests/ to check the hsba_scene_* loaders,
never by the product path.

* assemble(): BundleAdjustmentOptimizor::operator()
  (include/hs_sfm/sfm_pipeline/bundle_adjustment_optimizor.hpp:59-203):
  tracks with a point (track_point_map), views on images with an extrinsic
  that are not blunders, empty tracks dropped (:71-101); feature_maps =
  (extrinsic id, bundle track) in track order (:104-117); image_camera_map
  over extrinsics (:120-134); x = [points | (angle-axis, -R C) | 10
  intrinsics] (:146-187); y = the keys (:188-203).
* write_*(): the formats of sfm_file_io/*_saver.hpp (counts first, whitespace
  separated; object index maps store id + 1, 0 for none).
* bal_project(): the BAL camera model (Agarwal et al., "Bundle adjustment in
  the large", ECCV 2010): P = R X + t, p = -P / P_z,
  p' = f (1 + k1 |p|^2 + k2 |p|^4) p.
"""
from __future__ import annotations

import numpy as np


def rodrigues(r):
    r = np.asarray(r, np.float64)
    t = np.linalg.norm(r)
    if t < 1e-15:
        return np.eye(3)
    k = r / t
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * K @ K


def assemble(tracks, keysets, intrinsics, extrinsics, points, img_intr, img_extr, track_pt):
    """tracks: list of lists of (image, key, blunder); keysets: list of [n, 2]
    arrays; intrinsics [C, 5] (f skew px py pr); extrinsics [E, 6]
    (angle-axis, centre); points [P, 3]; index maps: lists with -1 = none.
    Returns feature_maps, image_camera_map, x (10 intrinsics per camera), y,
    bundle_point."""
    fm, y, bundle_point = [], [], []
    for t, views in enumerate(tracks):
        if t >= len(track_pt) or track_pt[t] < 0:
            continue
        kept = [(im, key) for im, key, bl in views
                if im < len(img_extr) and img_extr[im] >= 0 and bl != 1]
        if not kept:
            continue
        i = len(bundle_point)
        bundle_point.append(track_pt[t])
        for im, key in kept:
            fm.append((img_extr[im], i))
            y.extend(keysets[im][key])
    E = len(extrinsics)
    extr_img = [-1] * E
    for im, e in enumerate(img_extr):
        if e >= 0:
            extr_img[e] = im
    icm = [img_intr[extr_img[e]] for e in range(E)]
    x = [np.asarray(points)[bundle_point].reshape(-1)]
    for e in range(E):
        r, c = np.asarray(extrinsics[e][:3]), np.asarray(extrinsics[e][3:])
        x.append(np.concatenate([r, -rodrigues(r) @ c]))
    for q in intrinsics:
        f, skew, px, py, pr = q
        x.append(np.array([0, 0, 0, 0, 0, f, skew, px, py, pr], np.float64))
    return (np.asarray(fm, np.int32).reshape(-1, 2), np.asarray(icm, np.int32), np.concatenate(x),
            np.asarray(y, np.float64), bundle_point)


def _fmt(v):
    return repr(float(v))


def write_tracks(path, tracks):
    with open(path, "w") as f:
        f.write(f"{len(tracks)}\n")
        for views in tracks:
            f.write(str(len(views)) + "".join(f" {im} {key} {bl}" for im, key, bl in views) + "\n")


def write_keyset(path, keys):
    with open(path, "w") as f:
        f.write(f"{len(keys)}\n")
        for u, v in keys:
            f.write(f"{_fmt(u)} {_fmt(v)}\n")


def write_rows(path, rows, one_per_line=False):
    with open(path, "w") as f:
        f.write(f"{len(rows)}\n")
        for r in rows:
            f.write(("\n".join(_fmt(v) for v in r) if one_per_line else " ".join(_fmt(v) for v in r)) + "\n")


def write_index_map(path, m):
    with open(path, "w") as f:
        f.write(f"{len(m)}\n")
        for v in m:
            f.write(f"{0 if v < 0 else v + 1}\n")


def read_rows(path, width):
    tok = open(path).read().split()
    n = int(tok[0])
    return np.array([float(t) for t in tok[1:1 + n * width]]).reshape(n, width)


def bal_project(cam, X):
    """BAL camera [angle-axis, t, f, k1, k2] applied to points X [n, 3]."""
    R = rodrigues(cam[:3])
    P = X @ R.T + cam[3:6]
    p = -P[:, :2] / P[:, 2:3]
    n2 = np.sum(p * p, axis=1, keepdims=True)
    return cam[6] * (1 + cam[7] * n2 + cam[8] * n2 * n2) * p


def write_bal(path, cams, pts, obs):
    """obs: list of (camera, point, u, v)."""
    with open(path, "w") as f:
        f.write(f"{len(cams)} {len(pts)} {len(obs)}\n")
        for c, p, u, v in obs:
            f.write(f"{c} {p} {_fmt(u)} {_fmt(v)}\n")
        for cam in cams:
            f.write("\n".join(_fmt(v) for v in cam) + "\n")
        for X in pts:
            f.write("\n".join(_fmt(v) for v in X) + "\n")


def read_bal(path):
    tok = open(path).read().split()
    C, P, N = int(tok[0]), int(tok[1]), int(tok[2])
    o = 3
    obs = []
    for _ in range(N):
        obs.append((int(tok[o]), int(tok[o + 1]), float(tok[o + 2]), float(tok[o + 3])))
        o += 4
    cams = np.array([float(t) for t in tok[o:o + 9 * C]]).reshape(C, 9)
    o += 9 * C
    pts = np.array([float(t) for t in tok[o:o + 3 * P]]).reshape(P, 3)
    return cams, pts, obs
