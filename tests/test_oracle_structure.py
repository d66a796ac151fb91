"""Pins the oracle's sparsity restatements (coupling sets in the reference's
block-id order, the pair structure of the reduced system) against objects
built independently of them:

* every coupling block equals the matching block of the dense N = J^T W J
  assembled from the dense Jacobian (brute force), and the sets hold exactly
  the nonzero blocks of that N;
* the block ids follow first appearance in key order (the std::set insertion
  numbering of camera_shared_normal_equation_builder.hpp:157-204);
* an image pair is listed exactly when the reference's literal Schur loops
  (camera_shared_normal_equation_solver.hpp:198-268) emit a nonzero block for
  it, its incidences are keys of one point on the pair's two images, and the
  incidence count is sum_p k_p (k_p + 1) / 2.
"""
import numpy as np
import pytest

from hs_sfm_amd import synthetic as syn
from oracle.oracle import Oracle

FIX_MASKS = [0, syn.FIX_POINTS, syn.FIX_IMAGES, syn.FIX_CAMERAS,
             syn.FIX_POINTS | syn.FIX_IMAGES, syn.FIX_POINTS | syn.FIX_CAMERAS,
             syn.FIX_IMAGES | syn.FIX_CAMERAS]


def layout(s, d):
    P, I = s.num_points, s.num_images
    pts = not s.fix_mask & syn.FIX_POINTS
    imgs = not s.fix_mask & syn.FIX_IMAGES
    x_img = 3 * P if pts else 0
    x_cam = x_img + (6 * I if imgs else 0)
    return pts, imgs, not s.fix_mask & syn.FIX_CAMERAS and d > 0, x_img, x_cam


def block(N, r0, nr, c0, nc):
    return N[r0:r0 + nr, c0:c0 + nc]


@pytest.mark.parametrize("fix", FIX_MASKS)
def test_coupling_sets_and_blocks_match_dense_normal(fix):
    s = syn.tiny_scene(fix_mask=fix, constraints=True, random_cov=True, num_images=6, num_points=20)
    o = Oracle(s)
    o.build_normal(s.x0)
    _, N, _ = o.dense_normal(s.x0)
    d = o.cam_d
    pts, imgs, cams, x_img, x_cam = layout(s, d)
    fm = s.feature_maps
    icm = s.image_camera_map
    scale = np.abs(N).max()
    kinds = [(0, pts and imgs, 3, lambda r: 3 * r, 6, lambda c: x_img + 6 * c),
             (1, pts and cams, 3, lambda r: 3 * r, d, lambda c: x_cam + d * c),
             (2, imgs and cams, 6, lambda r: x_img + 6 * r, d, lambda c: x_cam + d * c)]
    for which, present, nr, roff, nc, coff in kinds:
        rc, W = o.coupling(which)
        if not present:
            assert len(rc) == 0
            continue
        # each block equals the dense N block
        for q, (r, c) in enumerate(rc):
            np.testing.assert_allclose(W[q], block(N, roff(r), nr, coff(c), nc), rtol=0, atol=1e-11 * scale)
        # the set is exactly the nonzero blocks of N (one per pair)
        rows = s.num_points if which < 2 else s.num_images
        cols = s.num_images if which == 0 else s.num_cameras
        nz = {(r, c) for r in range(rows) for c in range(cols)
              if np.abs(block(N, roff(r), nr, coff(c), nc)).max() > 1e-12 * scale}
        assert set(map(tuple, rc.tolist())) == nz and len(nz) == len(rc)
        # block ids follow the first appearance of each pair in key order
        key_rc = [(p, i) if which == 0 else ((p, icm[i]) if which == 1 else (i, icm[i])) for i, p in fm]
        first = [key_rc.index(tuple(x)) for x in rc.tolist()]
        assert first == sorted(first)
        if which == 0:
            assert [tuple(x) for x in rc.tolist()] == [(p, i) for i, p in fm]


@pytest.mark.parametrize("fix", [0, syn.FIX_IMAGES, syn.FIX_CAMERAS])
def test_pairs_match_literal_schur_blocks(fix):
    s = syn.tiny_scene(fix_mask=fix, constraints=True, random_cov=True, num_images=7, num_points=25)
    o = Oracle(s)
    o.build_normal(s.x0)
    pr, ptr, inc = o.pairs()
    fm = s.feature_maps
    k = np.bincount(fm[:, 1], minlength=s.num_points)
    assert ptr[-1] == len(inc) == int(np.sum(k * (k + 1) // 2))
    for q, (a, b) in enumerate(pr):
        assert a <= b
        prev = -1
        for ka, kb in inc[ptr[q]:ptr[q + 1]]:
            assert fm[ka, 0] == a and fm[kb, 0] == b and fm[ka, 1] == fm[kb, 1]
            assert fm[ka, 1] > prev
            prev = fm[ka, 1]
    # nonzero blocks of the literal reduced system (reference loops)
    mu = 1e-3 * o.max_diagonal()
    rc, S, _ = o.reduced_system(mu, literal=True)
    assert rc == 0
    listed = set(map(tuple, pr.tolist()))
    if not fix & syn.FIX_IMAGES:
        for a in range(s.num_images):
            for b in range(a + 1, s.num_images):
                nz = np.abs(S[6 * a:6 * a + 6, 6 * b:6 * b + 6]).max() > 0
                assert nz == ((a, b) in listed), (a, b)
    else:
        d = o.cam_d
        icm = s.image_camera_map
        cam_pairs = {tuple(sorted((icm[a], icm[b]))) for a, b in listed}
        for c1 in range(s.num_cameras):
            for c2 in range(c1 + 1, s.num_cameras):
                nz = np.abs(S[d * c1:d * c1 + d, d * c2:d * c2 + d]).max() > 0
                assert nz == ((c1, c2) in cam_pairs)


def test_pairs_absent_with_fixed_points():
    s = syn.tiny_scene(fix_mask=syn.FIX_POINTS)
    pr, ptr, inc = Oracle(s).pairs()
    assert len(pr) == 0 and len(inc) == 0 and list(ptr) == [0]
