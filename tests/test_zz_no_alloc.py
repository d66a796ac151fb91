"""hsba_covariance_static (no device allocation, fixed per-GPU scratch) against
hsba_covariance on the same problems: two calls in a row, the second with a
smaller problem, both bit-identical to the allocating entry point (same
kernels, same inputs).  Also: the limits are refused with HSBA_ERR_TOO_LARGE.
"""
import numpy as np
import pytest

from hs_sfm_amd import abi, synthetic as syn
from hs_sfm_amd.hsba import Problem

pytestmark = pytest.mark.gpu


def gauge_fixed_ring(num_cameras, num_points):
    """A ring scene whose similarity gauge is fixed by three points observed
    in all three coordinates (weight 100): N is nonsingular."""
    s = syn.ring_scene(num_cameras=num_cameras, num_points=num_points)
    ids = [0, 1, 2]
    s.point_constraints = np.array([[i, 7] for i in ids], np.int64)
    pts_true = s.x_true[:3 * s.num_points].reshape(-1, 3)
    s.y_obs = np.concatenate([s.y_obs[:2 * s.num_keys], pts_true[ids].reshape(-1)])
    s.constraint_weights = np.full(9, 100.0)
    return s


def same(a, b):
    return all(np.array_equal(a[k], b[k]) for k in ("V", "Ui", "Uc", "reduced"))


@pytest.mark.parametrize("prec", [abi.PRECISION_FP64, abi.PRECISION_MIXED])
def test_static_covariance_matches_allocating(gpu, prec):
    big = gauge_fixed_ring(60, 150)     # reduced system 360 (tiled factor, ld 384)
    small = syn.tiny_scene(fix_mask=syn.FIX_CAMERAS, constraints=True, random_cov=True)
    for s in (big, small):              # the second call has the smaller input
        p = Problem(s, precision=prec)
        print(f"covariance_static prec={prec} n_reduced={p.n_reduced} points={s.num_points}", flush=True)
        rc_a, ref = p.covariance()
        rc_s, got = p.covariance(static=True)
        assert rc_a == 0 and rc_s == 0
        assert same(got, ref)
        p.close()


def test_static_covariance_limits(gpu):
    """A reduced system beyond 512 padded rows is refused, nothing runs."""
    s = gauge_fixed_ring(90, 100)       # 540 reduced rows -> ld 576 > 512
    p = Problem(s, precision=abi.PRECISION_FP64)
    rc, _ = p.covariance(static=True)
    assert rc == -10  # HSBA_ERR_TOO_LARGE
    rc, _ = p.covariance()
    assert rc == 0                      # the allocating entry point still serves it
    p.close()
