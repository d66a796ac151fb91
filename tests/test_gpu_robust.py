"""Huber loss (HSBA_LOSS_HUBER) on the device against the oracle's Huber
path (itself pinned in test_oracle_robust.py).

Tolerances: FP64 builder 1e-9 relative, cost 1e-12 relative (summation order
only); mixed builder 2e-5; LM final cost 1e-7 relative in FP64 and
reprojection RMSE within 1e-3 px in mixed precision."""
import numpy as np
import pytest

from hs_sfm_amd import abi, synthetic as syn
from hs_sfm_amd.hsba import Problem
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def median_scale(s):
    """A Huber scale with about half the keys of x0 beyond it."""
    o = Oracle(s)
    N = s.num_keys
    r = (o.evaluate(s.x0) - s.y_obs)[:2 * N].reshape(N, 2)
    lam = s.key_blocks if s.key_blocks is not None else np.broadcast_to(np.eye(2), (N, 2, 2))
    v = np.sort(np.einsum("ka,kab,kb->k", r, lam, r))
    m = len(v) // 2
    return float(np.sqrt(np.sqrt(v[m - 1] * v[m])))  # no key sits on the threshold


@pytest.mark.parametrize("fix,prec,tol", [(syn.FIX_CAMERAS, abi.PRECISION_FP64, 1e-9),
                                          (0, abi.PRECISION_MIXED, 2e-5)])
def test_huber_builder_and_cost(gpu, fix, prec, tol):
    s = syn.tiny_scene(fix_mask=fix, constraints=True, random_cov=True)
    a = median_scale(s)
    o = Oracle(s, loss=abi.LOSS_HUBER, loss_scale=a)
    p = Problem(s, precision=prec, loss=abi.LOSS_HUBER, loss_scale=a)
    print(f"huber builder fix={fix} prec={prec} a={a:.4g}", flush=True)
    bo, bg = o.build_normal(s.x0), p.build_normal(s.x0)
    assert rel(bg["V"], bo["V"]) < tol and rel(bg["Ui"], bo["Ui"]) < tol and rel(bg["g"], bo["g"]) < tol
    if not fix & syn.FIX_CAMERAS:
        assert rel(bg["Uc"], bo["Uc"]) < tol
    fo = o.mahalanobis(s.x0)
    assert abs(p.mahalanobis(s.x0) - fo) <= 1e-12 * fo
    # the robust cost is below the plain one whenever a key is beyond the scale
    assert fo < Oracle(s).mahalanobis(s.x0)


def outlier_ring(seed=5):
    s = syn.ring_scene()
    rng = np.random.default_rng(seed)
    N = s.num_keys
    bad = rng.choice(N, N // 10, replace=False)
    y = s.y_obs.copy()
    y[2 * bad] += 40.0 * rng.choice([-1, 1], bad.size)
    y[2 * bad + 1] += 40.0 * rng.choice([-1, 1], bad.size)
    s.y_obs = y
    return s


@pytest.mark.parametrize("prec", [abi.PRECISION_FP64, abi.PRECISION_MIXED])
def test_huber_lm_ring_with_outliers(gpu, prec):
    s = outlier_ring()
    o = Oracle(s, loss=abi.LOSS_HUBER, loss_scale=2.0)
    p = Problem(s, precision=prec, loss=abi.LOSS_HUBER, loss_scale=2.0)
    print(f"huber lm prec={prec}", flush=True)
    rc_o, xo, st_o, _ = o.lm(max_iterations=20)
    rc_g, xg, st_g = p.lm(max_iterations=20)
    assert rc_o == 0 and rc_g == 0
    assert abs(st_g.initial_cost - st_o.initial_cost) < 1e-9 * st_o.initial_cost
    if prec == abi.PRECISION_FP64:
        assert abs(st_g.final_cost - st_o.final_cost) < 1e-7 * st_o.final_cost
        assert rel(xg, xo) < 1e-5
    else:
        assert abs(syn.reprojection_rmse(s, xg) - syn.reprojection_rmse(s, xo)) < 1e-3
    # the plain fit is pulled by the outliers: its cost under the Huber loss is higher
    _, xp, _, _ = Oracle(s).lm(max_iterations=20)
    assert o.mahalanobis(xg) < o.mahalanobis(xp)
