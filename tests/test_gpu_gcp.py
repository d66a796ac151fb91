"""hsba_gcp_* entry points (ba_gcp_constrained family) against the FP64
oracle in camera_shared form (identity intrinsics, GCPs as point constraints
with 3x3 weight blocks) and the numpy restatement in oracle/naive.py.
Tolerances: FP64 path, 1e-10 relative (summation order only); LM final cost
1e-9 relative, final x 1e-6 relative."""
import numpy as np
import pytest

from hs_sfm_amd import synthetic as syn
from hs_sfm_amd.hsba import GCPProblem
from oracle.naive import (gcp_evaluate, gcp_mahalanobis, gcp_shared_scene, make_gcp_scene, naive_from_scene,
                          naive_to_shared, shared_to_naive)
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def case():
    g = make_gcp_scene(naive_from_scene(syn.ring_scene()), 6, seed=11, gcp_noise=0.01)
    return g, gcp_shared_scene(g)


def problem(g):
    n = g.naive
    return GCPProblem(n.num_cameras, n.num_points, n.feature_maps, g.num_gcps)


def test_gcp_evaluate_and_mahalanobis(gpu, case):
    g, _ = case
    p = problem(g)
    x = g.naive.x0
    assert rel(p.evaluate(x), gcp_evaluate(x, g)) < 1e-12
    ref = gcp_mahalanobis(x, g)
    assert abs(p.mahalanobis(x, g.y_obs, gcp_blocks=g.gcp_blocks) - ref) <= 1e-10 * ref


def test_gcp_normal_and_solve(gpu, case):
    g, t = case
    n = g.naive
    p = problem(g)
    U, V, gr = p.build_normal(n.x0, g.y_obs, gcp_blocks=g.gcp_blocks)
    o = Oracle(t)
    b = o.build_normal(t.x0)
    assert rel(V, b["V"]) < 1e-10 and rel(U, b["Ui"]) < 1e-10
    assert rel(gr, shared_to_naive(b["g"], n.num_cameras, n.num_points)) < 1e-10
    mu = 1e-3 * o.max_diagonal()
    rc, dx = p.solve(mu)
    assert rc == 0
    _, do = o.solve(mu)
    assert rel(dx, shared_to_naive(do, n.num_cameras, n.num_points)) < 1e-8


def test_gcp_rejects_asymmetric_blocks(gpu, case):
    g, _ = case
    p = problem(g)
    B = g.gcp_blocks.copy()
    B[0, 0, 1] += 1e-3
    with pytest.raises(RuntimeError):
        p.build_normal(g.naive.x0, g.y_obs, gcp_blocks=B)


def test_gcp_lm(gpu, case):
    g, t = case
    n = g.naive
    p = problem(g)
    rc, x, st = p.lm(n.x0, g.y_obs, gcp_blocks=g.gcp_blocks, max_iterations=20)
    assert rc == 0 and st.status == 0
    rc2, xo, sto, _ = Oracle(t).lm(max_iterations=20)
    assert rc2 == 0
    assert abs(st.final_cost - sto.final_cost) <= 1e-9 * sto.final_cost
    # once converged, accept / reject turns on cost differences at rounding
    # level, so the accepted counts may differ; the optimum may not
    assert st.iterations == sto.iterations
    assert rel(x, shared_to_naive(xo, n.num_cameras, n.num_points)) < 1e-6
    # the GCPs fix the gauge, so the points move towards the truth itself
    # (no similarity alignment): RMS point error at least halves from x0
    P = n.num_points

    def rms_points(xx):
        return np.sqrt(np.mean((naive_to_shared(xx, n.num_cameras, P) - t.x_true)[:3 * P] ** 2))

    assert rms_points(x) < 0.5 * rms_points(n.x0)
