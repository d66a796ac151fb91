"""The ba_gcp_constrained restatement pinned independently of itself:

* y and the Mahalanobis distance against numpy closed forms
  (oracle/naive.py gcp_evaluate / gcp_mahalanobis);
* the oracle's N and g with 3x3 point-constraint blocks against J^T W J and
  J^T W r assembled here in numpy from the oracle's dense J (identity rows for
  the GCPs) and a block-diagonal W;
* the reference builder's form: N and g of the GCP problem are those of the
  plain ba_naive problem plus B on each GCP's point block and B r on its
  gradient segment (ba_gcp_constrained_normal_equation_builder.hpp:59-72);
* LM: with exact features and exact GCPs the similarity gauge is fixed, so
  the optimum is x_true itself."""
import numpy as np

from hs_sfm_amd import synthetic as syn
from oracle.naive import (gcp_evaluate, gcp_mahalanobis, gcp_shared_scene, make_gcp_scene, naive_evaluate,
                          naive_from_scene, naive_to_shared)
from oracle.oracle import Oracle


def small(num_gcps=4, seed=7):
    s = syn.ring_scene(num_cameras=5, num_points=40)
    n = naive_from_scene(s)
    return make_gcp_scene(n, num_gcps, seed=seed, gcp_noise=0.01)


def test_evaluate_layout_and_oracle():
    g = small()
    n = g.naive
    y = gcp_evaluate(n.x0, g)
    P, C, G = n.num_points, n.num_cameras, g.num_gcps
    assert y.size == 2 * n.feature_maps.shape[0] + 3 * G
    pts = n.x0[6 * C:].reshape(P, 3)
    assert np.array_equal(y[-3 * G:], pts[P - G:].reshape(-1))
    assert np.array_equal(y[:-3 * G], naive_evaluate(n.x0, C, P, n.feature_maps))
    t = gcp_shared_scene(g)
    assert np.allclose(Oracle(t).evaluate(t.x0), y, rtol=0, atol=1e-15)


def test_mahalanobis_closed_form():
    g = small()
    n = g.naive
    t = gcp_shared_scene(g)
    r = gcp_evaluate(n.x0, g) - g.y_obs
    F = n.feature_maps.shape[0]
    W = np.zeros((r.size, r.size))
    W[:2 * F, :2 * F] = np.eye(2 * F)
    for i in range(g.num_gcps):
        o = 2 * F + 3 * i
        W[o:o + 3, o:o + 3] = g.gcp_blocks[i]
    ref = r @ W @ r
    assert abs(gcp_mahalanobis(n.x0, g) - ref) <= 1e-13 * ref
    assert abs(Oracle(t).mahalanobis(t.x0) - ref) <= 1e-13 * ref


def test_block_weights_against_dense_brute_force():
    g = small()
    t = gcp_shared_scene(g)
    o = Oracle(t)
    J, _, _ = o.dense_normal(t.x0)
    F = t.num_keys
    G = g.num_gcps
    # identity rows on the GCP points
    for i in range(G):
        for a in range(3):
            row = J[2 * F + 3 * i + a]
            assert row[3 * (t.num_points - G + i) + a] == 1.0 and np.count_nonzero(row) == 1
    W = np.zeros((J.shape[0], J.shape[0]))
    W[:2 * F, :2 * F] = np.eye(2 * F)
    for i in range(G):
        s0 = 2 * F + 3 * i
        W[s0:s0 + 3, s0:s0 + 3] = g.gcp_blocks[i]
    r = o.evaluate(t.x0) - t.y_obs
    N = J.T @ W @ J
    gr = J.T @ W @ r
    b = o.build_normal(t.x0)
    P = t.num_points
    V = np.stack([N[3 * p:3 * p + 3, 3 * p:3 * p + 3] for p in range(P)])
    U = np.stack([N[3 * P + 6 * i:3 * P + 6 * i + 6, 3 * P + 6 * i:3 * P + 6 * i + 6] for i in range(t.num_images)])
    assert np.allclose(b["V"], V, rtol=1e-12, atol=1e-12 * np.abs(V).max())
    assert np.allclose(b["Ui"], U, rtol=1e-12, atol=1e-12 * np.abs(U).max())
    assert np.allclose(b["g"], gr, rtol=1e-12, atol=1e-12 * np.abs(gr).max())
    # the oracle's own dense path agrees too
    _, Nd, gd = o.dense_normal(t.x0)
    assert np.allclose(Nd, N, rtol=1e-12, atol=1e-12 * np.abs(N).max())
    assert np.allclose(gd, gr, rtol=1e-12, atol=1e-12 * np.abs(gr).max())


def test_builder_is_naive_plus_gcp_blocks():
    g = small()
    t = gcp_shared_scene(g)
    plain = gcp_shared_scene(g)
    plain.point_constraints = None
    plain.point_constraint_blocks = None
    plain.y_obs = g.y_obs[:2 * t.num_keys].copy()
    b = Oracle(t).build_normal(t.x0)
    b0 = Oracle(plain).build_normal(plain.x0)
    P, G = t.num_points, g.num_gcps
    assert np.array_equal(b["V"][:P - G], b0["V"][:P - G])
    assert np.allclose(b["V"][P - G:] - b0["V"][P - G:], g.gcp_blocks, rtol=0, atol=1e-12)
    r = (t.x0[:3 * P].reshape(P, 3)[P - G:] - g.y_obs[-3 * G:].reshape(G, 3))
    dg = (b["g"] - b0["g"])[:3 * P].reshape(P, 3)
    assert np.array_equal(dg[:P - G], np.zeros((P - G, 3)))
    assert np.allclose(dg[P - G:], np.einsum("gij,gj->gi", g.gcp_blocks, r), rtol=1e-12, atol=1e-15)
    assert np.array_equal(b["Ui"], b0["Ui"])


def test_non_full_or_missing_blocks_rejected():
    g = small()
    t = gcp_shared_scene(g)
    t.point_constraints = t.point_constraints.copy()
    t.point_constraints[0, 1] = 3  # X and Y only: a 3x3 block cannot apply
    try:
        Oracle(t)
    except ValueError:
        return
    raise AssertionError("a partial point constraint with a 3x3 block was accepted")


def test_lm_recovers_truth_with_exact_gcps():
    s = syn.ring_scene(num_cameras=5, num_points=40)
    n = naive_from_scene(s)
    n.y_obs = naive_evaluate(n.x_true, n.num_cameras, n.num_points, n.feature_maps)  # noise-free
    g = make_gcp_scene(n, 4, gcp_noise=0.0)
    t = gcp_shared_scene(g)
    rc, x, st, tr = Oracle(t).lm(max_iterations=50)
    assert rc == 0 and st.status == 0
    assert st.final_cost < 1e-20 * max(st.initial_cost, 1.0)
    assert np.abs(x - naive_to_shared(n.x_true, n.num_cameras, n.num_points)).max() < 1e-8
    costs = [e["cost"] for e in tr]
    assert all(b <= a for a, b in zip(costs, costs[1:]))
