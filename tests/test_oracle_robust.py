"""The oracle's Huber loss pinned by what the mathematics fixes:

* with every key inside the threshold (s <= a^2) the robust cost, N and g
  are exactly the plain ones;
* g = 1/2 grad F_robust: central differences of the oracle's robust cost
  (no 1/2 in F, r = f(x) - y) against the gradient the builder forms from the
  reweighted Jacobian — a wrong weight (rho' vs sqrt(rho'), a missing factor,
  the wrong side of the threshold) breaks it;
* N = J^T W J assembled here in numpy from the dense J and per-key weights
  rho'(s) Lambda_k computed here from the residuals;
* a robust fit ignores gross outliers: on a scene with 10 % of the keys moved
  by 40 px, the Huber LM's inlier RMSE beats the plain LM's."""
import numpy as np

from hs_sfm_amd import synthetic as syn
from oracle.oracle import Oracle


def scene(**kw):
    return syn.tiny_scene(fix_mask=0, constraints=True, random_cov=True, **kw)


def threshold_between_keys(sk):
    """A Huber scale with about half the keys beyond it and none within 1 % of
    it (central differences across the kink of rho'' would not be smooth)."""
    v = np.sort(sk)
    m = len(v) // 2
    t = np.sqrt(v[m - 1] * v[m])
    assert v[m - 1] < 0.99 * t and v[m] > 1.01 * t
    return float(np.sqrt(t))


def whitened_s(o, s, x):
    """s_k = r_k^T Lambda_k r_k for every key, from the oracle's y."""
    N = s.num_keys
    r = (o.evaluate(x) - s.y_obs)[:2 * N].reshape(N, 2)
    lam = s.key_blocks if s.key_blocks is not None else np.broadcast_to(np.eye(2), (N, 2, 2))
    return np.einsum("ka,kab,kb->k", r, lam, r), r, lam


def test_huber_with_large_scale_is_plain():
    s = scene()
    plain, rob = Oracle(s), Oracle(s, loss=1, loss_scale=1e6)
    assert whitened_s(plain, s, s.x0)[0].max() < 1e12
    assert rob.mahalanobis(s.x0) == plain.mahalanobis(s.x0)
    b0, b1 = plain.build_normal(s.x0), rob.build_normal(s.x0)
    for k in ("V", "Ui", "Uc", "g"):
        assert np.array_equal(b0[k], b1[k]), k


def test_gradient_is_half_the_robust_cost_gradient():
    s = scene()
    sk, _, _ = whitened_s(Oracle(s), s, s.x0)
    a = threshold_between_keys(sk)
    o = Oracle(s, loss=1, loss_scale=a)
    g = o.build_normal(s.x0)["g"]
    rng = np.random.default_rng(3)
    cols = rng.choice(o.x_size, 24, replace=False)
    for j in cols:
        h = 1e-6 * max(1.0, abs(s.x0[j]))
        xp, xm = s.x0.copy(), s.x0.copy()
        xp[j] += h
        xm[j] -= h
        fd = (o.mahalanobis(xp) - o.mahalanobis(xm)) / (2 * h)
        assert abs(0.5 * fd - g[j]) <= 1e-5 * max(abs(g).max(), 1e-12), j


def test_normal_matrix_against_reweighted_dense():
    s = scene()
    sk, r, lam = whitened_s(Oracle(s), s, s.x0)
    a = threshold_between_keys(sk)
    o = Oracle(s, loss=1, loss_scale=a)
    J, Nd, gd = o.dense_normal(s.x0)
    N = s.num_keys
    w = np.where(sk > a * a, a / np.sqrt(sk), 1.0)
    assert (w < 1).sum() > 0 and (w == 1).sum() > 0
    ny = J.shape[0]
    W = np.zeros((ny, ny))
    for k in range(N):
        W[2 * k:2 * k + 2, 2 * k:2 * k + 2] = w[k] * lam[k]
    cw = s.constraint_weights
    for q in range(ny - 2 * N):
        W[2 * N + q, 2 * N + q] = 1.0 if cw is None else cw[q]
    rr = o.evaluate(s.x0) - s.y_obs
    Nref, gref = J.T @ W @ J, J.T @ W @ rr
    assert np.allclose(Nd, Nref, rtol=1e-12, atol=1e-12 * np.abs(Nref).max())
    assert np.allclose(gd, gref, rtol=1e-12, atol=1e-12 * np.abs(gref).max())
    b = o.build_normal(s.x0)
    P = s.num_points
    V = np.stack([Nref[3 * p:3 * p + 3, 3 * p:3 * p + 3] for p in range(P)])
    assert np.allclose(b["V"], V, rtol=1e-12, atol=1e-12 * np.abs(V).max())
    assert np.allclose(b["g"], gref, rtol=1e-12, atol=1e-12 * np.abs(gref).max())
    # the robust cost itself: sum rho(s) plus the (never robustified) constraint rows
    rho = np.where(sk > a * a, 2 * a * np.sqrt(sk) - a * a, sk)
    rc = rr[2 * N:]
    cons = float(np.sum(rc * rc * (1.0 if cw is None else cw)))
    assert abs(o.mahalanobis(s.x0) - (rho.sum() + cons)) <= 1e-12 * (rho.sum() + cons)


def outlier_scene(seed=5):
    s = syn.ring_scene(num_cameras=10, num_points=300)
    rng = np.random.default_rng(seed)
    N = s.num_keys
    bad = rng.choice(N, N // 10, replace=False)
    y = s.y_obs.copy()
    y[2 * bad] += 40.0 * rng.choice([-1, 1], bad.size)
    y[2 * bad + 1] += 40.0 * rng.choice([-1, 1], bad.size)
    s.y_obs = y
    inl = np.ones(N, bool)
    inl[bad] = False
    return s, inl


def inlier_rmse(o, s, x, inl):
    N = s.num_keys
    y_clean = o.evaluate(s.x_true)[:2 * N].reshape(N, 2)
    r = o.evaluate(x)[:2 * N].reshape(N, 2) - y_clean
    return float(np.sqrt(np.mean(np.sum(r[inl] ** 2, axis=1))))


def test_huber_lm_resists_outliers():
    s, inl = outlier_scene()
    plain = Oracle(s)
    _, xp, stp, trp = plain.lm(max_iterations=30)
    rob = Oracle(s, loss=1, loss_scale=2.0)
    _, xr, str_, trr = rob.lm(max_iterations=30)
    assert stp.status == 0 and str_.status == 0
    # accepted robust costs never increase
    costs = [t["cost"] for t in trr]
    assert all(b <= a for a, b in zip(costs, costs[1:]))
    assert inlier_rmse(plain, s, xr, inl) < 0.5 * inlier_rmse(plain, s, xp, inl)
