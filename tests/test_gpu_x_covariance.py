"""hsba_x_covariance: the full dense covariance (every block, x layout)
against the reference's output, restated in oracle/covariance.py: the dense
N's SVD pseudo-inverse with singular values at or below 1e-6 dropped
(camera_shared_analytical_x_covariance_calculator.hpp:45-90).  FP64.

* Free gauge (the seven similarity directions are null): the automatic
  choice is the dense route (Jacobi eigen-decomposition on the device), and
  the number of kept eigenvalues equals numpy's count of singular values
  above 1e-6.
* Gauge fixed by constrained points: the automatic choice is the Schur route
  (N^-1 from S^-1 and matrix-core GEMMs); the dense route, forced, gives the
  same matrix.  Cases include a reduced system of 360 rows (tiled
  factorization and inverse) and more point rows than one GEMM tile.
* The diagonal blocks agree with hsba_covariance.
* Errors: a free gauge beyond the dense route's size limit returns
  HSBA_ERR_TOO_LARGE, a PCG problem HSBA_ERR_UNSUPPORTED.

Tolerance: 1e-8 of the largest entry (LAPACK's SVD in the oracle and the
device's Jacobi / Schur arithmetic differ by about eps * cond; the condition
numbers of the kept spectrum here are 4e2 .. 8e6).
"""
import numpy as np
import pytest

from hs_sfm_amd import abi, synthetic as syn
from hs_sfm_amd.hsba import Problem
from oracle.covariance import dense_normal, x_covariance
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def constrained(s, ids=(0, 1, 2), weight=100.0):
    """s with its similarity gauge fixed: points ids observed in all three
    coordinates."""
    ids = list(ids)
    s.point_constraints = np.array([[i, 7] for i in ids], np.int64)
    pts_true = s.x_true[:3 * s.num_points].reshape(-1, 3)
    s.y_obs = np.concatenate([s.y_obs[:2 * s.num_keys], pts_true[ids].reshape(-1)])
    s.constraint_weights = np.full(3 * len(ids), weight)
    return s


FREE = {
    "tiny": lambda: syn.tiny_scene(fix_mask=syn.FIX_CAMERAS, constraints=False, random_cov=True),
    "ring8": lambda: syn.ring_scene(num_cameras=8, num_points=60),
    "ring60": lambda: syn.ring_scene(num_cameras=60, num_points=150),
}


@pytest.mark.parametrize("case", list(FREE))
def test_free_gauge_dense_route(gpu, case):
    s = FREE[case]()
    o = Oracle(s)
    sv = np.linalg.svd(dense_normal(o, s.x0), compute_uv=False)
    p = Problem(s, precision=abi.PRECISION_FP64)
    print(f"x_covariance free gauge {case} x_size={p.x_size}", flush=True)
    rc, G, info = p.x_covariance()
    assert rc == 0
    assert info["method"] == abi.COV_DENSE
    assert info["rank"] == int((sv > 1e-6).sum()) == p.x_size - 7
    assert abs(info["smallest_kept"] - sv[sv > 1e-6].min()) < 1e-8 * sv.max()
    C = x_covariance(o, s.x0)
    assert rel(G, C) < 1e-8, rel(G, C)
    assert rel(G, G.T) < 1e-10


FIXED = {
    "tiny": lambda: syn.tiny_scene(fix_mask=syn.FIX_CAMERAS, constraints=True, random_cov=True),
    "ring_small": lambda: constrained(syn.ring_scene(num_cameras=8, num_points=60)),     # reduced 48
    "ring_large": lambda: constrained(syn.ring_scene(num_cameras=60, num_points=150)),   # reduced 360: tiled
}


@pytest.mark.parametrize("case", list(FIXED))
def test_fixed_gauge_schur_and_dense_routes(gpu, case):
    s = FIXED[case]()
    o = Oracle(s)
    p = Problem(s, precision=abi.PRECISION_FP64)
    print(f"x_covariance fixed gauge {case} x_size={p.x_size} n_reduced={p.n_reduced}", flush=True)
    C = x_covariance(o, s.x0)
    rc, Gs, info = p.x_covariance()
    assert rc == 0 and info["method"] == abi.COV_SCHUR and info["rank"] == p.x_size
    assert rel(Gs, C) < 1e-8, rel(Gs, C)
    rc, Gd, info_d = p.x_covariance(method=abi.COV_DENSE)
    assert rc == 0 and info_d["method"] == abi.COV_DENSE and info_d["rank"] == p.x_size
    assert rel(Gd, C) < 1e-8, rel(Gd, C)
    # the block entry point's diagonal blocks and reduced inverse
    rc, blk = p.covariance()
    assert rc == 0
    P3 = 3 * s.num_points
    Vd = np.stack([Gs[3 * q:3 * q + 3, 3 * q:3 * q + 3] for q in range(s.num_points)])
    assert rel(Vd, blk["V"]) < 1e-10
    assert rel(Gs[P3:, P3:], blk["reduced"]) < 1e-12


def test_dense_route_size_limit(gpu):
    s = syn.ring_scene(num_cameras=20, num_points=1400)   # free gauge, x_size 4320
    p = Problem(s, precision=abi.PRECISION_FP64)
    assert p.x_size > abi.COV_DENSE_MAX
    print(f"x_covariance limit x_size={p.x_size}", flush=True)
    rc, _, _ = p.x_covariance()
    assert rc == -10  # HSBA_ERR_TOO_LARGE


def test_pcg_problem_unsupported(gpu):
    s = constrained(syn.ring_scene(num_cameras=8, num_points=60))
    p = Problem(s, precision=abi.PRECISION_FP64, solver=abi.SOLVER_PCG)
    rc, _, _ = p.x_covariance()
    assert rc == -8  # HSBA_ERR_UNSUPPORTED
