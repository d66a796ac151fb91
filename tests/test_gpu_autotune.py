"""The autotuner and the two camera-pair kernels it chooses between.

* The three pair kernels (one wavefront per pair on register tiles, one
  workgroup per pair on the matrix cores, one wavefront per pair on the
  matrix cores) give the oracle's reduced system for camera blocks E = 6
  (windows, ring) and E = 9, mixed precision (2e-5 of the largest entry).
* The tuner records a decision per window class (precision, E, incidences
  per pair) and a window solved alone and inside a batch gets the same
  kernel, so the results stay bit-identical.
* The whole LM run is one graph launch; a run of more iterations than one
  graph holds still matches the same run done in pieces.
"""
import numpy as np
import pytest

from hs_sfm_amd import abi, synthetic as syn
from hs_sfm_amd.hsba import Batch, Problem
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def scenes():
    rng = np.random.default_rng(11)
    return {"window_e6": syn.window_scene(rng, 12, 300, name="w"),
            "tiny_e9": syn.tiny_scene(comp_mask=syn.COMPUTE_RADIAL | syn.COMPUTE_INTRINSIC,
                                      free_mask=(1 << syn.K1) | (1 << syn.K2) | (1 << syn.FOCAL),
                                      constraints=True, random_cov=True, num_images=8, num_points=40,
                                      num_cameras=3)}


@pytest.mark.parametrize("kind", ["window_e6", "tiny_e9"])
@pytest.mark.parametrize("kernel", [abi.PAIRS_VALU, abi.PAIRS_MFMA, abi.PAIRS_MFMA_WAVE])
def test_pair_kernels_match_oracle(gpu, kind, kernel):
    s = scenes()[kind]
    print(f"pair kernel {kernel} on {kind}", flush=True)
    o, p = Oracle(s), Problem(s, precision=abi.PRECISION_MIXED, pair_kernel=kernel)
    o.build_normal(s.x0)
    p.build_normal(s.x0)
    mu = 1e-3 * o.max_diagonal()
    rc_o, So, ro = o.reduced_system(mu)
    rc, Sg, rg = p.reduced_system(mu)
    assert rc_o == 0 and rc == 0
    assert rel(Sg, So) < 2e-5 and rel(rg, ro) < 2e-5


def test_tuner_records_classes_and_keeps_alone_equal_to_batch(gpu):
    abi.check(abi.load_library().hsba_tuning_reset(), "hsba_tuning_reset")
    rng = np.random.default_rng(5)
    ws = [syn.window_scene(rng, int(c), int(p), name=f"t{j}")
          for j, (c, p) in enumerate([(8, 200), (30, 2000), (12, 400), (40, 4000)])]
    b = Batch(ws)
    table = abi.tuning_table()
    print("tuning table:", table, flush=True)
    assert table and all(e["camera_block"] == 6 and e["precision"] == "mixed" for e in table)
    assert all(min(e["ms"].values()) > 0 for e in table)  # every candidate was timed
    rc, xs, _ = b.lm(max_iterations=4)
    assert rc == 0
    for s, x in zip(ws, xs):
        p = Problem(s)
        rc1, xa, _ = p.lm(max_iterations=4)
        assert rc1 == 0 and np.array_equal(x, xa), s.name
        p.close()
    assert len(abi.tuning_table()) == len(table)  # the single windows met known classes only


def test_long_run_in_graph_pieces(gpu):
    """More iterations than one graph holds (64): replays plus eager
    remainder give the run of the same length done 5 + 5 + ... at a time."""
    s = syn.ring_scene()
    p = Problem(s, precision=abi.PRECISION_FP64)
    abi.check(p.lm_begin(max_iterations=70), "lm_begin")
    abi.check(p.lm_iterate(70)[0], "lm_iterate")
    rc, x1, st1 = p.lm_end()
    abi.check(p.lm_begin(max_iterations=70), "lm_begin")
    for _ in range(14):
        abi.check(p.lm_iterate(5)[0], "lm_iterate")
    rc, x2, st2 = p.lm_end()
    # (after convergence the LM may stop by itself before 70 iterations; both
    # runs must stop at the same point with the same x)
    assert np.array_equal(x1, x2)
    assert (st1.iterations, st1.accepted, st1.status, st1.stop_reason) == \
        (st2.iterations, st2.accepted, st2.status, st2.stop_reason)
    assert st1.iterations > 5
