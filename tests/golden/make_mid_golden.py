"""Writes tests/golden/mid_fp64_oracle.json: the FP64 oracle's 20-iteration LM
run on the 200-camera / 50k-point / ~1M-key scene (about 200 s on one core),
and tests/golden/mid_fp64_oracle_x.npy: its final x (reference x layout).

The bench compares the GPU run's final cost, RMSE and parameters with these.
Run: python tests/golden/make_mid_golden.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from hs_sfm_amd import synthetic as syn  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402


def main():
    s = syn.mid_scene()
    o = Oracle(s)
    t = time.time()
    rc, x, st, _ = o.lm(max_iterations=20)
    dt = time.time() - t
    out = {"scene": "mid (seed %d)" % syn.SEED_MID, "iterations": 20, "rc": rc,
           "final_cost": st.final_cost, "initial_cost": st.initial_cost,
           "accepted": st.accepted, "rmse_px": syn.reprojection_rmse(s, x),
           "mean_error_px": syn.mean_reprojection_error(s, x), "oracle_seconds": dt}
    with open(os.path.join(HERE, "mid_fp64_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)
    np.save(os.path.join(HERE, "mid_fp64_oracle_x.npy"), np.asarray(x, np.float64))
    print(out)


if __name__ == "__main__":
    main()
