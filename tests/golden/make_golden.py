"""Writes tests/golden/*.json from the FP64 oracle.

Each fixture's values are pinned independently by the checks in
tests/test_oracle_*.py (complex-step derivatives, dense brute force, closed
forms); the files guard against regressions of the oracle itself.
Run: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from hs_sfm_amd import synthetic as syn  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402
from tests.test_oracle_projection import single_key_scene  # noqa: E402


def main():
    s = single_key_scene(syn.COMPUTE_RADIAL | syn.COMPUTE_DECENTERING | syn.COMPUTE_INTRINSIC)
    o = Oracle(s)
    Jp, Ji, Jc = o.jacobian(s.x0)
    out = {"key": o.evaluate(s.x0).tolist(), "point_block": Jp.reshape(-1).tolist(),
           "image_block": Ji.reshape(-1).tolist(), "camera_block": Jc.reshape(-1).tolist()}
    with open(os.path.join(HERE, "fixed_input.json"), "w") as f:
        json.dump(out, f, indent=1)
    # tiny general problem: normal equations, step and a 3-iteration LM trace
    t = syn.tiny_scene(constraints=True, random_cov=True)
    o = Oracle(t)
    b = o.build_normal(t.x0)
    mu = 1e-3 * o.max_diagonal()
    rc, delta = o.solve(mu)
    rc2, x, st, tr = o.lm(max_iterations=3)
    out = {"x0": t.x0.tolist(), "mu": mu, "g": b["g"].tolist(), "V": b["V"].reshape(-1).tolist(),
           "Ui": b["Ui"].reshape(-1).tolist(), "Uc": b["Uc"].reshape(-1).tolist(),
           "delta": delta.tolist(), "lm_x": x.tolist(), "lm_trace": tr,
           "lm_final_cost": st.final_cost}
    with open(os.path.join(HERE, "tiny_problem.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote golden fixtures")


if __name__ == "__main__":
    main()
