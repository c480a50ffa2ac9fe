"""The oracle reproduces the committed golden rollouts (tests/golden/,
written by tests/golden/make_golden.py): a regression pin of the oracle
itself.  The same fixtures pin the HIP path in test_gpu_golden.py."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as orc

FIXTURES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "rollout_material*.npz")))


def test_fixtures_present():
    assert len(FIXTURES) == 8


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_reproduces_golden_rollout(path):
    d = np.load(path)
    N = d["x0"].shape[0]
    P = orc.make_params(N, int(d["n_grid"]), material=int(d["material"]), gravity=tuple(d["gravity"]))
    P.alpha = float(d["alpha"])
    cols = [orc.surface_collider(p, q) for p, q in zip(d["col_point"], d["col_normal"])]
    st = orc.OracleState(d["x0"], d["v0"], np.zeros((N, 3, 3)), np.tile(np.eye(3), (N, 1, 1)), d["mu"], d["lam"],
                         d["ys"], d["mass"], d["vol"], precision="f64")
    orc.rollout(P, cols, 0.0, float(d["dt"]), int(d["steps"]), st)
    for k in ("x", "v", "C", "F"):
        ref = d[k]
        assert np.abs(getattr(st, k) - ref).max() <= 1e-10 * max(np.abs(ref).max(), 1.0), k
    assert np.abs(st.ys - d["ys_out"]).max() <= 1e-10 * np.abs(d["ys_out"]).max()
