"""The HIP rollout through the C-ABI against the committed FP64 golden
rollouts of tests/golden/ (every material id, 128 particles, 16^3, 20
substeps).  Bar: x and F_trial within 1e-4 relative (max abs error over
max abs value); v and C are reported."""
import glob
import os

import numpy as np
import pytest
import torch

from empm_amd import runtime as rt
from empm_amd._lib import Collider
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

FIXTURES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "rollout_material*.npz")))


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_hip_rollout_matches_golden(dev, path):
    d = np.load(path)
    N, n, m = d["x0"].shape[0], int(d["n_grid"]), int(d["material"])
    print(f"[golden] {os.path.basename(path)} N={N} n={n} material={m}", flush=True)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev).clone()
    model = rt.ModelArrays(material=m, n_grid=n, grid_lim=1.0, dx=float(np.float32(1.0 / n)),
                           inv_dx=float(np.float32(n)), mu=t(d["mu"]), lam=t(d["lam"]), yield_stress=t(d["ys"]),
                           mass=t(d["mass"]), vol=t(d["vol"]), gravity=tuple(float(g) for g in d["gravity"]),
                           alpha=float(d["alpha"]))
    cols = []
    for p, q in zip(d["col_point"], d["col_normal"]):
        c = Collider()
        c.kind, c.start_time, c.end_time = 0, 0.0, 999.0
        for i in range(3):
            c.point[i], c.normal[i] = float(p[i]), float(q[i])
        cols.append(c)
    s0 = rt.ParticleState(t(d["x0"]), t(d["v0"]), torch.zeros((N, 3, 3), device=dev),
                          torch.eye(3, device=dev).repeat(N, 1, 1).contiguous())
    ctx = rt.Context(n, N, dev)
    ro = rt.Rollout(ctx, model, N)
    final, _, _, _ = ro.forward(s0, cols, 0.0, float(d["dt"]), int(d["steps"]))
    torch.cuda.synchronize()
    err = {k: rel_err(getattr(final, k).cpu().numpy(), d[k]) for k in ("x", "v", "C", "F")}
    print("  rel err vs FP64 golden " + " ".join(f"{k}={v:.2e}" for k, v in err.items()), flush=True)
    assert err["x"] < 1e-4 and err["F"] < 1e-4
    if m == 5:
        assert rel_err(model.yield_stress.cpu().numpy(), d["ys_out"]) < 1e-4
