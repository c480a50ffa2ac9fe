"""Long-horizon embodied gradients (empm_amd.long_horizon) on the GPU: a
slab pressed and dragged by a kinematic box gripper whose velocity changes
every frame, a position loss at every frame, checkpoint-recompute in the
reverse sweep, and a batch of parameter sets.

- the batched gradients of each parameter set equal that set's gradient
  computed alone (bar 1e-3 relative on the summed gradients);
- with targets 1 cm off a nominal rollout, the loss of one set matches the
  FP64 oracle's (bar 1e-4 relative) and d loss / d log E matches central
  differences of FP64 oracle rollouts that replay the same gripper schedule
  (bar 1e-3 relative).
"""
import numpy as np
import pytest
import torch

from empm_amd import long_horizon as lh
from empm_amd import scenes
from empm_amd.population import Candidate
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

NOMINAL = Candidate(3e4, 0.3, 500.0)
SETS = [Candidate(2e4, 0.3, 300.0), Candidate(5e4, 0.35, 800.0), Candidate(1e4, 0.25, 200.0)]


def slab(n=2000, grid=24, frames=4, steps=30):
    return scenes.slab_c4(n_particles=n, n_grid=grid, n_frames=frames, frame_steps=steps, dt=1e-4)


def test_batched_sets_match_each_set_alone(dev):
    sc = slab()
    b = lh.LongHorizonBatch(sc, len(SETS), dev, ckpt_every=10)
    b.set_targets([NOMINAL] * len(SETS))
    loss, grad = b.loss_and_grad(SETS)
    torch.cuda.synchronize()
    for c, cand in enumerate(SETS):
        a = lh.LongHorizonBatch(sc, 1, dev, ckpt_every=10)
        a.set_targets([NOMINAL])
        l1, g1 = a.loss_and_grad([cand])
        el = abs(float(loss[c] - l1[0])) / float(l1[0])
        eg = float((grad[c] - g1[0]).abs().max() / g1[0].abs().max())
        print(f"  set {c}: loss {float(loss[c]):.6e} rel err {el:.1e}; grad {grad[c].tolist()} rel err {eg:.1e}",
              flush=True)
        assert el < 1e-4 and eg < 1e-3
        assert float(g1[0].abs().max()) > 0


def test_log_E_gradient_matches_oracle_central_differences(dev):
    sc = slab(n=600, grid=16, frames=4, steps=40)
    cand = Candidate(2e4, 0.3, 300.0)
    b = lh.LongHorizonBatch(sc, 1, dev, ckpt_every=10)
    b.set_targets([NOMINAL])
    # targets 1 cm off the nominal frames: the loss is then not a difference
    # of nearly equal rollouts, which float32 round-off would dominate
    b.targets = [t + 0.01 for t in b.targets]
    loss, grad = b.loss_and_grad([cand])
    targets = [t.double().cpu().numpy() for t in b.targets]
    g = sc.params["gripper"]
    N = sc.n

    def oracle_loss(E):
        mu, lam = orc.mu_lam(np.full(N, E), np.full(N, cand.nu))
        P = orc.make_params(N, sc.n_grid, material=7, gravity=(0, 0, -9.8))
        cols = [orc.surface_collider(c["point"], c["normal"]) for c in sc.colliders]
        cols.append(orc.cuboid(g["point"], g["half"], g["velocity"][0]))
        arr, _ = orc.colliders_array(cols)
        o = orc.OracleState(sc.x, np.zeros((N, 3)), np.zeros((N, 3, 3)), np.tile(np.eye(3), (N, 1, 1)), mu, lam,
                            np.full(N, cand.yield_stress), sc.vol * sc.rho, sc.vol)
        t, total = 0.0, 0.0
        for f in range(sc.params["n_frames"]):
            for i in range(3):
                arr[len(cols) - 1].velocity[i] = float(np.float32(g["velocity"][f][i]))
            t = orc.rollout(P, arr, t, sc.dt, sc.params["frame_steps"], o)
            total += np.mean(np.sum((o.x - targets[f]) ** 2, axis=1))
        return total

    h = 1e-3
    fd = (oracle_loss(cand.E * np.exp(h)) - oracle_loss(cand.E * np.exp(-h))) / (2 * h)
    gE = float(grad[0, 0])
    print(f"[long horizon] loss {float(loss[0]):.6e} (oracle {oracle_loss(cand.E):.6e}); d/dlogE {gE:.6e} vs FD "
          f"{fd:.6e}", flush=True)
    assert abs(float(loss[0]) - oracle_loss(cand.E)) < 1e-4 * abs(oracle_loss(cand.E))
    assert abs(gE - fd) < 1e-3 * abs(fd)
