"""Ragged and tiny particle counts through the C-ABI.

A particle launch covers whole waves of 64 lanes, the binning and the
block-owned scatters walk cells of a few particles each, and a single
particle leaves one active block and one chunk of one lane.  For N = 1, 7,
63, 65 and 129:
- forward rollouts (NH + von Mises, 30 substeps) against the FP32 oracle
  (bar 1e-4 relative on x and F_trial, 1e-3 on the displacement), and the
  lone particle against its closed-form ballistic path;
- for N = 7 and 65 falling onto the sticky ground, the reverse sweep's
  d loss / d(log E, nu, log Y) against central differences of FP64 oracle
  rollouts (bar 1e-3 relative).
"""
import numpy as np
import pytest
import torch

from empm_amd import runtime as rt
from empm_amd import scenes
from oracle import oracle as orc
from tests.helpers import fmt_errs, gpu_collider, state_errors
from tests.test_gpu_rollout import _loss_grad_gpu, _loss_oracle, setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 7, 63, 65, 129])
def test_ragged_particle_counts_match_oracle(dev, n):
    sc = scenes.small_scene(n_particles=n, n_grid=16, material=7, seed=20 + n, E=3e4, ys=200.0)
    assert sc.n == n
    steps = 30
    print(f"[ragged] forward N={n} steps={steps}", flush=True)
    model, s0, P, ocols, ostate = setup(sc, dev, v0=sc.params["v0"])
    ctx = rt.Context(sc.n_grid, n, dev)
    ro = rt.Rollout(ctx, model, n)
    final, _, _, _ = ro.forward(s0, [gpu_collider(c) for c in ocols], 0.0, sc.dt, steps)
    torch.cuda.synchronize()
    o = ostate("f32")
    orc.rollout(P, list(ocols), 0.0, sc.dt, steps, o)
    e = state_errors({"x": final.x.cpu().numpy(), "v": final.v.cpu().numpy(), "C": final.C.cpu().numpy(),
                      "F": final.F.cpu().numpy()}, o, x0=sc.x)
    print(f"  {fmt_errs(e)}", flush=True)
    assert e["x"] < 1e-4 and e["F"] < 1e-4 and e["disp"] < 1e-3, e
    if n == 1:
        # a lone particle far from the walls moves ballistically: its grid
        # velocities are its own (C = 0 stays 0), so semi-implicit Euler gives
        # x_T = x_0 + T dt v_0 + dt^2 T (T + 1) / 2 g, with F = I and C = 0
        # (the FP32 oracle's C is round-off of 2e-6, hence no relative bar)
        g = np.asarray(sc.gravity, np.float64)
        xt = sc.x[0] + steps * sc.dt * sc.params["v0"][0] + sc.dt ** 2 * steps * (steps + 1) / 2 * g
        dx = np.abs(final.x.cpu().numpy()[0] - xt).max()
        print(f"  lone particle: |x - closed form| {dx:.2e}, max|C| {final.C.abs().max().item():.2e}", flush=True)
        assert dx < 2e-6
        assert final.C.abs().max().item() < 1e-4 and (final.F[0] - torch.eye(3, device=dev)).abs().max().item() < 1e-6


@pytest.mark.parametrize("n,radii,cz", [(7, (0.03, 0.025, 0.02), 0.213), (65, (0.06, 0.05, 0.04), 0.235)])
def test_ragged_particle_counts_gradients_vs_central_differences(dev, n, radii, cz):
    steps = 60
    # a small body just above the sticky ground (3 cells, z = 0.1875) falling
    # onto it: its particles share nodes, deform and yield against it, so the
    # loss depends on (E, nu, Y) at about 1e-3 of its value
    sc = scenes.small_scene(n_particles=n, n_grid=16, material=7, seed=40 + n, E=3e3, ys=60.0, dt=2e-4,
                            center=(0.5, 0.5, cz), radii=radii, vel=(0.5, 0.0, -2.0), spin=1.0)
    base = {"E": 3e3, "nu": 0.3, "ys": 60.0}
    target = sc.x.astype(np.float64) + 0.01 * np.random.default_rng(12).standard_normal((n, 3))
    print(f"[ragged] gradients N={n} steps={steps}", flush=True)
    loss, g = _loss_grad_gpu(sc, dev, target, base, steps, ckpt=7)
    lo = lambda prm: _loss_oracle(sc, target, prm, steps)
    fd = []
    for key, h in (("E", 1e-4), ("nu", 1e-5), ("ys", 1e-4)):
        pp, pm = dict(base), dict(base)
        if key == "nu":
            pp[key] += h
            pm[key] -= h
        else:
            pp[key] *= np.exp(h)
            pm[key] *= np.exp(-h)
        fd.append((lo(pp) - lo(pm)) / (2 * h))
    fd = np.array(fd)
    print(f"  loss {loss:.6e}\n  adjoint d/d(logE, nu, logY) {g}\n  central FD              {fd}", flush=True)
    assert abs(fd[0]) > 1e-4 * loss and abs(fd[2]) > 1e-5 * loss  # the parameters act within the horizon
    for i in range(3):
        assert abs(g[i] - fd[i]) <= 1e-3 * abs(fd[i]) + 1e-9 * np.abs(fd).max(), (i, g[i], fd[i])
