"""GPU parity of restart_and_compute_F_C (mpm_solver_diff.py:1214-1262)
through the C-ABI (empm_restart_deformation) and the drop-in solver method.

- Random particles, some of them unselected, against the FP32 and FP64
  oracle restatement (oracle_restart_F): F_trial within 1e-5 of FP32.
- The closed forms the oracle is pinned by: a uniform translation gives
  F_trial = I, and a linear displacement A x + b on a lattice symmetric
  about the grid nodes gives I + A for the interior particles.
- The solver method writes F_trial and sets x to the target, leaving C.
"""
import numpy as np
import pytest
import torch

from empm_amd import runtime as rt
from oracle import oracle as orc
from tests.helpers import random_state, rel_err

pytestmark = pytest.mark.gpu


def _model(d, n, dev, sel=None):
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev).clone()
    P = orc.make_params(len(d["x"]), n)
    return rt.ModelArrays(material=0, n_grid=n, grid_lim=1.0, dx=P.dx, inv_dx=P.inv_dx, mu=t(d["mu"]),
                          lam=t(d["lam"]), yield_stress=t(d["ys"]), mass=t(d["mass"]), vol=t(d["vol"]),
                          selection=None if sel is None else t(sel), gravity=(0.0, 0.0, -9.8),
                          grid_v_damping_scale=0.9), P


def _run(dev, d, target, n, sel=None):
    model, P = _model(d, n, dev, sel)
    ctx = rt.Context(n, len(d["x"]), dev)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev).clone()
    F = t(d["F"])
    rt.restart_F(ctx, model, t(d["x"]), t(target), F)
    torch.cuda.synchronize()
    return F.cpu().numpy(), P


def test_restart_matches_oracle_with_unselected_particles(dev):
    n, N = 24, 3000
    print(f"[restart] random particles N={N} n={n}", flush=True)
    d = random_state(N, n, seed=31, material=0)
    rng = np.random.Generator(np.random.PCG64(32))
    target = (d["x"] + rng.normal(0, 0.3 / n, (N, 3))).astype(np.float32)
    sel = np.zeros(N, np.int32)
    sel[::7] = 1
    # damping 0.9 on the model and gravity are ignored by the restart (dt = 0, no damping pass)
    F, P = _run(dev, d, target, n, sel)
    o32 = orc.restart_F(P, d["x"], target, d["mass"], d["F"], sel, precision="f32")
    o64 = orc.restart_F(P, d["x"], target, d["mass"], d["F"], sel, precision="f64")
    e32, e64 = rel_err(F, o32), rel_err(F, o64)
    print(f"[restart] rel err F vs FP32 {e32:.2e}, vs FP64 {e64:.2e}", flush=True)
    assert e32 < 1e-5 and e64 < 1e-5
    assert np.array_equal(F[sel == 1], d["F"][sel == 1])
    assert np.abs(F[sel == 0] - np.eye(3)).max() > 1e-3  # the displacement is seen


def test_restart_closed_forms(dev):
    n = 16
    dx = 1.0 / n
    c = (np.arange(4, 28) + 0.5) * dx / 2
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    N = len(x)
    d = random_state(N, n, seed=33, material=0)
    d["x"] = x.astype(np.float32)
    d["mass"] = np.full(N, 1e-3, np.float32)
    d["F"] = np.tile(np.eye(3, dtype=np.float32), (N, 1, 1))
    u = np.array([0.011, -0.017, 0.006], np.float32)
    F, _ = _run(dev, d, d["x"] + u, n)
    print(f"[restart] translation: max |F - I| {np.abs(F - np.eye(3)).max():.2e}", flush=True)
    assert np.abs(F - np.eye(3)).max() < 1e-5
    A = np.array([[0.01, 0.02, -0.005], [0.003, -0.01, 0.004], [-0.02, 0.006, 0.015]])
    b = np.array([0.001, -0.002, 0.0005])
    F, _ = _run(dev, d, (x + x @ A.T + b).astype(np.float32), n)
    inner = np.all((x > c.min() + 3 * dx) & (x < c.max() - 3 * dx), axis=1)
    err = np.abs(F[inner] - (np.eye(3) + A)).max()
    print(f"[restart] linear field: max |F - (I + A)| over {inner.sum()} interior particles {err:.2e}", flush=True)
    assert err < 2e-5


def test_solver_restart_sets_F_trial_and_positions(dev):
    from empm_amd.diff_simulator.warp_mpm.mpm_data_structure import MPMModelStruct, MPMStateStruct
    from empm_amd.diff_simulator.warp_mpm.mpm_solver_diff import MPMWARPDiff

    n, N = 16, 500
    d = random_state(N, n, seed=34, material=0)
    rng = np.random.Generator(np.random.PCG64(35))
    target = (d["x"] + rng.normal(0, 0.2 / n, (N, 3))).astype(np.float32)
    state, model = MPMStateStruct(), MPMModelStruct()
    state.init(N, device=str(dev), requires_grad=False)
    state.init_grid(grid_res=n, device=str(dev))
    model.init(N, device=str(dev), requires_grad=False)
    model.init_other_params(n_grid=n, grid_lim=1.0, device=str(dev))
    state.reset_state(torch.as_tensor(d["x"], device=dev), None, None, device=str(dev))
    state.particle_mass.copy_(torch.as_tensor(d["mass"], device=dev))
    C_before = torch.randn(N, 3, 3, device=dev)
    state.particle_C.copy_(C_before)
    solver = MPMWARPDiff()
    solver.initialize(N, n_grid=n, grid_lim=1.0, device=str(dev))
    solver.restart_and_compute_F_C(model, state, torch.as_tensor(target, device=dev), str(dev))
    torch.cuda.synchronize()
    P = orc.make_params(N, n)
    o = orc.restart_F(P, d["x"], target, d["mass"], np.tile(np.eye(3), (N, 1, 1)), precision="f32")
    e = rel_err(state.particle_F_trial.cpu().numpy(), o)
    print(f"[restart] solver method: rel err F_trial {e:.2e}", flush=True)
    assert e < 1e-5
    assert torch.equal(state.particle_x, torch.as_tensor(target, device=dev))
    assert torch.equal(state.particle_C, C_before)
