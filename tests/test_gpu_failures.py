"""Failure detection on the device: the per-simulation flags that every
forward substep sets and a rollout reads once (empm_read_status).

The conditions are forced from the host on the inputs, and no kernel faults
on them: a Neo-Hookean particle whose F_trial is inverted (det < 0) gets a
NaN stress from J^(-2/3) (mpm_utils.py:38; the reference's rollout goes on
silently), which spreads NaN through the grid velocities while positions
stay finite (G2P clamps them into [2 dx, 1 - 2 dx], and the clamp returns
the bound for a NaN); and a particle faster than a cell per substep (the
CFL condition), which G2P also clamps into the grid."""
import warnings

import numpy as np
import pytest
import torch

from empm_amd import _lib
from empm_amd import runtime as rt
from empm_amd import scenes
from empm_amd._lib import EmpmError
from oracle import oracle as orc
from tests.helpers import gpu_collider

pytestmark = pytest.mark.gpu


def _sim(dev, material=6, n=16, N=400, seed=3, inverted=0, fast=0, dt=1e-4):
    sc = scenes.small_scene(n_particles=N, n_grid=n, material=material, seed=seed, E=2e4, dt=dt)
    N = sc.n
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev).clone()
    mu, lam = orc.mu_lam(sc.E.astype(np.float64), sc.nu.astype(np.float64))
    P = orc.make_params(N, n, material=material, gravity=sc.gravity)
    m = rt.ModelArrays(material=material, n_grid=n, grid_lim=1.0, dx=P.dx, inv_dx=P.inv_dx, mu=t(mu), lam=t(lam),
                       yield_stress=t(sc.yield_stress), mass=t(sc.vol * sc.rho), vol=t(sc.vol), gravity=sc.gravity)
    F = np.tile(np.eye(3, dtype=np.float32), (N, 1, 1))
    F[:inverted] = np.diag([1.0, 1.0, -1.0]).astype(np.float32)
    v = sc.params["v0"].copy()
    if fast:
        v[:, 0] = 100.0  # the whole body: 0.1 m per substep at dt 1e-3, more than dx = 1/16
    s0 = rt.ParticleState(t(sc.x), t(v), torch.zeros((N, 3, 3), device=dev), t(F))
    cols = [gpu_collider(orc.surface_collider(c["point"], c["normal"])) for c in sc.colliders]
    return sc, m, s0, cols, P, F, v, mu, lam


def test_inverted_neo_hookean_particle_raises_nonfinite(dev):
    sc, m, s0, cols, P, F, v, mu, lam = _sim(dev, inverted=4)
    ctx = rt.Context(sc.n_grid, sc.n, dev)
    ro = rt.Rollout(ctx, m, sc.n)
    print(f"[failure] inverted NH, N={sc.n}", flush=True)
    with pytest.raises(EmpmError, match="non-finite") as ei:
        ro.forward(s0, list(cols), 0.0, sc.dt, 3, check=True)
    assert ei.value.code == _lib.EMPM_E_NONFINITE
    # the oracle (the reference's formulas) turns the same state NaN
    o = orc.OracleState(sc.x, v, np.zeros((sc.n, 3, 3)), F, mu, lam, sc.yield_stress, sc.vol * sc.rho, sc.vol)
    orc.rollout(P, [orc.surface_collider(c["point"], c["normal"]) for c in sc.colliders], 0.0, sc.dt, 3, o)
    assert not np.all(np.isfinite(o.v))
    # the flags were cleared by that read: a healthy rollout on the same context passes
    sc2, m2, s2, cols2, *_ = _sim(dev)
    rt.Rollout(ctx, m2, sc2.n).forward(s2, list(cols2), 0.0, sc2.dt, 3, check=True)
    # unchecked, the flag is there to read
    ro.forward(s0, list(cols), 0.0, sc.dt, 3, check=False)
    rc, st = ctx.read_status()
    assert rc == _lib.EMPM_E_NONFINITE and st[0] & _lib.EMPM_STATUS_NONFINITE


def test_particle_faster_than_a_cell_per_substep_warns(dev):
    sc, m, s0, cols, *_ = _sim(dev, fast=1, dt=1e-3)
    ctx = rt.Context(sc.n_grid, sc.n, dev)
    print(f"[failure] CFL, N={sc.n}", flush=True)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        final, _, _, _ = rt.Rollout(ctx, m, sc.n).forward(s0, list(cols), 0.0, sc.dt, 2, check=True)
    assert any("more than a cell" in str(x.message) for x in w), [str(x.message) for x in w]
    assert torch.isfinite(final.x).all()


def test_batch_flags_name_the_failing_simulation(dev):
    a = _sim(dev, seed=3)
    b = _sim(dev, seed=4, inverted=2)
    bm = rt.BatchModel.concat([a[1], b[1]])
    ctx = rt.Context(16, bm.n, dev, max_cand=2)
    s0 = rt.ParticleState(*(torch.cat([getattr(a[2], f), getattr(b[2], f)]).contiguous() for f in ("x", "v", "C", "F")))
    rt.Rollout(ctx, bm, bm.n).forward(s0, list(a[3]), 0.0, a[0].dt, 3, check=False)
    rc, st = ctx.read_status()
    print(f"[failure] batch status {st}", flush=True)
    assert rc == _lib.EMPM_E_NONFINITE
    assert st[0] == 0 and st[1] & _lib.EMPM_STATUS_NONFINITE
