"""Tuned launch parameters (empm_set_tuning, empm_amd.autotune) on the GPU:
the tuning is validated, each parameter's extreme values give the default
tuning's results (forward positions within 1e-6 relative, summed parameter
gradients within 1e-4 of the sum of their magnitudes, as for two eager runs:
the sums cancel, and float atomic order spreads them), including the dynamic hand-out of blocks (a scene over more
occupied blocks than scatter workgroups), and the autotuner leaves a
configuration of its search space set and caches it."""
import numpy as np
import pytest
import torch

from empm_amd import _lib
from empm_amd import autotune as at
from empm_amd import runtime as rt
from empm_amd import scenes
from empm_amd._lib import EmpmError
from oracle import oracle as orc
from tests.helpers import gpu_collider, rel_err

pytestmark = pytest.mark.gpu


def _setup(dev, n=32, N=6000, material=7):
    # a wide body: more occupied 2x2x2-cell blocks than one scatter workgroup per CU
    sc = scenes.small_scene(n_particles=N, n_grid=n, material=material, seed=5, E=2e4, ys=150.0,
                            center=(0.5, 0.5, 0.45), radii=(0.3, 0.28, 0.25), vel=(0.3, 0.0, -1.0))
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev).clone()
    mu, lam = orc.mu_lam(sc.E.astype(np.float64), sc.nu.astype(np.float64))
    P = orc.make_params(sc.n, n, material=material, gravity=sc.gravity)
    m = rt.ModelArrays(material=material, n_grid=n, grid_lim=1.0, dx=P.dx, inv_dx=P.inv_dx, mu=t(mu), lam=t(lam),
                       yield_stress=t(sc.yield_stress), mass=t(sc.vol * sc.rho), vol=t(sc.vol), gravity=sc.gravity)
    s0 = rt.ParticleState(t(sc.x), t(sc.params["v0"]), torch.zeros((sc.n, 3, 3), device=dev),
                          torch.eye(3, device=dev).repeat(sc.n, 1, 1).contiguous())
    cols = [gpu_collider(orc.surface_collider(c["point"], c["normal"])) for c in sc.colliders]
    return sc, m, s0, cols


def _run(ctx, sc, m, s0, cols, steps=12):
    ro = rt.Rollout(ctx, m, sc.n)
    final, _, _, tape = ro.forward(s0, list(cols), 0.0, sc.dt, steps, tape=True)
    d = rt.ParticleState.empty(sc.n, ctx.device)
    d.zero_()
    d.x.copy_(torch.sin(final.x * 9.0))
    d0, g = ro.backward(tape, d, flags=_lib.EMPM_ADJ_EXACT)
    return final.x.clone(), [(float(t.double().sum()), float(t.double().abs().sum())) for t in (g.mu, g.lam, g.ys)]


def test_tuning_is_validated_and_read_back(dev):
    ctx = rt.Context(16, 100, dev)
    assert ctx.get_tuning() == at.DEFAULT
    for field, bad in (("scatter_wg", 64), ("scatter_per_cu", -1), ("heavy_chunks", -2), ("dynamic_rounds", 0),
                       ("node_grid_cap", 0)):
        with pytest.raises(EmpmError, match=field) as ei:
            ctx.set_tuning(dict(at.DEFAULT, **{field: bad}))
        assert ei.value.code == _lib.EMPM_E_INVALID_ARG
    ctx.set_tuning(dict(at.DEFAULT, scatter_wg=512, node_grid_cap=300))
    assert ctx.get_tuning()["scatter_wg"] == 512 and ctx.get_tuning()["node_grid_cap"] == 300


@pytest.mark.parametrize("change", [{"scatter_wg": 128}, {"scatter_wg": 512}, {"scatter_per_cu": 1},
                                    {"heavy_chunks": 0}, {"heavy_chunks": 1},
                                    {"scatter_per_cu": 1, "dynamic_rounds": 1}, {"node_grid_cap": 3}])
def test_results_do_not_depend_on_the_tuning(dev, change):
    sc, m, s0, cols = _setup(dev)
    ref_ctx = rt.Context(sc.n_grid, sc.n, dev)
    x_ref, g_ref = _run(ref_ctx, sc, m, s0, cols)
    ctx = rt.Context(sc.n_grid, sc.n, dev).set_tuning(dict(at.DEFAULT, **change))
    x, g = _run(ctx, sc, m, s0, cols)
    e = rel_err(x.cpu().numpy(), x_ref.cpu().numpy())
    print(f"[tuning {change}] rel err x {e:.1e}", flush=True)
    assert e < 1e-6
    for (s, a), (s0_, a0) in zip(g, g_ref):
        assert abs(s - s0_) <= 1e-4 * a0 + 1e-30, (change, s, s0_, a0)


def test_autotune_sets_and_caches_a_configuration_of_its_space(dev):
    sc, m, s0, cols = _setup(dev, N=3000)
    ctx = rt.Context(sc.n_grid, sc.n, dev)
    ro = rt.Rollout(ctx, m, sc.n)
    cache = {}
    r = ro.autotune(s0, cols, sc.dt, n_steps=10, backward=True, cache=cache, reps=2)
    print(f"[autotune] {r['tuning']} {r['seconds'] * 1e3:.3f} ms (default {r['default_seconds'] * 1e3:.3f} ms), "
          f"{r['timings']} timings", flush=True)
    assert not r["cached"] and ctx.get_tuning() == r["tuning"]
    for k, v in r["tuning"].items():
        assert v in at.SPACE[k]
    assert r["seconds"] <= r["default_seconds"]
    ctx2 = rt.Context(sc.n_grid, sc.n, dev)
    r2 = rt.Rollout(ctx2, m, sc.n).autotune(s0, cols, sc.dt, n_steps=10, backward=True, cache=cache)
    assert r2["cached"] and ctx2.get_tuning() == r["tuning"]
