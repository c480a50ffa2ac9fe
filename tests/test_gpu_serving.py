"""Serving engine on the GPU (empm_amd.serving.ServingEngine): each
request's result is the same alone and in a mix of concurrent requests of
other sizes, grids, horizons, materials and modes, including requests that
arrive while others are mid-horizon; and a served request matches the plain
single-rollout path.

Bars as for batched rollouts: 1e-5 relative on final positions (the order of
float atomics is the only difference), 1e-4 on losses, and 1e-3 relative on
the summed parameter gradients (sand, whose return map splits each
particle into cases, 1e-2)."""
import numpy as np
import pytest
import torch

from empm_amd import runtime as rt
from empm_amd import scenes
from empm_amd import serving as sv
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

# (material, grid, particles, substeps, wants gradient, seed)
MIX = [(0, 16, 900, 37, True, 1), (2, 24, 1500, 50, True, 2), (7, 32, 2500, 61, True, 3), (6, 16, 700, 23, False, 4),
       (5, 24, 1200, 44, False, 5), (7, 16, 1000, 16, False, 6), (0, 32, 3000, 29, False, 7),
       (5, 16, 800, 35, True, 10)]
LATE = [(6, 24, 1100, 41, True, 8), (2, 16, 600, 33, False, 9)]


def make_request(material, grid, n, steps, grad, seed):
    dx = 1.0 / grid
    radii = (0.12, 0.1, 0.08)
    # resting on the bounding box's 3-cell floor band, moving down at 2 m/s
    sc = scenes.small_scene(n_particles=n, n_grid=grid, material=material, seed=seed, E=3e4, ys=300.0,
                            center=(0.5, 0.5, 3 * dx + radii[2] + 0.25 * dx), radii=radii, vel=(0.3, -0.2, -2.0))
    return sv.RolloutRequest(x=sc.x, v=sc.params["v0"], vol=sc.vol, material=material, E=3e4 * (1 + 0.1 * seed),
                             nu=0.3, yield_stress=300.0, n_grid=grid, n_steps=steps, want_grad=grad,
                             target=(sc.x + np.float32(0.01)) if grad else None)


def engine(dev, chunk=16):
    return sv.ServingEngine(dev, max_particles=20_000, max_cand=16, max_grid=32, dt=1e-4, chunk=chunk,
                            hbm_budget_bytes=8 << 30)


def alone(dev, spec):
    e = engine(dev)
    rid = e.submit(make_request(*spec))
    res = e.run_until_idle()[rid]
    e.close()
    return res


def compare(spec, a, b):
    ex = rel_err(a.x.cpu().numpy(), b.x.cpu().numpy())
    msg = f"material {spec[0]} grid {spec[1]} N {spec[2]} T {spec[3]}: rel err x {ex:.1e}"
    assert ex < 1e-5, msg
    if spec[4]:
        el = abs(a.loss - b.loss) / abs(b.loss)
        eg = float((a.grad - b.grad).abs().max() / b.grad.abs().max())
        msg += f" loss {el:.1e} grad {eg:.1e}"
        assert el < 1e-4 and eg < (1e-2 if spec[0] == 2 else 1e-3), msg
    print("  " + msg, flush=True)


def test_each_result_is_the_same_alone_and_in_the_mix(dev):
    solo = [alone(dev, s) for s in MIX + LATE]
    e = engine(dev)
    ids = [e.submit(make_request(*s)) for s in MIX]
    e.step()
    e.step()
    # these arrive while the first requests are mid-horizon
    ids += [e.submit(make_request(*s)) for s in LATE]
    res = e.run_until_idle()
    print(f"[serving mix] {len(ids)} requests in {e.ticks} ticks, {e.calls} batched calls", flush=True)
    for spec, rid, ref in zip(MIX + LATE, ids, solo):
        compare(spec, res[rid], ref)
    # co-scheduled: fewer batched calls than the requests' segments
    segs = sum(-(-s[3] // 16) * (2 if s[4] else 1) for s in MIX + LATE)
    assert e.calls < segs
    e.close()


def test_served_request_matches_the_plain_rollout(dev):
    spec = (7, 24, 1500, 40, False, 11)
    r = make_request(*spec)
    res = alone(dev, spec)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)
    E, nu = t(np.full(r.n, r.E)), t(np.full(r.n, r.nu))
    mu, lam = torch.empty_like(E), torch.empty_like(E)
    rt.mu_lam(E, nu, mu, lam)
    from empm_amd.diff_simulator.warp_mpm.mpm_data_structure import friction_alpha
    model = rt.ModelArrays(material=r.material, n_grid=r.n_grid, grid_lim=1.0, dx=float(np.float32(1.0 / r.n_grid)),
                           inv_dx=float(np.float32(r.n_grid)), mu=mu, lam=lam,
                           yield_stress=t(np.full(r.n, r.yield_stress)), mass=t(r.vol * r.rho), vol=t(r.vol),
                           gravity=(0.0, 0.0, -9.8), alpha=friction_alpha(r.friction_angle))
    s0 = rt.ParticleState(t(r.x), t(r.v), torch.zeros((r.n, 3, 3), device=dev),
                          torch.eye(3, device=dev).repeat(r.n, 1, 1).contiguous())
    ctx = rt.Context(r.n_grid, r.n, dev)
    final, _, _, _ = rt.Rollout(ctx, model, r.n).forward(s0, [sv.bounding_box_collider()], 0.0, 1e-4, r.n_steps)
    e = rel_err(res.x.cpu().numpy(), final.x.cpu().numpy())
    print(f"  served vs plain rollout: rel err x {e:.1e}", flush=True)
    assert e < 1e-5


def test_engine_rejects_what_cannot_be_served(dev):
    e = engine(dev)
    with pytest.raises(ValueError):
        e.submit(make_request(2, 64, 500, 10, False, 1))  # grid above the engine's maximum
    e.close()


def test_plasticine_gradient_request_matches_the_plain_rollout(dev):
    """A gradient request of material 5, whose return map rewrites its yield
    stress, mu and lam: the engine re-runs each segment from that
    segment's checkpointed material state and carries the parameter
    cotangents across segments; its loss and gradient equal the plain
    rollout's recorded reverse sweep."""
    spec = (5, 16, 900, 45, True, 12)
    r = make_request(*spec)
    res = alone(dev, spec)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)
    E, nu = t(np.full(r.n, r.E)), t(np.full(r.n, r.nu))
    mu, lam = torch.empty_like(E), torch.empty_like(E)
    rt.mu_lam(E, nu, mu, lam)
    from empm_amd.diff_simulator.warp_mpm.mpm_data_structure import friction_alpha
    ys0 = t(np.full(r.n, r.yield_stress))
    model = rt.ModelArrays(material=5, n_grid=r.n_grid, grid_lim=1.0, dx=float(np.float32(1.0 / r.n_grid)),
                           inv_dx=float(np.float32(r.n_grid)), mu=mu, lam=lam, yield_stress=ys0.clone(),
                           mass=t(r.vol * r.rho), vol=t(r.vol), gravity=(0.0, 0.0, -9.8),
                           alpha=friction_alpha(r.friction_angle), softening=r.softening)
    s0 = rt.ParticleState(t(r.x), t(r.v), torch.zeros((r.n, 3, 3), device=dev),
                          torch.eye(3, device=dev).repeat(r.n, 1, 1).contiguous())
    ctx = rt.Context(r.n_grid, r.n, dev)
    ro = rt.Rollout(ctx, model, r.n)
    final, _, _, tape = ro.forward(s0, [sv.bounding_box_collider()], 0.0, 1e-4, r.n_steps, tape=True)
    diff = final.x - t(r.target)
    loss = float((diff * diff).sum(1).double().mean())
    d = rt.ParticleState.empty(r.n, dev)
    d.zero_()
    d.x.copy_(diff * (2.0 / r.n))
    _, g = ro.backward(tape, d, flags=0)
    dE, dnu = torch.zeros_like(E), torch.zeros_like(E)
    rt.mu_lam_adj(E, nu, g.mu, g.lam, dE, dnu)
    grad = torch.stack([(dE * E).double().sum(), dnu.double().sum(), (g.ys * ys0).double().sum()])
    el = abs(res.loss - loss) / abs(loss)
    eg = float((res.grad - grad).abs().max() / grad.abs().max())
    print(f"  plasticine served vs plain: loss {el:.1e} grad {eg:.1e} (grad {grad.tolist()})", flush=True)
    assert float(model.yield_stress.min()) < 300.0  # it did soften
    assert el < 1e-4 and eg < 1e-3


def test_a_failing_request_is_flagged_and_the_others_are_unaffected(dev):
    """nu = 0.5 makes lambda infinite: the Neo-Hookean stress of that request
    is NaN from its first substep (inf * (J^2 - 1) at J = 1), and its
    velocities stay NaN; the requests batched with it keep their results."""
    good = [(7, 16, 900, 30, False, 21), (0, 24, 1200, 25, True, 22)]
    solo = [alone(dev, s) for s in good]
    e = engine(dev)
    ids = [e.submit(make_request(*s)) for s in good]
    bad = make_request(6, 16, 800, 30, False, 23)
    bad.nu = 0.5
    bad_id = e.submit(bad)
    res = e.run_until_idle()
    print(f"[serving failure] failed flags {[res[i].failed for i in ids + [bad_id]]}", flush=True)
    assert res[bad_id].failed and not torch.isfinite(res[bad_id].v).all()
    for spec, rid, ref in zip(good, ids, solo):
        assert not res[rid].failed
        compare(spec, res[rid], ref)
    e.close()


def test_ticks_are_queued_ahead_of_the_device(dev):
    """step() returns before the device has run the tick (at most
    max_inflight ticks are queued), and the results appear after drain."""
    e = engine(dev)
    ids = [e.submit(make_request(*s)) for s in MIX[:4]]
    e.step()
    assert len(e._inflight) >= 1
    while not e.sched.idle():
        e.step()
        assert len(e._inflight) <= e.max_inflight
    e.drain()
    assert not e._inflight and all(i in e.results for i in ids)
    assert all(e.results[i].finished >= e.results[i].arrival for i in ids)
    e.close()
