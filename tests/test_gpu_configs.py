"""The workloads at their own sizes, against the CPU oracle.

- Single rollout (200k particles in an ellipsoid, 64^3, Neo-Hookean + von
  Mises, dt 5e-5, 400 substeps): forward against the FP32 oracle with the
  FP64 error reported; the gradient w.r.t. (log E, nu, log Y) on the first
  200 substeps against central differences of FP64 oracle rollouts.
- Batched population (32 candidates of 50k-150k particles on their own 64^3
  grids, 300 substeps): two candidates of the batched run (Drucker-Prager
  sand and von Mises plasticine) against the FP32 oracle.
- Serving: requests served in a mix, each checked against the FP32 oracle
  of its own rollout.

Errors are reported as max |d| / max |ref| and, for positions, relative to
the displacement max |x_ref - x0| and as the 99.9th percentile over
particles (plastic materials branch per particle, so a few particles near
the yield surface can take the other branch under FP32 round-off).
"""
import numpy as np
import pytest
import torch

from empm_amd import runtime as rt
from empm_amd import scenes
from empm_amd.diff_simulator.warp_mpm.mpm_data_structure import friction_alpha
from oracle import oracle as orc
from tests.helpers import fmt_errs, gpu_collider, state_errors

pytestmark = pytest.mark.gpu


def _p999(a, b):
    d = np.linalg.norm(np.asarray(a, np.float64).reshape(len(a), -1) - np.asarray(b, np.float64).reshape(len(b), -1),
                       axis=1)
    return float(np.percentile(d, 99.9))


def _c2_inputs(sc, dev, E, nu, ys):
    N = sc.n
    E32, nu32 = np.float32(E), np.float32(nu)
    mu = np.full(N, E32 / (np.float32(2) * (np.float32(1) + nu32)), np.float32)
    lam = np.full(N, E32 * nu32 / ((np.float32(1) + nu32) * (np.float32(1) - np.float32(2) * nu32)), np.float32)
    mass = (sc.vol * sc.rho).astype(np.float32)
    P = orc.make_params(N, sc.n_grid, material=sc.material, gravity=sc.gravity)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev).clone()
    model = rt.ModelArrays(material=sc.material, n_grid=sc.n_grid, grid_lim=1.0, dx=P.dx, inv_dx=P.inv_dx,
                           mu=t(mu), lam=t(lam), yield_stress=t(np.full(N, ys, np.float32)), mass=t(mass),
                           vol=t(sc.vol), gravity=sc.gravity, alpha=P.alpha)
    s0 = rt.ParticleState(t(sc.x), t(sc.params["v0"]), torch.zeros((N, 3, 3), device=dev),
                          torch.eye(3, device=dev).repeat(N, 1, 1).contiguous())
    ocols = [orc.surface_collider(c["point"], c["normal"]) for c in sc.colliders]
    ost = lambda prec, mu_=mu, lam_=lam, ys_=None: orc.OracleState(
        sc.x, sc.params["v0"], np.zeros((N, 3, 3)), np.tile(np.eye(3), (N, 1, 1)), mu_, lam_,
        np.full(N, ys if ys_ is None else ys_), mass, sc.vol, precision=prec)
    return model, s0, P, ocols, ost


@pytest.mark.timeout(600)
def test_single_rollout_workload_full_size_forward(dev):
    sc = scenes.ellipsoid_c2()
    prm = sc.params
    print(f"[c2 forward] N={sc.n} n={sc.n_grid} steps={sc.n_steps} material={sc.material} params={prm['E']:.4g}, "
          f"{prm['nu']:.4g}, {prm['yield_stress']:.4g}", flush=True)
    model, s0, P, ocols, ost = _c2_inputs(sc, dev, prm["E"], prm["nu"], prm["yield_stress"])
    ctx = rt.Context(sc.n_grid, sc.n, dev)
    ro = rt.Rollout(ctx, model, sc.n)
    final, _, _, _ = ro.forward(s0, [gpu_collider(c) for c in ocols], 0.0, sc.dt, sc.n_steps)
    torch.cuda.synchronize()
    g = {k: getattr(final, k).cpu().numpy() for k in ("x", "v", "C", "F")}
    F0 = np.tile(np.eye(3), (sc.n, 1, 1))
    res = {}
    for prec in ("f32", "f64"):
        o = ost(prec)
        orc.rollout(P, list(ocols), 0.0, sc.dt, sc.n_steps, o)
        res[prec] = state_errors(g, o, x0=sc.x, F0=F0)
        res[prec]["x_p999"] = _p999(g["x"], o.x) / float(np.max(np.abs(o.x - sc.x)))
        print(f"  vs {prec} oracle: {fmt_errs(res[prec])}", flush=True)
    e = res["f32"]
    # measured: x 1.1e-6, disp 1.5e-5, F 3.2e-5, x_p999 1.2e-5 of the displacement
    assert e["x"] < 1e-4 and e["F"] < 1e-4, e
    assert e["x_p999"] < 1e-4 and e["disp"] < 1e-3, e


@pytest.mark.timeout(600)
def test_single_rollout_workload_gradient_truncated_horizon(dev):
    """d loss / d(log E, nu, log Y) at full particle count on the first 200
    substeps (half the horizon): loss = mean |x_T - target|^2, target = the
    FP64 oracle rollout with the target parameters moved 5 mm down (a smooth
    offset keeps the residual far above FP32 round-off without making the
    gradient a cancelling sum of random terms); central differences of FP64
    oracle rollouts (bar 1e-3).  If no particle yields in the FP64 run, the
    yield-stress derivative is zero and the adjoint's must vanish with it."""
    steps = 200
    sc = scenes.ellipsoid_c2(n_steps=steps)
    prm = sc.params
    tprm = scenes.sample_params_c2(2)
    N = sc.n
    _, _, P, ocols, ost = _c2_inputs(sc, dev, prm["E"], prm["nu"], prm["yield_stress"])
    mu_lam = lambda E, nu: (np.full(N, E / (2 * (1 + nu))), np.full(N, E * nu / ((1 + nu) * (1 - 2 * nu))))
    mt, lt = mu_lam(tprm["E"], tprm["nu"])
    o = ost("f64", mt, lt, tprm["yield_stress"])
    orc.rollout(P, list(ocols), 0.0, sc.dt, steps, o)
    target = o.x - np.array([0.0, 0.0, 0.005])

    def oracle_loss(E, nu, ys):
        m_, l_ = mu_lam(E, nu)
        s = ost("f64", m_, l_, ys)
        orc.rollout(P, list(ocols), 0.0, sc.dt, steps, s)
        return float(np.mean(np.sum((s.x - target) ** 2, axis=1)))

    model, s0, _, _, _ = _c2_inputs(sc, dev, prm["E"], prm["nu"], prm["yield_stress"])
    ctx = rt.Context(sc.n_grid, N, dev)
    ro = rt.Rollout(ctx, model, N)
    final, _, _, tape = ro.forward(s0, [gpu_collider(c) for c in ocols], 0.0, sc.dt, steps, tape=True, ckpt_every=10)
    tgt = torch.as_tensor(target, dtype=torch.float32, device=dev)
    d = rt.ParticleState.empty(N, dev)
    d.x.copy_(2.0 * (final.x - tgt) / N)
    _, grads = ro.backward(tape, d, flags=0)
    Et = torch.full((N,), float(np.float32(prm["E"])), device=dev)
    nut = torch.full((N,), float(np.float32(prm["nu"])), device=dev)
    dE, dnu = torch.zeros(N, device=dev), torch.zeros(N, device=dev)
    rt.mu_lam_adj(Et, nut, grads.mu, grads.lam, dE, dnu)
    g = np.array([float((dE.double() * Et.double()).sum()), float(dnu.double().sum()),
                  float((grads.ys.double() * prm["yield_stress"]).sum())])
    E, nu, ys = prm["E"], prm["nu"], prm["yield_stress"]
    h = 1e-4
    fd = np.array([
        (oracle_loss(E * np.exp(h), nu, ys) - oracle_loss(E * np.exp(-h), nu, ys)) / (2 * h),
        (oracle_loss(E, nu + 1e-5, ys) - oracle_loss(E, nu - 1e-5, ys)) / 2e-5,
        (oracle_loss(E, nu, ys * np.exp(h)) - oracle_loss(E, nu, ys * np.exp(-h))) / (2 * h),
    ])
    rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-300)
    print(f"[c2 gradient, {steps} substeps] N={N}\n  adjoint d/d(logE, nu, logY) {g}\n  central FD "
          f"             {fd}\n  rel {rel}", flush=True)
    assert rel[0] < 1e-3 and rel[1] < 1e-3, rel
    if fd[2] == 0.0:
        assert abs(g[2]) < 1e-6 * np.abs(g).max(), g
    else:
        assert rel[2] < 1e-3, rel


def _candidate_oracle(ev, bm, c, cands, prec):
    sc = ev.scenes[c]
    a, b = int(ev.offsets[c]), int(ev.offsets[c + 1])
    N = b - a
    P = orc.make_params(N, sc.n_grid, material=sc.material, gravity=sc.gravity, softening=sc.softening)
    P.alpha = friction_alpha(cands[c].friction_angle)
    get = lambda t: t[a:b].double().cpu().numpy()
    o = orc.OracleState(sc.x, sc.params["v0"], np.zeros((N, 3, 3)), np.tile(np.eye(3), (N, 1, 1)), get(bm.mu),
                        get(bm.lam), get(bm.yield_stress), get(bm.mass), get(bm.vol), precision=prec)
    cols = [orc.surface_collider(cc["point"], cc["normal"]) for cc in sc.colliders]
    orc.rollout(P, cols, 0.0, sc.dt, sc.n_steps, o)
    return o


@pytest.mark.timeout(600)
def test_population_candidates_match_oracle(dev):
    """Two candidates of the batched 32-candidate population run (ids 2 and
    5: sand and plasticine, whose return map rewrites its yield stress)
    against the FP32 oracle of each candidate alone."""
    from empm_amd.population import Candidate, PopulationEvaluator

    pop = scenes.population_c3()
    ev = PopulationEvaluator(pop, dev)
    cands = [Candidate.of_scene(sc) for sc in pop]
    bm0, _, _ = ev.model(cands)  # the parameter arrays as the batch sees them (before plasticine rewrites Y)
    bm0 = rt.BatchModel(bm0.sims, bm0.counts, bm0.mu.clone(), bm0.lam.clone(), bm0.yield_stress.clone(), bm0.mass,
                        bm0.vol)
    bm, _, _ = ev.model(cands)
    ro = rt.Rollout(ev.ctx, bm, ev.n)
    final, _, _, _ = ro.forward(ev.initial_state(), list(ev.cols), 0.0, ev.dt, ev.n_steps)
    torch.cuda.synchronize()
    print(f"[population] {ev.n_cand} candidates, {ev.n} particles, {ev.n_steps} substeps", flush=True)
    for c in (0, 1):
        sc = pop[c]
        a, b = int(ev.offsets[c]), int(ev.offsets[c + 1])
        o = _candidate_oracle(ev, bm0, c, cands, "f32")
        g = {k: getattr(final, k)[a:b].cpu().numpy() for k in ("x", "v", "F")}
        e = state_errors(g, o, x0=sc.x)
        e["x_p999"] = _p999(g["x"], o.x) / float(np.max(np.abs(o.x - sc.x)))
        if sc.material == 5:
            e["ys"] = float(np.max(np.abs(bm.yield_stress[a:b].double().cpu().numpy() - o.ys)) / np.max(o.ys))
        print(f"  candidate {c} ({sc.name}): {fmt_errs(e)}", flush=True)
        assert e["x"] < 1e-4 and e["F"] < 1e-4 and e["x_p999"] < 1e-3 and e["disp"] < 1e-3, e
        if sc.material == 5:
            assert e["ys"] < 1e-4, e


@pytest.mark.timeout(600)
def test_served_requests_in_a_mix_match_oracle(dev):
    """Requests of the serving mix (forward and gradient requests, grids 32^3
    and 64^3, different horizons and materials) served together with
    continuous batching; each forward result against the FP32 oracle of its
    own rollout, each gradient request's final state too."""
    from empm_amd.serving import RolloutRequest, ServingEngine

    reqs = scenes.serving_requests_c5(n_requests=6, seed=16, grids=(32, 64), n_range=(10_000, 40_000),
                                      steps_range=(100, 300))
    dt = 5e-5
    eng = ServingEngine(dev, max_particles=400_000, max_cand=8, max_grid=64, dt=dt, chunk=50)
    ids = []
    for r in reqs:
        kw = {k: v for k, v in r.items() if k != "shape"}
        ids.append(eng.submit(RolloutRequest(**kw)))
    eng.step()  # the first tick runs with all six; a late arrival joins the next
    late = dict(reqs[0])
    late.pop("shape")
    late["want_grad"], late["target"] = False, None
    ids.append(eng.submit(RolloutRequest(**late)))
    reqs.append(dict(late, shape=reqs[0]["shape"]))
    res = eng.run_until_idle()
    torch.cuda.synchronize()
    print(f"[serving] {len(ids)} requests, {eng.ticks} ticks, {eng.calls} batched calls", flush=True)
    for rid, r in zip(ids, reqs):
        N = r["x"].shape[0]
        E32, nu32 = np.float32(r["E"]), np.float32(r["nu"])
        mu = np.full(N, E32 / (np.float32(2) * (np.float32(1) + nu32)), np.float32)
        lam = np.full(N, E32 * nu32 / ((np.float32(1) + nu32) * (np.float32(1) - np.float32(2) * nu32)), np.float32)
        P = orc.make_params(N, r["n_grid"], material=r["material"], gravity=(0, 0, -9.8))
        P.alpha = friction_alpha(30.0)
        o = orc.OracleState(r["x"], r["v"], np.zeros((N, 3, 3)), np.tile(np.eye(3), (N, 1, 1)), mu, lam,
                            np.full(N, r["yield_stress"], np.float32), (r["vol"] * np.float32(1000.0)), r["vol"],
                            precision="f32")
        orc.rollout(P, [orc.bounding_box()], 0.0, dt, r["n_steps"], o)
        g = {"x": res[rid].x.cpu().numpy(), "v": res[rid].v.cpu().numpy()}
        e = state_errors(g, o, x0=r["x"])
        e["x_p999"] = _p999(g["x"], o.x) / float(np.max(np.abs(o.x - r["x"])))
        print(f"  request {rid}: N={N} grid {r['n_grid']} steps {r['n_steps']} material {r['material']} "
              f"grad={r['want_grad']}: {fmt_errs(e)}", flush=True)
        assert e["x"] < 1e-4 and e["x_p999"] < 1e-3 and e["disp"] < 1e-3, e
