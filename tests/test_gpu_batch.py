"""Batched simulations through the C-ABI (empm_batch_rollout_*): several
scenes with different particle counts, grid sizes and materials advanced by
the same launches must give what each gives alone.

The per-simulation work is the same in both; what differs is the order in
which float atomics land (binning ranks, LDS tile sums), which is not fixed
from one run to the next either.  Bars: 1e-5 relative (max abs difference
over max abs value) on x and F_trial; 1e-4 on dL/dv0; 1e-3 on per-particle
mu / lam gradients (two runs of one simulation alone differ by up to ~1e-4)
and 1e-4 on their sums over the simulation.  Sand (Drucker-Prager) splits
each particle's return map into cases, so round-off moves its per-particle
gradients by up to ~4e-3 and their sums by ~1e-4 between two solo runs: its
bars are 2e-2 per particle and 1e-3 summed."""
import numpy as np
import pytest
import torch

from empm_amd import runtime as rt
from empm_amd import scenes
from empm_amd._lib import EmpmError
from oracle import oracle as orc
from tests.helpers import gpu_collider, rel_err

pytestmark = pytest.mark.gpu


def ground(z=0.1):
    return gpu_collider(orc.surface_collider((0.0, 0.0, z), (0.0, 0.0, 1.0)))


def sim(dev, material, n_grid, n_particles, seed, E=3e4, ys=200.0):
    # bodies hitting the sticky ground at 2 m/s, so that stress carries the gradients
    sc = scenes.small_scene(n_particles=n_particles, n_grid=n_grid, material=material, seed=seed, E=E, ys=ys,
                            center=(0.5, 0.5, 0.22), vel=(0.3, -0.2, -2.0))
    N = sc.n
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev).clone()
    E64, nu64 = sc.E.astype(np.float64), sc.nu.astype(np.float64)
    mu = (E64 / (2 * (1 + nu64))).astype(np.float32)
    lam = (E64 * nu64 / ((1 + nu64) * (1 - 2 * nu64))).astype(np.float32)
    P = orc.make_params(N, n_grid, material=material, gravity=sc.gravity, friction_angle=sc.friction_angle)
    model = rt.ModelArrays(material=material, n_grid=n_grid, grid_lim=1.0, dx=P.dx, inv_dx=P.inv_dx, mu=t(mu),
                           lam=t(lam), yield_stress=t(sc.yield_stress), mass=t(sc.vol * sc.rho), vol=t(sc.vol),
                           gravity=sc.gravity, alpha=P.alpha)
    s0 = rt.ParticleState(t(sc.x), t(sc.params["v0"]), torch.zeros((N, 3, 3), device=dev),
                          torch.eye(3, device=dev).repeat(N, 1, 1).contiguous())
    return model, s0, sc.dt


def cat_states(states):
    return rt.ParticleState(*(torch.cat([getattr(s, f) for s in states]).contiguous() for f in ("x", "v", "C", "F")))


def run_alone(dev, model, s0, dt, steps, tape=False, ckpt=1):
    ctx = rt.Context(model.n_grid, int(model.mu.shape[0]), dev)
    ro = rt.Rollout(ctx, model, int(model.mu.shape[0]))
    return ro, ro.forward(s0, [ground()], 0.0, dt, steps, tape=tape, ckpt_every=ckpt)


def run_batch(dev, models, states, dt, steps, tape=False, ckpt=1):
    bm = rt.BatchModel.concat(models)
    ctx = rt.Context(max(m.n_grid for m in models), bm.n, dev, max_cand=len(models))
    ro = rt.Rollout(ctx, bm, bm.n)
    return ro, bm, ro.forward(cat_states(states), [ground()], 0.0, dt, steps, tape=tape, ckpt_every=ckpt)


SIMS = [(0, 16, 300, 1), (2, 16, 257, 2), (7, 24, 420, 3), (6, 12, 130, 4), (5, 16, 200, 5)]


def test_batch_forward_matches_each_simulation_alone(dev):
    steps = 25
    built = [sim(dev, *a) for a in SIMS]
    models = [b[0] for b in built]
    states = [b[1] for b in built]
    dt = built[0][2]
    print(f"[batch fwd] {len(models)} simulations, N={[int(m.mu.shape[0]) for m in models]}", flush=True)
    _, bm, (final, t1, _, _) = run_batch(dev, [m for m in models], states, dt, steps)
    torch.cuda.synchronize()
    for c, (m, s0) in enumerate(zip(models, states)):
        # material 5 rewrote its arrays in the batch run: rebuild them for the solo run
        m_solo, s_solo, _ = sim(dev, *SIMS[c])
        _, (f1, _, _, _) = run_alone(dev, m_solo, s_solo, dt, steps)
        a, b = bm.offsets[c], bm.offsets[c + 1]
        ex = rel_err(final.x[a:b].cpu().numpy(), f1.x.cpu().numpy())
        eF = rel_err(final.F[a:b].cpu().numpy(), f1.F.cpu().numpy())
        print(f"  simulation {c} material {SIMS[c][0]} n={SIMS[c][1]}: rel err x {ex:.1e} F {eF:.1e}", flush=True)
        assert ex < 1e-5 and eF < 1e-5
    assert abs(t1 - steps * dt) < 1e-12


@pytest.mark.parametrize("ckpt", [1, 6])
def test_batch_gradients_match_each_simulation_alone(dev, ckpt):
    steps = 18
    spec = list(SIMS)  # material 5's (yield stress, mu, lam) are state of the tape
    built = [sim(dev, *a) for a in spec]
    models = [b[0] for b in built]
    states = [b[1] for b in built]
    dt = built[0][2]
    ro, bm, (final, _, _, tape) = run_batch(dev, models, states, dt, steps, tape=True, ckpt=ckpt)
    d = rt.ParticleState.empty(bm.n, dev)
    d.x.copy_(torch.sin(final.x * 7.0))
    d.v.copy_(torch.cos(final.x * 5.0) * 0.1)
    d0, g = ro.backward(tape, d, flags=0)
    torch.cuda.synchronize()
    for c, (m, s0) in enumerate(zip(models, states)):
        a, b = bm.offsets[c], bm.offsets[c + 1]
        r1, (f1, _, _, tp1) = run_alone(dev, m, s0, dt, steps, tape=True, ckpt=ckpt)
        d1 = rt.ParticleState.empty(b - a, dev)
        d1.x.copy_(torch.sin(f1.x * 7.0))
        d1.v.copy_(torch.cos(f1.x * 5.0) * 0.1)
        e0, g1 = r1.backward(tp1, d1, flags=0)
        torch.cuda.synchronize()
        errs, sums = {}, {}
        for k in ("mu", "lam"):
            A, B = getattr(g, k)[a:b].double().cpu().numpy(), getattr(g1, k).double().cpu().numpy()
            errs[k] = rel_err(A, B)
            sums[k] = abs(A.sum() - B.sum()) / np.abs(B).sum()
        errs["dv0"] = rel_err(d0.v[a:b].cpu().numpy(), e0.v.cpu().numpy())
        print(f"  [ckpt {ckpt}] simulation {c} material {spec[c][0]}: " +
              " ".join(f"{k} {v:.1e}" for k, v in errs.items()) +
              " | summed " + " ".join(f"{k} {v:.1e}" for k, v in sums.items()), flush=True)
        bar, bar_sum = (2e-2, 1e-3) if spec[c][0] == 2 else (1e-3, 1e-4)
        assert errs["mu"] < bar and errs["lam"] < bar and errs["dv0"] < 1e-4, errs
        assert all(v < bar_sum for v in sums.values()), sums


def test_batch_backward_needs_the_material_state_of_material_5(dev):
    """Material 5 rewrites (yield stress, mu, lam) in every yielding substep:
    a reverse sweep over states without their recorded material state is
    refused rather than differentiated from the latest values."""
    import ctypes as C

    built = [sim(dev, *a) for a in SIMS[:1] + SIMS[-1:]]
    ro, bm, (final, _, _, tape) = run_batch(dev, [b[0] for b in built], [b[1] for b in built], built[0][2], 4,
                                            tape=True)
    states = [rt.ParticleState(s.x, s.v, s.C, s.F, s.Fe) for s in tape.rec_states]  # mstate dropped
    adj = [rt.ParticleState.empty(bm.n, dev), rt.ParticleState.empty(bm.n, dev)]
    g = rt.ParamGrads.zeros(bm.n, dev)
    with pytest.raises(EmpmError, match="mstate"):
        ro._adj(tape.model.struct(), tape.cols_hist, tape.ncol, tape.times_hist, C.c_double(tape.dt), 4,
                rt._states_array(states), rt._grid_array(tape.rec_grids), rt._states_array(adj), *g.ptrs(), 0,
                rt.stream_handle(dev))


EDGE_SIMS = [(0, 16, 150, 31), (7, 24, 190, 32), (6, 12, 120, 33)]


def _edge_sim(dev, material, n, N, seed):
    """Random particles of one simulation with stencils across the grid's
    faces: base -1 in x (x = -0.7 dx, base = trunc(-1.2)) and base n - 2 in z
    (node n is outside)."""
    from tests.helpers import random_state

    d = random_state(N, n, seed=seed, material=material, deform=0.1 if material == 7 else 0.05)
    d["x"][:5, 0] = np.float32(-0.7 / n)
    d["x"][5:10, 2] = np.float32((n - 1.3) / n)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev).clone()
    P = orc.make_params(N, n, material=material, gravity=(0.0, 0.0, -9.8))
    m = rt.ModelArrays(material=material, n_grid=n, grid_lim=1.0, dx=P.dx, inv_dx=P.inv_dx, mu=t(d["mu"]),
                       lam=t(d["lam"]), yield_stress=t(d["ys"]), mass=t(d["mass"]), vol=t(d["vol"]),
                       gravity=(0.0, 0.0, -9.8), alpha=P.alpha)
    s = rt.ParticleState(t(d["x"]), t(d["v"]), t(d["C"]), t(d["F"]))
    return d, P, m, s


def test_batch_adjoint_with_edge_stencils_matches_alone_and_fd(dev):
    """Simulations on different grids (16^3, 24^3, 12^3) in one batch, each
    with stencils across its grid's faces: the adjoint kernels pick the
    bounds-tested node loads per wave, from the wave's own simulation (its
    grid size and node offset).  One substep's adjoint in the batch must
    equal each simulation's alone (1e-4 on the state cotangents, 1e-3 on
    per-particle mu gradients) and central differences of the FP64
    oracle in a random direction (1e-3)."""
    from tests.helpers import oracle_state

    col = orc.surface_collider((0.0, 0.0, 0.3), (0.0, 0.0, 1.0))
    dt = 1e-4
    built = [_edge_sim(dev, *a) for a in EDGE_SIMS]
    models = [b[2] for b in built]
    bm = rt.BatchModel.concat(models)
    ctx = rt.Context(max(m.n_grid for m in models), bm.n, dev, max_cand=len(models))
    ro = rt.Rollout(ctx, bm, bm.n)
    s0 = cat_states([b[3] for b in built])
    _, _, _, tape = ro.forward(s0, [gpu_collider(col)], 0.0, dt, 1, tape=True)
    rng = np.random.Generator(np.random.PCG64(77))
    w = {k: rng.normal(0, 1, (bm.n,) + sh) for k, sh in (("x", (3,)), ("v", (3,)), ("C", (3, 3)), ("F", (3, 3)))}
    tt = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)
    d_out = rt.ParticleState(tt(w["x"]), tt(w["v"]), tt(w["C"]), tt(w["F"]))
    d0, g = ro.backward(tape, d_out, flags=0)
    torch.cuda.synchronize()
    for c, (d, P, m, s) in enumerate(built):
        a, b = bm.offsets[c], bm.offsets[c + 1]
        ctx1 = rt.Context(m.n_grid, b - a, dev)
        r1 = rt.Rollout(ctx1, m, b - a)
        _, _, _, tp1 = r1.forward(s, [gpu_collider(col)], 0.0, dt, 1, tape=True)
        d1 = rt.ParticleState(*(getattr(d_out, f)[a:b].clone() for f in ("x", "v", "C", "F")))
        e0, g1 = r1.backward(tp1, d1, flags=0)
        torch.cuda.synchronize()
        errs = {f: rel_err(getattr(d0, f)[a:b].cpu().numpy(), getattr(e0, f).cpu().numpy())
                for f in ("x", "v", "C", "F")}
        errs["mu"] = rel_err(g.mu[a:b].cpu().numpy(), g1.mu.cpu().numpy())
        # directional derivative of <w, step(s)> against the FP64 oracle
        wc = {k: v[a:b] for k, v in w.items()}
        u = {k: rng.normal(0, 1, d[k].shape) for k in ("x", "v", "mu")}
        sc_ = {"x": 1e-3 / m.n_grid, "v": 1e-2, "mu": 1e-3 * d["mu"][:, None]}
        an = (np.sum(d0.x[a:b].double().cpu().numpy() * u["x"] * sc_["x"])
              + np.sum(d0.v[a:b].double().cpu().numpy() * u["v"] * sc_["v"])
              + np.sum(g.mu[a:b].double().cpu().numpy() * u["mu"] * sc_["mu"][:, 0]))
        eps = 1e-4

        def J(sign):
            dd = {k: np.asarray(v, np.float64).copy() for k, v in d.items()}
            dd["x"] = dd["x"] + sign * eps * u["x"] * sc_["x"]
            dd["v"] = dd["v"] + sign * eps * u["v"] * sc_["v"]
            dd["mu"] = dd["mu"] + sign * eps * u["mu"] * sc_["mu"][:, 0]
            o = orc.substep(P, [col], 0.0, dt, oracle_state(dd, "f64"))
            return sum(float(np.sum(wc[k] * getattr(o, k))) for k in ("x", "v", "C", "F"))

        fd = (J(1) - J(-1)) / (2 * eps)
        rel = abs(an - fd) / abs(fd)
        print(f"  [edge batch] simulation {c} n={m.n_grid} material {m.material}: batch vs alone "
              + " ".join(f"{k} {v:.1e}" for k, v in errs.items()) + f" | adjoint {an:+.6e} fd {fd:+.6e} "
              f"rel {rel:.1e}", flush=True)
        # per-particle mu gradients carry the order of the float atomics (see the module docstring)
        assert max(errs[f] for f in ("x", "v", "C", "F")) < 1e-4 and errs["mu"] < 1e-3, errs
        assert rel < 1e-3, (an, fd, rel)
