"""The drop-in Python API (empm_amd.diff_simulator, mirroring
phys_sim.diff_simulator) on the GPU: the reference's own call sequences
(warp_mpm_wrapper.py:135-306, interface.py:20-986, trainer_warp_mpm.py:332-400)
against the FP32 oracle and central differences."""
import numpy as np
import pytest
import torch

from empm_amd import runtime as rt
from empm_amd import scenes
from empm_amd.diff_simulator import WarpMPMWrapper
from empm_amd.diff_simulator.warp_mpm.interface import (MPMDifferentiableSimulation,
                                                        MPMDifferentiableSimulationClean,
                                                        MPMDifferentiableSimulationTrail,
                                                        MPMDifferentiableSimulationWCheckpoint)
from empm_amd.diff_simulator.warp_mpm.warp_mpm_wrapper import material_gallery
from oracle import oracle as orc
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _wrapper(sc, material="jelly", requires_grad=False, frame_dt=None, colliders=True, params=None):
    mpm = WarpMPMWrapper(grid_size=sc.n_grid, grid_lim=1.0, frame_dt=frame_dt or sc.dt * 10, sim_dt=sc.dt,
                         device="cuda:0", requires_grad=requires_grad)
    mpm.material_params = dict(material_gallery[material])
    mpm.material_params.update(params or {})
    mpm.add_particles(sc.x, sc.params.get("v0"), sc.vol.astype(np.float32), sc.E.astype(np.float32),
                      sc.nu.astype(np.float32), sc.rho.astype(np.float32))
    if colliders:
        for c in sc.colliders:
            mpm.surface_colliders.append([c["point"], c["normal"], "sticky", 0.0])
    mpm.initialize()
    return mpm


def _oracle_rollout(sc, steps, material_id, v0=None, prec="f32", cols=None):
    N = sc.n
    mu, lam = orc.mu_lam(sc.E.astype(np.float64), sc.nu.astype(np.float64))
    P = orc.make_params(N, sc.n_grid, material=material_id, gravity=(0, 0, -9.8))
    ocols = cols if cols is not None else [orc.surface_collider(c["point"], c["normal"]) for c in sc.colliders]
    o = orc.OracleState(sc.x, np.zeros((N, 3)) if v0 is None else v0, np.zeros((N, 3, 3)),
                        np.tile(np.eye(3), (N, 1, 1)), mu, lam, sc.yield_stress, sc.vol * sc.rho, sc.vol,
                        precision=prec)
    orc.rollout(P, ocols, 0.0, sc.dt, steps, o)
    return o


def test_wrapper_advance_to_matches_oracle():
    sc = scenes.cube_c1()
    mpm = _wrapper(sc)
    frames = mpm.advance_to(n_frames=5)  # 50 substeps
    assert frames.shape == (6, sc.n, 3)
    o = _oracle_rollout(sc, 50, 0)
    err = rel_err(frames[-1].cpu().numpy(), o.x)
    print(f"[wrapper advance_to] N={sc.n} rel err x {err:.2e}", flush=True)
    assert err < 1e-4
    assert abs(mpm.mpm_solver.time - 50 * sc.dt) < 1e-12


def test_wrapper_step_callback_path_equals_native_loop():
    sc = scenes.small_scene(n_particles=300, n_grid=16, material=0)
    a = _wrapper(sc)
    fa = a.advance_to(3)
    b = _wrapper(sc)
    seen = []
    b.step_callback = lambda m, k: seen.append(k)
    fb = b.advance_to(3)
    assert seen == list(range(30))
    assert torch.equal(fa, fb) or rel_err(fa.cpu().numpy(), fb.cpu().numpy()) < 1e-6


def test_moving_cuboid_frame_callback_and_point_advance():
    """trainer_warp_mpm.py:365-397: ground plane + a velocity cuboid whose
    velocity a frame callback sets; the cuboid point advances on the host."""
    sc = scenes.small_scene(n_particles=400, n_grid=16, material=0)
    mpm = WarpMPMWrapper(grid_size=16, frame_dt=sc.dt * 5, sim_dt=sc.dt)
    mpm.material_params = dict(material_gallery["jelly"])
    mpm.material_params["grid_v_damping_scale"] = 0.98
    mpm.add_particles(sc.x, None, float(sc.vol[0]), float(sc.E[0]), float(sc.nu[0]), float(sc.rho[0]))
    mpm.surface_colliders.append([(0.0, 0.0, 0.3), (0.0, 0.0, 1.0), "sticky", 0.0])
    mpm.moving_boundaries.append([(0.5, 0.5, 0.55), 0.06, (0.0, 0.0, 0.0), 0.0])

    def cb(m, f):
        m.mpm_solver.collider_params[1].velocity = (0.1 * (f + 1), 0.0, -0.2)

    mpm.frame_callback = cb
    mpm.initialize()
    p0 = mpm.mpm_solver.collider_params[1].point
    frames = mpm.advance_to(4)
    p1 = mpm.mpm_solver.collider_params[1].point
    exp_dx = sum(0.1 * (f + 1) for f in range(4)) * 5 * sc.dt
    assert abs((p1[0] - p0[0]) - exp_dx) < 1e-6
    # oracle replay of the same schedule
    N = sc.n
    cols = [orc.surface_collider((0, 0, 0.3), (0, 0, 1)), orc.cuboid((0.5, 0.5, 0.55), 0.06, (0, 0, 0))]
    mu, lam = orc.mu_lam(sc.E.astype(np.float64), sc.nu.astype(np.float64))
    P = orc.make_params(N, 16, material=0, gravity=(0, 0, -9.8), grid_v_damping_scale=0.98)
    o = orc.OracleState(sc.x, np.zeros((N, 3)), np.zeros((N, 3, 3)), np.tile(np.eye(3), (N, 1, 1)), mu, lam,
                        sc.yield_stress, sc.vol * sc.rho, sc.vol, precision="f32")
    arr, nc = orc.colliders_array(cols)
    t = 0.0
    for f in range(4):
        arr[1].velocity[0] = float(np.float32(0.1 * (f + 1)))
        arr[1].velocity[2] = float(np.float32(-0.2))
        t = orc.rollout(P, arr, t, sc.dt, 5, o)
    err = rel_err(frames[-1].cpu().numpy(), o.x)
    print(f"[cuboid + frame callback] rel err {err:.2e}", flush=True)
    assert err < 1e-4


def _trail_setup(sc, material="neo-hookean"):
    mpm = _wrapper(sc, material=material, requires_grad=True)
    return mpm


def test_trail_gradients_exact_mode_match_fd():
    """Scalar E (x1e6 inside the Trail, interface.py:70) and nu; the returned
    grads are means over particles (interface.py:207-232)."""
    sc = scenes.small_scene(n_particles=250, n_grid=16, material=6, seed=2, E=3e3, dt=2e-4, center=(0.5, 0.5, 0.27),
                            vel=(0.5, 0.0, -2.0), spin=1.0)
    mpm = _trail_setup(sc)
    mpm.mpm_solver.adjoint_mode = "exact"
    mpm.mpm_solver.checkpoint_every = 4
    N = sc.n
    x0 = torch.as_tensor(sc.x, device="cuda:0")
    v0 = torch.as_tensor(sc.params["v0"], device="cuda:0")
    target = (x0 + torch.tensor([0.01, -0.02, 0.0], device="cuda:0")).detach()
    rho = torch.full((N,), 1000.0, device="cuda:0")
    mask = torch.ones(N, dtype=torch.bool, device="cuda:0")

    def loss_of(E_val, nu_val, need_grad):
        E = torch.tensor(E_val, device="cuda:0", requires_grad=need_grad)
        nu = torch.tensor(nu_val, device="cuda:0", requires_grad=need_grad)
        mpm.mpm_solver.time = 0.0
        pos = MPMDifferentiableSimulationTrail.apply(mpm.mpm_solver, mpm.mpm_state, mpm.mpm_model, 10, sc.dt, 30, x0,
                                                    v0, E, nu, rho, mask, None, "cuda:0", True, 0)
        loss = ((pos - target) ** 2).sum(-1).mean()
        if need_grad:
            loss.backward()
            return float(loss), float(E.grad), float(nu.grad), pos
        return float(loss)

    loss, gE, gnu, pos = loss_of(0.003, 0.3, True)
    assert pos.shape == (3, N, 3)
    # FD on the GPU forward in FP32 is too noisy; use the FP64 oracle with the
    # same frames (states after substeps 1, 11, 21)
    def oracle_loss(Ephys, nu):
        mu = np.full(N, Ephys / (2 * (1 + nu)))
        lam = np.full(N, Ephys * nu / ((1 + nu) * (1 - 2 * nu)))
        P = orc.make_params(N, 16, material=6, gravity=(0, 0, -9.8))
        cols = [orc.surface_collider(c["point"], c["normal"]) for c in sc.colliders]
        o = orc.OracleState(sc.x, sc.params["v0"], np.zeros((N, 3, 3)), np.tile(np.eye(3), (N, 1, 1)), mu, lam,
                            sc.yield_stress, sc.vol * sc.rho, sc.vol)
        arr, _ = orc.colliders_array(cols)
        tg = target.cpu().numpy().astype(np.float64)
        total, t, done = 0.0, 0.0, 0
        for k in (1, 11, 21):
            t = orc.rollout(P, arr, t, sc.dt, k - done, o)
            done = k
            total += np.mean(np.sum((o.x - tg) ** 2, axis=1))
        return total / 3  # the Trail loss: mean over (frames, particles)

    h = 1e-4
    Ep = 0.003e6
    fdE = (oracle_loss(Ep * (1 + h), 0.3) - oracle_loss(Ep * (1 - h), 0.3)) / (2 * h * Ep)
    fdnu = (oracle_loss(Ep, 0.3 + 1e-5) - oracle_loss(Ep, 0.3 - 1e-5)) / 2e-5
    # Trail returns the mean over particles of per-particle grads
    print(f"[trail exact] loss {loss:.6e}  E.grad {gE:.6e} vs FD/N {fdE / N:.6e}  nu.grad {gnu:.6e} vs FD/N "
          f"{fdnu / N:.6e}", flush=True)
    assert abs(gE - fdE / N) < 1e-3 * abs(fdE / N)
    assert abs(gnu - fdnu / N) < 1e-3 * abs(fdnu / N)


def test_trail_reference_compat_matches_frozen_mass_fd():
    """The default adjoint mode (reference_compat): the Trail sums its frames'
    output gradients onto the last state (interface.py:176-197) and the grid
    mass carries no adjoint (mpm_data_structure.py:100-105,
    warp_utils.py:116-120).  With a loss sum_f <G_f, pos_f> of fixed G_f, the
    gradient the reference's Tape returns is d/dE of <sum_f G_f, x_T(E)> for
    the rollout whose every normalisation divides by the grid mass of the
    unperturbed run: central differences of FP64 oracle rollouts fed that
    mass history (bar 1e-3).  Scalar E enters x1e6 and comes back as the mean
    of per-particle physical-E gradients (interface.py:68-70,207-232)."""
    sc = scenes.small_scene(n_particles=250, n_grid=16, material=6, seed=2, E=3e3, dt=2e-4, center=(0.5, 0.5, 0.27),
                            vel=(0.5, 0.0, -2.0), spin=1.0)
    mpm = _trail_setup(sc)
    assert mpm.mpm_solver.adjoint_mode == "reference_compat"
    mpm.mpm_solver.checkpoint_every = 4
    N, steps, every = sc.n, 30, 10
    dev = "cuda:0"
    x0 = torch.as_tensor(sc.x, device=dev)
    v0 = torch.as_tensor(sc.params["v0"], device=dev)
    G = torch.as_tensor(np.random.default_rng(5).standard_normal((3, N, 3)), dtype=torch.float32, device=dev)
    rho = torch.full((N,), 1000.0, device=dev)
    E = torch.tensor(0.003, device=dev, requires_grad=True)
    nu = torch.tensor(0.3, device=dev, requires_grad=True)
    mpm.mpm_solver.time = 0.0
    pos = MPMDifferentiableSimulationTrail.apply(mpm.mpm_solver, mpm.mpm_state, mpm.mpm_model, every, sc.dt, steps,
                                                x0, v0, E, nu, rho, None, None, dev, True, 0)
    (pos * G).sum().backward()
    gE, gnu = float(E.grad), float(nu.grad)
    Gsum = G.sum(0).double().cpu().numpy()
    P = orc.make_params(N, 16, material=6, gravity=(0, 0, -9.8))
    cols = [orc.surface_collider(c["point"], c["normal"]) for c in sc.colliders]

    def J(Ephys, nu_, m_hist_in=None, m_hist_out=None):
        mu = np.full(N, Ephys / (2 * (1 + nu_)))
        lam = np.full(N, Ephys * nu_ / ((1 + nu_) * (1 - 2 * nu_)))
        o = orc.OracleState(sc.x, sc.params["v0"], np.zeros((N, 3, 3)), np.tile(np.eye(3), (N, 1, 1)), mu, lam,
                            sc.yield_stress, sc.vol * sc.rho, sc.vol)
        orc.rollout(P, cols, 0.0, sc.dt, steps, o, m_hist_in=m_hist_in, m_hist_out=m_hist_out)
        return float(np.sum(Gsum * o.x))

    Ep = 0.003e6
    mh = np.zeros((steps, 16 ** 3))
    J(Ep, 0.3, m_hist_out=mh)
    h = 1e-4
    fdE = (J(Ep * (1 + h), 0.3, mh) - J(Ep * (1 - h), 0.3, mh)) / (2 * h * Ep)
    fdnu = (J(Ep, 0.3 + 1e-5, mh) - J(Ep, 0.3 - 1e-5, mh)) / 2e-5
    exE = (J(Ep * (1 + h), 0.3) - J(Ep * (1 - h), 0.3)) / (2 * h * Ep)
    print(f"[trail compat] E.grad {gE:.6e} vs frozen-mass FD/N {fdE / N:.6e} (exact FD/N {exE / N:.6e}); "
          f"nu.grad {gnu:.6e} vs {fdnu / N:.6e}", flush=True)
    assert abs(gE - fdE / N) < 1e-3 * abs(fdE / N)
    assert abs(gnu - fdnu / N) < 1e-3 * abs(fdnu / N)


def test_warmup_steps_leave_initial_velocity_and_density_gradients_zero():
    """extra_no_grad_steps > 0: the reference's tape starts at the warm-up's
    last clone, so init_velocity's and the density's gradients (read from the
    caller's state, interface.py:201-204,236-241) stay zero, while E and nu
    still get theirs."""
    sc = scenes.small_scene(n_particles=200, n_grid=16, material=6, seed=7, E=3e3, center=(0.5, 0.5, 0.3),
                            vel=(0.2, 0.0, -1.5))
    mpm = _trail_setup(sc)
    N, dev = sc.n, "cuda:0"
    x0 = torch.as_tensor(sc.x, device=dev)
    for fn, extra in ((MPMDifferentiableSimulationTrail, (5,)), (MPMDifferentiableSimulation, (1,))):
        for warm in (0, 3):
            v0 = torch.as_tensor(sc.params["v0"], device=dev).clone().requires_grad_()
            E = torch.tensor(0.003 if fn is MPMDifferentiableSimulationTrail else 3e3, device=dev,
                             requires_grad=True)
            nu = torch.tensor(0.3, device=dev, requires_grad=True)
            rho = torch.full((N,), 1000.0, device=dev, requires_grad=True)
            mpm.mpm_solver.time = 0.0
            out = fn.apply(mpm.mpm_solver, mpm.mpm_state, mpm.mpm_model, *extra, sc.dt, 10, x0, v0, E, nu, rho, None,
                           None, dev, True, warm)
            (out * out).sum().backward()
            print(f"[warm-up] {fn.__name__} extra={warm}: |v0.grad| {float(v0.grad.abs().max()):.3e} "
                  f"|rho.grad| {float(rho.grad.abs().max()):.3e} E.grad {float(E.grad):.3e}", flush=True)
            assert torch.isfinite(E.grad) and float(E.grad) != 0.0
            if warm:
                assert float(v0.grad.abs().max()) == 0.0 and float(rho.grad.abs().max()) == 0.0
            else:
                assert float(v0.grad.abs().max()) > 0.0


def test_clean_reference_compat_seeds_only_the_position():
    """The reference's Clean backward seeds the position alone
    (interface.py:907-927): a loss on the returned velocity has zero
    gradient there in reference_compat mode, and its true gradient in exact
    mode."""
    sc = scenes.small_scene(n_particles=200, n_grid=16, material=0, seed=8, E=2e4, center=(0.5, 0.5, 0.3),
                            vel=(0.2, 0.0, -1.5))
    mpm = _trail_setup(sc, material="jelly")
    N, dev = sc.n, "cuda:0"
    x0 = torch.as_tensor(sc.x, device=dev)
    v0 = torch.as_tensor(sc.params["v0"], device=dev)
    F0 = torch.eye(3, device=dev).repeat(N, 1, 1)
    C0 = torch.zeros((N, 3, 3), device=dev)
    rho = torch.full((N,), 1000.0, device=dev)
    res = {}
    for mode in ("reference_compat", "exact"):
        mpm.mpm_solver.adjoint_mode = mode
        E = torch.full((N,), 2e4, device=dev, requires_grad=True)
        nu = torch.full((N,), 0.3, device=dev)
        mpm.mpm_solver.time = 0.0
        outs = MPMDifferentiableSimulationClean.apply(mpm.mpm_solver, mpm.mpm_state, mpm.mpm_model, sc.dt, 10, x0,
                                                      v0, F0, C0, E, nu, rho, None, dev, True, 0)
        (outs[1] * outs[1]).sum().backward()
        res[mode] = float(E.grad.abs().max())
    print(f"[clean] max |E.grad| of a velocity loss: {res}", flush=True)
    assert res["reference_compat"] == 0.0 and res["exact"] > 0.0


def test_function_variants_shapes_and_consistency():
    sc = scenes.small_scene(n_particles=200, n_grid=16, material=0, seed=5)
    mpm = _trail_setup(sc, material="jelly")
    N = sc.n
    dev = "cuda:0"
    x0 = torch.as_tensor(sc.x, device=dev)
    v0 = torch.as_tensor(sc.params["v0"], device=dev).requires_grad_()
    E = torch.full((N,), 2e4, device=dev, requires_grad=True)
    nu = torch.full((N,), 0.3, device=dev, requires_grad=True)
    rho = torch.full((N,), 1000.0, device=dev)
    x_last = MPMDifferentiableSimulation.apply(mpm.mpm_solver, mpm.mpm_state, mpm.mpm_model, 1, sc.dt, 10, x0, v0, E,
                                               nu, rho, None, None, dev, True, 0)
    assert x_last.shape == (N, 3)
    x_last.sum().backward()
    assert E.grad.shape == (N,) and nu.grad.shape == (N,) and v0.grad.shape == (N, 3)
    F0 = torch.eye(3, device=dev).repeat(N, 1, 1)
    C0 = torch.zeros((N, 3, 3), device=dev)
    xw = x0.clone().requires_grad_()
    outs = MPMDifferentiableSimulationWCheckpoint.apply(mpm.mpm_solver, mpm.mpm_state, mpm.mpm_model, sc.dt, 10, xw,
                                                        v0.detach(), F0, C0, E.detach(), nu.detach(), rho, None, dev,
                                                        True, 0)
    assert [o.shape for o in outs] == [(N, 3), (N, 3), (N, 3, 3), (N, 3, 3)]
    assert rel_err(outs[0].detach().cpu().numpy(), x_last.detach().cpu().numpy()) < 1e-6
    outs[0].sum().backward()
    assert xw.grad is not None and torch.isfinite(xw.grad).all()
    outs_c = MPMDifferentiableSimulationClean.apply(mpm.mpm_solver, mpm.mpm_state, mpm.mpm_model, sc.dt, 10, x0,
                                                    v0.detach(), F0, C0, E.detach(), nu.detach(), rho, None, dev,
                                                    True, 0)
    assert outs_c[4].shape == (N, 6)


@pytest.mark.parametrize("ckpt", [1, 3])
def test_material_5_forward_rewrites_yield_and_backward_differentiates_it(ckpt):
    """Plasticine rewrites the model's yield stress during the forward
    (persisting across rollouts, as the reference's in-place write at
    mpm_utils.py:133-186 does: only mu and lam are recomputed per rollout);
    the backward differentiates
    through those updates without touching the model's arrays again, and its
    E gradient matches central differences of FP64 oracle rollouts that start
    from the same softened yield stress."""
    sc = scenes.small_scene(n_particles=200, n_grid=16, material=5, seed=6, E=3e4, center=(0.5, 0.5, 0.3),
                            vel=(0.3, 0.0, -2.0))
    mpm = _wrapper(sc, material="plasticine", requires_grad=True, params={"yield_stress": 50.0})
    mpm.mpm_solver.adjoint_mode = "exact"
    mpm.mpm_solver.checkpoint_every = ckpt
    ys0 = mpm.mpm_model.yield_stress.clone()
    mpm.advance_to(2)
    assert torch.all(mpm.mpm_model.yield_stress <= ys0) and torch.any(mpm.mpm_model.yield_stress < ys0)
    N, dev, steps = sc.n, "cuda:0", 9
    ys_start = mpm.mpm_model.yield_stress.double().cpu().numpy().copy()
    x0 = torch.as_tensor(sc.x, device=dev)
    v0 = torch.as_tensor(sc.params["v0"], device=dev)
    E = torch.tensor(0.03, device=dev, requires_grad=True)
    nu = torch.tensor(0.3, device=dev)
    mpm.mpm_solver.time = 0.0
    pos = MPMDifferentiableSimulation.apply(mpm.mpm_solver, mpm.mpm_state, mpm.mpm_model, 1, sc.dt, steps, x0, v0,
                                            E * 1e6, nu, torch.full((N,), 1000.0, device=dev), None, None, dev,
                                            True, 0)
    ys_after_fwd = mpm.mpm_model.yield_stress.clone()
    assert float((ys_after_fwd - torch.as_tensor(ys_start, device=dev, dtype=torch.float32)).abs().max()) > 0
    # smooth weights (height plus spread about the centre): a random-sign
    # weighting would make the gradient a cancelling sum at FP32 round-off
    Gn = np.zeros((N, 3))
    Gn[:, 2] = 1.0
    Gn += 2.0 * (sc.x - sc.x.mean(0))
    G = torch.as_tensor(Gn, dtype=torch.float32, device=dev)
    (pos * G).sum().backward()
    assert torch.equal(mpm.mpm_model.yield_stress, ys_after_fwd)
    gE = float(E.grad) / 1e6  # d/dE_phys of the mean-aggregated scalar gradient

    def J(Ephys):
        mu = np.full(N, Ephys / (2 * 1.3))
        lam = np.full(N, Ephys * 0.3 / (1.3 * 0.4))
        P = orc.make_params(N, 16, material=5, gravity=(0, 0, -9.8), softening=0.1)
        cols = [orc.surface_collider(c["point"], c["normal"]) for c in sc.colliders]
        o = orc.OracleState(sc.x, sc.params["v0"], np.zeros((N, 3, 3)), np.tile(np.eye(3), (N, 1, 1)), mu, lam,
                            ys_start, sc.vol * sc.rho, sc.vol)
        orc.rollout(P, cols, 0.0, sc.dt, steps, o)
        return float(np.sum(G.double().cpu().numpy() * o.x))

    h = 1e-4
    fd = (J(3e4 * (1 + h)) - J(3e4 * (1 - h))) / (2 * h * 3e4) / N
    print(f"[plasticine backward, ckpt {ckpt}] E.grad {gE:.6e} vs FD/N {fd:.6e}", flush=True)
    assert abs(gE - fd) < 1e-3 * abs(fd)


def test_errors_mirror_reference():
    sc = scenes.small_scene(n_particles=100, n_grid=16)
    with pytest.raises(TypeError):
        m = _wrapper(sc, material="fluid")
    m = _wrapper(sc)
    with pytest.raises(ValueError):
        m.mpm_solver.add_surface_collider((0, 0, 0), (0, 0, 1), "sticky", friction=0.5)
    with pytest.raises(ValueError):
        m.mpm_solver.add_geometry_collider({"center": (0.5, 0.5, 0.5)})
    w = WarpMPMWrapper(grid_size=16)
    with pytest.raises(AssertionError):
        w.add_particles(np.array([[0.5, 0.5, 1.5]]))


def test_reference_compat_masks_the_reverse_sweep_with_the_colliders_last_state():
    """The reference's Tape records each collider launch with the solver's
    collider_params object (mpm_solver_diff.py:265-278) and modify_bc moves
    that object's point on after the launch (:860-866): the replay reads the
    object's last state.  In reference_compat mode the mirror's Trail,
    Simulation and Clean sweep with the table as it stands at backward time
    (tape.adj_cols_hist), while a checkpointed tape's replays keep each
    substep's own table.  Checks: the swept table is the solver's current one
    at every substep (exact mode: none); with a cuboid that moves through
    the body the gradient differs from the sweep masked by each substep's
    own cuboid; and it is the same with a recorded tape and with
    checkpoint-and-replay (K = 4).  Parity unpinned: what the reference's
    Tape returns here is read from its code, not run."""
    import types

    from empm_amd import _lib
    from empm_amd.diff_simulator.warp_mpm import interface as itf

    sc = scenes.small_scene(n_particles=600, n_grid=16, material=6, seed=4, E=3e3, dt=2e-4, center=(0.5, 0.5, 0.3),
                            vel=(0.0, 0.0, -1.0), spin=0.0)
    mpm = _trail_setup(sc)
    solver = mpm.mpm_solver
    # a cuboid pushing through the body along x over the whole rollout
    solver.set_velocity_on_cuboid((0.38, 0.5, 0.3), (0.05, 0.2, 0.2), (2.0, 0.0, 0.0))
    arrays = solver.model_arrays(mpm.mpm_model, mpm.mpm_state)
    ctx = solver.context(arrays.n_grid, sc.n, "cuda:0")
    st = mpm.mpm_state
    s0 = rt.ParticleState(st.particle_x.clone(), st.particle_v.clone(), torch.zeros((sc.n, 3, 3), device="cuda:0"),
                          torch.eye(3, device="cuda:0").repeat(sc.n, 1, 1).contiguous())
    steps = 24
    grads = {}
    for K in (1, 4):
        solver.time = 0.0
        solver.collider_params[-1].point = (0.38, 0.5, 0.3)
        ro = rt.Rollout(ctx, arrays, sc.n)
        cols0 = solver.collider_table()
        cols_out = []
        final, _, _, tape = ro.forward(s0, cols0, 0.0, sc.dt, steps, tape=True, ckpt_every=K, cols_out=cols_out)
        solver.sync_colliders_from(cols_out)
        assert abs(solver.collider_params[-1].point[0] - (0.38 + 2.0 * sc.dt * steps)) < 1e-5
        for mode in ("exact", "reference_compat"):
            solver.adjoint_mode = mode
            fake = types.SimpleNamespace(mpm_solver=solver, tape=tape)
            tape.adj_cols_hist = None
            itf._compat_colliders(fake)
            if mode == "exact":
                assert tape.adj_cols_hist is None
                continue
            cur = solver.collider_table()
            for k in (0, steps // 2, steps - 1):
                for c in range(tape.ncol):
                    got = tape.adj_cols_hist[k * tape.ncol + c]
                    assert tuple(got.point) == tuple(cur[c].point) and got.kind == cur[c].kind
        d = rt.ParticleState.empty(sc.n, "cuda:0")
        d.zero_()
        d.x.copy_(torch.sin(final.x * 11.0))
        _, g_last = ro.backward(tape, d, flags=_lib.EMPM_ADJ_DROP_GRID_MASS)
        tape.adj_cols_hist = None
        _, g_own = ro.backward(tape, d, flags=_lib.EMPM_ADJ_DROP_GRID_MASS)
        grads[K] = [float(t.double().sum()) for t in (g_last.mu, g_last.lam)]
        own = [float(t.double().sum()) for t in (g_own.mu, g_own.lam)]
        print(f"[compat colliders K={K}] last-state masks {grads[K]} own masks {own}", flush=True)
        assert abs(grads[K][0] - own[0]) > 1e-3 * abs(own[0])
    for a, b in zip(grads[1], grads[4]):
        assert abs(a - b) <= 1e-4 * abs(a)
