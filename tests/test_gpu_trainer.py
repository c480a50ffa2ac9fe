"""The trainer's CMA-ES evaluation (empm_amd.engine.InvPhyTrainerWarpMPM) on
synthetic tracking data: a jelly block above the sticky ground, pulled by two
controller cuboids (one following the mean of its control points, one a
single control point) whose velocities change every frame.

A generation scored in one batched rollout equals each candidate scored by
its own simulator (error_func_cma) within 1e-4 relative (the two runs sum
their scatters in different orders: ~1e-7 of a position, small beside the
loss's offsets), and one candidate's loss equals the oracle's rollout of the
same scene scored by the oracle's frame loss within 1e-3 relative."""
import numpy as np
import pytest

from empm_amd.engine import InvPhyTrainerWarpMPM
from empm_amd.engine import utils as eu
from oracle import chamfer as och
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GRID = 48


def trainer_data(seed=0, n_frames=4, N=1500, M=6):
    """object points in data coordinates (simulation = 0.5 + data / 5): a
    block on [0.42, 0.58]^2 x [0.52, 0.62] drifting up and sideways with
    per-point noise; random visibility and motion masks; control points
    over the block's top moving up at 0.6 m/s and sideways."""
    rng = np.random.default_rng(seed)
    sim0 = rng.uniform([0.42, 0.42, 0.52], [0.58, 0.58, 0.62], (N, 3))
    data0 = (sim0 - 0.5) * 5.0
    obj = [data0 + np.array([0.004, 0.0, 0.02]) * f + rng.normal(0.0, 0.003, (N, 3)) * (f > 0)
           for f in range(n_frames + 1)]
    vis = rng.random((n_frames + 1, N)) < 0.7
    valid = rng.random((n_frames + 1, N)) < 0.8
    c0 = (rng.uniform([0.47, 0.47, 0.6], [0.53, 0.53, 0.62], (M, 3)) - 0.5) * 5.0
    ctrl = np.stack([c0 + np.array([0.01 * np.sin(f), 0.006 * f, 0.03 * f]) for f in range(n_frames + 1)])
    ctrl_param = {"csize": [(0.03, 0.03, 0.02), (0.02, 0.02, 0.02)],
                  "handler_pos": [(0.47, 0.5, 0.61), (0.54, 0.5, 0.61)],
                  "ctrl_index": [-1, 4], "ctrl_mask": np.array([0, 0, 0, 0, 1, 1]),
                  "vscale": [(20.0, 20.0, 20.0), (20.0, 10.0, 20.0)]}
    return np.stack(obj).astype(np.float32), vis, valid, ctrl.astype(np.float32), ctrl_param


def make_trainer(dev, **kw):
    obj, vis, valid, ctrl, cp = trainer_data(**kw)
    return InvPhyTrainerWarpMPM(obj, vis, valid, ctrl, cp, device=str(dev),
                                wrapper_kwargs={"grid_size": GRID})


PARAMS = {"type": "jelly", "E": 0.008, "nu": 0.3, "rho": 2e3}
SOLUTIONS = [[0.008, 0.3], [0.02, 0.25], [0.004, 0.35], [0.012, 0.2]]


def test_generation_matches_each_candidate_alone(dev):
    tr = make_trainer(dev)
    print(f"[trainer] generation of {len(SOLUTIONS)}, N={tr.object_points.shape[1]}, "
          f"{tr.train_frame} frames, grid {GRID}", flush=True)
    batched = tr.evaluate_generation(SOLUTIONS, dict(PARAMS), ["E", "nu"])
    alone = [tr.error_func_cma(list(s), dict(PARAMS), ["E", "nu"]) for s in SOLUTIONS]
    print(f"  batched {batched}\n  alone   {alone}", flush=True)
    assert all(np.isfinite(batched)) and len(set(round(b, 3) for b in batched)) == len(batched)
    np.testing.assert_allclose(batched, alone, rtol=1e-4)
    # fewer frames: the loss covers frames 0 .. n - 1 only
    b2 = tr.evaluate_generation(SOLUTIONS[:2], dict(PARAMS), ["E", "nu"], n_train_frames=2)
    a2 = [tr.error_func_cma(list(s), dict(PARAMS), ["E", "nu"], n_train_frames=2) for s in SOLUTIONS[:2]]
    np.testing.assert_allclose(b2, a2, rtol=1e-4)
    assert b2[0] < batched[0]


def test_candidate_loss_matches_the_oracle(dev):
    tr = make_trainer(dev)
    sol = SOLUTIONS[1]
    loss = tr.error_func_cma(list(sol), dict(PARAMS), ["E", "nu"])
    # the same scene on the oracle: the trainer's simulator gives the inputs
    mpm = tr.initialize_mpm(dict(PARAMS, E=sol[0], nu=sol[1]), update_ctrl=True)
    st, md = mpm.mpm_state, mpm.mpm_model
    N = mpm.n_particles
    h = lambda t: t.detach().cpu().numpy().astype(np.float64)
    mu, lam = orc.mu_lam(h(md.E), h(md.nu))
    P = orc.make_params(N, GRID, material=0, grid_v_damping_scale=0.98)
    cols, nc = orc.colliders_array([orc.surface_collider((0.0, 0.0, 0.5), (0.0, 0.0, 1.0))]
                                   + [orc.cuboid(tr.ctrl_param["handler_pos"][i], tr.ctrl_param["csize"][i],
                                                 (0.0, 0.0, 0.0)) for i in range(2)])
    s = orc.OracleState(h(st.particle_x), np.zeros((N, 3)), np.zeros((N, 3, 3)), np.tile(np.eye(3), (N, 1, 1)),
                        mu, lam, h(md.yield_stress), h(st.particle_mass), h(st.particle_vol))
    frames = [s.x.copy()]
    t = 0.0
    n_sub = int(mpm.frame_dt / mpm.sim_dt)
    for f in range(tr.train_frame):
        for i in range(2):
            v = tr._ctrl_velocity(f, i)
            for j in range(3):
                cols[i + 1].velocity[j] = float(np.float32(v[j]))
        t = orc.rollout(P, cols, t, mpm.sim_dt, n_sub, s)
        frames.append(s.x.copy())
    pts = (np.stack(frames) - 0.5) * 5.0 / eu.CSCALE
    T = tr.train_frame
    obj = tr.object_points.cpu().numpy().astype(np.float64)
    ref = och.frame_loss(pts[:T], obj[:T], tr.object_visibilities.cpu().numpy()[:T],
                         tr.object_motions_valid.cpu().numpy()[:T])
    print(f"[trainer] loss {loss:.6f} oracle {ref:.6f}", flush=True)
    assert abs(loss - ref) <= 1e-3 * abs(ref)
    # the controllers moved the block: frames differ from the start
    assert np.abs(pts[-1] - pts[0]).max() > 1e-3
