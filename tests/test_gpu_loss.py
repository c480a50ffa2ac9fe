"""The training loss on the GPU (empm_amd.loss) against the float64
restatement (oracle/chamfer.py), and its gradient through the simulation:
per-frame Chamfer + tracking losses seed the reverse sweep at their frames,
and d loss / d log E matches central differences of FP64 oracle rollouts
scored by the restated loss (bar 1e-3 relative)."""
import numpy as np
import pytest
import torch

from empm_amd import loss
from empm_amd import runtime as rt
from empm_amd import scenes
from oracle import chamfer as och
from oracle import oracle as orc
from tests.helpers import gpu_collider

pytestmark = pytest.mark.gpu


def test_nearest_neighbours_match_brute_force(dev):
    rng = np.random.default_rng(3)
    a, b = rng.random((3000, 3)).astype(np.float32), rng.random((2500, 3)).astype(np.float32)
    d, i = loss.nn_sqdist(torch.as_tensor(a, device=dev), torch.as_tensor(b, device=dev))
    d, i = d.cpu().numpy(), i.cpu().numpy()
    rd, ri = och.nn_sqdist(a, b)
    # indices must agree wherever the runner-up is not a float32 tie
    dd = ((a[:, None, :].astype(np.float64) - b[None]) ** 2).sum(-1)
    dd.sort(1)
    clear = (dd[:, 1] - dd[:, 0]) > 1e-5 * dd[:, 1]
    print(f"[nn] {clear.sum()} of {len(a)} points without a near tie; index mismatches "
          f"{np.sum(i[clear] != ri[clear])}", flush=True)
    assert np.all(i[clear] == ri[clear])
    assert np.max(np.abs(d - rd) / np.maximum(rd, 1e-12)) < 1e-4


def test_nearest_neighbour_ties_take_the_smallest_index(dev):
    b = torch.tensor([[1.0, 0, 0], [0, 1.0, 0], [1.0, 0, 0], [-1.0, 0, 0]], device=dev)
    a = torch.tensor([[0.0, 0, 0], [2.0, 0, 0]], device=dev)
    d, i = loss.nn_sqdist(a, b)
    assert i.tolist() == [0, 0] and d.tolist() == [1.0, 1.0]


def test_nearest_neighbour_ties_across_splits_of_the_search(dev):
    # 300 points of A are too few threads, so B is searched in several
    # splits; equal nearest points in different splits: the smaller index wins
    from empm_amd import _lib
    m = 4096
    assert int(_lib.load().empm_nn_sqdist_workspace(300, m)) > 0
    rng = np.random.default_rng(8)
    b = rng.random((m, 3)).astype(np.float32) + 2.0
    a = rng.random((300, 3)).astype(np.float32)
    b[100] = a[7] + np.float32(0.001)
    b[3900] = b[100]
    d, i = loss.nn_sqdist(torch.as_tensor(a, device=dev), torch.as_tensor(b, device=dev))
    rd, ri = och.nn_sqdist(a, b)
    assert int(i[7]) == 100
    np.testing.assert_array_equal(i.cpu().numpy(), ri)
    np.testing.assert_allclose(d.cpu().numpy(), rd, rtol=1e-5)


def test_segmented_search_matches_each_segment(dev):
    rng = np.random.default_rng(9)
    na = [700, 0, 1500, 3, 2600]
    nb = [900, 50, 1200, 0, 2600]
    a = rng.random((sum(na), 3)).astype(np.float32)
    b = rng.random((sum(nb), 3)).astype(np.float32)
    ao = np.concatenate([[0], np.cumsum(na)]).astype(np.int32)
    bo = np.concatenate([[0], np.cumsum(nb)]).astype(np.int32)
    d, i = loss.nn_sqdist_segments(torch.as_tensor(a, device=dev), torch.as_tensor(b, device=dev),
                                   torch.as_tensor(ao, device=dev), torch.as_tensor(bo, device=dev), max(na), max(nb))
    d, i = d.cpu().numpy(), i.cpu().numpy()
    for g in range(len(na)):
        sa, sb = slice(ao[g], ao[g + 1]), slice(bo[g], bo[g + 1])
        if na[g] == 0:
            continue
        if nb[g] == 0:
            assert np.all(np.isinf(d[sa])) and np.all(i[sa] == -1)
            continue
        rd, ri = och.nn_sqdist(a[sa], b[sb])
        np.testing.assert_array_equal(i[sa], ri + bo[g])
        np.testing.assert_allclose(d[sa], rd, rtol=1e-5, atol=1e-9)


def test_batched_frame_loss_matches_per_frame_calls(dev):
    """frame_loss evaluates all frames together; the reference's per-frame
    calls (chamfer_distance of each frame's visible points, tracking of its
    valid points) give the same value and gradient, and so does the float64
    restatement."""
    rng = np.random.default_rng(10)
    T, N = 6, 3000
    p = rng.random((T, N, 3)).astype(np.float32)
    q = (p + 0.01 * rng.standard_normal((T, N, 3))).astype(np.float32)
    vis = rng.random((T, N)) < 0.7
    valid = rng.random((T, N)) < 0.8
    valid[2] = True
    q[2][valid[2]] = p[2][valid[2]]           # a frame with zero tracking error: the gradient stays finite
    pt = torch.as_tensor(p, device=dev).requires_grad_()
    qt, vt, mt = torch.as_tensor(q, device=dev), torch.as_tensor(vis, device=dev), torch.as_tensor(valid, device=dev)
    L = loss.frame_loss(pt, qt, vt, mt)
    L.backward()
    g_batched = pt.grad.clone()
    pt.grad = None
    L2 = sum(loss.chamfer_distance(pt[t][vt[t]].unsqueeze(0), qt[t][vt[t]].unsqueeze(0))[0] for t in range(T)) * 1e3
    L2 = L2 + sum(torch.linalg.vector_norm(pt[t][mt[t]] - qt[t][mt[t]]) / int(mt[t].sum()) for t in range(T)) * 1e3
    L2.backward()
    ref = och.frame_loss(p, q, vis, valid)
    print(f"[frame loss] batched {float(L):.8e}, per frame {float(L2):.8e}, oracle {ref:.8e}", flush=True)
    assert abs(float(L) - float(L2)) < 1e-5 * abs(float(L2))
    assert abs(float(L) - ref) < 1e-5 * abs(ref)
    assert torch.isfinite(g_batched).all()
    torch.testing.assert_close(g_batched, pt.grad, rtol=1e-4, atol=1e-6)


def test_frame_loss_tracks_the_first_points(dev):
    """n_track keeps the first points of the simulated frames (the trainer's
    points[:, :n_track_points]); the targets and masks cover those points."""
    rng = np.random.default_rng(11)
    T, N, n_track = 4, 2500, 1800
    p = rng.random((T, N, 3)).astype(np.float32)
    q = (p[:, :n_track] + 0.01 * rng.standard_normal((T, n_track, 3))).astype(np.float32)
    vis = rng.random((T, n_track)) < 0.75
    valid = rng.random((T, n_track)) < 0.8
    pt = torch.as_tensor(p, device=dev).requires_grad_()
    L = loss.frame_loss(pt, torch.as_tensor(q, device=dev), torch.as_tensor(vis, device=dev),
                        torch.as_tensor(valid, device=dev), n_track=n_track)
    L.backward()
    ref = och.frame_loss(p[:, :n_track], q, vis, valid)
    assert abs(float(L) - ref) < 1e-5 * abs(ref)
    assert torch.all(pt.grad[:, n_track:] == 0) and torch.any(pt.grad[:, :n_track] != 0)


def test_chamfer_value_and_gradient(dev):
    rng = np.random.default_rng(4)
    x, y = rng.random((2, 700, 3)).astype(np.float32), rng.random((2, 500, 3)).astype(np.float32)
    xt = torch.as_tensor(x, device=dev).requires_grad_()
    val, none = loss.chamfer_distance(xt, torch.as_tensor(y, device=dev))
    assert none is None
    val.backward()
    ref = 0.5 * (och.chamfer(x[0], y[0]) + och.chamfer(x[1], y[1]))
    g_ref = 0.5 * np.stack([och.chamfer_grad(x[b], y[b]) for b in range(2)])
    eg = np.max(np.abs(xt.grad.cpu().numpy() - g_ref)) / np.max(np.abs(g_ref))
    print(f"[chamfer] value {float(val):.8e} vs {ref:.8e}; grad rel err {eg:.1e}", flush=True)
    assert abs(float(val) - ref) < 1e-5 * ref and eg < 1e-4


def test_frame_loss_gradient_through_the_rollout_matches_oracle_fd(dev):
    n_grid, dx = 16, 1.0 / 16
    radii = (0.15, 0.12, 0.1)
    sc = scenes.small_scene(n_particles=300, n_grid=n_grid, material=6, seed=5, E=2e4,
                            center=(0.5, 0.5, 3 * dx + radii[2] + 0.25 * dx), radii=radii, vel=(0.2, -0.1, -2.0))
    N, frames, dt = sc.n, (10, 20, 30), sc.dt
    rng = np.random.default_rng(6)
    vis = rng.random((3, N)) < 0.8
    valid = rng.random((3, N)) < 0.7
    ocols = [orc.surface_collider(c["point"], c["normal"]) for c in sc.colliders]
    P = orc.make_params(N, n_grid, material=6, gravity=sc.gravity)

    def oracle_frames(E, prec="f64"):
        mu, lam = orc.mu_lam(np.full(N, E), np.full(N, 0.3))
        o = orc.OracleState(sc.x, sc.params["v0"], np.zeros((N, 3, 3)), np.tile(np.eye(3), (N, 1, 1)), mu, lam,
                            sc.yield_stress, sc.vol * sc.rho, sc.vol, precision=prec)
        arr, _ = orc.colliders_array(ocols)
        t, done, out = 0.0, 0, []
        for k in frames:
            t = orc.rollout(P, arr, t, dt, k - done, o)
            done = k
            out.append(o.x.copy())
        return np.stack(out)

    # targets: a nominal rollout, moved 1 cm so the loss is not a difference of nearly equal frames
    target = oracle_frames(3e4) + 0.01
    E = 2e4
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)
    Et, nut = t(np.full(N, E)), t(np.full(N, 0.3))
    mu, lam = torch.empty_like(Et), torch.empty_like(Et)
    rt.mu_lam(Et, nut, mu, lam)
    model = rt.ModelArrays(material=6, n_grid=n_grid, grid_lim=1.0, dx=P.dx, inv_dx=P.inv_dx, mu=mu, lam=lam,
                           yield_stress=t(sc.yield_stress), mass=t(sc.vol * sc.rho), vol=t(sc.vol), gravity=sc.gravity)
    s0 = rt.ParticleState(t(sc.x), t(sc.params["v0"]), torch.zeros((N, 3, 3), device=dev),
                          torch.eye(3, device=dev).repeat(N, 1, 1).contiguous())
    ro = rt.Rollout(rt.Context(n_grid, N, dev), model, N)
    final, _, xs, tape = ro.forward(s0, [gpu_collider(c) for c in ocols], 0.0, dt, frames[-1], record=frames[:-1],
                                    tape=True)
    pts = torch.stack(xs + [final.x]).clone().requires_grad_()
    L = loss.frame_loss(pts, t(target), torch.as_tensor(vis, device=dev), torch.as_tensor(valid, device=dev))
    L.backward()
    g = pts.grad
    seeds = {}
    for f, k in enumerate(frames[:-1]):
        sd = rt.ParticleState.empty(N, dev)
        sd.zero_()
        sd.x.copy_(g[f])
        seeds[k] = sd
    d = rt.ParticleState.empty(N, dev)
    d.zero_()
    d.x.copy_(g[-1])
    _, grads = ro.backward(tape, d, seeds=seeds, flags=0)
    dE, dnu = torch.zeros_like(Et), torch.zeros_like(Et)
    rt.mu_lam_adj(Et, nut, grads.mu, grads.lam, dE, dnu)
    gE = float((dE * Et).double().sum())

    oloss = lambda E_: och.frame_loss(oracle_frames(E_), target, vis, valid)
    h = 1e-3
    fd = (oloss(E * np.exp(h)) - oloss(E * np.exp(-h))) / (2 * h)
    l0 = oloss(E)
    print(f"[frame loss] GPU {float(L):.8e} vs oracle {l0:.8e}; d/dlogE {gE:.6e} vs FD {fd:.6e}", flush=True)
    assert abs(float(L) - l0) < 1e-4 * abs(l0)
    assert abs(gE - fd) < 1e-3 * abs(fd)
