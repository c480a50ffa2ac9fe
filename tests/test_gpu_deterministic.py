"""Deterministic mode (empm_set_deterministic) through the C-ABI.

- A recorded forward rollout and its checkpointed reverse sweep run twice in
  deterministic mode give the same bits (states, dL/d(state 0), per-particle
  parameter gradients), also when the second run replays HIP graphs; and
  they agree with the default mode to its atomic-order spread (1e-5 relative
  on x, v, F; 1e-4 on C and dL/d(state 0); 2e-2 per particle and 1e-3 summed
  on the parameter gradients).
- A block of 6000 particles (beyond the cell ordering's LDS copy) repeats
  bit for bit and matches the FP32 oracle.
- A batch of simulations with different materials, particle counts and
  grids gives, bit for bit, what each simulation gives alone: per particle
  the work is the same, and deterministic mode fixes the order of every
  float sum the batch could change.
"""
import numpy as np
import pytest
import torch

from empm_amd import _lib
from empm_amd import runtime as rt
from empm_amd import scenes
from oracle import oracle as orc
from tests.helpers import fmt_errs, gpu_collider, rel_err, state_errors
from tests.test_gpu_batch import cat_states, ground, sim
from tests.test_gpu_rollout import setup

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "C", "F", "d0_x", "d0_v", "d_mu", "d_lam", "d_ys")


def _fwd_bwd(dev, det, graphs=False, steps=30):
    sc = scenes.small_scene(n_particles=6000, n_grid=32, material=7, seed=8, E=3e4, ys=200.0,
                            center=(0.5, 0.5, 0.22), vel=(0.3, -0.2, -2.0))
    model, s0, P, ocols, _ = setup(sc, dev, v0=sc.params["v0"])
    ctx = rt.Context(sc.n_grid, sc.n, dev)
    ctx.set_deterministic(det)
    if graphs:
        ctx.set_graphs(True)
    ro = rt.Rollout(ctx, model, sc.n)
    final, _, _, tape = ro.forward(s0, [gpu_collider(c) for c in ocols], 0.0, sc.dt, steps, tape=True,
                                   ckpt_every=5)
    d = rt.ParticleState.empty(sc.n, dev)
    d.x.copy_(torch.sin(final.x * 7.0))
    d0, g = ro.backward(tape, d, flags=_lib.EMPM_ADJ_EXACT)
    torch.cuda.synchronize()
    out = dict(x=final.x, v=final.v, C=final.C, F=final.F, d0_x=d0.x, d0_v=d0.v, d_mu=g.mu, d_lam=g.lam, d_ys=g.ys)
    return {k: t.detach().clone() for k, t in out.items()}


def test_deterministic_mode_repeats_bit_for_bit(dev):
    print("[deterministic] 6000 particles, 32^3, 30 substeps fwd + adjoint", flush=True)
    a = _fwd_bwd(dev, True)
    b = _fwd_bwd(dev, True)
    c = _fwd_bwd(dev, True, graphs=True)
    ref = _fwd_bwd(dev, False)
    for k in FIELDS:
        print(f"  {k}: repeat equal {torch.equal(a[k], b[k])}, graph equal {torch.equal(a[k], c[k])}, "
              f"vs default mode {rel_err(a[k].cpu().numpy(), ref[k].cpu().numpy()):.1e}", flush=True)
    for k in FIELDS:
        assert torch.equal(a[k], b[k]), k
        assert torch.equal(a[k], c[k]), k
    # the default mode sums in atomic arrival order: within its run-to-run
    # spread (C and the per-particle (mu, lam) gradients carry cancellation:
    # 1.6e-5 and 5e-3 measured; their sums over particles agree closer)
    for k, bar in (("x", 1e-5), ("v", 1e-5), ("F", 1e-5), ("C", 1e-4), ("d0_x", 1e-4), ("d0_v", 1e-4),
                   ("d_mu", 2e-2), ("d_lam", 2e-2), ("d_ys", 2e-2)):
        assert rel_err(a[k].cpu().numpy(), ref[k].cpu().numpy()) < bar, k
    for k in ("d_mu", "d_lam", "d_ys"):
        sa, sr = float(a[k].double().sum()), float(ref[k].double().sum())
        print(f"  sum {k}: deterministic {sa:.6e}, default {sr:.6e}", flush=True)
        assert abs(sa - sr) <= 1e-3 * max(abs(sr), 1e-30), k


def test_deterministic_batch_equals_each_simulation_alone(dev):
    specs = [(0, 16, 300, 1), (2, 16, 257, 2), (7, 24, 420, 3), (6, 12, 130, 4)]
    steps = 25
    built = [sim(dev, *a) for a in specs]
    models, states, dt = [b[0] for b in built], [b[1] for b in built], built[0][2]
    print(f"[deterministic] batch of {len(specs)} simulations vs each alone, {steps} substeps", flush=True)
    bm = rt.BatchModel.concat(models)
    ctx = rt.Context(max(m.n_grid for m in models), bm.n, dev, max_cand=len(models))
    ctx.set_deterministic(True)
    final, _, _, _ = rt.Rollout(ctx, bm, bm.n).forward(cat_states(states), [ground()], 0.0, dt, steps)
    torch.cuda.synchronize()
    for c, (m, s0) in enumerate(zip(models, states)):
        n = int(m.mu.shape[0])
        solo = rt.Context(m.n_grid, n, dev)
        solo.set_deterministic(True)
        f1, _, _, _ = rt.Rollout(solo, m, n).forward(s0, [ground()], 0.0, dt, steps)
        torch.cuda.synchronize()
        a, b = bm.offsets[c], bm.offsets[c + 1]
        same = {k: torch.equal(getattr(final, k)[a:b], getattr(f1, k)) for k in ("x", "v", "C", "F")}
        print(f"  simulation {c} material {specs[c][0]} n={specs[c][1]}: equal {same}", flush=True)
        assert all(same.values()), (c, same)


def test_deterministic_block_beyond_the_lds_copy(dev):
    """6000 particles in one 2^3-cell block (more than the 4096 the cell
    ordering copies to LDS, so its copy goes through global memory): two
    deterministic rollouts give the same bits and match the FP32 oracle."""
    n, N, steps = 16, 6000, 5
    rng = np.random.Generator(np.random.PCG64(77))
    # stencil base cells 8 and 9 on every axis (x / dx - 0.5 in [8, 10)): block 4
    x = rng.uniform(8.6 / n, 10.4 / n, (N, 3)).astype(np.float32)
    sc = scenes.small_scene(n_particles=N, n_grid=n, material=7, seed=78, E=3e4, ys=200.0)
    sc.x = x
    sc.vol = np.full(N, (2.0 / n) ** 3 / N, np.float32)
    sc.rho = np.full(N, 1000.0, np.float32)
    sc.E = np.full(N, 3e4, np.float32)
    sc.nu = np.full(N, 0.3, np.float32)
    sc.yield_stress = np.full(N, 200.0, np.float32)
    v0 = (0.05 * rng.standard_normal((N, 3))).astype(np.float32)
    print(f"[deterministic] {N} particles in one block, {steps} substeps", flush=True)
    outs = []
    for _ in range(2):
        model, s0, P, ocols, ostate = setup(sc, dev, v0=v0)
        ctx = rt.Context(n, N, dev)
        ctx.set_deterministic(True)
        f, _, _, _ = rt.Rollout(ctx, model, N).forward(s0, [gpu_collider(c) for c in ocols], 0.0, sc.dt, steps)
        torch.cuda.synchronize()
        outs.append({k: getattr(f, k).clone() for k in ("x", "v", "C", "F")})
    o = ostate("f32")
    orc.rollout(P, list(ocols), 0.0, sc.dt, steps, o)
    e = state_errors({k: t.cpu().numpy() for k, t in outs[0].items()}, o, x0=sc.x)
    same = {k: torch.equal(outs[0][k], outs[1][k]) for k in outs[0]}
    print(f"  repeat equal {same}; vs FP32 oracle {fmt_errs(e)}", flush=True)
    assert all(same.values()), same
    assert e["x"] < 1e-4 and e["F"] < 1e-4, e


def _check_cell_runs(perm, x, n):
    """perm holds each stencil-base cell's particles contiguously; returns
    whether every cell's run is in ascending index order."""
    base = np.clip(np.floor(x[perm] * n - 0.5).astype(np.int64), 0, n - 1)
    key = (base[:, 0] * n + base[:, 1]) * n + base[:, 2]
    starts = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    assert len(np.unique(key[starts])) == len(starts), "a cell's particles are not contiguous"
    ascending = True
    for a, b in zip(starts, np.r_[starts[1:], len(perm)]):
        ascending &= bool(np.all(np.diff(perm[a:b]) > 0))
    return ascending


@pytest.mark.parametrize("dense", [False, True])
def test_deterministic_binning_orders_each_cell_by_index(dev, dense):
    """empm_sort_by_cell in deterministic mode: every cell's particles in
    ascending index order, for blocks that fit the LDS copy (an ellipsoid of
    6000 particles on 32^3) and for one block of 6000 (the global copy)."""
    if dense:
        n, N = 16, 6000
        x = np.random.Generator(np.random.PCG64(77)).uniform(8.6 / n, 10.4 / n, (N, 3)).astype(np.float32)
    else:
        sc = scenes.small_scene(n_particles=6000, n_grid=32, material=7, seed=8)
        n, N, x = 32, sc.n, sc.x
    print(f"[deterministic] cell order N={N} n={n} dense={dense}", flush=True)
    sc = scenes.small_scene(n_particles=N, n_grid=n, material=7, seed=8)
    sc.x = x
    model, _, _, _, _ = setup(sc, dev)
    ctx = rt.Context(n, N, dev)
    ro = rt.Rollout(ctx, model, N)
    xt = torch.as_tensor(x, device=dev)
    p0 = ro.cell_order(xt).cpu().numpy()
    print(f"  default mode: cells in index order {_check_cell_runs(p0, x, n)}", flush=True)
    ctx.set_deterministic(True)
    p1 = ro.cell_order(xt).cpu().numpy()
    p2 = ro.cell_order(xt).cpu().numpy()
    assert np.array_equal(np.sort(p1), np.arange(N))
    assert np.array_equal(p1, p2)
    assert _check_cell_runs(p1, x, n)
