"""Long-horizon embodied gradients: a batch of material-parameter sets, each
rolled out over thousands of substeps with a kinematic gripper and a
position loss at every frame, differentiated with checkpoint-recompute.

The reference's training loop advances one parameter set frame by frame,
setting the controller collider's velocity before each frame, and seeds a
position loss at every frame of its Tape
(phys_sim/engine/trainer_warp_mpm.py:379-394,714-718;
phys_sim/diff_simulator/warp_mpm/interface.py:167-251).  Here the parameter
sets are one batch (each its own copy of the object and its own grid, one
launch per stage for all of them), every frame is one forward call that
keeps a checkpoint every `ckpt_every` substeps, and the reverse sweep runs
frame by frame from the last, adding each frame's position seed as it
reaches that frame's end state.

The loss of set c is sum over frames f of mean over its particles of
|x_f - x*_f|^2, with x* the frames of a target rollout; the gradient is
taken with respect to (log E, nu, log yield stress).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import runtime as rt
from .population import Candidate, surface_colliders

F32 = torch.float32


def gripper_collider(point, half, velocity):
    """set_velocity_on_cuboid (mpm_solver_diff.py:814-869): nodes within
    `half` of `point` take `velocity`; the point moves with it."""
    c = _lib.Collider()
    c.kind = 1
    c.active = 1
    c.start_time, c.end_time = 0.0, 999.0
    for i in range(3):
        c.point[i] = float(point[i])
        c.size[i] = float(half[i])
        c.velocity[i] = float(velocity[i])
    return c


class LongHorizonBatch:
    """One object (scenes.Scene with params["gripper"], ["n_frames"],
    ["frame_steps"]) replicated for n_sets parameter sets on one device."""

    def __init__(self, scene, n_sets: int, device="cuda:0", ckpt_every: int = 50):
        self.sc = scene
        self.n_sets = int(n_sets)
        self.device = torch.device(device)
        self.ckpt_every = int(ckpt_every)
        self.n1 = scene.n
        self.n = self.n1 * self.n_sets
        self.n_frames = int(scene.params["n_frames"])
        self.frame_steps = int(scene.params["frame_steps"])
        self.dt = float(scene.dt)
        g = scene.params["gripper"]
        self.grip_point, self.grip_half = g["point"], g["half"]
        self.grip_vel = np.asarray(g["velocity"], np.float64)
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=self.device)
        rep = lambda a: torch.cat([a] * self.n_sets).contiguous()
        self.x0 = rep(t(scene.x))
        self.vol = rep(t(scene.vol))
        self.mass = rep(t(scene.vol * scene.rho))
        self.walls = surface_colliders(scene.colliders)
        self.ctx = rt.Context(scene.n_grid, self.n, self.device, max_cand=self.n_sets)
        self.targets = None
        self._ends = torch.arange(1, self.n_sets + 1, device=self.device) * self.n1 - 1

    def model(self, cands: list[Candidate]):
        if len(cands) != self.n_sets:
            raise ValueError(f"{len(cands)} parameter sets for a batch of {self.n_sets}")
        per = lambda vals: torch.repeat_interleave(torch.as_tensor(np.asarray(vals, np.float32), device=self.device),
                                                   self.n1).contiguous()
        E, nu, ys = per([c.E for c in cands]), per([c.nu for c in cands]), per([c.yield_stress for c in cands])
        mu, lam = torch.empty_like(E), torch.empty_like(E)
        rt.mu_lam(E, nu, mu, lam)
        sc = self.sc
        sims = []
        for c in range(self.n_sets):
            a, b = c * self.n1, (c + 1) * self.n1
            sims.append(rt.ModelArrays(
                material=sc.material, n_grid=sc.n_grid, grid_lim=sc.grid_lim,
                dx=float(np.float32(sc.grid_lim / sc.n_grid)), inv_dx=float(np.float32(sc.n_grid / sc.grid_lim)),
                mu=mu[a:b], lam=lam[a:b], yield_stress=ys[a:b], mass=self.mass[a:b], vol=self.vol[a:b],
                gravity=sc.gravity, softening=sc.softening))
        return rt.BatchModel(sims, [self.n1] * self.n_sets, mu, lam, ys, self.mass, self.vol), E, nu

    def _seg_sum(self, v):
        cs = torch.cumsum(v.double(), 0)[self._ends]
        return torch.diff(cs, prepend=torch.zeros(1, dtype=torch.float64, device=self.device))

    def _initial(self):
        N = self.n
        return rt.ParticleState(self.x0.clone(), torch.zeros((N, 3), device=self.device),
                                torch.zeros((N, 3, 3), device=self.device),
                                torch.eye(3, device=self.device).repeat(N, 1, 1).contiguous())

    def rollout(self, cands, tape: bool):
        """Frames of the batch: (x after each frame, tapes per frame or None,
        BatchModel, E, nu)."""
        bm, E, nu = self.model(cands)
        ro = rt.Rollout(self.ctx, bm, self.n)
        state, t = self._initial(), 0.0
        grip = gripper_collider(self.grip_point, self.grip_half, self.grip_vel[0])
        xs, tapes = [], []
        for f in range(self.n_frames):
            for i in range(3):
                grip.velocity[i] = float(self.grip_vel[f][i])
            out_cols: list = []
            state, t, _, tp = ro.forward(state, self.walls + [grip], t, self.dt, self.frame_steps, tape=tape,
                                         ckpt_every=self.ckpt_every, cols_out=out_cols)
            grip = out_cols[-1]  # the gripper point advanced over the frame
            xs.append(state.x.clone())
            tapes.append(tp)
        return xs, (tapes if tape else None), ro, E, nu, bm

    def set_targets(self, cands):
        self.targets = [x.clone() for x in self.rollout(cands, tape=False)[0]]

    def loss_and_grad(self, cands):
        """(loss per set (n_sets,), d loss / d(log E, nu, log Y) per set
        (n_sets, 3)), float64."""
        xs, tapes, ro, E, nu, bm = self.rollout(cands, tape=True)
        loss = torch.zeros(self.n_sets, dtype=torch.float64, device=self.device)
        seeds = []
        for x, tx in zip(xs, self.targets):
            diff = x - tx
            loss += self._seg_sum((diff * diff).sum(1)) / self.n1
            seeds.append(diff * (2.0 / self.n1))
        d = rt.ParticleState.empty(self.n, self.device)
        d.zero_()
        grads = rt.ParamGrads.zeros(self.n, self.device)
        for f in range(self.n_frames - 1, -1, -1):
            d.x.add_(seeds[f])
            d, _ = ro.backward(tapes[f], d, flags=_lib.EMPM_ADJ_EXACT, grads=grads)
            tapes[f] = None  # release the frame's checkpoints
        dE, dnu = torch.zeros_like(E), torch.zeros_like(E)
        rt.mu_lam_adj(E, nu, grads.mu, grads.lam, dE, dnu)
        grad = torch.stack([self._seg_sum(dE * E), self._seg_sum(dnu), self._seg_sum(grads.ys * bm.yield_stress)], 1)
        return loss, grad
