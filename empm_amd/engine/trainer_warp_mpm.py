"""The MPM-calling part of the reference's trainer
(phys_sim/engine/trainer_warp_mpm.py: InvPhyTrainerWarpMPM), on this
library's simulator mirror.

Mirrored: `denormalize_param` (:314-330), `initialize_mpm` (:332-400: the
simulator of one parameter set with its ground, controller cuboids and
per-frame controller velocities) and the CMA-ES evaluation of
`train_cma` (:598-663: one rollout per candidate scored by the frame loss).
Added: `evaluate_generation`, which scores a whole CMA-ES generation with
one batched rollout per frame (every candidate's particles in one arena,
every stage one launch) where the reference builds and rolls out one
simulator per candidate.

The trainer's data (tracked points, visibility and motion masks, controller
trajectories and parameters) come in through the constructor; the
reference loads them from its data set's pickles (:192-312), which are not
part of this path.  The special-case set-up of `initialize_mpm_test`
(:402-500, two named experiments) is not mirrored.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from .. import _lib
from .. import runtime as rt
from ..diff_simulator.warp_mpm.warp_mpm_wrapper import WarpMPMWrapper, material_gallery
from ..loss import frame_loss
from . import utils as eu


class InvPhyTrainerWarpMPM:
    """object_points (F+1, N, 3), object_visibilities and object_motions_valid
    (F, N) booleans, in the data's coordinates; ctrl_points (F+1, M, 3);
    ctrl_param with the reference's keys ("csize", "handler_pos",
    "ctrl_index", "ctrl_mask", "vscale"); wrapper_kwargs go to
    WarpMPMWrapper (the reference uses its defaults: grid 100, frame_dt
    0.01, sim_dt 1e-3)."""

    def __init__(self, object_points, object_visibilities, object_motions_valid, ctrl_points, ctrl_param,
                 exp_name="", num_original_points=None, train_frame=None, device="cuda:0", wrapper_kwargs=None):
        self.device = torch.device(device)
        t = lambda a, dt=torch.float32: torch.as_tensor(np.asarray(a), dtype=dt, device=self.device)
        self.object_points = t(object_points)
        self.object_visibilities = t(object_visibilities, torch.bool)
        self.object_motions_valid = t(object_motions_valid, torch.bool)
        self.ctrl_points = t(ctrl_points)
        self.ctrl_param = ctrl_param
        self.n_ctrl_parts = len(ctrl_param["csize"])
        self.exp_name = exp_name
        self.num_original_points = int(self.object_points.shape[1] if num_original_points is None
                                       else num_original_points)
        self.train_frame = int(self.object_points.shape[0] - 1 if train_frame is None else train_frame)
        self.wrapper_kwargs = dict(wrapper_kwargs or {})
        self.current_vertices = None

    # trainer_warp_mpm.py:314-330
    def denormalize_param(self, param):
        new_param = {"type": "jelly", "E": 5e4, "nu": 0.3, "rho": 2e3, "ys": 1e5, "vol": 1e-5}
        for key, value in param.items():
            if key == "E":
                new_param[key] = value * 1e6
            else:
                new_param[key] = value
        return new_param

    def _initial_positions(self):
        init_x = self.object_points[0].clone()
        init_x = eu.transform_points(x=init_x, name=self.exp_name)
        if init_x[:, 2].min() < 0.5 + 1e-4:
            init_x[:, 2] += 0.001  # lift up a bit to avoid initial penetration (:348-349)
        return init_x

    def _material_params(self, parameters):
        """material_gallery[type] with the trainer's overrides (:353-356).
        The reference assigns the gallery's own dict and then writes into it
        (so its changes persist in the module); here a copy is written.  Its
        test for the yield stress key reads "y_s" while the parameter dicts
        carry "ys", so the gallery's yield stress is always used: kept."""
        mp = dict(material_gallery[parameters["type"]])
        if "y_s" in parameters:
            mp["yield_stress"] = parameters["ys"]
        mp["grid_v_damping_scale"] = 0.98
        return mp

    def _ctrl_velocity(self, f, i):
        """Velocity of controller part i during frame f (:379-391): the
        displacement of its handle (or the mean of its points) from frame f to
        f + 1, in data coordinates, scaled by vscale."""
        indices = self.ctrl_param["ctrl_index"]
        if indices[i] == -1:
            mask = torch.as_tensor(np.asarray(self.ctrl_param["ctrl_mask"]) == i, device=self.device)
            vel = self.ctrl_points[f + 1, mask, :].mean(dim=0) - self.ctrl_points[f, mask, :].mean(dim=0)
        else:
            vel = self.ctrl_points[f + 1, indices[i], :] - self.ctrl_points[f, indices[i], :]
        vscale = self.ctrl_param["vscale"][i]
        return tuple(float(vel[j]) * float(vscale[j]) for j in range(3))

    # trainer_warp_mpm.py:332-400
    def initialize_mpm(self, parameters, diff_sim=False, update_ctrl=False):
        if self.exp_name in ("double_tear_pita_test", "single_poke_bluey_test"):
            raise NotImplementedError(f"the special initialisation of {self.exp_name} (initialize_mpm_test) "
                                      f"is not mirrored")
        parameters = self.denormalize_param(parameters)
        init_x = self._initial_positions()
        mpm = WarpMPMWrapper(requires_grad=diff_sim, device=str(self.device), **self.wrapper_kwargs)
        mpm.material_params = self._material_params(parameters)
        mpm.add_particles(init_x, None, parameters["vol"], parameters["E"], parameters["nu"], parameters["rho"])
        mpm.surface_colliders.append([(0.0, 0.0, 0.5), (0.0, 0.0, 1.0), "sticky", 0.0])
        for i in range(self.n_ctrl_parts):
            mpm.moving_boundaries.append([self.ctrl_param["handler_pos"][i], self.ctrl_param["csize"][i],
                                          (0.0, 0.0, 0.0), 0.0])

        def sim_callback(mpm, f):
            if f < self.ctrl_points.shape[0]:
                for i in range(self.n_ctrl_parts):
                    mpm.mpm_solver.collider_params[i + 1].velocity = self._ctrl_velocity(f, i)

        if update_ctrl:
            mpm.frame_callback = sim_callback
        mpm.initialize()
        return mpm

    # the loss of train_cma's error_func_cma (:615-657)
    def cma_loss(self, vertices_data, n_train_frames):
        """vertices_data: (>= n_train_frames, N, 3) frames in data coordinates
        (frame 0 = the initial positions).  Sum over frames t < n_train_frames
        of the visible points' Chamfer distance and the valid points'
        tracking term, times 1e3."""
        T = int(n_train_frames)
        return frame_loss(vertices_data[:T], self.object_points[:T], self.object_visibilities[:T],
                          self.object_motions_valid[:T], n_track=self.num_original_points)

    # train_cma's per-candidate evaluation, error_func_cma (:598-659)
    def error_func_cma(self, param_list, params, param_keys, n_train_frames=None):
        n_train_frames = self.train_frame if n_train_frames is None else int(n_train_frames)
        for i in range(len(param_list)):
            params[param_keys[i]] = param_list[i]
        mpm = self.initialize_mpm(parameters=params, diff_sim=False, update_ctrl=True)
        try:
            vertices = mpm.advance_to(n_frames=n_train_frames)
        except _lib.EmpmError as e:
            # a candidate whose state turned non-finite (an inverted element,
            # say) is ranked last; the reference scores its clamped positions
            if e.code == _lib.EMPM_E_NONFINITE:
                return float("inf")
            raise
        self.current_vertices = vertices[-1].cpu().numpy()
        vertices = eu.inv_transform_points(vertices, name=self.exp_name)
        return float(self.cma_loss(vertices, n_train_frames))

    def evaluate_generation(self, solutions, params, param_keys, n_train_frames=None):
        """The losses of a CMA-ES generation (es.ask()'s list of parameter
        vectors; returns the list es.tell takes), each equal to
        error_func_cma's for that candidate, from one batched rollout: every
        candidate's particles in one arena, one launch per stage and frame
        for all of them.  The candidates share the objects, colliders and
        controller schedule; they differ in the parameters named by
        param_keys."""
        n_train_frames = self.train_frame if n_train_frames is None else int(n_train_frames)
        cands = []
        for sol in solutions:
            p = dict(params)
            for i in range(len(sol)):
                p[param_keys[i]] = float(sol[i])
            cands.append(p)
        # the first candidate's simulator sets the configuration every
        # candidate shares (material, colliders, controllers, time step)
        mpm = self.initialize_mpm(parameters=cands[0], diff_sim=False, update_ctrl=True)
        solver = mpm.mpm_solver
        base = solver.model_arrays(mpm.mpm_model, mpm.mpm_state)
        N = mpm.n_particles
        n_sub = int(mpm.frame_dt / mpm.sim_dt)
        models = []
        for p in cands:
            q = self.denormalize_param(p)
            E = torch.full((N,), float(q["E"]), dtype=torch.float32, device=self.device)
            nu = torch.full((N,), float(q["nu"]), dtype=torch.float32, device=self.device)
            mu, lam = torch.empty_like(E), torch.empty_like(E)
            rt.mu_lam(E, nu, mu, lam)
            vol = torch.full((N,), float(q["vol"]), dtype=torch.float32, device=self.device)
            rho = torch.full((N,), float(q["rho"]), dtype=torch.float32, device=self.device)
            mass = torch.empty_like(vol)
            rt.mass_from_density(rho, vol, mass)
            models.append(dataclasses.replace(base, mu=mu, lam=lam, yield_stress=base.yield_stress.clone(),
                                              mass=mass, vol=vol, selection=None))
        bm = rt.BatchModel.concat(models)
        ctx = rt.Context(int(base.n_grid), bm.n, self.device, max_cand=len(cands))
        ro = rt.Rollout(ctx, bm, bm.n)
        x0 = mpm.mpm_state.particle_x
        state = rt.ParticleState(x0.repeat(len(cands), 1).contiguous(),
                                 mpm.mpm_state.particle_v.repeat(len(cands), 1).contiguous(),
                                 torch.zeros((bm.n, 3, 3), device=self.device),
                                 torch.eye(3, device=self.device).repeat(bm.n, 1, 1).contiguous())
        frames = [state.x.clone()]
        t = float(solver.time)
        for f in range(n_train_frames):
            if mpm.frame_callback:
                mpm.frame_callback(mpm, f)
            cols_out = []
            state, t, _, _ = ro.forward(state, solver.collider_table(), t, mpm.sim_dt, n_sub, cols_out=cols_out)
            solver.sync_colliders_from(cols_out)
            frames.append(state.x.clone())
        _, status = ctx.read_status()
        verts = torch.stack(frames)  # (F+1, n_cand * N, 3)
        losses = []
        for c in range(len(cands)):
            if status[c] & _lib.EMPM_STATUS_NONFINITE:
                losses.append(float("inf"))
                continue
            vc = eu.inv_transform_points(verts[:, c * N:(c + 1) * N], name=self.exp_name)
            losses.append(float(self.cma_loss(vc, n_train_frames)))
        self.current_vertices = verts[-1, :N].cpu().numpy()
        ctx.close()
        return losses
