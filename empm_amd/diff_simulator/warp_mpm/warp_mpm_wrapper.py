"""WarpMPMWrapper: the user-facing simulator of the reference
(phys_sim/diff_simulator/warp_mpm/warp_mpm_wrapper.py:26-499) on libempm.

Same constructor, attributes, material gallery and method semantics.
Differences: no pyvista/matplotlib (PLY files are written directly, loss
plots are skipped), and advance_to runs each frame's substeps in one native
call unless a per-substep step_callback is installed.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from .interface import (MPMDifferentiableSimulation, MPMDifferentiableSimulationClean,  # noqa: F401
                        MPMDifferentiableSimulationTrail, MPMDifferentiableSimulationWCheckpoint)
from .mpm_data_structure import MPMModelStruct, MPMStateStruct
from .mpm_solver_diff import MPMWARPDiff

DTYPE = torch.float32

# warp_mpm_wrapper.py:26-54
material_gallery = {
    "jelly": {"material": "jelly"},
    "metal": {"material": "metal"},
    "sand": {"material": "sand", "friction_angle": 35},
    "foam": {"material": "foam"},
    "snow": {"material": "snow"},
    "plasticine": {"material": "plasticine", "yield_stress": 5000.0, "softening": 0.1},
    "neo-hookean": {"material": "neo-hookean"},
    "fluid": {"material": "fluid"},
}


def sample_cube(lower_corner, cube_size, N):
    """warp_mpm_wrapper.py:57-61"""
    return np.random.rand(N, 3) * np.asarray(cube_size) + np.asarray(lower_corner)


def sample_sphere(center, radius, N):
    """warp_mpm_wrapper.py:64-70"""
    d = np.random.normal(size=(N, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = np.random.rand(N) ** (1 / 3) * radius
    return d * r[:, None] + np.asarray(center)


def _write_ply(points: np.ndarray, filename: str):
    pts = np.asarray(points, dtype=np.float32).reshape(-1, 3)
    with open(filename, "w") as f:
        f.write(f"ply\nformat ascii 1.0\nelement vertex {pts.shape[0]}\n"
                "property float x\nproperty float y\nproperty float z\nend_header\n")
        np.savetxt(f, pts, fmt="%.7g")


def save_ply_seq(vertices, save_path="./outputs/sim", prefix="sim_"):
    """warp_mpm_wrapper.py:73-78 (ASCII PLY, one file per frame)"""
    os.makedirs(save_path, exist_ok=True)
    for f in range(len(vertices)):
        _write_ply(vertices[f].detach().cpu().numpy(), f"{save_path}/{prefix}{f:04d}.ply")


def save_ply_file(vertices, filename=""):
    """warp_mpm_wrapper.py:81-87"""
    x = vertices.detach().cpu().numpy() if isinstance(vertices, torch.Tensor) else np.array(vertices)
    _write_ply(x, filename)


class WarpMPMWrapper:
    """warp_mpm_wrapper.py:90-499"""

    def __init__(self, grid_size=100, grid_lim=1.0, frame_dt=0.01, sim_dt=1e-3, device="cuda:0", requires_grad=False,
                 frame_callback=None, step_callback=None, optim_callback=None):
        self.n_particles = 0
        self.grid_size = grid_size
        self.grid_lim = grid_lim
        self.frame_dt = frame_dt
        self.sim_dt = sim_dt
        self.device = device
        self.requires_grad = requires_grad
        self.gravity = [0.0, 0.0, -9.8]
        self.surface_colliders = []
        self.moving_boundaries = []
        self.geometry_colliders = []
        self.grid_geometry_movers = []
        self.particle_geometry_movers = []
        self.frame_callback = frame_callback
        self.step_callback = step_callback
        self.optim_callback = optim_callback
        empty3 = torch.empty((0, 3), dtype=DTYPE, device=device)
        empty1 = torch.empty(0, dtype=DTYPE, device=device)
        self.x_t, self.v_t = empty3, empty3.clone()
        self.vol_t, self.E_t, self.nu_t, self.rho_t = empty1, empty1.clone(), empty1.clone(), empty1.clone()
        self.initialized = False
        self.material_params = material_gallery["jelly"]

    # warp_mpm_wrapper.py:135-177
    def add_particles(self, x, v=None, vol=1e-5, E=5e4, nu=0.3, rho=2000.0):
        n_p = x.shape[0]
        x_t = torch.as_tensor(x).to(dtype=DTYPE, device=self.device)
        self.x_t = torch.cat((self.x_t, x_t))
        v_t = torch.zeros((n_p, 3)) if v is None else torch.as_tensor(v)
        self.v_t = torch.cat((self.v_t, v_t.to(dtype=DTYPE, device=self.device)))

        def per_particle(q):
            if isinstance(q, float):
                t = torch.ones(n_p) * q
            elif isinstance(q, np.ndarray):
                t = torch.tensor(q)
            elif isinstance(q, torch.Tensor):
                t = torch.ones(n_p) * q.item() if q.ndim == 0 else q
            else:
                t = torch.ones(n_p) * float(q)
            return t.to(dtype=DTYPE, device=self.device)

        self.vol_t = torch.cat((self.vol_t, per_particle(vol)))
        self.E_t = torch.cat((self.E_t, per_particle(E)))
        self.nu_t = torch.cat((self.nu_t, per_particle(nu)))
        self.rho_t = torch.cat((self.rho_t, per_particle(rho)))
        assert x_t.min().item() > 0.0 and x_t.max().item() < self.grid_lim
        self.n_particles += n_p

    # warp_mpm_wrapper.py:179-192 (the reference's spelling)
    def set_paramters(self, params):
        assert self.initialized
        if "E" in params:
            self.E_t = torch.ones(self.n_particles, dtype=DTYPE, device=self.device) * params["E"]
        if "E" in params or "nu" in params:
            self.mpm_solver.set_E_nu_from_torch(self.mpm_model, self.E_t, self.nu_t, device=self.device)
            self.mpm_solver.prepare_mu_lam(self.mpm_model, self.mpm_state, self.device)

    def add_grid_geometry_mover(self, geo_args):
        self.grid_geometry_movers.append(geo_args)
        return len(self.grid_geometry_movers) - 1

    def add_particle_mover(self, geo_args):
        self.mpm_solver.enforce_particle_velocity_rotation(self.mpm_state, point=geo_args["center"],
                                                           normal=geo_args["axis"])

    # warp_mpm_wrapper.py:211-270
    def initialize(self):
        self.mpm_state = MPMStateStruct()
        self.mpm_state.init(self.n_particles, device=self.device, requires_grad=self.requires_grad)
        self.mpm_state.from_torch(self.x_t, self.vol_t, None, self.v_t, n_grid=self.grid_size, grid_lim=self.grid_lim,
                                  device=self.device, requires_grad=self.requires_grad)
        self.mpm_state.reset_density(tensor_density=self.rho_t, selection_mask=None, device=self.device,
                                     update_mass=True)
        self.mpm_model = MPMModelStruct()
        self.mpm_model.init(self.n_particles, device=self.device, requires_grad=self.requires_grad)
        self.mpm_model.init_other_params(n_grid=self.grid_size, grid_lim=self.grid_lim, device=self.device)
        self.mpm_model.gravitational_accelaration = tuple(float(g) for g in self.gravity)
        self.mpm_solver = MPMWARPDiff()
        self.mpm_solver.initialize(self.n_particles, n_grid=self.grid_size, grid_lim=self.grid_lim,
                                   device=self.device)
        self.mpm_solver.set_parameters_dict(self.mpm_model, self.mpm_state, self.material_params)
        self.mpm_solver.set_E_nu_from_torch(self.mpm_model, self.E_t, self.nu_t, device=self.device)
        self.mpm_solver.prepare_mu_lam(self.mpm_model, self.mpm_state, self.device)
        for sc in self.surface_colliders:
            self.mpm_solver.add_surface_collider(sc[0], sc[1], sc[2], sc[3])
        for mb in self.moving_boundaries:
            self.mpm_solver.set_velocity_on_cuboid(mb[0], mb[1], mb[2], mb[3])
        for gc in self.geometry_colliders:
            self.mpm_solver.add_geometry_collider(geo_args=gc)
        for gm in self.grid_geometry_movers:
            self.mpm_solver.set_grid_velocity_on_geometry(geo_args=gm)
        self.initialized = True

    def advance_to_static(self, n_frames=10):
        cb = self.frame_callback
        self.frame_callback = None
        self.advance_to(n_frames=n_frames)
        self.frame_callback = cb

    # warp_mpm_wrapper.py:278-306
    def advance_to(self, n_frames):
        """Positions at every frame boundary, shape (n_frames+1, N, 3)."""
        from .interface import _run_forward, _state_from

        n_sub = int(self.frame_dt / self.sim_dt)
        state = self.mpm_state
        frames = [state.particle_x.clone()]
        step = 0
        solver = self.mpm_solver
        for frame in range(n_frames):
            if self.frame_callback:
                self.frame_callback(self, frame)
            if self.step_callback:
                for _ in range(n_sub):
                    self.step_callback(self, step)
                    nxt = state.partial_clone(requires_grad=False)
                    solver.p2g2p_differentiable(self.mpm_model, state, nxt, self.sim_dt, device=self.device)
                    state = nxt
                    step += 1
            else:
                final, _, _, _ = _run_forward(solver, state, self.mpm_model, self.sim_dt, n_sub, (), False,
                                              step_callback=False)
                state = _state_from(state, final)
                step += n_sub
            frames.append(state.particle_x.clone())
        self.mpm_state = state
        return torch.stack(frames, dim=0)

    # warp_mpm_wrapper.py:308-326
    def prepare_online_optimization(self, x, parameters):
        if not self.initialized:
            self.initialize()
        E_t = torch.tensor(1.0, dtype=DTYPE, device=self.device) * parameters["E"]
        E_t.requires_grad_()
        self.optim = {"trainable_params": [E_t]}

    # warp_mpm_wrapper.py:328-389
    def run_online_optimization_step(self, x, parameters, n_train_frames=1, n_iters=1, loss_func=None):
        x_t = x.clone()
        rho_mask = torch.ones(self.n_particles, dtype=torch.bool, device=self.device)
        E_t = torch.tensor(1.0, dtype=DTYPE, device=self.device) * parameters["E"]
        E_t.requires_grad_()
        losses = []
        n_sub = int(self.frame_dt / self.sim_dt)
        optimizer = torch.optim.AdamW([E_t], lr=1e-4, weight_decay=0.0)
        mpm_state = self.mpm_state.partial_clone(requires_grad=True)
        particle_pos = None
        for _ in range(n_iters):
            optimizer.zero_grad()
            if self.optim_callback:
                self.optim_callback(self)
            particle_pos = MPMDifferentiableSimulationTrail.apply(
                self.mpm_solver, mpm_state, self.mpm_model, n_sub, self.sim_dt, n_train_frames * n_sub, x_t, self.v_t,
                E_t, self.nu_t, self.rho_t, rho_mask, None, self.device, True, 0)
            loss = loss_func(particle_pos)
            losses.append(loss.item())
            loss.backward()
            optimizer.step()
        return {"param": {"E": E_t.item()}, "loss": np.mean(losses), "points": particle_pos.clone().detach()}

    # warp_mpm_wrapper.py:391-499
    def run_diff_optimization(self, parameters, n_train_frames, max_iters, loss_func=None, out_dir="outputs"):
        if not self.initialized:
            self.initialize()
        n_p = self.n_particles
        x_t, v_t = self.x_t.clone(), self.v_t.clone()
        E_t, nu_t, rho_t = self.E_t.clone(), self.nu_t.clone(), self.rho_t.clone()
        rho_mask = torch.ones(n_p, dtype=torch.bool, device=self.device)
        trainable, optimal = [], {}
        if "E" in parameters:
            E_t = torch.tensor(1.0, dtype=DTYPE, device=self.device) * parameters["E"]
            E_t.requires_grad_()
            trainable.append(E_t)
            optimal["E"] = E_t
        if "nu" in parameters:
            nu_t = torch.tensor(1.0, dtype=DTYPE, device=self.device) * parameters["nu"]
            nu_t.requires_grad_()
            trainable.append(nu_t)
            optimal["nu"] = nu_t
        if "rho" in parameters:
            rho_t = torch.ones(n_p, dtype=DTYPE, device=self.device) * parameters["rho"]
            rho_t.requires_grad_()
            trainable.append(rho_t)
            optimal["rho"] = rho_t
        optimizer = torch.optim.AdamW(trainable, lr=1e-3, weight_decay=0.0)
        scheduler = torch.optim.lr_scheduler.ReduceLROnPlateau(optimizer, mode="min", factor=0.1, patience=5)
        n_sub = int(self.frame_dt / self.sim_dt)
        self.losses = []
        os.makedirs(f"{out_dir}/optim", exist_ok=True)
        for it in range(max_iters):
            optimizer.zero_grad()
            if self.optim_callback:
                self.optim_callback(self)
            particle_pos = MPMDifferentiableSimulationTrail.apply(
                self.mpm_solver, self.mpm_state, self.mpm_model, n_sub, self.sim_dt, n_train_frames * n_sub, x_t, v_t,
                E_t, nu_t, rho_t, rho_mask, None, self.device, True, 0)
            loss = loss_func(particle_pos)
            self.losses.append(loss.item())
            loss.backward()
            optimizer.step()
            scheduler.step(loss.item())
            save_ply_file(particle_pos[-1], f"{out_dir}/optim/optim_{it:04d}.ply")
            if optimizer.param_groups[0]["lr"] < 1e-5:
                break
        return optimal
