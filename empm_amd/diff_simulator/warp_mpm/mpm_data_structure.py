"""Particle state and model containers with the field names, methods and
semantics of the reference's MPMStateStruct / MPMModelStruct
(phys_sim/diff_simulator/warp_mpm/mpm_data_structure.py:13-498) and its
collider structs (:502-621).

Fields are torch tensors on the GPU (the reference holds Warp arrays).  The
grid is not part of the state here: it lives in the libempm context and is
never materialised per state (the reference allocates a full grid in every
partial_clone, mpm_data_structure.py:356-387).
"""
from __future__ import annotations

from typing import Optional, Sequence, Union

import numpy as np
import torch
from torch import Tensor

from ...runtime import ParticleState

F32 = torch.float32


def _eye(n, device):
    return torch.eye(3, dtype=F32, device=device).repeat(n, 1, 1).contiguous()


def _dev(device, fallback="cuda:0"):
    return torch.device(device if device is not None else fallback)


class MPMStateStruct:
    """mpm_data_structure.py:13-387"""

    def __init__(self):
        self.particle_x = None
        self.particle_v = None
        self.particle_F = None
        self.particle_cov = None
        self.particle_F_trial = None
        self.particle_stress = None
        self.particle_C = None
        self.particle_vol = None
        self.particle_mass = None
        self.particle_density = None
        self.particle_selection = None
        self.grid_res = None
        self.requires_grad = False
        self.device = None

    # mpm_data_structure.py:41-95
    def init(self, shape: Union[Sequence[int], int], device=None, requires_grad=False) -> None:
        n = int(shape if np.ndim(shape) == 0 else shape[0])
        dev = _dev(device)
        self.device = dev
        self.requires_grad = requires_grad
        z3 = lambda: torch.zeros((n, 3), dtype=F32, device=dev)
        z9 = lambda: torch.zeros((n, 3, 3), dtype=F32, device=dev)
        z1 = lambda: torch.zeros(n, dtype=F32, device=dev)
        self.particle_x = z3()
        self.particle_v = z3()
        self.particle_F = z9()
        self.particle_cov = torch.zeros(n * 6, dtype=F32, device=dev)
        self.particle_F_trial = z9()
        self.particle_stress = z9()
        self.particle_C = z9()
        self.particle_vol = z1()
        self.particle_mass = z1()
        self.particle_density = z1()
        self.particle_selection = torch.zeros(n, dtype=torch.int32, device=dev)
        self.grid_res = 10

    # mpm_data_structure.py:97-117
    def init_grid(self, grid_res: int, device=None, requires_grad=False):
        self.grid_res = int(grid_res)

    @property
    def n_particles(self):
        return int(self.particle_x.shape[0])

    # mpm_data_structure.py:119-172
    def from_torch(self, tensor_x: Tensor, tensor_volume: Tensor, tensor_cov: Optional[Tensor] = None,
                   tensor_velocity: Optional[Tensor] = None, n_grid: int = 100, grid_lim=1.0, device="cuda:0",
                   requires_grad=True):
        assert tensor_x.shape[0] == tensor_volume.shape[0]
        dev = _dev(device)
        self.init_grid(grid_res=n_grid, device=dev, requires_grad=requires_grad)
        if tensor_x is not None:
            self.particle_x = tensor_x.detach().to(dev, F32).contiguous().clone()
        if tensor_volume is not None:
            self.particle_vol = tensor_volume.detach().to(dev, F32).contiguous().clone()
        if tensor_cov is not None:
            self.particle_cov = tensor_cov.detach().reshape(-1).to(dev, F32).contiguous().clone()
        if tensor_velocity is not None:
            self.particle_v = tensor_velocity.detach().to(dev, F32).contiguous().clone()
        self.particle_F_trial = _eye(tensor_x.shape[0], dev)

    # mpm_data_structure.py:174-255
    def reset_state(self, tensor_x: Tensor, tensor_cov: Optional[Tensor] = None,
                    tensor_velocity: Optional[Tensor] = None, tensor_density: Optional[Tensor] = None,
                    selection_mask: Optional[Tensor] = None, device="cuda:0", requires_grad=True):
        n = tensor_x.shape[0]
        dev = _dev(device)
        if tensor_x is not None:
            # the reference aliases the passed tensor (from_torch_safe)
            self.particle_x = tensor_x.detach().to(dev, F32).contiguous()
        if tensor_cov is not None:
            self.particle_cov = tensor_cov.detach().reshape(-1).to(dev, F32).contiguous().clone()
        if tensor_velocity is not None:
            self.particle_v = tensor_velocity.detach().to(dev, F32).contiguous().clone()
        if tensor_density is not None and selection_mask is not None:
            m = selection_mask.to(dev).to(torch.int32) == 1
            self.particle_density = torch.where(m, tensor_density.detach().to(dev, F32), self.particle_density)
        self.particle_F_trial = _eye(n, dev)
        self.particle_F = _eye(n, dev)
        self.particle_C = torch.zeros((n, 3, 3), dtype=F32, device=dev)
        self.particle_stress = torch.zeros((n, 3, 3), dtype=F32, device=dev)

    # mpm_data_structure.py:257-292
    def continue_from_torch(self, tensor_x: Tensor, tensor_velocity: Optional[Tensor] = None,
                            tensor_F: Optional[Tensor] = None, tensor_C: Optional[Tensor] = None, device="cuda:0",
                            requires_grad=True):
        dev = _dev(device)
        if tensor_x is not None:
            self.particle_x = tensor_x.detach().to(dev, F32).contiguous()
        if tensor_velocity is not None:
            self.particle_v = tensor_velocity.detach().to(dev, F32).contiguous().clone()
        if tensor_F is not None:
            self.particle_F_trial = tensor_F.detach().reshape(-1, 3, 3).to(dev, F32).contiguous().clone()
        if tensor_C is not None:
            self.particle_C = tensor_C.detach().reshape(-1, 3, 3).to(dev, F32).contiguous().clone()

    # mpm_data_structure.py:294-303
    def set_require_grad(self, requires_grad=True):
        self.requires_grad = requires_grad

    # mpm_data_structure.py:305-354
    def reset_density(self, tensor_density: Tensor, selection_mask: Optional[Tensor] = None, device="cuda:0",
                      requires_grad=True, update_mass=False):
        from ...runtime import mass_from_density

        dev = _dev(device)
        if tensor_density is not None:
            rho = tensor_density.detach().to(dev, F32).contiguous()
            if rho.ndim == 0:
                rho = rho.expand(self.n_particles).contiguous()
            if selection_mask is not None:
                m = selection_mask.to(dev).to(torch.int32) == 1
                self.particle_density = torch.where(m, rho, self.particle_density).contiguous()
            else:
                self.particle_density = rho.clone()
        if update_mass:
            self.particle_mass = torch.empty_like(self.particle_vol)
            mass_from_density(self.particle_density, self.particle_vol, self.particle_mass)

    # mpm_data_structure.py:356-387
    def partial_clone(self, device="cuda:0", requires_grad=True):
        new = MPMStateStruct()
        n = self.n_particles
        new.init(n, device=device, requires_grad=requires_grad)
        new.particle_vol = self.particle_vol.clone()
        new.particle_density = self.particle_density.clone()
        new.particle_mass = self.particle_mass.clone()
        new.particle_selection = self.particle_selection.clone()
        new.init_grid(grid_res=self.grid_res, device=device, requires_grad=requires_grad)
        new.particle_F_trial = _eye(n, new.device)
        return new

    def particles(self, with_fe=True) -> ParticleState:
        """The substep-facing view: x, v, C, F_trial (+ particle_F as the
        return-mapped F output)."""
        return ParticleState(self.particle_x, self.particle_v, self.particle_C, self.particle_F_trial,
                             self.particle_F if with_fe else None)


class MPMModelStruct:
    """mpm_data_structure.py:390-498"""

    def __init__(self):
        self.grid_lim = 1.0
        self.n_particles = 0
        self.n_grid = 100
        self.dx = 0.01
        self.inv_dx = 100.0
        self.grid_dim_x = self.grid_dim_y = self.grid_dim_z = 100
        self.mu = self.lam = self.E = self.nu = None
        self.material = 0
        self.yield_stress = None
        self.friction_angle = 25.0
        self.alpha = 0.0
        self.gravitational_accelaration = (0.0, 0.0, 0.0)
        self.hardening = 0.0
        self.xi = 0.0
        self.plastic_viscosity = 0.0
        self.softening = 0.1
        self.rpic_damping = 0.0
        self.grid_v_damping_scale = 1.1
        self.update_cov_with_F = False

    # mpm_data_structure.py:424-448
    def init(self, shape, device=None, requires_grad=False) -> None:
        n = int(shape if np.ndim(shape) == 0 else shape[0])
        dev = _dev(device)
        self.n_particles = n
        z1 = lambda: torch.zeros(n, dtype=F32, device=dev)
        self.E, self.nu, self.mu, self.lam, self.yield_stress = z1(), z1(), z1(), z1(), z1()

    # mpm_data_structure.py:450-454
    def finalize_mu_lam(self, n_particles, device="cuda:0"):
        from ...runtime import mu_lam

        mu_lam(self.E, self.nu, self.mu, self.lam)

    # mpm_data_structure.py:456-484 (dx, inv_dx and alpha are float32 there)
    def init_other_params(self, n_grid=100, grid_lim=1.0, device="cuda:0"):
        self.grid_lim = grid_lim
        self.n_grid = n_grid
        self.grid_dim_x = self.grid_dim_y = self.grid_dim_z = n_grid
        self.dx = float(np.float32(grid_lim / n_grid))
        self.inv_dx = float(np.float32(n_grid / grid_lim))
        self.update_cov_with_F = False
        self.material = 0
        self.plastic_viscosity = 0.0
        self.softening = 0.1
        self.friction_angle = 25.0
        self.alpha = friction_alpha(self.friction_angle)
        self.gravitational_accelaration = (0.0, 0.0, 0.0)
        self.rpic_damping = 0.0
        self.grid_v_damping_scale = 1.1

    # mpm_data_structure.py:486-491
    def from_torch(self, tensor_E: Tensor, tensor_nu: Tensor, device="cuda:0", requires_grad=False):
        self.E = tensor_E.detach().to(_dev(device), F32).contiguous().clone()
        self.nu = tensor_nu.detach().to(_dev(device), F32).contiguous().clone()
        n = self.E.shape[0]
        if self.mu is None or self.mu.shape[0] != n:
            self.mu = torch.zeros_like(self.E)
            self.lam = torch.zeros_like(self.E)
        if self.yield_stress is None or self.yield_stress.shape[0] != n:
            self.yield_stress = torch.zeros_like(self.E)
        self.finalize_mu_lam(n, device)

    def set_require_grad(self, requires_grad=True):
        pass


def friction_alpha(friction_angle_deg: float) -> float:
    """Drucker-Prager alpha from the friction angle, in float32 as the
    reference computes it (mpm_solver_diff.py:69-72)."""
    sin_phi = np.float32(np.sin(np.float32(friction_angle_deg / 180.0 * 3.14159265)))
    return float(np.float32(np.sqrt(np.float32(2.0 / 3.0)) * np.float32(2.0) * sin_phi / (np.float32(3.0) - sin_phi)))


class Dirichlet_collider:
    """Mutable collider parameters (mpm_data_structure.py:502-537).  Callers
    assign .point / .velocity / .active between substeps, as with the Warp
    struct; zero-initialised like it."""

    _VEC3 = ("point", "normal", "direction", "velocity", "rot_axis", "x_unit", "y_unit", "size",
             "horizontal_axis_1", "horizontal_axis_2")

    def __init__(self):
        for f in self._VEC3:
            setattr(self, f, (0.0, 0.0, 0.0))
        self.start_time = 0.0
        self.end_time = 0.0
        self.active = False
        self.friction = 0.0
        self.surface_type = 0
        self.angular_velocity = 0.0
        self.threshold = 0.0
        self.reset = 0
        self.index = 0
        self.radius = 0.0
        self.v_scale = 0.0
        self.width = 0.0
        self.height = 0.0
        self.length = 0.0
        self.R = 0.0
        self.half_height_and_radius = (0.0, 0.0)
        self.half_height = 0.0


class GridCollider:
    """mpm_data_structure.py:540-547"""

    def __init__(self):
        self.point = (0.0, 0.0, 0.0)
        self.normal = (0.0, 0.0, 0.0)
        self.direction = (0.0, 0.0, 0.0)
        self.start_time = 0.0
        self.end_time = 0.0
        self.mask = None
